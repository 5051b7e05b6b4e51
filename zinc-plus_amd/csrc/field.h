// field.h -- arithmetic in F_p for a 256-bit odd prime p, Montgomery form
// with R = 2^256, usable in host and device code.
//
// This is the field of the opening path: MontyField<4> of the reference's
// benches (zip-plus/benches/zip_common.rs:35, modulus from
// Transcript::get_random_field_cfg, transcript/src/traits.rs:371-382).
// Elements are four little-endian u64 limbs; a value v is held as
// v * 2^256 mod p (the element's "inner" representation, the bytes the proof
// stream carries).  Multiplication is word-serial Montgomery reduction
// (Koc, Acar, Kaliski, "Analyzing and comparing Montgomery multiplication
// algorithms", IEEE Micro 1996: CIOS), which needs p odd and inputs < p.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zpi {

struct Fp {
  uint64_t w[4];
};

struct FpParams {
  uint64_t p[4];
  uint64_t pinv;   // -p^-1 mod 2^64
  uint64_t r1[4];  // 2^256 mod p (one, in Montgomery form)
  uint64_t r2[4];  // 2^512 mod p
};

typedef unsigned __int128 u128;

__host__ __device__ inline bool fp_geq(const uint64_t a[4], const uint64_t b[4]) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] != b[i]) return a[i] > b[i];
  }
  return true;
}

// a -= b, returns the borrow.
__host__ __device__ inline uint64_t fp_sub_raw(uint64_t a[4], const uint64_t b[4]) {
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    const uint64_t d = a[i] - b[i];
    const uint64_t b1 = a[i] < b[i];
    const uint64_t d2 = d - br;
    br = b1 | (d < br);
    a[i] = d2;
  }
  return br;
}

// a += b, returns the carry.
__host__ __device__ inline uint64_t fp_add_raw(uint64_t a[4], const uint64_t b[4]) {
  uint64_t c = 0;
  for (int i = 0; i < 4; i++) {
    const u128 s = (u128)a[i] + b[i] + c;
    a[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  return c;
}

__host__ __device__ inline Fp fp_add(const FpParams& P, Fp a, const Fp& b) {
  const uint64_t c = fp_add_raw(a.w, b.w);
  if (c || fp_geq(a.w, P.p)) fp_sub_raw(a.w, P.p);
  return a;
}

__host__ __device__ inline Fp fp_sub(const FpParams& P, Fp a, const Fp& b) {
  if (fp_sub_raw(a.w, b.w)) fp_add_raw(a.w, P.p);
  return a;
}

__host__ __device__ inline Fp fp_neg(const FpParams& P, const Fp& a) {
  Fp z = {{0, 0, 0, 0}};
  return fp_sub(P, z, a);
}

__host__ __device__ inline bool fp_is_zero(const Fp& a) {
  return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) == 0;
}

// a * b * 2^-256 mod p (CIOS).
__host__ __device__ inline Fp fp_mul(const FpParams& P, const Fp& a, const Fp& b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    uint64_t c = 0;
    for (int j = 0; j < 4; j++) {
      const u128 s = (u128)a.w[j] * b.w[i] + t[j] + c;
      t[j] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
    u128 s = (u128)t[4] + c;
    t[4] = (uint64_t)s;
    t[5] = (uint64_t)(s >> 64);
    const uint64_t m = t[0] * P.pinv;
    s = (u128)m * P.p[0] + t[0];
    c = (uint64_t)(s >> 64);
    for (int j = 1; j < 4; j++) {
      s = (u128)m * P.p[j] + t[j] + c;
      t[j - 1] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
    s = (u128)t[4] + c;
    t[3] = (uint64_t)s;
    t[4] = t[5] + (uint64_t)(s >> 64);
  }
  Fp r = {{t[0], t[1], t[2], t[3]}};
  if (t[4] || fp_geq(r.w, P.p)) fp_sub_raw(r.w, P.p);
  return r;
}

// Montgomery form of a canonical value v < p.
__host__ __device__ inline Fp fp_to_mont(const FpParams& P, const Fp& v) {
  Fp r2 = {{P.r2[0], P.r2[1], P.r2[2], P.r2[3]}};
  return fp_mul(P, v, r2);
}

// Canonical value of a Montgomery-form element.
__host__ __device__ inline Fp fp_from_mont(const FpParams& P, const Fp& a) {
  Fp one = {{1, 0, 0, 0}};
  return fp_mul(P, a, one);
}

__host__ __device__ inline Fp fp_one(const FpParams& P) {
  return Fp{{P.r1[0], P.r1[1], P.r1[2], P.r1[3]}};
}

// Montgomery form of a two's complement integer of n little-endian u64
// limbs (Horner over the limbs, from the top).  Needs p > 2^64.
__host__ __device__ inline Fp fp_from_signed_limbs(const FpParams& P, const uint64_t* limbs, int n) {
  const bool neg = (int64_t)limbs[n - 1] < 0;
  // 2^64 * 2^256 mod p: the Montgomery form of 2^64.
  const Fp m64 = fp_to_mont(P, Fp{{0, 1, 0, 0}});
  Fp acc = {{0, 0, 0, 0}};
  uint64_t borrow = 1;  // magnitude of a negative value: ~v + 1, limb by limb
  uint64_t mag[16];
  for (int i = 0; i < n; i++) {
    uint64_t l = limbs[i];
    if (neg) {
      const uint64_t x = ~l;
      l = x + borrow;
      borrow = borrow && l == 0;
    }
    mag[i] = l;
  }
  for (int i = n - 1; i >= 0; i--) {
    acc = fp_mul(P, acc, m64);
    acc = fp_add(P, acc, fp_to_mont(P, Fp{{mag[i], 0, 0, 0}}));
  }
  return neg ? fp_neg(P, acc) : acc;
}

// Montgomery parameters of an odd modulus (host).
inline FpParams fp_params(const uint64_t p[4]) {
  FpParams P;
  for (int i = 0; i < 4; i++) P.p[i] = p[i];
  uint64_t inv = 1;  // Newton iteration for p^-1 mod 2^64
  for (int i = 0; i < 6; i++) inv *= 2 - p[0] * inv;
  P.pinv = 0 - inv;
  // 2^256 mod p and 2^512 mod p by doubling 1.
  uint64_t x[4] = {1, 0, 0, 0};
  for (int k = 1; k <= 512; k++) {
    const uint64_t c = fp_add_raw(x, x);
    if (c || fp_geq(x, P.p)) fp_sub_raw(x, P.p);
    if (k == 256)
      for (int i = 0; i < 4; i++) P.r1[i] = x[i];
  }
  for (int i = 0; i < 4; i++) P.r2[i] = x[i];
  return P;
}

}  // namespace zpi
