// wide.hip -- LinearCode::encode_wide on the device: the codes applied to
// rows of CombR = Int<L> (L little-endian u64 limbs, two's complement), the
// verifier's encoding of the combined row.
//
// Reference: zip-plus/src/code.rs:55; IprsCode::encode_wide
// code/iprs.rs:148-150 -> encode_inner :54-87 (pntt::radix8 with
// Int<L>::mul_by_scalar(&i64), utils/src/mul_by_scalar.rs:31-74);
// RaaCode::encode_wide code/raa.rs:151-153 and RaaSignFlippingCode
// code/raa_sign_flip.rs:97-99 -> encode_inner.  All arithmetic is modulo
// 2^(64L), as in the unchecked build (the checked build panics where this
// wraps, and agrees everywhere else).
//
// IPRS: the base layer runs one workgroup per (row, chunk, 256 base rows)
// with the chunk's inputs in LDS; each radix-8 stage is one launch with a
// thread per butterfly.  RAA: one workgroup per row; each accumulate is a
// per-thread segment sum, a block scan of the segment sums and a second
// pass that emits the prefix values.
#include "common.h"

namespace zpi {

// The base layer runs on the matrix cores when the code has the int8 slices
// of its base matrix (depth 1 or 2, base length <= 256), unless the
// per-element kernel is forced.
bool wide_mfma_eligible(const zp_code* cd) {
  return cd->d_afrag_c != nullptr && (cd->depth == 1 || cd->depth == 2) && g_force_wide != 1;
}

namespace {

template <int L>
struct Big {
  uint64_t w[L];
};

template <int L>
__device__ __forceinline__ void big_add(Big<L>& a, const Big<L>& b) {
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const uint64_t s = a.w[i] + b.w[i];
    const uint64_t c1 = s < a.w[i];
    const uint64_t s2 = s + c;
    c = c1 | (s2 < s);
    a.w[i] = s2;
  }
}

template <int L>
__device__ __forceinline__ void big_sub(Big<L>& a, const Big<L>& b) {
  uint64_t br = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const uint64_t d = a.w[i] - b.w[i];
    const uint64_t b1 = a.w[i] < b.w[i];
    const uint64_t d2 = d - br;
    br = b1 | (d < br);
    a.w[i] = d2;
  }
}

template <int L>
__device__ __forceinline__ void big_neg(Big<L>& a) {
  uint64_t c = 1;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const uint64_t s = ~a.w[i] + c;
    c = c && s == 0;
    a.w[i] = s;
  }
}

// acc += x * m modulo 2^(64L), for a table entry |m| < 2^31.
template <int L>
__device__ __forceinline__ void big_mac(Big<L>& acc, const Big<L>& x, int32_t m) {
  const uint64_t um = (uint64_t)(m < 0 ? -(int64_t)m : (int64_t)m);
  Big<L> p;
  uint64_t hi = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const uint64_t lo = x.w[i] * um;
    const uint64_t s = lo + hi;
    hi = __umul64hi(x.w[i], um) + (s < lo);
    p.w[i] = s;
  }
  if (m < 0)
    big_sub(acc, p);
  else
    big_add(acc, p);
}

template <int L>
__device__ __forceinline__ Big<L> big_load(const uint64_t* p) {
  Big<L> a;
#pragma unroll
  for (int i = 0; i < L; i++) a.w[i] = p[i];
  return a;
}

// An input element: L limbs, or one signed u64 limb sign-extended to L.
template <int L>
__device__ __forceinline__ Big<L> in_load(const uint64_t* p, bool one_limb) {
  if (!one_limb) return big_load<L>(p);
  Big<L> a;
  a.w[0] = p[0];
  const uint64_t ext = (int64_t)p[0] < 0 ? ~0ull : 0ull;
#pragma unroll
  for (int i = 1; i < L; i++) a.w[i] = ext;
  return a;
}

template <int L>
__device__ __forceinline__ void big_store(uint64_t* p, const Big<L>& a) {
#pragma unroll
  for (int i = 0; i < L; i++) p[i] = a.w[i];
}

// ---- Checked arithmetic (the CHECKED codes) -----------------------------
//
// The checked build folds with num_traits checked_mul / checked_add and
// panics at the first result outside Int<L> = [-2^(64L-1), 2^(64L-1))
// (code/iprs.rs:76-84, pntt/radix8.rs:150-157, butterfly.rs:34-42,
// raa.rs:205-218, raa_sign_flip.rs:146-155).  These helpers compute the
// same wrapping values as the unchecked ones and report whether the exact
// result left that range.  Overflow bits reported to the host:
constexpr uint32_t kOvfMul = 1, kOvfAdd = 2, kOvfNeg = 4;

template <int L>
__device__ __forceinline__ bool big_is_neg(const Big<L>& a) {
  return (int64_t)a.w[L - 1] < 0;
}

// acc += x * m, with the product and the sum each checked.  |x| is formed
// as an unsigned L-limb magnitude (Int::MIN maps to 2^(64L-1)), multiplied
// by |m| into L limbs plus a high limb; the product fits iff the high limb
// is zero and the magnitude is below 2^(64L-1), or equal to it for a
// negative product.  The sum overflows iff both operands have one sign and
// the result the other.
template <int L>
__device__ __forceinline__ uint32_t big_mac_checked(Big<L>& acc, const Big<L>& x, int32_t m) {
  const bool xneg = big_is_neg(x);
  Big<L> p = x;
  if (xneg) big_neg(p);
  const uint64_t um = (uint64_t)(m < 0 ? -(int64_t)m : (int64_t)m);
  uint64_t hi = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const uint64_t lo = p.w[i] * um;
    const uint64_t s = lo + hi;
    hi = __umul64hi(p.w[i], um) + (s < lo);
    p.w[i] = s;
  }
  const bool neg = xneg != (m < 0);
  uint32_t ovf = hi != 0 ? kOvfMul : 0;
  if (big_is_neg(p)) {  // magnitude >= 2^(64L-1): only -2^(64L-1) fits
    bool is_min = p.w[L - 1] == (1ull << 63);
#pragma unroll
    for (int i = 0; i < L - 1; i++) is_min = is_min && p.w[i] == 0;
    if (!(neg && is_min)) ovf |= kOvfMul;
  }
  if (neg) big_neg(p);
  const bool as = big_is_neg(acc), ps = big_is_neg(p);
  big_add(acc, p);
  if (as == ps && big_is_neg(acc) != as) ovf |= kOvfAdd;
  return ovf;
}

// a = -a, checked: only Int::MIN has no negation.
template <int L>
__device__ __forceinline__ uint32_t big_neg_checked(Big<L>& a) {
  bool is_min = a.w[L - 1] == (1ull << 63);
#pragma unroll
  for (int i = 0; i < L - 1; i++) is_min = is_min && a.w[i] == 0;
  big_neg(a);
  return is_min ? kOvfNeg : 0;
}

__device__ __forceinline__ uint32_t octrev_w(uint32_t c, uint32_t d) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < d; i++) {
    r = (r << 3) | (c & 7u);
    c >>= 3;
  }
  return r;
}

// Where the elements of a (row, plane) sequence live, in u64 units: the
// sequence rp = row * planes + plane starts at row * row + plane * plane_off,
// and its elements are `elem` apart.  Input elements have L limbs or one
// signed limb (sign-extended).  Output elements always have L limbs.
struct WideLayout {
  uint32_t planes;
  bool in_one_limb;
  uint64_t in_row, in_plane, in_elem;
  uint64_t out_row, out_plane, out_elem;
  // Input elements are i32 (sign-extended; offsets in 4-byte units).  Only
  // the matrix-core base layer reads such inputs.
  bool in_i32 = false;
  // Input elements are bits: sequence rp reads bit (rp % planes) of the
  // u64 word at in_at(rp, k) (a bit-packed BinaryPoly whose coefficient
  // planes are encoded one by one; in_plane is 0).  Only the per-element
  // base kernel reads such inputs.
  bool in_bits = false;
  __device__ __forceinline__ uint64_t in_at(uint64_t rp, uint64_t k) const {
    return (rp / planes) * in_row + (rp % planes) * in_plane + k * in_elem;
  }
  __device__ __forceinline__ uint64_t out_at(uint64_t rp, uint64_t i) const {
    return (rp / planes) * out_row + (rp % planes) * out_plane + i * out_elem;
  }
};

constexpr uint32_t kWideTile = 256;

// One input element of sequence rp at index k, as L limbs.
template <int L>
__device__ __forceinline__ Big<L> lay_load(const uint64_t* in, const WideLayout& lay,
                                           uint64_t rp, uint64_t k) {
  if (lay.in_bits) {
    Big<L> a;
    a.w[0] = (in[lay.in_at(rp, k)] >> (rp % lay.planes)) & 1u;
#pragma unroll
    for (int i = 1; i < L; i++) a.w[i] = 0;
    return a;
  }
  return in_load<L>(in + lay.in_at(rp, k), lay.in_one_limb);
}

// Block (tile, chunk, sequence): out[c*bd + r] = x[o] + sum_{k>=1} f(r*k mod bd)
// x[o + k 8^d] (radix8.rs:117-161), r in this block's tile of base rows.
// kChecked: every product and partial sum is checked in the reference's
// fold order (k = 1, 2, ...; the k = 0 term x[o] * f(0) = x[o] is
// from_ref and cannot overflow) and overflow bits are or-ed into *flag.
template <int L, bool kChecked>
__global__ void __launch_bounds__(kWideTile)
wide_base_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                 const int32_t* __restrict__ ftab, WideLayout lay, uint32_t bd, uint32_t bl,
                 uint32_t depth, uint32_t nchunks, uint32_t ntiles, uint32_t* __restrict__ flag) {
  __shared__ uint64_t xs[kWideTile * L];  // the chunk's inputs, kWideTile at a time
  const uint32_t tile = blockIdx.x % ntiles;
  const uint32_t c = (blockIdx.x / ntiles) % nchunks;
  const uint64_t rp = blockIdx.x / ntiles / nchunks;
  const uint32_t o = octrev_w(c, depth);
  const uint32_t r = tile * kWideTile + threadIdx.x;
  Big<L> acc;
#pragma unroll
  for (int l = 0; l < L; l++) acc.w[l] = 0;
  uint32_t ovf = 0;
  for (uint32_t k0 = 0; k0 < bl; k0 += kWideTile) {
    const uint32_t nk = min(kWideTile, bl - k0);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nk; k += blockDim.x)
      big_store(xs + k * L, lay_load<L>(in, lay, rp, (uint64_t)o + (uint64_t)(k0 + k) * nchunks));
    __syncthreads();
    if (r < bd) {  // f(0) = 1 is from_ref for k = 0
      for (uint32_t k = 0; k < nk; k++) {
        const int32_t m = ftab[(r * (k0 + k)) & (bd - 1)];
        if (kChecked)
          ovf |= big_mac_checked(acc, big_load<L>(xs + k * L), m);
        else
          big_mac(acc, big_load<L>(xs + k * L), m);
      }
    }
  }
  if (r < bd) big_store(out + lay.out_at(rp, (uint64_t)c * bd + r), acc);
  if (kChecked && ovf) atomicOr(flag, ovf);
}

// ---- Base layer on the matrix cores (depth 1 or 2, bl <= 256) ----------
//
// out[c*bd + r] = sum_k f(r*k mod bd) x[o + k 8^d] is an integer matrix
// product F (bd x bl) times X (bl x columns), columns n = (sequence, chunk).
// F is held as three int8 slices (the binary encoder's A fragments:
// f = s0 + 128 s1 + 16384 s2).  X is split into signed base-256 digits:
// with C = 0x80 in every byte, the bytes y_b of y = x + C (modulo 2^(64L))
// give d_b = y_b - 128 in [-128, 127] and x = sum_b d_b 256^b modulo
// 2^(64L).  Every slice product F_a D_b is one run of
// v_mfma_i32_16x16x64_i8 per 64 k (|sum| < 2^22), and
//   out = sum_{a,b} 2^(7a + 8b) (F_a D_b)   modulo 2^(64L).
// A one-limb input (|x| < 2^63, sign-extended) needs only 9 digits (8 for
// L = 1), an i32 input 5; a full input needs 8L.
typedef int wv4i __attribute__((ext_vector_type(4)));

inline uint32_t wide_digit_count(const WideLayout& lay, int L) {
  if (lay.in_i32) return 5;  // |x| < 2^31: x + C stays below 2^40
  return lay.in_one_limb ? (L == 1 ? 8u : 9u) : 8u * (uint32_t)L;
}

// Thread per (column n, group of 16 k): writes the digit bytes of the 16
// inputs x[n][16 kg + j], j < 16, in the B-operand layout: block
// ((n / 16) * S + s) * nd + b holds 64 lanes x 16 bytes, lane 16 g + n % 16
// with k = 64 s + 16 g + j at byte j (kg = 4 s + g).  Columns n >= ncols
// and k >= bl are zero.
template <int L>
__global__ void __launch_bounds__(256)
wide_digits_kernel(const uint64_t* __restrict__ in, uint8_t* __restrict__ frag, WideLayout lay,
                   uint32_t bl, uint32_t depth, uint32_t nchunks, uint32_t S, uint32_t nd,
                   uint64_t ncols, uint64_t total) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t nkg = 4 * S;
  const uint64_t n = t / nkg;
  const uint32_t kg = (uint32_t)(t % nkg), s = kg / 4, g = kg % 4;
  const bool live = n < ncols;
  const uint64_t rp = live ? n / nchunks : 0;
  const uint32_t o = live ? octrev_w((uint32_t)(n % nchunks), depth) : 0;
  constexpr uint64_t C = 0x8080808080808080ull;
  uint64_t carry[16];
#pragma unroll
  for (int j = 0; j < 16; j++) carry[j] = 0;
  uint8_t* dst = frag + (((n / 16) * S + s) * nd) * 1024 + (16 * g + n % 16) * 16;
  for (uint32_t limb = 0; limb * 8 < nd; limb++) {
    uint64_t y[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const uint32_t k = 16 * kg + j;
      uint64_t x = 0;
      if (live && k < bl) {
        const uint64_t at = lay.in_at(rp, (uint64_t)o + (uint64_t)k * nchunks);
        if (lay.in_i32) {
          const int64_t v = ((const int32_t*)in)[at];
          x = limb == 0 ? (uint64_t)v : (v < 0 ? ~0ull : 0ull);
        } else {
          const uint64_t* p = in + at;
          x = lay.in_one_limb ? (limb == 0 ? p[0] : ((int64_t)p[0] < 0 ? ~0ull : 0ull)) : p[limb];
        }
      }
      const uint64_t sum = x + C;
      const uint64_t c1 = sum < x;
      y[j] = sum + carry[j];
      carry[j] = c1 | (y[j] < sum);
    }
    const uint32_t nb = min(8u, nd - limb * 8);
    for (uint32_t bb = 0; bb < nb; bb++) {
      uint32_t w[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        w[q] = 0;
#pragma unroll
        for (int z = 0; z < 4; z++)
          w[q] |= (uint32_t)(((y[4 * q + z] >> (8 * bb)) & 0xff) ^ 0x80) << (8 * z);
      }
      *(uint4*)(dst + (limb * 8 + bb) * 1024) = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
}

// Wave per 16 x 16 output tile (r-block rb, column tile ct); a workgroup is
// four column tiles of one r-block.  Digits are taken four at a time (one
// 32-bit word of the result): their slice products are added, shifted, into
// a 64-bit column sum (each term is below 2^56, twelve of them fit), and
// the column sums are carried into 32-bit result words in increasing order,
// so the sign of every partial sum reaches the top word.
template <int L, int S>
__global__ void __launch_bounds__(256)
wide_mfma_kernel(const wv4i* __restrict__ afrag, const wv4i* __restrict__ bfrag,
                 uint64_t* __restrict__ out, WideLayout lay, uint32_t bd, uint32_t nd,
                 uint32_t nchunks, uint64_t ncols, uint64_t nct, uint32_t nrb) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t rb = blockIdx.x % nrb;
  const uint64_t ct = (uint64_t)(blockIdx.x / nrb) * 4 + wave;
  if (ct >= nct) return;
  wv4i A[S][3];
#pragma unroll
  for (int s = 0; s < S; s++)
#pragma unroll
    for (int a = 0; a < 3; a++) A[s][a] = afrag[(((uint64_t)rb * S + s) * 3 + a) * 64 + lane];
  const wv4i* bt = bfrag + ct * S * nd * 64 + lane;
  const uint64_t n = ct * 16 + (lane & 15);
  const bool live = n < ncols;
  const uint64_t rp = live ? n / nchunks : 0, c = live ? n % nchunks : 0;
  uint32_t* dst[4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const uint32_t r = rb * 16 + 4 * (lane >> 4) + v;  // consecutive rows (d_afrag_c)
    dst[v] = (uint32_t*)(out + lay.out_at(rp, c * bd + r));
  }
  // The B fragments of the next four digits are loaded while the current
  // four are multiplied (with few waves per SIMD, nothing else hides them).
  auto load_word = [&](uint32_t w, wv4i (&Bw)[4][S]) {
#pragma unroll
    for (uint32_t bb = 0; bb < 4; bb++) {
      const uint32_t b = 4 * w + bb;
#pragma unroll
      for (int s = 0; s < S; s++)
        Bw[bb][s] = b < nd ? bt[((uint64_t)s * nd + b) * 64] : wv4i{0, 0, 0, 0};
    }
  };
  wv4i Bn[4][S];
  load_word(0, Bn);
  int64_t carry[4] = {0, 0, 0, 0};
  for (uint32_t w = 0; w < 2 * (uint32_t)L; w++) {
    int64_t col[4] = {0, 0, 0, 0};
    wv4i Bc[4][S];
#pragma unroll
    for (uint32_t bb = 0; bb < 4; bb++)
#pragma unroll
      for (int s = 0; s < S; s++) Bc[bb][s] = Bn[bb][s];
    if (w + 1 < 2 * (uint32_t)L && 4 * (w + 1) < nd) load_word(w + 1, Bn);
#pragma unroll
    for (uint32_t bb = 0; bb < 4; bb++) {
      const uint32_t b = 4 * w + bb;
      if (b < nd) {
        wv4i acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0}, acc2 = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < S; s++) {
          const wv4i B = Bc[bb][s];
          acc0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[s][0], B, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[s][1], B, acc1, 0, 0, 0);
          acc2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[s][2], B, acc2, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; v++)
          col[v] += ((int64_t)acc0[v] << (8 * bb)) + ((int64_t)acc1[v] << (8 * bb + 7)) +
                    ((int64_t)acc2[v] << (8 * bb + 14));
      }
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int64_t tot = col[v] + carry[v];
      if (live) dst[v][w] = (uint32_t)tot;
      carry[v] = tot >> 32;
    }
  }
}

template <int L, int S>
int launch_wide_mfma_s(const zp_code* cd, const uint64_t* in, uint64_t seqs,
                       const WideLayout& lay, uint64_t* out, hipStream_t s) {
  const uint32_t nchunks = 1u << (3 * cd->depth);
  const uint32_t nd = wide_digit_count(lay, L);
  const uint64_t ncols = seqs * nchunks, nct = (ncols + 15) / 16;
  const uint64_t bytes = nct * S * nd * 1024;
  uint8_t* frag = nullptr;
  ZP_TRY(dev_alloc((void**)&frag, bytes, s));
  const uint64_t total = nct * 16 * 4 * S;
  const uint32_t nrb = cd->bd / 16;
  const uint64_t blocks = nrb * ((nct + 3) / 4);
  int rc = ZP_OK;
  if ((total + 255) / 256 > 0x7fffffffull || blocks > 0x7fffffffull)
    rc = fail(ZP_ERR_INVALID_PARAM, "too many rows");
  if (rc == ZP_OK) {
    hipLaunchKernelGGL((wide_digits_kernel<L>), dim3((uint32_t)((total + 255) / 256)), dim3(256),
                       0, s, in, frag, lay, cd->bl, cd->depth, nchunks, (uint32_t)S, nd, ncols,
                       total);
    hipLaunchKernelGGL((wide_mfma_kernel<L, S>), dim3((uint32_t)blocks), dim3(256), 0, s,
                       (const wv4i*)cd->d_afrag_c, (const wv4i*)frag, out, lay, cd->bd, nd, nchunks,
                       ncols, nct, nrb);
    if (hipGetLastError() != hipSuccess) rc = fail(ZP_ERR_DEVICE, "launch failed");
  }
  dev_free(frag, s);
  return rc;
}

template <int L>
int launch_wide_mfma(const zp_code* cd, const uint64_t* in, uint64_t seqs, const WideLayout& lay,
                     uint64_t* out, hipStream_t s) {
  switch (cd->mfma_ksteps) {
    case 1: return launch_wide_mfma_s<L, 1>(cd, in, seqs, lay, out, s);
    case 2: return launch_wide_mfma_s<L, 2>(cd, in, seqs, lay, out, s);
    case 3: return launch_wide_mfma_s<L, 3>(cd, in, seqs, lay, out, s);
    case 4: return launch_wide_mfma_s<L, 4>(cd, in, seqs, lay, out, s);
  }
  return fail(ZP_ERR_INVALID_PARAM, "unsupported base length for the matrix-core base layer");
}


// Stage k (radix8.rs:58-113): thread per (sequence, block, i); the 8 values
// j*sub + i of a block of 8*sub are replaced by their radix-8 butterfly.
// The plane of a sequence is the fastest thread index: the planes of one
// element are adjacent in memory, so a wave's loads and stores coalesce.
template <int L, bool kChecked>
__global__ void __launch_bounds__(256)
wide_stage_kernel(uint64_t* __restrict__ data, const int32_t* __restrict__ tw, WideLayout lay,
                  uint64_t N, uint64_t sub, uint64_t total, uint32_t* __restrict__ flag) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint64_t per_row = N / 8;
  const uint64_t plane = t % lay.planes, rest = t / lay.planes;
  const uint64_t rem = rest % per_row, rp = (rest / per_row) * lay.planes + plane;
  const uint64_t blk = rem / sub, i = rem % sub;
  const uint64_t first = blk * 8 * sub + i;
  Big<L> xs[8];
#pragma unroll
  for (int j = 0; j < 8; j++) xs[j] = big_load<L>(data + lay.out_at(rp, first + (uint64_t)j * sub));
  // All 56 coefficients of this position are loaded at once (a stage has
  // few threads, so one load round trip per output would be exposed); the
  // coefficient of input j in output r is tw[(j-1)*8 + (r*j mod 8)]
  // (butterfly.rs:8-17).  kChecked: the fold over j = 1..7 is checked in
  // that order (butterfly.rs:34-42).
  const int32_t* twi = tw + i * 56;
  int32_t T[56];
#pragma unroll
  for (int e = 0; e < 56; e++) T[e] = twi[e];
  uint32_t ovf = 0;
#pragma unroll
  for (int r = 0; r < 8; r++) {
    Big<L> a = xs[0];
#pragma unroll
    for (int jm1 = 0; jm1 < 7; jm1++) {
      const int32_t m = T[jm1 * 8 + ((r * (jm1 + 1)) & 7)];
      if (kChecked)
        ovf |= big_mac_checked(a, xs[jm1 + 1], m);
      else
        big_mac(a, xs[jm1 + 1], m);
    }
    big_store(data + lay.out_at(rp, first + (uint64_t)r * sub), a);
  }
  if (kChecked && ovf) atomicOr(flag, ovf);
}

constexpr uint32_t kScanThreads = 1024;

// Exclusive scan of one Big<L> per thread over the block (part: kScanThreads
// x L words of LDS); returns this thread's prefix and the block total.
template <int L>
__device__ Big<L> block_exscan_big(const Big<L>& v, uint64_t* part, Big<L>* total) {
  const uint32_t t = threadIdx.x;
  big_store(part + t * L, v);
  __syncthreads();
  for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
    Big<L> cur = big_load<L>(part + t * L);
    if (t >= off) big_add(cur, big_load<L>(part + (t - off) * L));
    __syncthreads();
    big_store(part + t * L, cur);
    __syncthreads();
  }
  Big<L> incl = big_load<L>(part + t * L);
  *total = big_load<L>(part + (kScanThreads - 1) * L);
  __syncthreads();
  big_sub(incl, v);
  return incl;
}

// raa.rs:93-126 over Int<L>: y1[i] = s(p1[i]) x[p1[i] mod row_len],
// a = s * prefix(y1), out = prefix(a[p2[i]]), s(i) = -1 for odd i under the
// sign-flipping code (raa_sign_flip.rs:45-57).  One workgroup per sequence;
// a goes through the sequence's scratch row in `amid` (N x L words).
template <int L, bool kFlip>
__global__ void __launch_bounds__(kScanThreads)
wide_raa_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                uint64_t* __restrict__ amid, const uint32_t* __restrict__ perm1,
                const uint32_t* __restrict__ perm2, WideLayout lay, uint32_t row_len, uint32_t N) {
  __shared__ uint64_t part[kScanThreads * L];
  const uint64_t rp = blockIdx.x;
  uint64_t* a = amid + rp * N * L;
  const uint32_t per = (N + kScanThreads - 1) / kScanThreads;
  const uint32_t i0 = min(N, threadIdx.x * per), i1 = min(N, i0 + per);
  auto y1 = [&](uint32_t i) {
    const uint32_t s = perm1[i];
    Big<L> v = in_load<L>(in + lay.in_at(rp, s & (row_len - 1)), lay.in_one_limb);
    if (kFlip && (s & 1u)) big_neg(v);
    return v;
  };
  Big<L> sum, total;
#pragma unroll
  for (int l = 0; l < L; l++) sum.w[l] = 0;
  for (uint32_t i = i0; i < i1; i++) big_add(sum, y1(i));
  Big<L> run = block_exscan_big(sum, part, &total);
  for (uint32_t i = i0; i < i1; i++) {
    big_add(run, y1(i));
    Big<L> v = run;
    if (kFlip && (i & 1u)) big_neg(v);
    big_store(a + (uint64_t)i * L, v);
  }
  __threadfence_block();
  __syncthreads();
#pragma unroll
  for (int l = 0; l < L; l++) sum.w[l] = 0;
  for (uint32_t i = i0; i < i1; i++) big_add(sum, big_load<L>(a + (uint64_t)perm2[i] * L));
  run = block_exscan_big(sum, part, &total);
  for (uint32_t i = i0; i < i1; i++) {
    big_add(run, big_load<L>(a + (uint64_t)perm2[i] * L));
    big_store(out + lay.out_at(rp, i), run);
  }
}

// The checked encoding (raa.rs:205-218 accumulate with add!, and
// raa_sign_flip.rs:146-155 with neg!).  Sums run in L + 1 limbs, exact for
// any N < 2^64: every prefix of the accumulate must fit Int<L> (the first
// prefix that does not is where the sequential checked fold panics), and
// only Int::MIN fails to negate.  Values are stored in L limbs; they equal
// the unchecked ones wherever nothing overflowed.
template <int L>
__device__ __forceinline__ Big<L + 1> big_widen(const Big<L>& a) {
  Big<L + 1> w;
#pragma unroll
  for (int i = 0; i < L; i++) w.w[i] = a.w[i];
  w.w[L] = big_is_neg(a) ? ~0ull : 0ull;
  return w;
}

template <int L>
__device__ __forceinline__ Big<L> big_narrow(const Big<L + 1>& a) {
  Big<L> n;
#pragma unroll
  for (int i = 0; i < L; i++) n.w[i] = a.w[i];
  return n;
}

template <int L>
__device__ __forceinline__ bool big_fits(const Big<L + 1>& a) {
  return a.w[L] == ((int64_t)a.w[L - 1] < 0 ? ~0ull : 0ull);
}

template <int L, bool kFlip>
__global__ void __launch_bounds__(kScanThreads)
wide_raa_checked_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                        uint64_t* __restrict__ amid, const uint32_t* __restrict__ perm1,
                        const uint32_t* __restrict__ perm2, WideLayout lay, uint32_t row_len,
                        uint32_t N, uint32_t* __restrict__ flag) {
  constexpr int LS = L + 1;
  __shared__ uint64_t part[kScanThreads * LS];
  const uint64_t rp = blockIdx.x;
  uint64_t* a = amid + rp * N * L;
  const uint32_t per = (N + kScanThreads - 1) / kScanThreads;
  const uint32_t i0 = min(N, threadIdx.x * per), i1 = min(N, i0 + per);
  uint32_t ovf = 0;
  auto y1 = [&](uint32_t i) {
    const uint32_t s = perm1[i];
    Big<L> v = lay_load<L>(in, lay, rp, s & (row_len - 1));
    if (kFlip && (s & 1u)) ovf |= big_neg_checked(v);
    return big_widen(v);
  };
  Big<LS> sum, total;
#pragma unroll
  for (int l = 0; l < LS; l++) sum.w[l] = 0;
  for (uint32_t i = i0; i < i1; i++) big_add(sum, y1(i));
  Big<LS> run = block_exscan_big(sum, part, &total);
  for (uint32_t i = i0; i < i1; i++) {
    big_add(run, y1(i));
    if (!big_fits<L>(run)) ovf |= kOvfAdd;
    Big<L> v = big_narrow<L>(run);
    if (kFlip && (i & 1u)) ovf |= big_neg_checked(v);
    big_store(a + (uint64_t)i * L, v);
  }
  __threadfence_block();
  __syncthreads();
#pragma unroll
  for (int l = 0; l < LS; l++) sum.w[l] = 0;
  for (uint32_t i = i0; i < i1; i++)
    big_add(sum, big_widen(big_load<L>(a + (uint64_t)perm2[i] * L)));
  run = block_exscan_big(sum, part, &total);
  for (uint32_t i = i0; i < i1; i++) {
    big_add(run, big_widen(big_load<L>(a + (uint64_t)perm2[i] * L)));
    if (!big_fits<L>(run)) ovf |= kOvfAdd;
    big_store(out + lay.out_at(rp, i), big_narrow<L>(run));
  }
  if (ovf) atomicOr(flag, ovf);
}

// Sign-normalised magnitude of every input element: for each limb position
// l, the or over elements x of (x < 0 ? ~x : x) limb l, into mag[l].  An
// element whose normalised value has its highest set bit at position b has
// |x| <= 2^(b+1).  Elements are kLimbs words apart, contiguous.
template <int kLimbs>
__global__ void __launch_bounds__(256)
magnitude_kernel(const uint64_t* __restrict__ in, uint64_t count,
                 unsigned long long* __restrict__ mag) {
  unsigned long long acc[kLimbs];
#pragma unroll
  for (int l = 0; l < kLimbs; l++) acc[l] = 0;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t* p = in + e * kLimbs;
    const uint64_t sx = (int64_t)p[kLimbs - 1] < 0 ? ~0ull : 0ull;
#pragma unroll
    for (int l = 0; l < kLimbs; l++) acc[l] |= p[l] ^ sx;
  }
#pragma unroll
  for (int l = 0; l < kLimbs; l++) {
    unsigned long long v = acc[l];
    for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0 && v) atomicOr(mag + l, v);
  }
}

// Words of the check buffer: the magnitude limbs, then the overflow flag.
constexpr uint32_t kCheckWords = 9;

__global__ void zero_words_kernel(unsigned long long* w, uint32_t n) {
  if (threadIdx.x < n) w[threadIdx.x] = 0;
}

// flag != nullptr: the checked kernels (per-element base layer, checked
// stages, checked accumulates), or-ing overflow bits into *flag.
template <int L>
int launch_wide(const zp_code* cd, const uint64_t* in, uint64_t seqs, const WideLayout& lay,
                uint64_t* out, hipStream_t s, uint32_t* flag) {
  const uint32_t row_len = (uint32_t)cd->row_len, N = (uint32_t)cd->N;
  if (cd->kind == ZP_CODE_IPRS) {
    const uint32_t nchunks = 1u << (3 * cd->depth);
    const uint32_t ntiles = (cd->bd + kWideTile - 1) / kWideTile;
    const uint64_t blocks = seqs * nchunks * ntiles;
    if (!flag && wide_mfma_eligible(cd) && !lay.in_bits) {
      ZP_TRY(launch_wide_mfma<L>(cd, in, seqs, lay, out, s));
    } else {
      if (blocks > 0x7fffffffull) return fail(ZP_ERR_INVALID_PARAM, "too many rows");
      if (flag)
        hipLaunchKernelGGL((wide_base_kernel<L, true>), dim3((uint32_t)blocks), dim3(kWideTile),
                           0, s, in, out, cd->d_ftab, lay, cd->bd, cd->bl, cd->depth, nchunks,
                           ntiles, flag);
      else
        hipLaunchKernelGGL((wide_base_kernel<L, false>), dim3((uint32_t)blocks), dim3(kWideTile),
                           0, s, in, out, cd->d_ftab, lay, cd->bd, cd->bl, cd->depth, nchunks,
                           ntiles, flag);
      ZP_HIP(hipGetLastError());
    }
    for (uint32_t k = 0; k < cd->depth; k++) {
      const uint64_t sub = (uint64_t)cd->bd << (3 * k);
      const uint64_t total = seqs * (N / 8);
      const dim3 grid((uint32_t)((total + 255) / 256));
      if (flag)
        hipLaunchKernelGGL((wide_stage_kernel<L, true>), grid, dim3(256), 0, s, out, cd->d_tw[k],
                           lay, (uint64_t)N, sub, total, flag);
      else
        hipLaunchKernelGGL((wide_stage_kernel<L, false>), grid, dim3(256), 0, s, out,
                           cd->d_tw[k], lay, (uint64_t)N, sub, total, flag);
      ZP_HIP(hipGetLastError());
    }
    return ZP_OK;
  }
  if (seqs > 0x7fffffffull) return fail(ZP_ERR_INVALID_PARAM, "too many rows");
  uint64_t* amid = nullptr;
  ZP_TRY(dev_alloc((void**)&amid, seqs * N * L * sizeof(uint64_t), s));
  const bool flip = cd->kind == ZP_CODE_RAA_SIGNFLIP;
  const dim3 grid((uint32_t)seqs), block(kScanThreads);
  if (flag && flip)
    hipLaunchKernelGGL((wide_raa_checked_kernel<L, true>), grid, block, 0, s, in, out, amid,
                       cd->d_perm1, cd->d_perm2, lay, row_len, N, flag);
  else if (flag)
    hipLaunchKernelGGL((wide_raa_checked_kernel<L, false>), grid, block, 0, s, in, out, amid,
                       cd->d_perm1, cd->d_perm2, lay, row_len, N, flag);
  else if (flip)
    hipLaunchKernelGGL((wide_raa_kernel<L, true>), grid, block, 0, s, in, out, amid, cd->d_perm1,
                       cd->d_perm2, lay, row_len, N);
  else
    hipLaunchKernelGGL((wide_raa_kernel<L, false>), grid, block, 0, s, in, out, amid,
                       cd->d_perm1, cd->d_perm2, lay, row_len, N);
  const int rc = hipGetLastError() == hipSuccess ? ZP_OK : fail(ZP_ERR_DEVICE, "launch failed");
  dev_free(amid, s);
  return rc;
}

int dispatch_wide(const zp_code* cd, uint32_t L, const uint64_t* in, uint64_t seqs,
                  const WideLayout& lay, uint64_t* out, hipStream_t s, uint32_t* flag = nullptr) {
  switch (L) {
    case 1: return launch_wide<1>(cd, in, seqs, lay, out, s, flag);
    case 2: return launch_wide<2>(cd, in, seqs, lay, out, s, flag);
    case 3: return launch_wide<3>(cd, in, seqs, lay, out, s, flag);
    case 4: return launch_wide<4>(cd, in, seqs, lay, out, s, flag);
    case 5: return launch_wide<5>(cd, in, seqs, lay, out, s, flag);
    case 6: return launch_wide<6>(cd, in, seqs, lay, out, s, flag);
    case 7: return launch_wide<7>(cd, in, seqs, lay, out, s, flag);
    case 8: return launch_wide<8>(cd, in, seqs, lay, out, s, flag);
  }
  return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 1..8");
}

int launch_magnitude(const uint64_t* in, uint64_t count, uint32_t limbs, unsigned long long* mag,
                     hipStream_t s) {
  if (count == 0) return ZP_OK;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(2048, (count + 255) / 256);
  switch (limbs) {
#define ZP_MAG(n) \
  case n: hipLaunchKernelGGL((magnitude_kernel<n>), dim3(blocks), dim3(256), 0, s, in, count, mag); break;
    ZP_MAG(1) ZP_MAG(2) ZP_MAG(3) ZP_MAG(4) ZP_MAG(5) ZP_MAG(6) ZP_MAG(7) ZP_MAG(8)
#undef ZP_MAG
    default: return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 1..8");
  }
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

// The checked encoding of `seqs` sequences (eval_type < 0: encode_wide
// rows of Int<L>).  Tier 1: if inputs of the type's full width cannot
// overflow the Cw type through this code (code_growth_log2), nothing is
// checked.  Tier 2 (integer inputs): a scan of the inputs' magnitudes
// decides the same for these inputs.  Tier 3: the checked kernels encode
// into `out` and the host reads their overflow bits.  *done tells the
// caller whether `out` holds the encoding (tiers 1 and 2 leave it to the
// ordinary encoder, whose result then equals the checked one).
int run_checked(const zp_code* cd, uint32_t L, const uint64_t* in, uint64_t seqs,
                const WideLayout& lay, uint64_t* out, hipStream_t s, double cw_bits,
                double in_bits, uint64_t scan_count, uint32_t scan_limbs, bool* done) {
  *done = false;
  const double growth = code_growth_log2(cd);
  if (in_bits + growth + 1e-9 < cw_bits - 1) return ZP_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  ZP_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone)
    return fail(ZP_ERR_INVALID_PARAM,
                "a checked code cannot encode these inputs under stream capture: its overflow "
                "check reads a result on the host");
  unsigned long long* buf = nullptr;
  ZP_TRY(dev_alloc((void**)&buf, kCheckWords * 8, s));
  unsigned long long host[kCheckWords] = {};
  int rc = ZP_OK;
  hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, s, buf, kCheckWords);
  if (hipGetLastError() != hipSuccess) rc = fail(ZP_ERR_DEVICE, "launch failed");
  bool small = false;
  if (rc == ZP_OK && scan_limbs) {
    rc = launch_magnitude(in, scan_count, scan_limbs, buf, s);
    if (rc == ZP_OK && (hipMemcpyAsync(host, buf, 8 * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess))
      rc = fail(ZP_ERR_DEVICE, "magnitude read failed");
    if (rc == ZP_OK) {
      double bits = 0;
      for (int l = (int)scan_limbs - 1; l >= 0; l--)
        if (host[l]) {
          bits = 64.0 * l + (64 - __builtin_clzll(host[l]));
          break;
        }
      small = bits + growth + 1e-9 < cw_bits - 1;
    }
  }
  if (rc == ZP_OK && !small) {
    uint32_t* flag = (uint32_t*)(buf + 8);
    rc = dispatch_wide(cd, L, in, seqs, lay, out, s, flag);
    if (rc == ZP_OK && (hipMemcpyAsync(&host[8], buf + 8, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess))
      rc = fail(ZP_ERR_DEVICE, "overflow flag read failed");
    if (rc == ZP_OK) {
      *done = true;
      const uint32_t f = (uint32_t)host[8];
      if (f)
        rc = fail(ZP_ERR_OVERFLOW,
                  std::string(f & kOvfMul ? "Multiplication by twiddle should not overflow"
                              : f & kOvfAdd ? (cd->kind == ZP_CODE_IPRS ? "Addition overflow"
                                                                        : "Accumulation overflow")
                                            : "Negation overflow"));
    }
  }
  dev_free(buf, s);
  return rc;
}

}  // namespace

int launch_encode_wide(const zp_code* cd, uint32_t L, const uint64_t* in, uint64_t rows,
                       uint64_t* out, hipStream_t s) {
  if (rows == 0) return ZP_OK;
  const WideLayout lay{1, false, cd->row_len * L, 0, L, cd->N * L, 0, L};
  if (cd->checked) {
    bool done = false;
    ZP_TRY(run_checked(cd, L, in, rows, lay, out, s, 64.0 * L, 64.0 * L - 1, rows * cd->row_len,
                       L, &done));
    if (done) return ZP_OK;
  }
  return dispatch_wide(cd, L, in, rows, lay, out, s);
}

bool checked_may_sync(const zp_code* cd, int eval_type, uint32_t D) {
  (void)D;
  if (!cd->checked) return false;
  const bool iprs = cd->kind == ZP_CODE_IPRS;
  const double g = code_growth_log2(cd) + 1e-9;
  switch (eval_type) {
    case ZP_EVAL_I64: return !(63 + g < 127);
    case ZP_EVAL_DENSE_I64: return !(63 + g < (iprs ? 63 : 255));
    case ZP_EVAL_BINPOLY: return iprs && !(g < 63);
  }
  return false;  // i32 rows: cleared by the bound, or refused, without a sync
}

int encode_checked(const zp_code* cd, int eval_type, uint32_t D, const void* in, uint64_t rows,
                   void* out, hipStream_t s, bool* done) {
  *done = false;
  if (rows == 0) return ZP_OK;
  const bool iprs = cd->kind == ZP_CODE_IPRS;
  const double growth = code_growth_log2(cd);
  const uint64_t* w = (const uint64_t*)in;
  uint64_t* o = (uint64_t*)out;
  switch (eval_type) {
    case ZP_EVAL_I32:  // |x| <= 2^31 into i128: no code here can overflow it
      if (31 + growth + 1e-9 < 127) return ZP_OK;
      return fail(ZP_ERR_INVALID_PARAM, "no checked encoder for i32 rows of this code");
    case ZP_EVAL_I64: {  // Cw = i128 under either code
      const WideLayout lay{1, true, cd->row_len, 1, 1, cd->N * 2, 2, 2};
      return run_checked(cd, 2, w, rows, lay, o, s, 128, 63, rows * cd->row_len, 1, done);
    }
    case ZP_EVAL_DENSE_I64: {  // Cw = [i64; D] (IPRS) or [Int<4>; D] (RAA)
      const uint32_t L = iprs ? 1 : 4;
      const WideLayout lay{D, true, cd->row_len * D, 1, D, cd->N * D * L, L, (uint64_t)D * L};
      return run_checked(cd, L, w, rows * D, lay, o, s, 64.0 * L, 63, rows * cd->row_len * D, 1,
                         done);
    }
    case ZP_EVAL_BINPOLY: {
      // Coefficients are bits (|x| <= 1): Cw = [i64; D] under IPRS, whose
      // planes the checked kernels encode one by one; Cw = [i32; D] under
      // RAA, which the code's width check at creation keeps from
      // overflowing (raa.rs:52-69).
      if (!iprs) {
        if (growth + 1e-9 < 31) return ZP_OK;
        return fail(ZP_ERR_INVALID_PARAM, "no checked encoder for binary rows of this code");
      }
      WideLayout lay{D, false, cd->row_len, 0, 1, cd->N * D, 1, D};
      lay.in_bits = true;
      return run_checked(cd, 1, w, rows * D, lay, o, s, 64, 0, 0, 0, done);
    }
  }
  return fail(ZP_ERR_INVALID_PARAM, "unknown eval type");
}

// Rows of i64 evals (Eval = i64, Cw = i128: L = 2) or of
// DensePolynomial<i64, D> evals (D coefficient planes, each encoded as its
// own sequence: Cw = [i64; D] under IPRS, L = 1, or [Int<4>; D] under RAA,
// L = 4).  Inputs are sign-extended on load; outputs are the Cw layout.
// The i32 -> i128 IPRS encoder (IntZipTypes: Eval = i32, Cw = i128) through
// the matrix-core base layer (5 digits per input) and the 2-limb stages.
int launch_encode_i32_mfma(const zp_code* cd, const int32_t* in, uint64_t rows, void* out,
                           hipStream_t s) {
  if (rows == 0) return ZP_OK;
  if (!wide_mfma_eligible(cd)) return fail(ZP_ERR_INVALID_PARAM, "code has no matrix-core tables");
  WideLayout lay{1, true, cd->row_len, 0, 1, cd->N * 2, 0, 2};
  lay.in_i32 = true;
  return dispatch_wide(cd, 2, (const uint64_t*)in, rows, lay, (uint64_t*)out, s);
}

int launch_encode_i64_planes(const zp_code* cd, uint32_t D, uint32_t L, const int64_t* in,
                             uint64_t rows, void* out, hipStream_t s) {
  if (rows == 0) return ZP_OK;
  const WideLayout lay{D, true, cd->row_len * D, 1, D, cd->N * D * L, L, (uint64_t)D * L};
  return dispatch_wide(cd, L, (const uint64_t*)in, rows * D, lay, (uint64_t*)out, s);
}

}  // namespace zpi
