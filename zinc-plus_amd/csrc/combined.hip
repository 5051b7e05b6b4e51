// combined.hip -- the prover's combined row for binary polynomials.
//
// Reference: zip-plus/src/pcs/phase_prove.rs:153-174 (per eval,
// w' = EvalDotChal(eval, alphas): the sum of alpha_c over the set bits c,
// poly/src/univariate/binary_u64.rs:393-432) and :205-244
// (comb[col] = sum_i sum_j s_j * w'_i[j*row_len + col] in CombR = Int<L>,
// with Int<L>::mul_by_scalar(&i128), utils/src/mul_by_scalar.rs:46-74).
//
// The device evaluates the same exact value with the sums reordered:
//   comb[col] = sum_b sum_c alpha_{b,c} * S_{b,c}[col],
//   S_{b,c}[col] = sum_j s_j * bit_c(x_b[j*row_len + col]).
// Every product and the final sum are taken mod 2^(64L) in two's
// complement, which is the reference's Int<L> result.
//
// Matrix-core path (comb_mfma_kernel).  For one column, S_{b,.}[col] is the
// product (19 x R slices of s) . (R x 64 bit matrix of the column): each
// i128 s_j is cut into 18 unsigned 7-bit slices and a signed top slice
// (s_j >> 126 in [-2, 1]), s_j = sum_m 2^(7m) slice_m(s_j).  The slices
// are the A operand of v_mfma_i32_16x16x64_i8 (two 16-row tiles, rows 19..31
// zero), shared by every column; the bits of one plane over 64 rows and 16
// columns are the B operand.  Every i32 partial is at most 127 * R in
// magnitude (R <= 2^24), so the slice sums are exact; the epilogue
// recombines them into S (192-bit two's complement; |S| < 2^151), multiplies
// by alpha mod 2^(64L) and reduces over planes.
//
// VALU path (comb_partial_kernel + comb_final_kernel, for more than 2^24
// rows or on request): one thread per (poly, 8 planes, 128-row segment,
// column) sums s_j in 64-bit words; the second kernel adds the segments,
// multiplies by alpha and reduces the planes.
#include <algorithm>

#include "common.h"

namespace zpi {
namespace {

constexpr uint32_t kSegRows = 128;

__device__ __forceinline__ void add3(uint64_t& s0, uint64_t& s1, uint64_t& s2, uint64_t a0,
                                     uint64_t a1, uint64_t a2) {
  const uint64_t t0 = s0 + a0;
  const uint64_t c0 = t0 < s0;
  const uint64_t t1a = s1 + a1;
  const uint64_t c1a = t1a < s1;
  const uint64_t t1 = t1a + c0;
  const uint64_t c1 = c1a | (t1 < t1a);
  s2 = s2 + a2 + c1;
  s1 = t1;
  s0 = t0;
}

// part[((b*nseg + seg)*D + c)*row_len + col][3]
//
// Each i128 coefficient s_j is taken as 32-bit words w0, w1, w2 (unsigned)
// and w3 (signed), s_j = w0 + 2^32 w1 + 2^64 w2 + 2^96 w3.  Per plane the
// thread keeps one 64-bit sum per word over its segment of at most 128 rows
// (each below 2^39 in magnitude, so no carry is needed inside the loop) and
// recombines the four sums into three two's-complement limbs at the end.
__global__ void __launch_bounds__(256)
comb_partial_kernel(const uint64_t* __restrict__ evals, const int64_t* __restrict__ coeffs,
                    uint64_t* __restrict__ part, uint64_t num_rows, uint64_t row_len, uint32_t D,
                    uint32_t ngroups, uint32_t nseg, uint32_t batch) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)batch * ngroups * nseg * row_len;
  if (idx >= total) return;
  const uint64_t col = idx % row_len;
  uint64_t rest = idx / row_len;
  const uint32_t seg = (uint32_t)(rest % nseg);
  rest /= nseg;
  const uint32_t g = (uint32_t)(rest % ngroups);
  const uint64_t b = rest / ngroups;
  const uint64_t j0 = (uint64_t)seg * kSegRows;
  const uint64_t j1 = min(num_rows, j0 + kSegRows);
  const uint64_t* x = evals + b * num_rows * row_len + col;
  uint64_t a0[8], a1[8], a2[8];
  int64_t a3[8];
#pragma unroll
  for (int k = 0; k < 8; k++) a0[k] = a1[k] = a2[k] = a3[k] = 0;
  for (uint64_t j = j0; j < j1; j++) {
    const uint32_t bits = (uint32_t)(x[j * row_len] >> (8 * g));
    const uint64_t lo = (uint64_t)coeffs[2 * j];
    const int64_t hi = coeffs[2 * j + 1];
    const uint32_t w0 = (uint32_t)lo, w1 = (uint32_t)(lo >> 32), w2 = (uint32_t)hi;
    const int64_t w3 = hi >> 32;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t m = 0u - ((bits >> k) & 1u);
      a0[k] += w0 & m;
      a1[k] += w1 & m;
      a2[k] += w2 & m;
      a3[k] += w3 & (int64_t)(int32_t)m;
    }
  }
  const uint32_t nc = min(8u, D - 8 * g);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if ((uint32_t)k >= nc) break;
    // a0 + 2^32 a1 + 2^64 a2 + 2^96 a3 as limbs s0, s1, s2 (a3 = 2^32 h + l).
    const uint64_t s0 = a0[k] + (a1[k] << 32);
    uint64_t s1 = (a1[k] >> 32) + a2[k] + (s0 < a0[k] ? 1u : 0u);
    const uint64_t t = s1 + ((uint64_t)a3[k] << 32);
    const uint64_t s2 = (uint64_t)(a3[k] >> 32) + (t < s1 ? 1u : 0u);
    s1 = t;
    uint64_t* dst = part + ((((b * nseg + seg) * D) + 8 * g + k) * row_len + col) * 3;
    dst[0] = s0;
    dst[1] = s1;
    dst[2] = s2;
  }
}

// acc += a * b mod 2^(64L), with a, b given as sign-extended limb arrays.
template <int L>
__device__ __forceinline__ void mac_mod(uint64_t acc[L], const uint64_t a[L], const uint64_t b[L]) {
  uint64_t r[L];
#pragma unroll
  for (int i = 0; i < L; i++) r[i] = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    uint64_t carry = 0;
#pragma unroll
    for (int j = 0; i + j < L; j++) {
      const uint64_t lo = a[i] * b[j];
      const uint64_t hi = __umul64hi(a[i], b[j]);
      uint64_t s = r[i + j] + lo;
      uint64_t c1 = s < lo;
      s += carry;
      c1 += s < carry;
      r[i + j] = s;
      carry = hi + c1;
    }
  }
  uint64_t carry = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    uint64_t s = acc[i] + r[i];
    uint64_t c1 = s < r[i];
    s += carry;
    c1 += s < carry;
    acc[i] = s;
    carry = c1;
  }
}

template <int L>
__device__ __forceinline__ void add_mod(uint64_t acc[L], const uint64_t a[L]) {
  uint64_t carry = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    uint64_t s = acc[i] + a[i];
    uint64_t c1 = s < a[i];
    s += carry;
    c1 += s < carry;
    acc[i] = s;
    carry = c1;
  }
}

// Block: 64 threads per column (planes c, c+64, ... for D > 64 is not
// needed: D <= 64), kColsPerBlock columns.
constexpr uint32_t kColsPerBlock = 4;

template <int L>
__global__ void __launch_bounds__(256)
comb_final_kernel(const uint64_t* __restrict__ part, const int64_t* __restrict__ alphas,
                  uint64_t* __restrict__ out, uint64_t row_len, uint32_t D, uint32_t nseg,
                  uint32_t batch) {
  __shared__ uint64_t red[kColsPerBlock][64][L];
  const uint32_t c = threadIdx.x & 63, lc = threadIdx.x >> 6;
  const uint64_t col = (uint64_t)blockIdx.x * kColsPerBlock + lc;
  uint64_t acc[L];
#pragma unroll
  for (int i = 0; i < L; i++) acc[i] = 0;
  if (col < row_len && c < D) {
    for (uint32_t b = 0; b < batch; b++) {
      uint64_t s0 = 0, s1 = 0, s2 = 0;
      for (uint32_t seg = 0; seg < nseg; seg++) {
        const uint64_t* p = part + ((((uint64_t)b * nseg + seg) * D + c) * row_len + col) * 3;
        add3(s0, s1, s2, p[0], p[1], p[2]);
      }
      uint64_t a[L], al[L];
      a[0] = s0;
      a[1] = s1;
      const uint64_t ea = (int64_t)s2 < 0 ? ~0ull : 0ull;
#pragma unroll
      for (int i = 2; i < L; i++) a[i] = i == 2 ? s2 : ea;
      const int64_t* av = alphas + 2 * ((uint64_t)b * D + c);
      al[0] = (uint64_t)av[0];
      al[1] = (uint64_t)av[1];
      const uint64_t eb = av[1] < 0 ? ~0ull : 0ull;
#pragma unroll
      for (int i = 2; i < L; i++) al[i] = eb;
      mac_mod<L>(acc, a, al);
    }
  }
#pragma unroll
  for (int i = 0; i < L; i++) red[lc][c][i] = acc[i];
  __syncthreads();
  for (uint32_t off = 32; off > 0; off >>= 1) {
    if (c < off) {
      uint64_t other[L];
#pragma unroll
      for (int i = 0; i < L; i++) other[i] = red[lc][c + off][i];
#pragma unroll
      for (int i = 0; i < L; i++) acc[i] = red[lc][c][i];
      add_mod<L>(acc, other);
#pragma unroll
      for (int i = 0; i < L; i++) red[lc][c][i] = acc[i];
    }
    __syncthreads();
  }
  if (c == 0 && col < row_len) {
#pragma unroll
    for (int i = 0; i < L; i++) out[col * L + i] = red[lc][0][i];
  }
}


// ---- Matrix-core path --------------------------------------------------

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr uint32_t kMaxMfmaRows = 1u << 24;  // keeps 127 * R below 2^31
constexpr uint32_t kZPitch = 20;             // words per (byte, column) row of the LDS tile

// The 7-bit slice m of an i128 (lo, hi): unsigned for m < 18, the signed top
// two bits for m = 18, zero above.
__device__ __forceinline__ int32_t s_slice(uint64_t lo, int64_t hi, uint32_t m) {
  if (m >= 19) return 0;
  if (m == 18) return (int32_t)(hi >> 62);
  const uint32_t b = 7 * m;
  uint64_t v;
  if (b + 7 <= 64)
    v = lo >> b;
  else if (b >= 64)
    v = (uint64_t)hi >> (b - 64);
  else
    v = (lo >> b) | ((uint64_t)hi << (64 - b));
  return (int32_t)(v & 127);
}

// acc (3 limbs, mod 2^192) += v * 2^sh for a signed 64-bit v.
__device__ __forceinline__ void add_shifted3(uint64_t acc[3], int64_t v, uint32_t sh) {
  if (sh >= 192) return;
  const uint64_t ext = v < 0 ? ~0ull : 0ull;
  uint64_t t[3];
  const uint32_t q = sh / 64, r = sh % 64;
  const uint64_t w0 = (uint64_t)v << r;
  const uint64_t w1 = r ? (((uint64_t)v >> (64 - r)) | (ext << r)) : ext;
#pragma unroll
  for (uint32_t i = 0; i < 3; i++) t[i] = i < q ? 0 : i == q ? w0 : i == q + 1 ? w1 : ext;
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const uint64_t s1 = acc[i] + t[i];
    const uint64_t c1 = s1 < t[i];
    const uint64_t s2 = s1 + c;
    c = c1 | (s2 < c);
    acc[i] = s2;
  }
}

template <int L>
__device__ __forceinline__ void addL(uint64_t acc[L], const uint64_t a[L]) {
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const uint64_t s1 = acc[i] + a[i];
    const uint64_t c1 = s1 < a[i];
    const uint64_t s2 = s1 + c;
    c = c1 | (s2 < c);
    acc[i] = s2;
  }
}

// acc (L limbs) += V * A mod 2^(64L), V a 192-bit and A a 128-bit two's
// complement value.  The unsigned 320-bit product of the patterns, minus
// 2^192 A if V < 0 and 2^128 V if A < 0, is the signed product mod 2^320;
// it is sign-extended to L limbs (or truncated).
template <int L>
__device__ __forceinline__ void mac_192x128(uint64_t acc[L], const uint64_t V[3],
                                            const uint64_t A[2]) {
  uint64_t p[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    uint64_t carry = 0;
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const uint64_t lo = V[i] * A[j], hi = __umul64hi(V[i], A[j]);
      uint64_t s1 = p[i + j] + lo;
      uint64_t c1 = s1 < lo;
      const uint64_t s2 = s1 + carry;
      c1 += s2 < carry;
      p[i + j] = s2;
      carry = hi + c1;
    }
    p[i + 2] = carry;  // p[i + 2] is still zero here
  }
  // p[3..4] -= A if V < 0; p[2..4] -= V if A < 0 (mod 2^320).
  if ((int64_t)V[2] < 0) {
    const uint64_t d3 = p[3] - A[0], b3 = p[3] < A[0];
    p[3] = d3;
    p[4] = p[4] - A[1] - b3;
  }
  if ((int64_t)A[1] < 0) {
    const uint64_t d2 = p[2] - V[0], b2 = p[2] < V[0];
    const uint64_t t3 = p[3] - V[1], b3a = p[3] < V[1];
    const uint64_t d3 = t3 - b2, b3 = b3a | (t3 < b2);
    p[2] = d2;
    p[3] = d3;
    p[4] = p[4] - V[2] - b3;
  }
  uint64_t r[L];
  const uint64_t ext = (int64_t)p[4] < 0 ? ~0ull : 0ull;
#pragma unroll
  for (int i = 0; i < L; i++) r[i] = i < 5 ? p[i] : ext;
  addL<L>(acc, r);
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}

// Grid: (poly b, column group, row segment).  A workgroup of 4 waves takes
// 16 T columns (T = 4 / QP tiles of 16) and the K-steps [kt0, kt1) of 64
// rows; wave (tile, q) owns planes 16 q .. 16 q + 15 of its tile (active for
// q < Q).
//
// Prologue: the workgroup writes the A operand of its K-steps to LDS: lane
// l of the fragment (K-step kt, row tile mt) holds slice 16 mt + l % 16 of
// s_j for j = 64 kt + 16 (l / 16) + e in byte e.  Rows 19..31 are left
// unwritten: they only reach accumulator rows the epilogue never reads.
//
// K-step: every thread loads 4 rows of one column per tile (a ring of PF
// steps in registers keeps PF steps of loads in flight), byte-transposes
// them (byte beta of the 4 rows -> one word) into LDS, and each active wave
// reads, per byte beta it owns, the 16 rows of its lane group as one 128-bit
// word: (z >> c') & 0x01010101 is the B operand of plane 8 beta + c'.  The
// eval tile is double-buffered: one barrier per K-step.
constexpr uint32_t kMaxSegSteps = 16;

template <int L, int QP>
__global__ void __launch_bounds__(256, 2)
comb_mfma_kernel(const uint64_t* __restrict__ evals, const int64_t* __restrict__ coeffs,
                 const int64_t* __restrict__ alphas, uint64_t* __restrict__ out,
                 uint64_t num_rows, uint64_t row_len, uint32_t D, uint32_t ngroups,
                 uint32_t steps_per_seg, uint32_t ksteps, uint32_t nseg) {
  constexpr uint32_t T = 4 / QP;   // column tiles of 16 per workgroup
  constexpr uint32_t PF = 4 / T;   // K-steps of loads in flight
  // One LDS pool: the double-buffered eval tile, then the s slices; the
  // epilogue reuses the whole pool.
  constexpr uint32_t kTileWords = 128 * kZPitch;
  constexpr uint32_t kPoolWords = 2 * kTileWords + kMaxSegSteps * 2 * 64 * 4;
  __shared__ __align__(16) uint32_t pool[kPoolWords];
  __shared__ uint64_t red[4][16][L];
  uint32_t(*zt)[kTileWords] = (uint32_t(*)[kTileWords])pool;
  v4i* afr = (v4i*)(pool + 2 * kTileWords);
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t Q = (D + 15) / 16;  // plane chunks in use
  const uint32_t NB = 2 * Q;         // eval bytes in use (byte beta = planes 8 beta ..)
  const uint32_t tile = wave / QP, q = wave % QP;
  const bool active = q < Q;
  uint64_t rest = blockIdx.x;
  const uint32_t seg = (uint32_t)(rest % nseg);
  rest /= nseg;
  const uint32_t grp = (uint32_t)(rest % ngroups);
  const uint64_t b = rest / ngroups;
  const uint64_t col0 = (uint64_t)grp * 16 * T;
  const uint32_t kt0 = seg * steps_per_seg, nst = min(ksteps, kt0 + steps_per_seg) - kt0;
  const uint64_t* xb = evals + b * num_rows * row_len;
  constexpr uint32_t cols = 16 * T;

  // Loader: task t = tid + 256 i (i < T): column t % cols, rows 4 (t / cols) + 0..3.
  uint64_t xr[PF][T][4];
  auto load = [&](uint64_t (&x)[T][4], uint32_t kt) {
#pragma unroll
    for (uint32_t i = 0; i < T; i++) {
      const uint32_t t = tid + 256 * i;
      const uint64_t col = col0 + t % cols;
      const uint64_t j = (uint64_t)kt * 64 + 4 * (t / cols);
#pragma unroll
      for (uint32_t r = 0; r < 4; r++)
        x[i][r] = (col < row_len && j + r < num_rows) ? xb[(j + r) * row_len + col] : 0ull;
    }
  };
  auto store = [&](const uint64_t (&x)[T][4], uint32_t buf) {
#pragma unroll
    for (uint32_t i = 0; i < T; i++) {
      const uint32_t t = tid + 256 * i;
      const uint32_t c = t % cols, jq = t / cols;
#pragma unroll
      for (uint32_t h = 0; h < 2; h++) {
        if (4 * h >= NB) break;
        const uint32_t e0 = (uint32_t)(x[i][0] >> (32 * h)), e1 = (uint32_t)(x[i][1] >> (32 * h));
        const uint32_t e2 = (uint32_t)(x[i][2] >> (32 * h)), e3 = (uint32_t)(x[i][3] >> (32 * h));
        // 4 x 4 byte transpose: z_beta = byte beta of e0, e1, e2, e3.
        const uint32_t u0 = __builtin_amdgcn_perm(e1, e0, 0x05010400u);
        const uint32_t u1 = __builtin_amdgcn_perm(e1, e0, 0x07030602u);
        const uint32_t w0 = __builtin_amdgcn_perm(e3, e2, 0x05010400u);
        const uint32_t w1 = __builtin_amdgcn_perm(e3, e2, 0x07030602u);
        const uint32_t z[4] = {__builtin_amdgcn_perm(w0, u0, 0x05040100u),
                               __builtin_amdgcn_perm(w0, u0, 0x07060302u),
                               __builtin_amdgcn_perm(w1, u1, 0x05040100u),
                               __builtin_amdgcn_perm(w1, u1, 0x07060302u)};
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
          const uint32_t beta = 4 * h + k;
          if (beta < NB) zt[buf][(beta * cols + c) * kZPitch + jq] = z[k];
        }
      }
    }
  };

#pragma unroll
  for (uint32_t u = 0; u < PF; u++)
    if (u < nst) load(xr[u], kt0 + u);

  // Prologue: the s slices of rows 64 kt0 .. 64 (kt0 + nst) - 1.
  {
    uint8_t* ab = (uint8_t*)afr;
    for (uint32_t jl = tid; jl < nst * 64; jl += 256) {
      const uint64_t j = (uint64_t)kt0 * 64 + jl;
      uint64_t lo = 0;
      int64_t hi = 0;
      if (j < num_rows) {
        lo = (uint64_t)coeffs[2 * j];
        hi = coeffs[2 * j + 1];
      }
      const uint32_t base = ((jl / 64) * 2 * 64 + ((jl % 64) / 16) * 16) * 16 + jl % 16;
#pragma unroll
      for (uint32_t m = 0; m < 19; m++)
        ab[base + ((m / 16) * 64 + m % 16) * 16] = (uint8_t)s_slice(lo, hi, m);
    }
  }

  v4i acc[16][2];
#pragma unroll
  for (int p = 0; p < 16; p++) acc[p][0] = acc[p][1] = v4i{0, 0, 0, 0};
  const uint32_t g = lane >> 4, n = lane & 15;
  for (uint32_t base = 0; base < nst; base += PF) {
#pragma unroll
    for (uint32_t u = 0; u < PF; u++) {
      const uint32_t st = base + u;
      if (st >= nst) break;
      const uint32_t buf = st & 1;
      store(xr[u], buf);
      if (st + PF < nst) load(xr[u], kt0 + st + PF);
      __syncthreads();
      if (active) {
        const v4i A0 = afr[(st * 2 + 0) * 64 + lane];
        const v4i A1 = afr[(st * 2 + 1) * 64 + lane];
#pragma unroll
        for (uint32_t bb = 0; bb < 2; bb++) {
          const uint32_t beta = 2 * q + bb;
          const uint4 z = *(const uint4*)&zt[buf][(beta * cols + 16 * tile + n) * kZPitch + 4 * g];
          // Plane c' of the byte: B = z & (0x01010101 << c') for c' = 0, and
          // (z >> 1) & (0x01010101 << (c' - 1)) for c' >= 1, i.e. the bits
          // scaled by 2^max(c' - 1, 0) <= 64 (an int8); the epilogue shifts
          // the exact multiple back.
          const uint4 z1 = make_uint4(z.x >> 1, z.y >> 1, z.z >> 1, z.w >> 1);
#pragma unroll
          for (uint32_t c = 0; c < 8; c++) {
            if (16 * q + 8 * bb + c >= D) continue;
            const uint4 zz = c == 0 ? z : z1;
            const uint32_t mk = 0x01010101u << (c == 0 ? 0 : c - 1);
            v4i bf;
            bf[0] = (int)(zz.x & mk);
            bf[1] = (int)(zz.y & mk);
            bf[2] = (int)(zz.z & mk);
            bf[3] = (int)(zz.w & mk);
            acc[8 * bb + c][0] =
                __builtin_amdgcn_mfma_i32_16x16x64_i8(A0, bf, acc[8 * bb + c][0], 0, 0, 0);
            acc[8 * bb + c][1] =
                __builtin_amdgcn_mfma_i32_16x16x64_i8(A1, bf, acc[8 * bb + c][1], 0, 0, 0);
          }
        }
      }
    }
  }

  // Epilogue, in two halves of 8 planes through LDS (the eval tile and the
  // slices are no longer needed): lane (g, n) of wave w holds slices 4g + v
  // (row tile 0) and, for g = 0, 16 + v (v < 3, row tile 1) of S_c[col n]
  // for its planes, scaled by 2^max(c' - 1, 0).  Each wave writes them to
  // E[w][p][n][m]; lane (g, n) then takes planes p = g and g + 4 of the half
  // for column n: it recombines all 19 slices into a 192-bit value,
  // multiplies by alpha and accumulates.
  uint64_t comb[L];
#pragma unroll
  for (int i = 0; i < L; i++) comb[i] = 0;
  constexpr uint32_t kEP = 20;  // words per (plane, column) entry
  int32_t* E = (int32_t*)pool + wave * (8 * 16 * kEP);
  static_assert(kPoolWords >= 4 * 8 * 16 * kEP, "epilogue LDS");
#pragma unroll
  for (uint32_t half = 0; half < 2; half++) {
    __syncthreads();  // the K loop (or the previous half) is done with this LDS
    if (active) {
#pragma unroll
      for (uint32_t pl = 0; pl < 8; pl++) {
        const uint32_t p = 8 * half + pl;
        const int sc = (p & 7) == 0 ? 0 : (int)(p & 7) - 1;
        int32_t* e = E + (pl * 16 + n) * kEP;
        *(int4*)(e + 4 * g) = make_int4(acc[p][0][0] >> sc, acc[p][0][1] >> sc,
                                        acc[p][0][2] >> sc, acc[p][0][3] >> sc);
        if (g == 0) {
          e[16] = acc[p][1][0] >> sc;
          e[17] = acc[p][1][1] >> sc;
          e[18] = acc[p][1][2] >> sc;
        }
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (uint32_t k = 0; k < 2; k++) {
        const uint32_t pl = g + 4 * k, c = 16 * q + 8 * half + pl;
        if (c >= D) continue;
        const int32_t* e = E + (pl * 16 + n) * kEP;
        const int4 s0 = *(const int4*)(e + 0), s1 = *(const int4*)(e + 4);
        const int4 s2 = *(const int4*)(e + 8), s3 = *(const int4*)(e + 12);
        const int4 s4 = *(const int4*)(e + 16);
        const int32_t sl[19] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w, s2.x, s2.y,
                                s2.z, s2.w, s3.x, s3.y, s3.z, s3.w, s4.x, s4.y, s4.z};
        // Groups of 5 slices (|slice sum| < 2^23, so each group is below
        // 2^52 in magnitude) at bit offsets 0, 35, 70, 105.
        uint64_t V[3] = {0, 0, 0};
#pragma unroll
        for (int gi = 0; gi < 4; gi++) {
          int64_t t = 0;
#pragma unroll
          for (int m = 4; m >= 0; m--) {
            const int idx = 5 * gi + m;
            if (idx < 19) t = t * 128 + sl[idx];
          }
          add_shifted3(V, t, 35 * gi);
        }
        const int64_t* av = alphas + 2 * (b * D + c);
        const uint64_t A[2] = {(uint64_t)av[0], (uint64_t)av[1]};
        mac_192x128<L>(comb, V, A);
      }
    }
  }
  // Sum over the lane groups, then over the plane waves of the tile.
#pragma unroll
  for (int m = 16; m <= 32; m *= 2) {
    uint64_t o[L];
#pragma unroll
    for (int i = 0; i < L; i++) o[i] = shfl_xor64(comb[i], m);
    addL<L>(comb, o);
  }
  if (g == 0) {
#pragma unroll
    for (int i = 0; i < L; i++) red[wave][n][i] = comb[i];
  }
  __syncthreads();
  if (q == 0 && g == 0) {
    for (uint32_t w = 1; w < QP; w++) addL<L>(comb, red[wave + w][n]);
    const uint64_t col = col0 + 16 * tile + n;
    if (col < row_len) {
      uint64_t* dst = out + ((b * nseg + seg) * row_len + col) * L;
#pragma unroll
      for (int i = 0; i < L; i++) dst[i] = comb[i];
    }
  }
}

// out[col] = sum over the n_part partial rows mod 2^(64L).
template <int L>
__global__ void __launch_bounds__(256)
comb_sum_kernel(const uint64_t* __restrict__ part, uint64_t* __restrict__ out, uint64_t row_len,
                uint32_t n_part) {
  const uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= row_len) return;
  uint64_t acc[L];
#pragma unroll
  for (int i = 0; i < L; i++) acc[i] = part[col * L + i];
  for (uint32_t k = 1; k < n_part; k++) {
    uint64_t a[L];
#pragma unroll
    for (int i = 0; i < L; i++) a[i] = part[((uint64_t)k * row_len + col) * L + i];
    addL<L>(acc, a);
  }
#pragma unroll
  for (int i = 0; i < L; i++) out[col * L + i] = acc[i];
}

struct MfmaShape {
  uint32_t QP, T, ngroups, ksteps, steps_per_seg, nseg;
};

MfmaShape mfma_shape(uint32_t D, uint64_t num_rows, uint64_t row_len, uint32_t batch) {
  MfmaShape m;
  const uint32_t Q = (D + 15) / 16;
  m.QP = Q <= 1 ? 1 : Q <= 2 ? 2 : 4;
  m.T = 4 / m.QP;
  m.ngroups = (uint32_t)((row_len + 16 * m.T - 1) / (16 * m.T));
  m.ksteps = (uint32_t)((num_rows + 63) / 64);
  // Split the rows until there are about 512 workgroups (two per CU), with
  // at most kMaxSegSteps K-steps (their s slices in LDS) per workgroup.
  const uint64_t base = (uint64_t)batch * m.ngroups;
  uint32_t want = (uint32_t)std::min<uint64_t>(m.ksteps, (512 + base - 1) / base);
  want = std::max(1u, want);
  m.steps_per_seg = std::min((m.ksteps + want - 1) / want, kMaxSegSteps);
  m.nseg = (m.ksteps + m.steps_per_seg - 1) / m.steps_per_seg;
  return m;
}

bool use_mfma(uint64_t num_rows) { return g_force_comb != 1 && num_rows <= kMaxMfmaRows; }

template <int L>
int launch_mfma(uint32_t D, uint64_t num_rows, uint64_t row_len, uint32_t batch,
                const uint64_t* evals, const int64_t* alphas, const int64_t* coeffs,
                uint64_t* out, void* scratch, hipStream_t s) {
  const MfmaShape m = mfma_shape(D, num_rows, row_len, batch);
  auto kern = m.QP == 4 ? comb_mfma_kernel<L, 4> : m.QP == 2 ? comb_mfma_kernel<L, 2>
                                                             : comb_mfma_kernel<L, 1>;
  uint64_t* part = (uint64_t*)scratch;
  const uint32_t n_part = batch * m.nseg;
  const uint64_t blocks = (uint64_t)batch * m.ngroups * m.nseg;
  hipLaunchKernelGGL(kern, dim3((uint32_t)blocks), dim3(256), 0, s, evals, coeffs, alphas,
                     n_part == 1 ? out : part, num_rows, row_len, D, m.ngroups, m.steps_per_seg,
                     m.ksteps, m.nseg);
  if (n_part > 1)
    hipLaunchKernelGGL((comb_sum_kernel<L>), dim3((uint32_t)((row_len + 255) / 256)), dim3(256), 0,
                       s, part, out, row_len, n_part);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

template <int L>
void launch_final(const uint64_t* part, const int64_t* alphas, uint64_t* out, uint64_t row_len,
                  uint32_t D, uint32_t nseg, uint32_t batch, hipStream_t s) {
  const uint32_t blocks = (uint32_t)((row_len + kColsPerBlock - 1) / kColsPerBlock);
  hipLaunchKernelGGL((comb_final_kernel<L>), dim3(blocks), dim3(64 * kColsPerBlock), 0, s, part,
                     alphas, out, row_len, D, nseg, batch);
}

uint32_t num_segments(uint64_t num_rows) {
  return (uint32_t)((num_rows + kSegRows - 1) / kSegRows);
}

}  // namespace

uint64_t combined_row_scratch_bytes(uint32_t D, uint64_t row_len, uint32_t batch,
                                    uint64_t num_rows) {
  if (use_mfma(num_rows)) {
    const MfmaShape m = mfma_shape(D, num_rows, row_len, batch);
    return (uint64_t)batch * m.nseg * row_len * 8 * sizeof(uint64_t);
  }
  return (uint64_t)batch * num_segments(num_rows) * D * row_len * 3 * sizeof(uint64_t);
}

int launch_combined_row(uint32_t D, uint32_t L, uint64_t num_rows, uint64_t row_len,
                        uint32_t batch, const uint64_t* evals, const int64_t* alphas,
                        const int64_t* coeffs, uint64_t* out, void* scratch, hipStream_t s) {
  if ((uint64_t)batch * D * row_len == 0) return ZP_OK;
  if (use_mfma(num_rows)) {
    switch (L) {
      case 3: return launch_mfma<3>(D, num_rows, row_len, batch, evals, alphas, coeffs, out, scratch, s);
      case 4: return launch_mfma<4>(D, num_rows, row_len, batch, evals, alphas, coeffs, out, scratch, s);
      case 5: return launch_mfma<5>(D, num_rows, row_len, batch, evals, alphas, coeffs, out, scratch, s);
      case 6: return launch_mfma<6>(D, num_rows, row_len, batch, evals, alphas, coeffs, out, scratch, s);
      case 7: return launch_mfma<7>(D, num_rows, row_len, batch, evals, alphas, coeffs, out, scratch, s);
      case 8: return launch_mfma<8>(D, num_rows, row_len, batch, evals, alphas, coeffs, out, scratch, s);
      default: return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 3..8");
    }
  }
  uint64_t* part = (uint64_t*)scratch;
  const uint32_t ngroups = (D + 7) / 8, nseg = num_segments(num_rows);
  const uint64_t total = (uint64_t)batch * ngroups * nseg * row_len;
  hipLaunchKernelGGL(comb_partial_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s,
                     evals, coeffs, part, num_rows, row_len, D, ngroups, nseg, batch);
  ZP_HIP(hipGetLastError());
  switch (L) {
    case 3: launch_final<3>(part, alphas, out, row_len, D, nseg, batch, s); break;
    case 4: launch_final<4>(part, alphas, out, row_len, D, nseg, batch, s); break;
    case 5: launch_final<5>(part, alphas, out, row_len, D, nseg, batch, s); break;
    case 6: launch_final<6>(part, alphas, out, row_len, D, nseg, batch, s); break;
    case 7: launch_final<7>(part, alphas, out, row_len, D, nseg, batch, s); break;
    case 8: launch_final<8>(part, alphas, out, row_len, D, nseg, batch, s); break;
    default: return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 3..8");
  }
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

}  // namespace zpi
