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
// |S| < 2^(127 + log2 num_rows) fits three signed 64-bit limbs; every
// product and the final sum are taken mod 2^(64L) in two's complement,
// which is the reference's Int<L> result.
//
// Kernel 1: one thread per (poly b, plane group of 8, row segment, column)
// sums s_j over its segment for the 8 planes (one eval load per row).
// Kernel 2: one thread per (column, plane) adds the segment partials,
// multiplies by alpha, and the block reduces the plane products.
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
  return (uint64_t)batch * num_segments(num_rows) * D * row_len * 3 * sizeof(uint64_t);
}

int launch_combined_row(uint32_t D, uint32_t L, uint64_t num_rows, uint64_t row_len,
                        uint32_t batch, const uint64_t* evals, const int64_t* alphas,
                        const int64_t* coeffs, uint64_t* out, void* scratch, hipStream_t s) {
  if ((uint64_t)batch * D * row_len == 0) return ZP_OK;
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
