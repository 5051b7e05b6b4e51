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
// |S| < 2^(127 + log2 num_rows) fits three signed limbs; every product and
// the final sum are taken mod 2^(64L) in two's complement, which is the
// reference's Int<L> result.
#include "common.h"

namespace zpi {
namespace {

// S_{b,c}[col] for every (b, c, col); adjacent threads take adjacent
// columns so the eval loads coalesce, and s_j is uniform per iteration.
__global__ void __launch_bounds__(256)
comb_partial_kernel(const uint64_t* __restrict__ evals, const int64_t* __restrict__ coeffs,
                    uint64_t* __restrict__ part, uint64_t num_rows, uint64_t row_len, uint32_t D,
                    uint32_t batch) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)batch * D * row_len;
  if (idx >= total) return;
  const uint64_t col = idx % row_len;
  const uint32_t c = (uint32_t)((idx / row_len) % D);
  const uint64_t b = idx / (row_len * D);
  const uint64_t* x = evals + b * num_rows * row_len + col;
  uint64_t s0 = 0, s1 = 0, s2 = 0;
  for (uint64_t j = 0; j < num_rows; j++) {
    const uint64_t w = x[j * row_len];
    if ((w >> c) & 1ull) {
      const uint64_t lo = (uint64_t)coeffs[2 * j], hi = (uint64_t)coeffs[2 * j + 1];
      const uint64_t ext = (int64_t)hi < 0 ? ~0ull : 0ull;
      const uint64_t t0 = s0 + lo;
      const uint64_t c0 = t0 < s0;
      const uint64_t t1a = s1 + hi;
      const uint64_t c1a = t1a < s1;
      const uint64_t t1 = t1a + c0;
      const uint64_t c1 = c1a | (t1 < t1a);
      s2 = s2 + ext + c1;
      s1 = t1;
      s0 = t0;
    }
  }
  uint64_t* dst = part + idx * 3;
  dst[0] = s0;
  dst[1] = s1;
  dst[2] = s2;
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
__global__ void __launch_bounds__(256)
comb_final_kernel(const uint64_t* __restrict__ part, const int64_t* __restrict__ alphas,
                  uint64_t* __restrict__ out, uint64_t row_len, uint32_t D, uint32_t batch) {
  const uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= row_len) return;
  uint64_t acc[L];
#pragma unroll
  for (int i = 0; i < L; i++) acc[i] = 0;
  for (uint32_t b = 0; b < batch; b++)
    for (uint32_t c = 0; c < D; c++) {
      const uint64_t* p = part + (((uint64_t)b * D + c) * row_len + col) * 3;
      uint64_t a[L], al[L];
      a[0] = p[0];
      a[1] = p[1];
      const uint64_t ea = (int64_t)p[2] < 0 ? ~0ull : 0ull;
#pragma unroll
      for (int i = 2; i < L; i++) a[i] = i == 2 ? p[2] : ea;
      const int64_t* av = alphas + 2 * ((uint64_t)b * D + c);
      al[0] = (uint64_t)av[0];
      al[1] = (uint64_t)av[1];
      const uint64_t eb = av[1] < 0 ? ~0ull : 0ull;
#pragma unroll
      for (int i = 2; i < L; i++) al[i] = eb;
      mac_mod<L>(acc, a, al);
    }
  uint64_t* dst = out + col * L;
#pragma unroll
  for (int i = 0; i < L; i++) dst[i] = acc[i];
}

template <int L>
void launch_final(const uint64_t* part, const int64_t* alphas, uint64_t* out, uint64_t row_len,
                  uint32_t D, uint32_t batch, hipStream_t s) {
  hipLaunchKernelGGL((comb_final_kernel<L>), dim3((uint32_t)((row_len + 255) / 256)), dim3(256),
                     0, s, part, alphas, out, row_len, D, batch);
}

}  // namespace

uint64_t combined_row_scratch_bytes(uint32_t D, uint64_t row_len, uint32_t batch) {
  return (uint64_t)batch * D * row_len * 3 * sizeof(uint64_t);
}

int launch_combined_row(uint32_t D, uint32_t L, uint64_t num_rows, uint64_t row_len,
                        uint32_t batch, const uint64_t* evals, const int64_t* alphas,
                        const int64_t* coeffs, uint64_t* out, void* scratch, hipStream_t s) {
  const uint64_t total = (uint64_t)batch * D * row_len;
  if (total == 0) return ZP_OK;
  uint64_t* part = (uint64_t*)scratch;
  hipLaunchKernelGGL(comb_partial_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s,
                     evals, coeffs, part, num_rows, row_len, D, batch);
  ZP_HIP(hipGetLastError());
  switch (L) {
    case 3: launch_final<3>(part, alphas, out, row_len, D, batch, s); break;
    case 4: launch_final<4>(part, alphas, out, row_len, D, batch, s); break;
    case 5: launch_final<5>(part, alphas, out, row_len, D, batch, s); break;
    case 6: launch_final<6>(part, alphas, out, row_len, D, batch, s); break;
    case 7: launch_final<7>(part, alphas, out, row_len, D, batch, s); break;
    case 8: launch_final<8>(part, alphas, out, row_len, D, batch, s); break;
    default: return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 3..8");
  }
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

}  // namespace zpi
