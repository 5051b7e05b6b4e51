// raa.hip -- RAA (repeat-accumulate-accumulate) row encoder for gfx950.
//
// Reference: zip-plus/src/code/raa.rs:93-126 encode_inner (repeat :185-195,
// accumulate :208-236, permutation by index :239-244) and the sign-flipping
// variant zip-plus/src/code/raa_sign_flip.rs:30-71 (odd indices negated
// after the repeat and after the first accumulate, :135-165).
//
// Per coefficient plane p of a row with codeword length N:
//   y1[i] = s1(perm1[i]) * bit_p(x[perm1[i] mod row_len])
//   a[i]  = s2(i) * sum_{l <= i} y1[l]
//   out[i] = sum_{l <= i} a[perm2[i]]
// with s1, s2 = -1 at odd indices for the sign-flipping code, else +1.
// The workgroup keeps a[] in LDS (the only randomly gathered array) and
// computes both scans without materialising y1 or the second gather.
// Values stay below 2^26 in magnitude (raa.rs:52-69 width check), and the
// arithmetic wraps like the reference's i32 accumulate_unchecked.
#include "common.h"

namespace zpi {
namespace {

// Inclusive-to-exclusive block scan of one value per thread (blockDim <= 1024).
__device__ int32_t block_exclusive_scan(int32_t v, int32_t* part) {
  const uint32_t t = threadIdx.x, T = blockDim.x;
  part[t] = v;
  __syncthreads();
  for (uint32_t off = 1; off < T; off <<= 1) {
    int32_t add = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] = (int32_t)((uint32_t)part[t] + (uint32_t)add);
    __syncthreads();
  }
  const int32_t incl = part[t];
  __syncthreads();
  return (int32_t)((uint32_t)incl - (uint32_t)v);
}

__global__ void __launch_bounds__(256)
raa_binary_kernel(const uint64_t* __restrict__ evals, int32_t* __restrict__ out,
                  const uint32_t* __restrict__ perm1, const uint32_t* __restrict__ perm2,
                  uint32_t row_len, uint32_t N, uint32_t D, int sign_flip, uint64_t row0) {
  extern __shared__ int32_t a[];  // N values
  __shared__ int32_t part[256];
  const uint32_t t = threadIdx.x, T = blockDim.x;
  const uint32_t p = blockIdx.x;
  const uint64_t row = row0 + blockIdx.y;
  const uint64_t* x = evals + row * row_len;
  const uint32_t S = (N + T - 1) / T;
  const uint32_t i0 = min(N, t * S), i1 = min(N, i0 + S);
  const uint32_t rmask = row_len - 1;

  // First accumulate.
  uint32_t sum = 0;
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t src = perm1[i];
    uint32_t v = (uint32_t)(x[src & rmask] >> p) & 1u;
    if (sign_flip && (src & 1u)) v = 0u - v;
    sum += v;
  }
  uint32_t run = (uint32_t)block_exclusive_scan((int32_t)sum, part);
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t src = perm1[i];
    uint32_t v = (uint32_t)(x[src & rmask] >> p) & 1u;
    if (sign_flip && (src & 1u)) v = 0u - v;
    run += v;
    a[i] = (int32_t)((sign_flip && (i & 1u)) ? 0u - run : run);
  }
  __syncthreads();

  // Second accumulate over the perm2 gather.
  sum = 0;
  for (uint32_t i = i0; i < i1; i++) sum += (uint32_t)a[perm2[i]];
  run = (uint32_t)block_exclusive_scan((int32_t)sum, part);
  int32_t* o = out + (row * N) * D + p;
  for (uint32_t i = i0; i < i1; i++) {
    run += (uint32_t)a[perm2[i]];
    o[(uint64_t)i * D] = (int32_t)run;
  }
}

}  // namespace

int launch_raa_binary(const zp_code* cd, const uint64_t* evals, uint64_t rows, uint32_t D,
                      int32_t* out, hipStream_t s) {
  if (rows == 0) return ZP_OK;
  const uint32_t N = (uint32_t)cd->N;
  const size_t lds = (size_t)N * sizeof(int32_t);
  if (lds > 128 * 1024)
    return fail(ZP_ERR_INVALID_PARAM, "RAA codeword length above 32768 is not supported");
  if (lds > 64 * 1024)
    ZP_HIP(hipFuncSetAttribute((const void*)raa_binary_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  for (uint64_t r0 = 0; r0 < rows; r0 += 65535) {
    const uint32_t nr = (uint32_t)std::min<uint64_t>(65535, rows - r0);
    hipLaunchKernelGGL(raa_binary_kernel, dim3(D, nr), dim3(256), lds, s, evals, out,
                       cd->d_perm1, cd->d_perm2, (uint32_t)cd->row_len, N, D,
                       cd->kind == ZP_CODE_RAA_SIGNFLIP ? 1 : 0, r0);
    ZP_HIP(hipGetLastError());
  }
  return ZP_OK;
}

}  // namespace zpi
