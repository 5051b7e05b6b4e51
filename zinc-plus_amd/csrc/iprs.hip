// iprs.hip -- IPRS radix-8 pseudo-NTT row encoders for gfx950.
//
// Reference: zip-plus/src/code/iprs/pntt/radix8.rs:21-161 (pntt,
// base_multiply_into_output, combine_stages) with butterfly.rs:8-43.
//
// Structure used here.  Write the output index as i = c*bd + r (chunk
// c < 8^depth, base row r < bd).  The base layer gives
//   B[c][r] = sum_k f(r*k mod bd) * x[octrev(c) + k*8^depth]
// and stage s combines, for the same r, the 8 chunks
//   c = cb + j*8^s + c_lo   (j = 0..7, c_lo < 8^s)
// with twiddles tw_s[c_lo*bd + r][j-1][(t*j) mod 8].  Every stage keeps r
// fixed, so a workgroup that owns a few r values and all 8^depth chunks of a
// row runs the whole transform on chip and writes each output once.
//
// Exactness: for binary inputs |B| < 2^23 (i32), stage 0 < 2^41 and stage 1
// < 2^59 (i64), so depths up to 2 (every new_with_optimal_depth code) never
// overflow.  An explicit depth-3 code can pass 2^63 at stage 2: there the
// i64 arithmetic wraps, exactly as the reference's unchecked build does (its
// checked build panics instead).  i32 inputs need i64 base sums (< 2^54)
// and i128 stages, which cannot overflow.
#include "common.h"

namespace zpi {
namespace {

// octet_reversal (octet_reversal.rs:6-29): reverse `d` base-8 digits.
__device__ __forceinline__ uint32_t octrev(uint32_t c, uint32_t d) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < d; i++) {
    r = (r << 3) | (c & 7u);
    c >>= 3;
  }
  return r;
}

// Binary polynomials: D coefficient planes, processed PG planes at a time.
// Threads: RB base rows x NC chunks; thread t owns (r = r0 + t / NC,
// c = t % NC).  LDS holds one i64 per (plane, thread) in [plane][thread]
// order.
template <int NC, int PG>
__global__ void __launch_bounds__(512)
iprs_binary_kernel(const uint64_t* __restrict__ evals, int64_t* __restrict__ out,
                   const int32_t* __restrict__ ftab, const int32_t* __restrict__ tw0,
                   const int32_t* __restrict__ tw1, const int32_t* __restrict__ tw2,
                   uint32_t row_len, uint32_t bd, uint32_t bl, uint32_t D, uint32_t depth,
                   uint32_t RB, uint64_t row0) {
  extern __shared__ int64_t lds[];
  const uint32_t BS = blockDim.x;
  const uint32_t t = threadIdx.x;
  const uint32_t rl = t / NC, c = t % NC;
  const uint32_t r = blockIdx.x * RB + rl;
  const uint64_t row = row0 + blockIdx.y;
  const uint64_t N = (uint64_t)bd * NC;
  const uint64_t* x = evals + row * row_len;
  int64_t* o = out + (row * N + (uint64_t)c * bd + r) * D;
  const uint32_t shift = 3 * depth;
  const uint32_t oc = octrev(c, depth);
  const uint32_t mask = bd - 1;

  for (uint32_t p0 = 0; p0 < D; p0 += PG) {
    int32_t acc[PG];
#pragma unroll
    for (int q = 0; q < PG; q++) acc[q] = 0;
    for (uint32_t k = 0; k < bl; k++) {
      const uint64_t w = x[oc | (k << shift)] >> p0;
      const int32_t f = ftab[(r * k) & mask];
      const uint32_t lo = (uint32_t)w;
      const uint32_t hi = (uint32_t)(w >> 32);
#pragma unroll
      for (int q = 0; q < PG; q++) {
        const uint32_t bit = q < 32 ? (lo >> q) & 1u : (hi >> (q - 32)) & 1u;
        acc[q] += (int32_t)bit * f;
      }
    }
#pragma unroll
    for (int q = 0; q < PG; q++) lds[q * BS + t] = acc[q];
    __syncthreads();

    uint32_t span = 1;
    for (uint32_t s = 0; s < depth; s++) {
      const int32_t* tw = s == 0 ? tw0 : (s == 1 ? tw1 : tw2);
      const uint32_t c_lo = c % span;
      const uint32_t jself = (c / span) & 7u;
      const uint32_t cb = c - c_lo - jself * span;
      const int32_t* twi = tw + ((uint64_t)c_lo * bd + r) * 56;
      const uint32_t base_t = rl * NC + cb + c_lo;
      uint64_t y[PG];
#pragma unroll
      for (int q = 0; q < PG; q++) y[q] = (uint64_t)lds[q * BS + base_t];
      for (uint32_t j = 1; j < 8; j++) {
        const uint64_t tv = (uint64_t)(int64_t)twi[(j - 1) * 8 + ((jself * j) & 7u)];
        const uint32_t tj = base_t + j * span;
#pragma unroll
        for (int q = 0; q < PG; q++) y[q] += (uint64_t)lds[q * BS + tj] * tv;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PG; q++) lds[q * BS + t] = (int64_t)y[q];
      __syncthreads();
      span *= 8;
    }
#pragma unroll
    for (int q = 0; q < PG; q++)
      if (p0 + q < D) o[p0 + q] = lds[q * BS + t];
    __syncthreads();
  }
}

// Eval = i32, Cw = i128 (i32 x i64 -> i128, utils/src/mul_by_scalar.rs:89-108).
template <int NC>
__global__ void __launch_bounds__(512)
iprs_i32_kernel(const int32_t* __restrict__ evals, __int128* __restrict__ out,
                const int32_t* __restrict__ ftab, const int32_t* __restrict__ tw0,
                const int32_t* __restrict__ tw1, const int32_t* __restrict__ tw2,
                uint32_t row_len, uint32_t bd, uint32_t bl, uint32_t depth, uint32_t RB,
                uint64_t row0) {
  extern __shared__ __int128 lds128[];
  const uint32_t t = threadIdx.x;
  const uint32_t rl = t / NC, c = t % NC;
  const uint32_t r = blockIdx.x * RB + rl;
  const uint64_t row = row0 + blockIdx.y;
  const uint64_t N = (uint64_t)bd * NC;
  const int32_t* x = evals + row * row_len;
  const uint32_t shift = 3 * depth;
  const uint32_t oc = octrev(c, depth);
  const uint32_t mask = bd - 1;

  int64_t acc = 0;
  for (uint32_t k = 0; k < bl; k++)
    acc += (int64_t)x[oc | (k << shift)] * (int64_t)ftab[(r * k) & mask];
  lds128[t] = (__int128)acc;
  __syncthreads();
  uint32_t span = 1;
  for (uint32_t s = 0; s < depth; s++) {
    const int32_t* tw = s == 0 ? tw0 : (s == 1 ? tw1 : tw2);
    const uint32_t c_lo = c % span;
    const uint32_t jself = (c / span) & 7u;
    const uint32_t cb = c - c_lo - jself * span;
    const int32_t* twi = tw + ((uint64_t)c_lo * bd + r) * 56;
    const uint32_t base_t = rl * NC + cb + c_lo;
    unsigned __int128 y = (unsigned __int128)lds128[base_t];
    for (uint32_t j = 1; j < 8; j++) {
      const __int128 tv = (__int128)twi[(j - 1) * 8 + ((jself * j) & 7u)];
      y += (unsigned __int128)(lds128[base_t + j * span] * tv);
    }
    __syncthreads();
    lds128[t] = (__int128)y;
    __syncthreads();
    span *= 8;
  }
  out[row * N + (uint64_t)c * bd + r] = lds128[t];
}

constexpr int kPlaneGroup = 16;

template <int NC>
int launch_binary_nc(const zp_code* cd, const uint64_t* evals, uint64_t rows, uint32_t D,
                     int64_t* out, hipStream_t s) {
  const uint32_t RB = std::min<uint32_t>(cd->bd, NC >= 256 ? 1 : 256 / NC);
  const uint32_t threads = RB * NC;
  const size_t lds = (size_t)threads * kPlaneGroup * sizeof(int64_t);
  for (uint64_t r0 = 0; r0 < rows; r0 += 65535) {
    const uint32_t nr = (uint32_t)std::min<uint64_t>(65535, rows - r0);
    dim3 grid(cd->bd / RB, nr);
    hipLaunchKernelGGL((iprs_binary_kernel<NC, kPlaneGroup>), grid, dim3(threads), lds, s, evals,
                       out, cd->d_ftab, cd->d_tw[0], cd->d_tw[1], cd->d_tw[2],
                       (uint32_t)cd->row_len, cd->bd, cd->bl, D, cd->depth, RB, r0);
    ZP_HIP(hipGetLastError());
  }
  return ZP_OK;
}

template <int NC>
int launch_i32_nc(const zp_code* cd, const int32_t* evals, uint64_t rows, void* out,
                  hipStream_t s) {
  const uint32_t RB = std::min<uint32_t>(cd->bd, NC >= 256 ? 1 : 256 / NC);
  const uint32_t threads = RB * NC;
  const size_t lds = (size_t)threads * sizeof(__int128);
  for (uint64_t r0 = 0; r0 < rows; r0 += 65535) {
    const uint32_t nr = (uint32_t)std::min<uint64_t>(65535, rows - r0);
    dim3 grid(cd->bd / RB, nr);
    hipLaunchKernelGGL((iprs_i32_kernel<NC>), grid, dim3(threads), lds, s, evals,
                       (__int128*)out, cd->d_ftab, cd->d_tw[0], cd->d_tw[1], cd->d_tw[2],
                       (uint32_t)cd->row_len, cd->bd, cd->bl, cd->depth, RB, r0);
    ZP_HIP(hipGetLastError());
  }
  return ZP_OK;
}

}  // namespace

int launch_iprs_binary(const zp_code* cd, const uint64_t* evals, uint64_t rows, uint32_t D,
                       int64_t* out, hipStream_t s) {
  if (rows == 0) return ZP_OK;
  switch (cd->depth) {
    case 0: return launch_binary_nc<1>(cd, evals, rows, D, out, s);
    case 1: return launch_binary_nc<8>(cd, evals, rows, D, out, s);
    case 2: return launch_binary_nc<64>(cd, evals, rows, D, out, s);
    case 3: return launch_binary_nc<512>(cd, evals, rows, D, out, s);
  }
  return fail(ZP_ERR_INVALID_PARAM, "IPRS depth > 3 is not supported on the device");
}

int launch_iprs_i32(const zp_code* cd, const int32_t* evals, uint64_t rows, void* out,
                    hipStream_t s) {
  if (rows == 0) return ZP_OK;
  switch (cd->depth) {
    case 0: return launch_i32_nc<1>(cd, evals, rows, out, s);
    case 1: return launch_i32_nc<8>(cd, evals, rows, out, s);
    case 2: return launch_i32_nc<64>(cd, evals, rows, out, s);
    case 3: return launch_i32_nc<512>(cd, evals, rows, out, s);
  }
  return fail(ZP_ERR_INVALID_PARAM, "IPRS depth > 3 is not supported on the device");
}

}  // namespace zpi
