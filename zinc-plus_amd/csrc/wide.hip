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
  __device__ __forceinline__ uint64_t in_at(uint64_t rp, uint64_t k) const {
    return (rp / planes) * in_row + (rp % planes) * in_plane + k * in_elem;
  }
  __device__ __forceinline__ uint64_t out_at(uint64_t rp, uint64_t i) const {
    return (rp / planes) * out_row + (rp % planes) * out_plane + i * out_elem;
  }
};

constexpr uint32_t kWideTile = 256;

// Block (tile, chunk, sequence): out[c*bd + r] = x[o] + sum_{k>=1} f(r*k mod bd)
// x[o + k 8^d] (radix8.rs:117-161), r in this block's tile of base rows.
template <int L>
__global__ void __launch_bounds__(kWideTile)
wide_base_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                 const int32_t* __restrict__ ftab, WideLayout lay, uint32_t bd, uint32_t bl,
                 uint32_t depth, uint32_t nchunks, uint32_t ntiles) {
  __shared__ uint64_t xs[kWideTile * L];  // the chunk's inputs, kWideTile at a time
  const uint32_t tile = blockIdx.x % ntiles;
  const uint32_t c = (blockIdx.x / ntiles) % nchunks;
  const uint64_t rp = blockIdx.x / ntiles / nchunks;
  const uint32_t o = octrev_w(c, depth);
  const uint32_t r = tile * kWideTile + threadIdx.x;
  Big<L> acc;
#pragma unroll
  for (int l = 0; l < L; l++) acc.w[l] = 0;
  for (uint32_t k0 = 0; k0 < bl; k0 += kWideTile) {
    const uint32_t nk = min(kWideTile, bl - k0);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nk; k += blockDim.x)
      big_store(xs + k * L,
                in_load<L>(in + lay.in_at(rp, (uint64_t)o + (uint64_t)(k0 + k) * nchunks),
                           lay.in_one_limb));
    __syncthreads();
    if (r < bd)  // f(0) = 1 is from_ref for k = 0
      for (uint32_t k = 0; k < nk; k++)
        big_mac(acc, big_load<L>(xs + k * L), ftab[(r * (k0 + k)) & (bd - 1)]);
  }
  if (r < bd) big_store(out + lay.out_at(rp, (uint64_t)c * bd + r), acc);
}

// butterfly.rs:8-17.
__constant__ uint8_t kBfly[8][7] = {
    {0, 0, 0, 0, 0, 0, 0}, {1, 2, 3, 4, 5, 6, 7}, {2, 4, 6, 0, 2, 4, 6}, {3, 6, 1, 4, 7, 2, 5},
    {4, 0, 4, 0, 4, 0, 4}, {5, 2, 7, 4, 1, 6, 3}, {6, 4, 2, 0, 6, 4, 2}, {7, 6, 5, 4, 3, 2, 1},
};

// Stage k (radix8.rs:58-113): thread per (sequence, block, i); the 8 values
// j*sub + i of a block of 8*sub are replaced by their radix-8 butterfly.
template <int L>
__global__ void __launch_bounds__(256)
wide_stage_kernel(uint64_t* __restrict__ data, const int32_t* __restrict__ tw, WideLayout lay,
                  uint64_t N, uint64_t sub, uint64_t total) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint64_t per_row = N / 8;
  const uint64_t rp = t / per_row, rem = t % per_row;
  const uint64_t blk = rem / sub, i = rem % sub;
  const uint64_t first = blk * 8 * sub + i;
  Big<L> xs[8];
#pragma unroll
  for (int j = 0; j < 8; j++) xs[j] = big_load<L>(data + lay.out_at(rp, first + (uint64_t)j * sub));
  const int32_t* twi = tw + i * 56;
  for (int r = 0; r < 8; r++) {
    Big<L> a = xs[0];
#pragma unroll
    for (int jm1 = 0; jm1 < 7; jm1++) big_mac(a, xs[jm1 + 1], twi[jm1 * 8 + kBfly[r][jm1]]);
    big_store(data + lay.out_at(rp, first + (uint64_t)r * sub), a);
  }
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

template <int L>
int launch_wide(const zp_code* cd, const uint64_t* in, uint64_t seqs, const WideLayout& lay,
                uint64_t* out, hipStream_t s) {
  const uint32_t row_len = (uint32_t)cd->row_len, N = (uint32_t)cd->N;
  if (cd->kind == ZP_CODE_IPRS) {
    const uint32_t nchunks = 1u << (3 * cd->depth);
    const uint32_t ntiles = (cd->bd + kWideTile - 1) / kWideTile;
    const uint64_t blocks = seqs * nchunks * ntiles;
    if (blocks > 0x7fffffffull) return fail(ZP_ERR_INVALID_PARAM, "too many rows");
    hipLaunchKernelGGL((wide_base_kernel<L>), dim3((uint32_t)blocks), dim3(kWideTile), 0, s, in,
                       out, cd->d_ftab, lay, cd->bd, cd->bl, cd->depth, nchunks, ntiles);
    ZP_HIP(hipGetLastError());
    for (uint32_t k = 0; k < cd->depth; k++) {
      const uint64_t sub = (uint64_t)cd->bd << (3 * k);
      const uint64_t total = seqs * (N / 8);
      hipLaunchKernelGGL((wide_stage_kernel<L>), dim3((uint32_t)((total + 255) / 256)), dim3(256),
                         0, s, out, cd->d_tw[k], lay, (uint64_t)N, sub, total);
      ZP_HIP(hipGetLastError());
    }
    return ZP_OK;
  }
  if (seqs > 0x7fffffffull) return fail(ZP_ERR_INVALID_PARAM, "too many rows");
  uint64_t* amid = nullptr;
  ZP_TRY(dev_alloc((void**)&amid, seqs * N * L * sizeof(uint64_t), s));
  if (cd->kind == ZP_CODE_RAA_SIGNFLIP)
    hipLaunchKernelGGL((wide_raa_kernel<L, true>), dim3((uint32_t)seqs), dim3(kScanThreads), 0, s,
                       in, out, amid, cd->d_perm1, cd->d_perm2, lay, row_len, N);
  else
    hipLaunchKernelGGL((wide_raa_kernel<L, false>), dim3((uint32_t)seqs), dim3(kScanThreads), 0,
                       s, in, out, amid, cd->d_perm1, cd->d_perm2, lay, row_len, N);
  const int rc = hipGetLastError() == hipSuccess ? ZP_OK : fail(ZP_ERR_DEVICE, "launch failed");
  dev_free(amid, s);
  return rc;
}

int dispatch_wide(const zp_code* cd, uint32_t L, const uint64_t* in, uint64_t seqs,
                  const WideLayout& lay, uint64_t* out, hipStream_t s) {
  switch (L) {
    case 1: return launch_wide<1>(cd, in, seqs, lay, out, s);
    case 2: return launch_wide<2>(cd, in, seqs, lay, out, s);
    case 3: return launch_wide<3>(cd, in, seqs, lay, out, s);
    case 4: return launch_wide<4>(cd, in, seqs, lay, out, s);
    case 5: return launch_wide<5>(cd, in, seqs, lay, out, s);
    case 6: return launch_wide<6>(cd, in, seqs, lay, out, s);
    case 7: return launch_wide<7>(cd, in, seqs, lay, out, s);
    case 8: return launch_wide<8>(cd, in, seqs, lay, out, s);
  }
  return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 1..8");
}

}  // namespace

int launch_encode_wide(const zp_code* cd, uint32_t L, const uint64_t* in, uint64_t rows,
                       uint64_t* out, hipStream_t s) {
  if (rows == 0) return ZP_OK;
  const WideLayout lay{1, false, cd->row_len * L, 0, L, cd->N * L, 0, L};
  return dispatch_wide(cd, L, in, rows, lay, out, s);
}

// Rows of i64 evals (Eval = i64, Cw = i128: L = 2) or of
// DensePolynomial<i64, D> evals (D coefficient planes, each encoded as its
// own sequence: Cw = [i64; D] under IPRS, L = 1, or [Int<4>; D] under RAA,
// L = 4).  Inputs are sign-extended on load; outputs are the Cw layout.
int launch_encode_i64_planes(const zp_code* cd, uint32_t D, uint32_t L, const int64_t* in,
                             uint64_t rows, void* out, hipStream_t s) {
  if (rows == 0) return ZP_OK;
  const WideLayout lay{D, true, cd->row_len * D, 1, D, cd->N * D * L, L, (uint64_t)D * L};
  return dispatch_wide(cd, L, (const uint64_t*)in, rows * D, lay, (uint64_t*)out, s);
}

}  // namespace zpi
