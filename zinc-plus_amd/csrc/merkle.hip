// merkle.hip -- BLAKE3 column leaves and Merkle layers for gfx950.
//
// Reference: zip-plus/src/merkle.rs:128-168 (leaf j = BLAKE3 of the
// little-endian bytes of column j over all rows, one Hasher::update per
// column) and :172-221 (parents via blake3::hazmat::merge_subtrees_non_root,
// the top layer via merge_subtrees_root).  BLAKE3 itself follows the public
// specification (7-round compression, 1 KiB chunks, 64-byte blocks, flags
// CHUNK_START/CHUNK_END/PARENT/ROOT, left subtree = largest power of two of
// chunks below the total).
#include "common.h"

namespace zpi {
namespace {

enum : uint32_t { kChunkStart = 1, kChunkEnd = 2, kParent = 4, kRoot = 8 };

__device__ __constant__ uint32_t kIV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                           0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};

// Message word order of each round: the spec's permutation applied r times.
struct Schedule {
  int s[7][16];
};
constexpr Schedule make_schedule() {
  Schedule sc{};
  constexpr int perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
  for (int i = 0; i < 16; i++) sc.s[0][i] = i;
  for (int r = 1; r < 7; r++)
    for (int i = 0; i < 16; i++) sc.s[r][i] = sc.s[r - 1][perm[i]];
  return sc;
}
constexpr Schedule kSched = make_schedule();

__device__ __forceinline__ uint32_t rotr(uint32_t x, uint32_t n) {
  return __builtin_amdgcn_alignbit(x, x, n);
}

#define B3G(a, b, c, d, mx, my)          \
  a = a + b + (mx); d = rotr(d ^ a, 16); \
  c = c + d;        b = rotr(b ^ c, 12); \
  a = a + b + (my); d = rotr(d ^ a, 8);  \
  c = c + d;        b = rotr(b ^ c, 7);

// Compression returning the 8-word chaining value.
__device__ __forceinline__ void compress(uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                                         uint32_t blen, uint32_t flags) {
  uint32_t v0 = cv[0], v1 = cv[1], v2 = cv[2], v3 = cv[3];
  uint32_t v4 = cv[4], v5 = cv[5], v6 = cv[6], v7 = cv[7];
  uint32_t v8 = kIV[0], v9 = kIV[1], v10 = kIV[2], v11 = kIV[3];
  uint32_t v12 = (uint32_t)counter, v13 = (uint32_t)(counter >> 32), v14 = blen, v15 = flags;
#pragma unroll
  for (int r = 0; r < 7; r++) {
    B3G(v0, v4, v8, v12, m[kSched.s[r][0]], m[kSched.s[r][1]])
    B3G(v1, v5, v9, v13, m[kSched.s[r][2]], m[kSched.s[r][3]])
    B3G(v2, v6, v10, v14, m[kSched.s[r][4]], m[kSched.s[r][5]])
    B3G(v3, v7, v11, v15, m[kSched.s[r][6]], m[kSched.s[r][7]])
    B3G(v0, v5, v10, v15, m[kSched.s[r][8]], m[kSched.s[r][9]])
    B3G(v1, v6, v11, v12, m[kSched.s[r][10]], m[kSched.s[r][11]])
    B3G(v2, v7, v8, v13, m[kSched.s[r][12]], m[kSched.s[r][13]])
    B3G(v3, v4, v9, v14, m[kSched.s[r][14]], m[kSched.s[r][15]])
  }
  cv[0] = v0 ^ v8; cv[1] = v1 ^ v9; cv[2] = v2 ^ v10; cv[3] = v3 ^ v11;
  cv[4] = v4 ^ v12; cv[5] = v5 ^ v13; cv[6] = v6 ^ v14; cv[7] = v7 ^ v15;
}

__device__ __forceinline__ void parent_cv(const uint32_t l[8], uint32_t r_cv[8], bool root) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    m[i] = l[i];
    m[8 + i] = r_cv[i];
  }
  uint32_t cv[8];
#pragma unroll
  for (int i = 0; i < 8; i++) cv[i] = kIV[i];
  compress(cv, m, 0, 64, kParent | (root ? kRoot : 0u));
#pragma unroll
  for (int i = 0; i < 8; i++) r_cv[i] = cv[i];
}

constexpr int kMaxStack = 40;

// One thread per column j: BLAKE3 of the column's byte stream, in which row
// r contributes `elem` bytes from rows + (r*width + j)*elem.
__global__ void __launch_bounds__(256)
leaf_hash_kernel(const uint8_t* __restrict__ rows, uint64_t nrows, uint64_t width, uint32_t elem,
                 uint8_t* __restrict__ leaves) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= width) return;
  const uint64_t total = nrows * elem;
  const uint64_t nchunks = total == 0 ? 1 : (total + 1023) / 1024;
  const uint64_t stride = width * elem;
  const uint8_t* base = rows + j * elem;
  uint64_t cur_row = 0;
  uint32_t cur_off = 0;
  uint32_t stack[kMaxStack][8];
  int sp = 0;
  uint32_t cv[8];
  for (uint64_t ci = 0; ci < nchunks; ci++) {
    const uint64_t clen = min((uint64_t)1024, total - ci * 1024);
    const uint32_t nblocks = clen == 0 ? 1 : (uint32_t)((clen + 63) / 64);
#pragma unroll
    for (int i = 0; i < 8; i++) cv[i] = kIV[i];
    for (uint32_t b = 0; b < nblocks; b++) {
      const uint32_t blen = (uint32_t)min((uint64_t)64, clen - (uint64_t)b * 64);
      uint32_t m[16];
#pragma unroll
      for (int w = 0; w < 16; w++) {
        if ((uint32_t)w * 4 < blen) {
          m[w] = *(const uint32_t*)(base + cur_row * stride + cur_off);
          cur_off += 4;
          if (cur_off == elem) {
            cur_off = 0;
            cur_row++;
          }
        } else {
          m[w] = 0;
        }
      }
      uint32_t flags = 0;
      if (b == 0) flags |= kChunkStart;
      if (b + 1 == nblocks) flags |= kChunkEnd | (nchunks == 1 ? kRoot : 0u);
      compress(cv, m, ci, blen, flags);
    }
    if (ci + 1 < nchunks) {
      uint64_t tc = ci + 1;
      while ((tc & 1) == 0) {
        --sp;
        parent_cv(stack[sp], cv, false);
        tc >>= 1;
      }
#pragma unroll
      for (int i = 0; i < 8; i++) stack[sp][i] = cv[i];
      sp++;
    } else {
      while (sp > 0) {
        --sp;
        parent_cv(stack[sp], cv, sp == 0);
      }
    }
  }
  uint32_t* dst = (uint32_t*)(leaves + j * 32);
#pragma unroll
  for (int i = 0; i < 8; i++) dst[i] = cv[i];
}

// One thread per parent node: cur[m] = merge(prev[2m], prev[2m+1]).
__global__ void __launch_bounds__(256)
merkle_layer_kernel(const uint8_t* __restrict__ prev, uint8_t* __restrict__ cur, uint64_t npar,
                    int root) {
  const uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npar) return;
  const uint32_t* src = (const uint32_t*)(prev + 64 * m);
  uint32_t l[8], r[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    l[i] = src[i];
    r[i] = src[8 + i];
  }
  parent_cv(l, r, root != 0);
  uint32_t* dst = (uint32_t*)(cur + 32 * m);
#pragma unroll
  for (int i = 0; i < 8; i++) dst[i] = r[i];
}

}  // namespace

int launch_leaf_hash(const uint8_t* rows, uint64_t nrows, uint64_t width, uint32_t elem,
                     uint8_t* leaves, hipStream_t s) {
  if (width == 0) return ZP_OK;
  if (elem % 4 != 0) return fail(ZP_ERR_SHAPE, "element size must be a multiple of 4 bytes");
  const uint64_t blocks = (width + 255) / 256;
  hipLaunchKernelGGL(leaf_hash_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, rows, nrows,
                     width, elem, leaves);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

// layers: (2n-1) x 32 bytes with the n leaves first (already written).
int launch_merkle_layers(uint8_t* layers, uint64_t n, hipStream_t s) {
  uint64_t off = 0, cnt = n;
  while (cnt > 1) {
    const uint64_t npar = cnt / 2;
    const uint64_t blocks = (npar + 255) / 256;
    hipLaunchKernelGGL(merkle_layer_kernel, dim3((uint32_t)blocks), dim3(256), 0, s,
                       layers + 32 * off, layers + 32 * (off + cnt), npar, cnt == 2 ? 1 : 0);
    ZP_HIP(hipGetLastError());
    off += cnt;
    cnt = npar;
  }
  return ZP_OK;
}

}  // namespace zpi
