// merkle.hip -- BLAKE3 column leaves and Merkle layers for gfx950.
//
// Reference: zip-plus/src/merkle.rs:128-168 (leaf j = BLAKE3 of the
// little-endian bytes of column j over all rows, one Hasher::update per
// column) and :172-221 (parents via blake3::hazmat::merge_subtrees_non_root,
// the top layer via merge_subtrees_root).  BLAKE3 itself follows the public
// specification (7-round compression, 1 KiB chunks, 64-byte blocks, flags
// CHUNK_START/CHUNK_END/PARENT/ROOT, left subtree = largest power of two of
// chunks below the total).
#include <type_traits>

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

__device__ __forceinline__ void load_block_words(const uint8_t* p, uint32_t m[16]) {
  const uint4* q = (const uint4*)p;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint4 v = q[i];
    m[4 * i] = v.x;
    m[4 * i + 1] = v.y;
    m[4 * i + 2] = v.z;
    m[4 * i + 3] = v.w;
  }
}

// One thread per (column j, chunk ci).  Column j's byte stream is row r's
// element j (elem bytes at rows + (r*width + j)*elem) for r = 0, 1, ...
// Chunk ci is stream bytes [1024 ci, 1024 ci + 1024).  With one chunk per
// column the chunk is the whole input: its last block carries ROOT and the
// output is the leaf; otherwise the chunk's chaining value goes to cvs.
// kAligned: elem is a multiple of 64, so every block is 64 contiguous bytes.
template <bool kAligned>
__device__ __forceinline__ void chunk_cv_one(const uint8_t* __restrict__ rows, uint64_t nrows,
                                             uint64_t width, uint32_t elem, uint64_t nchunks,
                                             uint64_t j, uint64_t ci, uint32_t* __restrict__ cvs,
                                             uint8_t* __restrict__ leaves) {
  const uint64_t total = nrows * elem;
  const uint64_t c0 = ci * 1024;
  const uint64_t clen = min((uint64_t)1024, total - c0);
  const uint32_t nblocks = clen == 0 ? 1 : (uint32_t)((clen + 63) / 64);
  const uint64_t rstride = width * elem;
  const uint8_t* col = rows + j * elem;
  uint32_t cv[8];
#pragma unroll
  for (int i = 0; i < 8; i++) cv[i] = kIV[i];
  // Stream position of the next word as (row, offset within the element).
  uint64_t row = c0 / elem;
  uint32_t off = (uint32_t)(c0 - row * elem);
  for (uint32_t b = 0; b < nblocks; b++) {
    const uint32_t blen = (uint32_t)min((uint64_t)64, clen - (uint64_t)b * 64);
    uint32_t m[16];
    if (kAligned) {
      load_block_words(col + row * rstride + off, m);
      off += 64;
      if (off == elem) {
        off = 0;
        row++;
      }
    } else {
#pragma unroll
      for (int w = 0; w < 16; w++) {
        if ((uint32_t)w * 4 < blen) {
          m[w] = *(const uint32_t*)(col + row * rstride + off);
          off += 4;
          if (off == elem) {
            off = 0;
            row++;
          }
        } else {
          m[w] = 0;
        }
      }
    }
    uint32_t flags = 0;
    if (b == 0) flags |= kChunkStart;
    if (b + 1 == nblocks) flags |= kChunkEnd | (nchunks == 1 ? kRoot : 0u);
    compress(cv, m, ci, blen, flags);
  }
  if (nchunks == 1) {
    uint32_t* dst = (uint32_t*)(leaves + j * 32);
#pragma unroll
    for (int i = 0; i < 8; i++) dst[i] = cv[i];
  } else {
    uint4* dst = (uint4*)(cvs + (ci * width + j) * 8);
    dst[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
    dst[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
  }
}

template <bool kAligned>
__global__ void __launch_bounds__(256)
chunk_cv_kernel(const uint8_t* __restrict__ rows, uint64_t nrows, uint64_t width, uint32_t elem,
                uint64_t nchunks, uint64_t ci_begin, uint64_t ci_end, uint32_t* __restrict__ cvs,
                uint8_t* __restrict__ leaves) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t j = idx % width, ci = ci_begin + idx / width;
  if (ci >= ci_end) return;
  chunk_cv_one<kAligned>(rows, nrows, width, elem, nchunks, j, ci, cvs, leaves);
}

// One level of a column's chunk tree: pairs (2i, 2i+1) merge, an odd last
// node moves up unchanged.  Pairing from the left level by level gives
// BLAKE3's tree (left subtree = largest power of two of chunks below the
// total).  The level with one node is the root: its merge carries ROOT and
// it is written to the leaf array.
__global__ void __launch_bounds__(256)
cv_level_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t width,
                uint64_t n_in, uint8_t* __restrict__ leaves) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n_out = (n_in + 1) / 2;
  const uint64_t j = idx % width, ci = idx / width;
  if (ci >= n_out) return;
  uint32_t l[8], r[8];
  const uint4* a = (const uint4*)(in + (2 * ci * width + j) * 8);
  uint4 a0 = a[0], a1 = a[1];
  r[0] = a0.x; r[1] = a0.y; r[2] = a0.z; r[3] = a0.w;
  r[4] = a1.x; r[5] = a1.y; r[6] = a1.z; r[7] = a1.w;
  if (2 * ci + 1 < n_in) {
#pragma unroll
    for (int i = 0; i < 8; i++) l[i] = r[i];
    const uint4* bq = (const uint4*)(in + ((2 * ci + 1) * width + j) * 8);
    uint4 b0 = bq[0], b1 = bq[1];
    r[0] = b0.x; r[1] = b0.y; r[2] = b0.z; r[3] = b0.w;
    r[4] = b1.x; r[5] = b1.y; r[6] = b1.z; r[7] = b1.w;
    parent_cv(l, r, n_out == 1);
  }
  if (n_out == 1) {
    uint32_t* dst = (uint32_t*)(leaves + j * 32);
#pragma unroll
    for (int i = 0; i < 8; i++) dst[i] = r[i];
  } else {
    uint4* dst = (uint4*)(out + (ci * width + j) * 8);
    dst[0] = make_uint4(r[0], r[1], r[2], r[3]);
    dst[1] = make_uint4(r[4], r[5], r[6], r[7]);
  }
}

// The whole chunk tree of G columns per workgroup, in LDS: the same
// level-by-level left pairing as cv_level_kernel, with the levels in two
// ping-pong areas (level 0 in A, level 1 in B, level 2 in A, ...) and one
// barrier per level; the root merge is written to the leaf array.
constexpr uint32_t kTreeLdsWords = 48 * 1024 / 4;

__global__ void __launch_bounds__(256)
cv_tree_kernel(const uint32_t* __restrict__ cvs, uint64_t width, uint32_t nchunks, uint32_t G,
               uint8_t* __restrict__ leaves) {
  __shared__ uint4 lds[kTreeLdsWords / 4];
  uint4* bufs[2] = {lds, lds + (size_t)nchunks * G * 2};
  const uint64_t j0 = (uint64_t)blockIdx.x * G;
  const uint32_t gh = (uint32_t)min((uint64_t)G, width - j0);
  for (uint32_t i = threadIdx.x; i < nchunks * gh * 2; i += blockDim.x) {
    const uint32_t ci = i / (gh * 2), rem = i % (gh * 2);
    bufs[0][(ci * G) * 2 + rem] = ((const uint4*)cvs)[((uint64_t)ci * width + j0) * 2 + rem];
  }
  __syncthreads();
  uint32_t n = nchunks, cur = 0;
  while (n > 1) {
    const uint32_t n_out = (n + 1) / 2;
    const uint4* in = bufs[cur];
    uint4* out = bufs[cur ^ 1];
    for (uint32_t t = threadIdx.x; t < n_out * gh; t += blockDim.x) {
      const uint32_t pi = t / gh, g = t % gh;
      const uint4* a = in + ((2 * pi) * G + g) * 2;
      uint32_t r[8];
      const uint4 a0 = a[0], a1 = a[1];
      r[0] = a0.x; r[1] = a0.y; r[2] = a0.z; r[3] = a0.w;
      r[4] = a1.x; r[5] = a1.y; r[6] = a1.z; r[7] = a1.w;
      if (2 * pi + 1 < n) {
        uint32_t l[8];
#pragma unroll
        for (int i = 0; i < 8; i++) l[i] = r[i];
        const uint4* b = in + ((2 * pi + 1) * G + g) * 2;
        const uint4 b0 = b[0], b1 = b[1];
        r[0] = b0.x; r[1] = b0.y; r[2] = b0.z; r[3] = b0.w;
        r[4] = b1.x; r[5] = b1.y; r[6] = b1.z; r[7] = b1.w;
        parent_cv(l, r, n_out == 1);
      }
      if (n_out == 1) {
        uint32_t* dst = (uint32_t*)(leaves + (j0 + g) * 32);
#pragma unroll
        for (int i = 0; i < 8; i++) dst[i] = r[i];
      } else {
        uint4* o = out + (pi * G + g) * 2;
        o[0] = make_uint4(r[0], r[1], r[2], r[3]);
        o[1] = make_uint4(r[4], r[5], r[6], r[7]);
      }
    }
    __syncthreads();
    n = n_out;
    cur ^= 1;
  }
}

// Columns per workgroup of cv_tree_kernel (0: the tree does not fit).
uint32_t cv_tree_columns(uint64_t nchunks) {
  const uint64_t per_col = (nchunks + (nchunks + 1) / 2) * 8;  // words, both areas
  const uint64_t g = kTreeLdsWords / per_col;
  return (uint32_t)std::min<uint64_t>(g, 64);
}

// One BLAKE3 compression spread over the 4 lanes of a quad: lane q holds
// column q of the state (v[q], v[4+q], v[8+q], v[12+q]) and runs that
// column's G; for the diagonal step b, c and d are rotated across the quad
// by 1, 2 and 3 lanes (DPP quad_perm), so lane q then holds
// (v[q], v[4+(q+1)%4], v[8+(q+2)%4], v[12+(q+3)%4]), and rotated back.
// Lane q's message words, in use order, are m[kQuadSched[q][i]]: per round,
// the column step takes words s[2q], s[2q+1] of the round's schedule and the
// diagonal step s[8+2q], s[9+2q].
struct QuadSchedule {
  uint8_t s[4][28];
};
constexpr QuadSchedule make_quad_schedule() {
  QuadSchedule qs{};
  for (int q = 0; q < 4; q++)
    for (int r = 0; r < 7; r++) {
      qs.s[q][4 * r + 0] = (uint8_t)kSched.s[r][2 * q];
      qs.s[q][4 * r + 1] = (uint8_t)kSched.s[r][2 * q + 1];
      qs.s[q][4 * r + 2] = (uint8_t)kSched.s[r][8 + 2 * q];
      qs.s[q][4 * r + 3] = (uint8_t)kSched.s[r][9 + 2 * q];
    }
  return qs;
}
__device__ __constant__ QuadSchedule kQuadSched = make_quad_schedule();

// quad_perm [1,2,3,0], [2,3,0,1], [3,0,1,2]: lane q takes lane (q + k) % 4.
constexpr int kQuadRot1 = 0x39, kQuadRot2 = 0x4e, kQuadRot3 = 0x93;
template <int kCtrl>
__device__ __forceinline__ uint32_t quad_rot(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, kCtrl, 0xf, 0xf, false);
}

// Parent node of the 16 message words at msg (left CV, right CV), computed
// by the quad; lane q returns output words q (lo) and 4 + q (hi).
__device__ __forceinline__ void parent_quad(const uint32_t* msg, uint32_t q, bool root,
                                            uint32_t& lo, uint32_t& hi) {
  uint32_t m[28];
#pragma unroll
  for (int i = 0; i < 28; i++) m[i] = msg[kQuadSched.s[q][i]];
  uint32_t a = kIV[q], b = kIV[4 + q], c = kIV[q];
  uint32_t d = q < 2 ? 0u : q == 2 ? 64u : (kParent | (root ? kRoot : 0u));
#pragma unroll
  for (int r = 0; r < 7; r++) {
    B3G(a, b, c, d, m[4 * r], m[4 * r + 1])
    b = quad_rot<kQuadRot1>(b);
    c = quad_rot<kQuadRot2>(c);
    d = quad_rot<kQuadRot3>(d);
    B3G(a, b, c, d, m[4 * r + 2], m[4 * r + 3])
    b = quad_rot<kQuadRot3>(b);
    c = quad_rot<kQuadRot2>(c);
    d = quad_rot<kQuadRot1>(d);
  }
  lo = a ^ c;
  hi = b ^ d;
}

// Merkle layers (merkle.rs:172-221) for n = 2^h leaves stored at layers[0..n),
// layer L at node offset 2n - 2(n >> L).  A workgroup of 512 threads (128
// quads, one parent per quad) reads 256 nodes of layer L0 and produces up
// to 8 layers above them in LDS, writing each node to its place; the layer
// with one node uses the root merge.  Every level is one compression
// deep, so a quad per node cuts the latency of the 16 levels of a 2^16-leaf
// tree about threefold against a thread per node.
__device__ __forceinline__ void merkle_quad_block(uint8_t* __restrict__ layers, uint64_t n,
                                                  uint32_t L0, uint32_t levels, uint64_t block,
                                                  uint32_t* buf) {
  const uint32_t t = threadIdx.x, quad = t >> 2, q = t & 3;
  const uint64_t cnt0 = n >> L0;
  const uint64_t base = block * 256;
  const uint32_t m = (uint32_t)min((uint64_t)256, cnt0 - base);
  const uint32_t* src = (const uint32_t*)(layers + 32 * (2 * n - 2 * cnt0));
  for (uint32_t i = t; i < m * 8; i += 512) buf[i] = src[base * 8 + i];
  __syncthreads();
  uint32_t cur = m;
  for (uint32_t lv = 1; lv <= levels && cur > 1; lv++) {
    const uint32_t npar = cur / 2;
    const uint64_t cnt = n >> (L0 + lv);
    uint32_t lo = 0, hi = 0;
    if (quad < npar) parent_quad(buf + quad * 16, q, cnt == 1, lo, hi);
    __syncthreads();
    if (quad < npar) {
      uint32_t* dst = (uint32_t*)(layers + 32 * (2 * n - 2 * cnt + (base >> lv) + quad));
      buf[quad * 8 + q] = lo;
      buf[quad * 8 + 4 + q] = hi;
      dst[q] = lo;
      dst[4 + q] = hi;
    }
    __syncthreads();
    cur = npar;
  }
}

__global__ void __launch_bounds__(512)
merkle_quad_kernel(uint8_t* __restrict__ layers, uint64_t n, uint32_t L0, uint32_t levels) {
  __shared__ uint32_t buf[256 * 8];
  merkle_quad_block(layers, n, L0, levels, blockIdx.x, buf);
}

// Chunk trees of columns of nchunks = C*M chunks (C in 4..256 and M in
// 1..8, powers of two), C lanes per column and no chunk values in LDS.
// Lane l folds chunks lM .. lM + M - 1, a perfect subtree, in registers:
// chunk i is merged into the stack ctz(i + 1) times (the binary-counter
// order, which for a power-of-two count is the level-by-level left
// pairing).  The column's C/4 quads then merge its C nodes level by level
// with parent_quad (two rounds at the first level, one after), reading each
// pair of nodes as one 16-word message in LDS and writing the parent in
// place after a barrier.  Columns past the last fold column 0 and store
// nothing, so every quad stays whole for the DPP rotations.
template <int M>
__device__ __forceinline__ void cv_quad_tree_block(const uint32_t* __restrict__ cvs,
                                                   uint64_t width, uint32_t C,
                                                   uint8_t* __restrict__ leaves, uint64_t block,
                                                   uint32_t* nodes) {
  static_assert(M >= 1 && M <= 8 && (M & (M - 1)) == 0, "M is 1, 2, 4 or 8");
  constexpr int kDepth = M >= 8 ? 4 : M >= 4 ? 3 : M >= 2 ? 2 : 1;
  const uint32_t tid = threadIdx.x, q = tid & 3;
  const uint32_t l = tid & (C - 1), col = tid / C;
  const uint64_t j = block * (256 / C) + col;
  const bool live = j < width;
  const uint64_t jc = live ? j : 0;
  uint32_t stk[kDepth][8];
#pragma unroll
  for (int i = 0; i < M; i++) {
    const uint4* src = (const uint4*)(cvs + ((uint64_t)(l * M + i) * width + jc) * 8);
    const uint4 a0 = src[0], a1 = src[1];
    uint32_t cur[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    const int height = __builtin_popcount(i), merges = __builtin_ctz(i + 1);
#pragma unroll
    for (int k = 0; k < kDepth; k++)
      if (k < merges) parent_cv(stk[height - 1 - k], cur, false);
#pragma unroll
    for (int w = 0; w < 8; w++) stk[height - merges][w] = cur[w];
  }
  uint32_t* area = nodes + col * C * 8;
#pragma unroll
  for (int w = 0; w < 8; w++) area[l * 8 + w] = stk[0][w];
  __syncthreads();
  const uint32_t qc = l >> 2, nq = C >> 2;
  uint32_t lo[2] = {0, 0}, hi[2] = {0, 0};
  for (uint32_t k = C; k > 1; k >>= 1) {
    const uint32_t merges = k >> 1;
#pragma unroll
    for (uint32_t r = 0; r < 2; r++) {
      const uint32_t idx = r * nq + qc;
      if (idx < merges) parent_quad(area + 16 * idx, q, k == 2, lo[r], hi[r]);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < 2; r++) {
      const uint32_t idx = r * nq + qc;
      if (idx < merges) {
        area[8 * idx + q] = lo[r];
        area[8 * idx + 4 + q] = hi[r];
      }
    }
    __syncthreads();
  }
  // The last merge ran on the column's first quad: lane q holds root words
  // q and 4 + q.
  if (live && l < 4) {
    uint32_t* dst = (uint32_t*)(leaves + j * 32);
    dst[q] = lo[0];
    dst[4 + q] = hi[0];
  }
}

template <int M>
__global__ void __launch_bounds__(256)
cv_quad_tree_kernel(const uint32_t* __restrict__ cvs, uint64_t width, uint32_t C,
                    uint8_t* __restrict__ leaves) {
  __shared__ uint32_t nodes[256 * 8];
  cv_quad_tree_block<M>(cvs, width, C, leaves, blockIdx.x, nodes);
}

// Merkle layers (merkle.rs:172-221) for n = 2^h leaves stored at layers[0..n).
// Layer L starts at node offset 2n - 2(n >> L).  A workgroup of 256 threads
// reads 512 nodes of layer L0 and produces up to 9 layers above them in LDS,
// writing each node to its place; the layer with one node uses the root merge.
__global__ void __launch_bounds__(256)
merkle_tree_kernel(uint8_t* __restrict__ layers, uint64_t n, uint32_t L0, uint32_t levels) {
  __shared__ uint32_t buf[512 * 8];
  const uint32_t t = threadIdx.x;
  const uint64_t cnt0 = n >> L0;
  const uint64_t base = (uint64_t)blockIdx.x * 512;
  const uint32_t m = (uint32_t)min((uint64_t)512, cnt0 - base);
  const uint32_t* src = (const uint32_t*)(layers + 32 * (2 * n - 2 * cnt0));
  for (uint32_t i = t; i < m * 8; i += 256) buf[i] = src[base * 8 + i];
  __syncthreads();
  uint32_t cur = m;
  for (uint32_t lv = 1; lv <= levels && cur > 1; lv++) {
    const uint32_t npar = cur / 2;
    const uint64_t cnt = n >> (L0 + lv);
    uint32_t l[8], r[8];
    if (t < npar) {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        l[i] = buf[(2 * t) * 8 + i];
        r[i] = buf[(2 * t + 1) * 8 + i];
      }
      parent_cv(l, r, cnt == 1);
    }
    __syncthreads();
    if (t < npar) {
      uint32_t* dst = (uint32_t*)(layers + 32 * (2 * n - 2 * cnt + (base >> lv) + t));
#pragma unroll
      for (int i = 0; i < 8; i++) {
        buf[t * 8 + i] = r[i];
        dst[i] = r[i];
      }
    }
    __syncthreads();
    cur = npar;
  }
}

// A transcript's framed byte stream: `count` entries of E bytes absorbed as
// 0x6 || entry || 0x7 each (zip-plus/src/pcs_transcript.rs absorb of
// public columns through absorb_slice, transcript/src/traits.rs:388-392).
// Thread per 16 stream bytes [q0 + 16 t, q0 + 16 t + 16), written to
// out[16 t ..]; the caller keeps the range inside the stream.
__global__ void __launch_bounds__(256)
framed_stream_kernel(const uint8_t* __restrict__ data, uint64_t E, uint64_t q0, uint64_t nwords,
                     uint4* __restrict__ out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nwords) return;
  const uint64_t q = q0 + 16 * t;
  uint64_t i = q / (E + 2), u = q % (E + 2);
  uint32_t w[4];
#pragma unroll
  for (int z = 0; z < 4; z++) {
    uint32_t word = 0;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const uint32_t v = u == 0 ? 0x6u : u == E + 1 ? 0x7u : (uint32_t)data[i * E + u - 1];
      word |= v << (8 * y);
      if (++u == E + 2) {
        u = 0;
        i++;
      }
    }
    w[z] = word;
  }
  out[t] = make_uint4(w[0], w[1], w[2], w[3]);
}

// Chaining values of n full chunks of a byte stream (chunk k at stream +
// 1024 k, counter c0 + k); none of them is the stream's last chunk.  The
// next block's 64 bytes are loaded while the current one is compressed.
__global__ void __launch_bounds__(64)
stream_chunk_cv_kernel(const uint8_t* __restrict__ stream, uint64_t c0, uint64_t n,
                       uint32_t* __restrict__ cvs) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint8_t* p = stream + 1024 * k;
  uint32_t cv[8];
#pragma unroll
  for (int w = 0; w < 8; w++) cv[w] = kIV[w];
  uint32_t m[16], nx[16];
  load_block_words(p, nx);
  for (uint32_t b = 0; b < 16; b++) {
#pragma unroll
    for (int w = 0; w < 16; w++) m[w] = nx[w];
    if (b < 15) load_block_words(p + 64 * (b + 1), nx);
    compress(cv, m, c0 + k, 64, (b == 0 ? kChunkStart : 0u) | (b == 15 ? kChunkEnd : 0u));
  }
  uint4* dst = (uint4*)(cvs + k * 8);
  dst[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
  dst[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

// Grouped launches for many commitments at once (batched commits).  Each
// launch carries its items by value (no device tables, so a captured graph
// holds no copies); an item owns a contiguous range of the launch's tiles or
// blocks, and a block finds its item by binary search over the range starts.
constexpr uint32_t kGroupMax = 48;

struct LeafTask {
  const uint8_t* rows;  // codeword rows (nrows x width elements of elem bytes)
  uint32_t* cvs;        // chunk chaining values (nchunks > 1)
  uint8_t* leaves;
  uint64_t nrows;
  uint32_t width, elem, nchunks, tile0;
};
struct LeafGroup {
  LeafTask t[kGroupMax];
  uint32_t n, tiles;
};
static_assert(sizeof(LeafGroup) < 4000, "kernel argument limit");

struct TreeTask {
  const uint32_t* cvs;
  uint8_t* leaves;
  uint32_t width, C, block0, pad;
};
struct TreeGroup {
  TreeTask t[kGroupMax];
  uint32_t n, blocks;
};

struct MerkleTask {
  uint8_t* layers;
  uint64_t n;
  uint32_t levels, block0, blocks, pad;
};
struct MerkleGroup {
  MerkleTask t[kGroupMax];
  uint32_t n, blocks, L0;
};

// The item whose range [start, next start) holds `x` (starts increase).
template <class G>
__device__ __forceinline__ uint32_t find_item(const G& g, uint32_t x, uint32_t (*start)(const G&, uint32_t)) {
  uint32_t lo = 0, hi = g.n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (start(g, mid) <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ uint32_t leaf_start(const LeafGroup& g, uint32_t i) { return g.t[i].tile0; }
__device__ __forceinline__ uint32_t tree_start(const TreeGroup& g, uint32_t i) { return g.t[i].block0; }
__device__ __forceinline__ uint32_t merkle_start(const MerkleGroup& g, uint32_t i) { return g.t[i].block0; }

// Chunk chaining values (or single-chunk leaves) of every item: a
// persistent grid walks the tiles of 256 (column, chunk) tasks, largest
// items first, so the small items' tails fill in behind the large ones.
__global__ void __launch_bounds__(256) grouped_chunk_cv_kernel(LeafGroup g) {
  for (uint32_t tile = blockIdx.x; tile < g.tiles; tile += gridDim.x) {
    const LeafTask& it = g.t[find_item(g, tile, leaf_start)];
    const uint64_t idx = (uint64_t)(tile - it.tile0) * 256 + threadIdx.x;
    const uint64_t j = idx % it.width, ci = idx / it.width;
    if (ci >= it.nchunks) continue;
    if ((it.elem & 63) == 0)
      chunk_cv_one<true>(it.rows, it.nrows, it.width, it.elem, it.nchunks, j, ci, it.cvs,
                         it.leaves);
    else
      chunk_cv_one<false>(it.rows, it.nrows, it.width, it.elem, it.nchunks, j, ci, it.cvs,
                          it.leaves);
  }
}

// Chunk trees of items with nchunks = C * M (the quad kernel's shapes).
template <int M>
__global__ void __launch_bounds__(256) grouped_cv_quad_tree_kernel(TreeGroup g) {
  __shared__ uint32_t nodes[256 * 8];
  const TreeTask& it = g.t[find_item(g, blockIdx.x, tree_start)];
  cv_quad_tree_block<M>(it.cvs, it.width, it.C, it.leaves, blockIdx.x - it.block0, nodes);
}

// Two-chunk columns: the root merge of the pair.  Thread per column.
__global__ void __launch_bounds__(256) grouped_cv_pair_kernel(TreeGroup g) {
  const TreeTask& it = g.t[find_item(g, blockIdx.x, tree_start)];
  const uint64_t j = (uint64_t)(blockIdx.x - it.block0) * 256 + threadIdx.x;
  if (j >= it.width) return;
  uint32_t l[8], r[8];
  const uint4* a = (const uint4*)(it.cvs + j * 8);
  const uint4* b = (const uint4*)(it.cvs + ((uint64_t)it.width + j) * 8);
  const uint4 a0 = a[0], a1 = a[1], b0 = b[0], b1 = b[1];
  l[0] = a0.x; l[1] = a0.y; l[2] = a0.z; l[3] = a0.w;
  l[4] = a1.x; l[5] = a1.y; l[6] = a1.z; l[7] = a1.w;
  r[0] = b0.x; r[1] = b0.y; r[2] = b0.z; r[3] = b0.w;
  r[4] = b1.x; r[5] = b1.y; r[6] = b1.z; r[7] = b1.w;
  parent_cv(l, r, true);
  uint32_t* dst = (uint32_t*)(it.leaves + j * 32);
#pragma unroll
  for (int i = 0; i < 8; i++) dst[i] = r[i];
}

// Up to 8 Merkle layers above layer L0 of every item that has one.
__global__ void __launch_bounds__(512) grouped_merkle_quad_kernel(MerkleGroup g) {
  __shared__ uint32_t buf[256 * 8];
  const MerkleTask& it = g.t[find_item(g, blockIdx.x, merkle_start)];
  merkle_quad_block(it.layers, it.n, g.L0, it.levels, blockIdx.x - it.block0, buf);
}

}  // namespace

int launch_framed_chunk_cvs(const uint8_t* data, uint64_t entry_bytes, uint64_t q0, uint64_t c0,
                            uint64_t n, uint32_t* cvs, hipStream_t s) {
  // In pieces of at most 2^16 chunks (64 MiB of staged stream).
  constexpr uint64_t kPiece = 1ull << 16;
  if (n == 0) return ZP_OK;
  uint8_t* stage = nullptr;
  ZP_TRY(dev_alloc((void**)&stage, std::min(n, kPiece) * 1024, s));
  int rc = ZP_OK;
  for (uint64_t k0 = 0; k0 < n && rc == ZP_OK; k0 += kPiece) {
    const uint64_t nk = std::min(kPiece, n - k0);
    const uint64_t nwords = nk * 64;
    hipLaunchKernelGGL(framed_stream_kernel, dim3((uint32_t)((nwords + 255) / 256)), dim3(256), 0,
                       s, data, entry_bytes, q0 + 1024 * k0, nwords, (uint4*)stage);
    hipLaunchKernelGGL(stream_chunk_cv_kernel, dim3((uint32_t)((nk + 63) / 64)), dim3(64), 0, s,
                       stage, c0 + k0, nk, cvs + k0 * 8);
    if (hipGetLastError() != hipSuccess) rc = fail(ZP_ERR_DEVICE, "launch failed");
  }
  dev_free(stage, s);
  return rc;
}

uint64_t leaf_chunks(uint64_t nrows, uint32_t elem) {
  const uint64_t total = nrows * elem;
  return total == 0 ? 1 : (total + 1023) / 1024;
}

// Chunk chaining values for chunks [ci_begin, ci_end) of every column (or the
// leaves themselves when each column is a single chunk).
int launch_chunk_cvs(const uint8_t* rows, uint64_t nrows, uint64_t width, uint32_t elem,
                     uint64_t ci_begin, uint64_t ci_end, uint32_t* cvs, uint8_t* leaves,
                     hipStream_t s) {
  if (elem % 4 != 0) return fail(ZP_ERR_SHAPE, "element size must be a multiple of 4 bytes");
  const uint64_t nchunks = leaf_chunks(nrows, elem);
  if (ci_end <= ci_begin || width == 0) return ZP_OK;
  const uint64_t threads = (ci_end - ci_begin) * width;
  const uint32_t blocks = (uint32_t)((threads + 255) / 256);
  if (elem % 64 == 0)
    hipLaunchKernelGGL((chunk_cv_kernel<true>), dim3(blocks), dim3(256), 0, s, rows, nrows, width,
                       elem, nchunks, ci_begin, ci_end, cvs, leaves);
  else
    hipLaunchKernelGGL((chunk_cv_kernel<false>), dim3(blocks), dim3(256), 0, s, rows, nrows,
                       width, elem, nchunks, ci_begin, ci_end, cvs, leaves);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

template <int M>
int launch_cv_quad_tree(const uint32_t* cvs, uint64_t width, uint32_t C, uint8_t* leaves,
                        hipStream_t s) {
  const uint64_t per_block = 256 / C;
  hipLaunchKernelGGL(cv_quad_tree_kernel<M>, dim3((uint32_t)((width + per_block - 1) / per_block)),
                     dim3(256), 0, s, cvs, width, C, leaves);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

// Merges each column's chunk chaining values (cvs, nchunks x width) up to
// the leaf; `tmp` holds ceil(nchunks/2) x width values.  Power-of-two counts
// from 4 to 2048 take the quad kernel (C lanes per column, M chunks each);
// others the LDS tree kernel, or one launch per level when a column's tree
// does not fit in LDS.
std::vector<int> ctree_fold_candidates(uint64_t nchunks) {
  std::vector<int> c;
  if (!(pow2(nchunks) && nchunks >= 4 && nchunks <= 2048)) return c;
  // M chunks per lane; C = nchunks / M lanes per column, a multiple of 4
  // (the quads) and at most 256 (one workgroup).
  for (int M = 1; M <= 8; M *= 2)
    if ((uint64_t)M * 4 <= nchunks && nchunks / M <= 256) c.push_back(M);
  return c;
}

// Chunks per lane of the quad chunk-tree kernel for this shape, 0 where it
// does not apply.  Up to 8 chunks per lane, fewer while that leaves under
// 2^18 lanes in all (4 waves per SIMD) and a column stays within 256 lanes.
// More lanes per column add quad levels: at 16 chunks x 2^16 columns, 4
// lanes of 4 chunks take 26.5 us and 8 lanes of 2 take 31.4 us.  A tuned
// fold (ctree_m) replaces the rule when the shape admits it.
uint32_t ctree_fold(uint64_t nchunks, uint64_t width, int ctree_m) {
  if (g_force_ctree == 1 || !pow2(nchunks) || nchunks < 4 || nchunks > 2048) return 0;
  uint64_t M = std::min<uint64_t>(8, nchunks / 4);
  while (M > 1 && width * (nchunks / M) < (1u << 18) && nchunks / M < 256) M /= 2;
  if (ctree_m > 0)
    for (int c : ctree_fold_candidates(nchunks))
      if (c == ctree_m) M = (uint64_t)c;
  return (uint32_t)M;
}

int launch_cv_merge(uint32_t* cvs, uint32_t* tmp, uint64_t nchunks, uint64_t width,
                    uint8_t* leaves, hipStream_t s, int ctree_m) {
  if (nchunks <= 1) return ZP_OK;
  if (const uint64_t M = ctree_fold(nchunks, width, ctree_m)) {
    const uint32_t C = (uint32_t)(nchunks / M);
    switch (M) {
      case 1: return launch_cv_quad_tree<1>(cvs, width, C, leaves, s);
      case 2: return launch_cv_quad_tree<2>(cvs, width, C, leaves, s);
      case 4: return launch_cv_quad_tree<4>(cvs, width, C, leaves, s);
      default: return launch_cv_quad_tree<8>(cvs, width, C, leaves, s);
    }
  }
  const uint32_t G = cv_tree_columns(nchunks);
  if (G > 0) {
    hipLaunchKernelGGL(cv_tree_kernel, dim3((uint32_t)((width + G - 1) / G)), dim3(256), 0, s, cvs,
                       width, (uint32_t)nchunks, G, leaves);
    ZP_HIP(hipGetLastError());
    return ZP_OK;
  }
  uint32_t* buf[2] = {cvs, tmp};
  uint64_t n = nchunks;
  int cur = 0;
  while (n > 1) {
    const uint64_t n_out = (n + 1) / 2;
    const uint64_t th = n_out * width;
    hipLaunchKernelGGL(cv_level_kernel, dim3((uint32_t)((th + 255) / 256)), dim3(256), 0, s,
                       buf[cur], buf[cur ^ 1], width, n, leaves);
    ZP_HIP(hipGetLastError());
    n = n_out;
    cur ^= 1;
  }
  return ZP_OK;
}

int launch_leaf_hash(const uint8_t* rows, uint64_t nrows, uint64_t width, uint32_t elem,
                     uint8_t* leaves, hipStream_t s, int ctree_m) {
  if (width == 0) return ZP_OK;
  const uint64_t nchunks = leaf_chunks(nrows, elem);
  uint32_t* cv[2] = {nullptr, nullptr};
  if (nchunks > 1) {
    ZP_TRY(dev_alloc((void**)&cv[0], nchunks * width * 32, s));
    int rc0 = dev_alloc((void**)&cv[1], ((nchunks + 1) / 2) * width * 32, s);
    if (rc0 != ZP_OK) {
      dev_free(cv[0], s);
      return rc0;
    }
  }
  int rc = launch_chunk_cvs(rows, nrows, width, elem, 0, nchunks, cv[0], leaves, s);
  if (rc == ZP_OK) rc = launch_cv_merge(cv[0], cv[1], nchunks, width, leaves, s, ctree_m);
  dev_free(cv[0], s);
  dev_free(cv[1], s);
  return rc;
}

namespace {

// Resident 256-thread blocks of the grouped chunk kernel on this device:
// the persistent grid's size.
uint32_t grouped_cv_capacity() {
  static int cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
  if (!cached[dev]) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grouped_chunk_cv_kernel, 256, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        per_cu <= 0 || cus <= 0) {
      (void)hipGetLastError();
      return 1024;
    }
    cached[dev] = per_cu * cus;
  }
  return (uint32_t)cached[dev];
}

template <class G, class T>
int launch_groups(const std::vector<T>& tasks, uint32_t (*units)(const T&), G proto,
                  void (*launch)(const G&, uint32_t, hipStream_t), hipStream_t s) {
  for (size_t i0 = 0; i0 < tasks.size(); i0 += kGroupMax) {
    G g = proto;
    g.n = 0;
    uint32_t total = 0;
    for (size_t i = i0; i < std::min(tasks.size(), i0 + kGroupMax); i++) {
      g.t[g.n] = tasks[i];
      uint32_t* start = nullptr;
      if constexpr (std::is_same<T, LeafTask>::value) start = &g.t[g.n].tile0;
      else start = &g.t[g.n].block0;
      *start = total;
      total += units(tasks[i]);
      g.n++;
    }
    if (total == 0) continue;
    if constexpr (std::is_same<T, LeafTask>::value) g.tiles = total;
    else g.blocks = total;
    launch(g, total, s);
    ZP_HIP(hipGetLastError());
  }
  return ZP_OK;
}

}  // namespace

// Leaves and Merkle layers of many commitments in grouped launches: one
// persistent launch for every item's chunk chaining values, one launch per
// chunk-tree shape, one per band of 8 Merkle levels.  Shapes the grouped
// kernels do not take (chunk counts that are not a power of two, or above
// 2048, and option merkle_tree = 1) run through the per-item launchers on
// the same stream.
int launch_grouped_hash(const std::vector<LeafJob>& jobs, hipStream_t s) {
  struct Plan {
    uint64_t nchunks = 0, cv_off = 0, tmp_off = 0;
    uint32_t M = 0;
  };
  std::vector<Plan> plan(jobs.size());
  uint64_t bytes = 0;
  auto take = [&](uint64_t b) {
    const uint64_t off = bytes;
    bytes += (b + 255) / 256 * 256;
    return off;
  };
  for (size_t i = 0; i < jobs.size(); i++) {
    const LeafJob& jb = jobs[i];
    if (jb.elem % 4 != 0) return fail(ZP_ERR_SHAPE, "element size must be a multiple of 4 bytes");
    if (jb.width == 0) continue;
    if (jb.width >= (1ull << 32) || jb.nrows >= (1ull << 40))
      return fail(ZP_ERR_SHAPE, "grouped hashing: shape too large");
    Plan& p = plan[i];
    p.nchunks = leaf_chunks(jb.nrows, jb.elem);
    if (p.nchunks > 1) {
      p.cv_off = take(p.nchunks * jb.width * 32);
      // One launch holds every item's columns, so no item needs its lanes
      // spread for occupancy: the largest fold (fewest quad levels, whose
      // merges cost about 1.6x the lane-operations of a folded one).
      p.M = ctree_fold(p.nchunks, ~0ull >> 8, -1);
      if (!p.M && p.nchunks != 2) p.tmp_off = take(((p.nchunks + 1) / 2) * jb.width * 32);
    }
  }
  uint8_t* scratch = nullptr;
  if (bytes) ZP_TRY(dev_alloc((void**)&scratch, bytes, s));
  int rc = ZP_OK;
  // Chunk chaining values, largest items first.
  std::vector<size_t> order;
  for (size_t i = 0; i < jobs.size(); i++)
    if (jobs[i].width) order.push_back(i);
  auto tiles_of = [&](size_t i) { return (plan[i].nchunks * jobs[i].width + 255) / 256; };
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return tiles_of(a) > tiles_of(b); });
  std::vector<LeafTask> leaf;
  for (size_t i : order) {
    const LeafJob& jb = jobs[i];
    if (tiles_of(i) >= (1ull << 31)) {
      rc = fail(ZP_ERR_SHAPE, "grouped hashing: too many chunks");
      break;
    }
    leaf.push_back({jb.rows, plan[i].nchunks > 1 ? (uint32_t*)(scratch + plan[i].cv_off) : nullptr,
                    jb.layers, jb.nrows, (uint32_t)jb.width, jb.elem, (uint32_t)plan[i].nchunks,
                    0});
  }
  if (rc == ZP_OK)
    rc = launch_groups<LeafGroup, LeafTask>(
        leaf,
        [](const LeafTask& t) { return (uint32_t)(((uint64_t)t.nchunks * t.width + 255) / 256); },
        LeafGroup{},
        [](const LeafGroup& g, uint32_t total, hipStream_t st) {
          const uint32_t grid = std::min(total, grouped_cv_capacity());
          hipLaunchKernelGGL(grouped_chunk_cv_kernel, dim3(grid), dim3(256), 0, st, g);
        },
        s);
  // Chunk trees.
  std::vector<TreeTask> quad[4], pair;
  for (size_t i = 0; rc == ZP_OK && i < jobs.size(); i++) {
    const LeafJob& jb = jobs[i];
    const Plan& p = plan[i];
    if (!jb.width || p.nchunks <= 1) continue;
    uint32_t* cvs = (uint32_t*)(scratch + p.cv_off);
    if (p.M) {
      const uint32_t C = (uint32_t)(p.nchunks / p.M);
      quad[__builtin_ctz(p.M)].push_back({cvs, jb.layers, (uint32_t)jb.width, C, 0, 0});
    } else if (p.nchunks == 2) {
      pair.push_back({cvs, jb.layers, (uint32_t)jb.width, 2, 0, 0});
    } else {
      rc = launch_cv_merge(cvs, (uint32_t*)(scratch + p.tmp_off), p.nchunks, jb.width, jb.layers,
                           s, jb.ctree_m);
    }
  }
  auto quad_blocks = [](const TreeTask& t) {
    const uint32_t per = 256 / t.C;
    return (t.width + per - 1) / per;
  };
  if (rc == ZP_OK)
    rc = launch_groups<TreeGroup, TreeTask>(quad[0], quad_blocks, TreeGroup{},
        [](const TreeGroup& g, uint32_t total, hipStream_t st) {
          hipLaunchKernelGGL(grouped_cv_quad_tree_kernel<1>, dim3(total), dim3(256), 0, st, g);
        }, s);
  if (rc == ZP_OK)
    rc = launch_groups<TreeGroup, TreeTask>(quad[1], quad_blocks, TreeGroup{},
        [](const TreeGroup& g, uint32_t total, hipStream_t st) {
          hipLaunchKernelGGL(grouped_cv_quad_tree_kernel<2>, dim3(total), dim3(256), 0, st, g);
        }, s);
  if (rc == ZP_OK)
    rc = launch_groups<TreeGroup, TreeTask>(quad[2], quad_blocks, TreeGroup{},
        [](const TreeGroup& g, uint32_t total, hipStream_t st) {
          hipLaunchKernelGGL(grouped_cv_quad_tree_kernel<4>, dim3(total), dim3(256), 0, st, g);
        }, s);
  if (rc == ZP_OK)
    rc = launch_groups<TreeGroup, TreeTask>(quad[3], quad_blocks, TreeGroup{},
        [](const TreeGroup& g, uint32_t total, hipStream_t st) {
          hipLaunchKernelGGL(grouped_cv_quad_tree_kernel<8>, dim3(total), dim3(256), 0, st, g);
        }, s);
  if (rc == ZP_OK)
    rc = launch_groups<TreeGroup, TreeTask>(
        pair, [](const TreeTask& t) { return (t.width + 255) / 256; }, TreeGroup{},
        [](const TreeGroup& g, uint32_t total, hipStream_t st) {
          hipLaunchKernelGGL(grouped_cv_pair_kernel, dim3(total), dim3(256), 0, st, g);
        }, s);
  // Merkle layers, 8 levels per band.
  if (g_force_tree == 1) {
    for (size_t i = 0; rc == ZP_OK && i < jobs.size(); i++)
      if (jobs[i].width) rc = launch_merkle_layers(jobs[i].layers, jobs[i].width, s);
  } else {
    for (uint32_t L0 = 0; rc == ZP_OK; L0 += 8) {
      std::vector<MerkleTask> mt;
      for (const LeafJob& jb : jobs) {
        const uint64_t cnt = L0 < 64 ? jb.width >> L0 : 0;
        if (cnt <= 1) continue;
        const uint32_t lg = 63 - (uint32_t)__builtin_clzll(cnt);
        mt.push_back({jb.layers, jb.width, std::min<uint32_t>(8, lg), 0,
                      (uint32_t)((cnt + 255) / 256), 0});
      }
      if (mt.empty()) break;
      MerkleGroup proto{};
      proto.L0 = L0;
      rc = launch_groups<MerkleGroup, MerkleTask>(
          mt, [](const MerkleTask& t) { return t.blocks; }, proto,
          [](const MerkleGroup& g, uint32_t total, hipStream_t st) {
            hipLaunchKernelGGL(grouped_merkle_quad_kernel, dim3(total), dim3(512), 0, st, g);
          }, s);
    }
  }
  if (scratch) dev_free(scratch, s);
  return rc;
}

// layers: (2n-1) x 32 bytes with the n leaves first (already written); n is
// a power of two.
int launch_merkle_layers(uint8_t* layers, uint64_t n, hipStream_t s) {
  uint32_t L = 0;
  while ((n >> L) > 1) {
    const uint64_t cnt = n >> L;
    const uint32_t lg = 63 - (uint32_t)__builtin_clzll(cnt);
    if (g_force_tree == 1) {
      const uint32_t levels = std::min<uint32_t>(9, lg);
      const uint64_t blocks = (cnt + 511) / 512;
      hipLaunchKernelGGL(merkle_tree_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, layers, n,
                         L, levels);
      ZP_HIP(hipGetLastError());
      L += levels;
    } else {
      const uint32_t levels = std::min<uint32_t>(8, lg);
      const uint64_t blocks = (cnt + 255) / 256;
      hipLaunchKernelGGL(merkle_quad_kernel, dim3((uint32_t)blocks), dim3(512), 0, s, layers, n,
                         L, levels);
      ZP_HIP(hipGetLastError());
      L += levels;
    }
  }
  return ZP_OK;
}

}  // namespace zpi
