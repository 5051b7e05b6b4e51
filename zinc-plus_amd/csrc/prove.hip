// prove.hip -- ZipPlus::prove: the opening proof of committed polys.
//
// Reference: zip-plus/src/pcs/phase_prove.rs:85-253 (prove / prove_f),
// :282-304 (open_merkle_trees_for_column), pcs/utils.rs:93-114
// (point_to_tensor with poly/src/utils.rs:140-190 build_eq_x_r),
// pcs/utils.rs:16-78 (validate_input).
//
// Device work (all exact):
//   q_1, q_0 = eq(x, r) tables of the two halves of the point, one thread
//     per entry (the product of r_k or 1 - r_k over the index bits);
//   b_j = sum_col F(w'_j[col]) q_1[col] with w'_j[col] = sum_c alpha_c
//     bit_c(x_j[col]) (an exact integer: |w'| < 2^134), computed as
//     b_j = sum_c alpha_c P_c[j],  P_c[j] = sum_{col: bit c set} q_1[col]:
//     one workgroup per row sums the Montgomery-form q_1 entries of every
//     plane exactly (288-bit accumulators) and reduces once;
//   the combined row (combined.hip), and the openings packed in
//   proof-stream order (columns, Merkle proof headers and paths).
// The host runs the transcript: alphas, the b elements, eval = <q_0, b>,
// the coefficients s, the column indices, and the proof bytes.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "common.h"
#include "field.h"

namespace zpi {

void ts_challenge(zp_transcript* t, uint8_t* out, size_t n);
uint64_t ts_squeeze_idx(zp_transcript* t, uint64_t cap);
void ts_write_field_elements(zp_transcript* t, const uint64_t mod[4], const Fp* v, size_t n);
void ts_put(zp_transcript* t, const void* p, size_t n);
void ts_reserve(zp_transcript* t, size_t more);
uint8_t* ts_grow_pinned(zp_transcript* t, size_t more);
void ts_unput(zp_transcript* t, size_t n);

namespace {

// out[i] = prod_k (bit k of i ? r_k : 1 - r_k); rr = [r_0, 1 - r_0, r_1, ...].
__global__ void __launch_bounds__(256)
eq_table_kernel(const Fp* __restrict__ rr, uint32_t nv, Fp* __restrict__ out, FpParams P) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1ull << nv)) return;
  Fp acc = fp_one(P);
  for (uint32_t k = 0; k < nv; k++) acc = fp_mul(P, acc, rr[2 * k + (((i >> k) & 1) ? 0 : 1)]);
  out[i] = acc;
}

constexpr int kSumWaves = 4;

// One workgroup per (poly row, block of columns): for every plane c the exact
// sum (nine 32-bit words) of q1[col] over the block's columns whose eval has
// bit c set.  Lane = (plane c, sub-column); lanes of one wave read 64/Dp
// adjacent columns, four columns per lane in flight.
// part[(row*nblk + blk)*D + c][9].
template <typename Word>
__global__ void __launch_bounds__(64 * kSumWaves)
plane_sums_kernel(const Word* __restrict__ evals, const Fp* __restrict__ q1,
                  uint32_t* __restrict__ part, uint64_t row_len, uint32_t cpb, uint32_t nblk,
                  uint32_t D, uint32_t dp_log2) {
  __shared__ uint32_t red[64 * kSumWaves][9];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t Dp = 1u << dp_log2;
  const uint32_t c = lane & (Dp - 1), sub = lane >> dp_log2;
  const uint32_t per_step = (64u >> dp_log2) * kSumWaves;
  const uint64_t row = blockIdx.x / nblk;
  const uint32_t blk = blockIdx.x % nblk;
  const uint64_t c0 = (uint64_t)blk * cpb, c1 = min(row_len, c0 + cpb);
  const Word* x = evals + row * row_len;
  uint32_t acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  auto add = [&](uint64_t col) {
    const uint32_t mask = 0u - (uint32_t)((x[col] >> c) & 1u);
    const uint32_t* q = (const uint32_t*)(q1 + col);
    uint64_t carry = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint64_t s = (uint64_t)acc[k] + (q[k] & mask) + carry;
      acc[k] = (uint32_t)s;
      carry = s >> 32;
    }
    acc[8] += (uint32_t)carry;
  };
  uint64_t col = c0 + (uint64_t)wave * (64u >> dp_log2) + sub;
  for (; col + 3 * per_step < c1; col += 4 * per_step) {
#pragma unroll
    for (int u = 0; u < 4; u++) add(col + u * per_step);
  }
  for (; col < c1; col += per_step) add(col);
#pragma unroll
  for (int k = 0; k < 9; k++) red[threadIdx.x][k] = acc[k];
  __syncthreads();
  if (threadIdx.x >= D) return;
  // Sum the partials of plane c: lanes c + m*Dp of every wave.
  uint32_t s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t t = threadIdx.x; t < 64 * kSumWaves; t += Dp) {
    uint64_t carry = 0;
    for (int k = 0; k < 9; k++) {
      const uint64_t v = (uint64_t)s[k] + red[t][k] + carry;
      s[k] = (uint32_t)v;
      carry = v >> 32;
    }
  }
  uint32_t* o = part + (((uint64_t)blockIdx.x) * D + threadIdx.x) * 9;
  for (int k = 0; k < 9; k++) o[k] = s[k];
}

// ---- Plane sums on the matrix cores ------------------------------------
//
// P_c[j] = sum_col q1[col] bit_c(x_j[col]) is, for one row j, the product
// (37 x row_len seven-bit slices of q1) . (row_len x D bit matrix of the row):
// q1 < 2^256 is cut into 37 unsigned slices (7 x 37 = 259 bits), the A
// operand of v_mfma_i32_16x16x64_i8 (three 16-row tiles, rows 37..47 zero),
// shared by every row; the B operand of one K-step is 64 consecutive
// columns x 16 planes of the row.  Lane (g, n) of a wave loads the 16 evals
// of its lane group (columns 16g .. 16g + 15 of the K-step) and
// byte-transposes the byte of its plane (v_perm); plane c' of the byte
// enters as bits scaled by 2^max(c' - 1, 0) <= 64, shifted back in the
// epilogue.  Every i32 partial is below 127 * 64 * 64 per K-step, and a
// block holds at most kMaxBlockSteps K-steps, so the sums are exact.
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr uint32_t kMaxBlockSteps = 2048;  // 2048 * 520192 < 2^31

// frag[(kt * 3 + mt) * 64 + lane]: slice 16 mt + lane % 16 of q1[col] for
// col = 64 kt + 16 (lane / 16) + e in byte e.
__global__ void __launch_bounds__(256)
q1_afrag_kernel(const Fp* __restrict__ q1, v4i* __restrict__ frag, uint64_t row_len,
                uint32_t ksteps) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ksteps * 192) return;
  const uint32_t lane = idx % 64, mt = (idx / 64) % 3, kt = idx / 192;
  const uint32_t m = 16 * mt + lane % 16;
  uint32_t w[4] = {0, 0, 0, 0};
  if (m < 37) {
    const uint32_t b0 = 7 * m, wi = b0 / 64, sh = b0 % 64;
#pragma unroll
    for (uint32_t e = 0; e < 16; e++) {
      const uint64_t col = (uint64_t)kt * 64 + 16 * (lane / 16) + e;
      if (col >= row_len) continue;
      const uint64_t* q = q1[col].w;
      uint64_t v = q[wi] >> sh;
      if (sh > 57 && wi < 3) v |= q[wi + 1] << (64 - sh);
      w[e / 4] |= (uint32_t)(v & 127) << (8 * (e % 4));
    }
  }
  frag[idx] = v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
}

// Grid: (row, plane chunk q of 16, column block).  The 4 waves of a
// workgroup take every fourth K-step of the block; their slice sums are
// added in LDS and wave 0 recombines them into the nine 32-bit words of
// part[(row * nblk + blk) * D + c].
template <typename Word>
__global__ void __launch_bounds__(256)
plane_sums_mfma_kernel(const Word* __restrict__ evals, const v4i* __restrict__ frag,
                       uint32_t* __restrict__ part, uint64_t row_len, uint32_t ksteps,
                       uint32_t spb, uint32_t nblk, uint32_t D) {
  __shared__ int32_t red[4][64][12];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t g = lane >> 4, n = lane & 15;
  const uint32_t Q = (D + 15) / 16;
  uint64_t rest = blockIdx.x;
  const uint32_t blk = (uint32_t)(rest % nblk);
  rest /= nblk;
  const uint32_t q = (uint32_t)(rest % Q);
  const uint64_t row = rest / Q;
  const uint32_t c = 16 * q + n;               // this lane's plane (output column)
  const uint32_t beta = c / 8, cp = c % 8;     // its byte and bit within the byte
  const uint32_t h = beta / 4, k = beta % 4;   // 32-bit half of the eval, byte in it
  const uint32_t sel = ((4 + k) << 8) | k;     // bytes k of the two operands
  const uint32_t mk = 0x01010101u << (cp == 0 ? 0 : cp - 1);
  const Word* x = evals + row * row_len;
  const uint32_t kt0 = blk * spb, kt1 = min(ksteps, kt0 + spb);
  v4i acc[3] = {v4i{0, 0, 0, 0}, v4i{0, 0, 0, 0}, v4i{0, 0, 0, 0}};
  for (uint32_t kt = kt0 + wave; kt < kt1; kt += 4) {
    const uint64_t col0 = (uint64_t)kt * 64 + 16 * g;
    uint32_t e[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint64_t col = col0 + i;
      const Word v = col < row_len ? x[col] : (Word)0;
      e[i] = (uint32_t)((uint64_t)v >> (32 * h));
    }
    v4i bf;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const uint32_t u = __builtin_amdgcn_perm(e[4 * v + 1], e[4 * v + 0], sel);
      const uint32_t w = __builtin_amdgcn_perm(e[4 * v + 3], e[4 * v + 2], sel);
      const uint32_t z = __builtin_amdgcn_perm(w, u, 0x05040100u);
      bf[v] = (int)((cp == 0 ? z : z >> 1) & mk);
    }
#pragma unroll
    for (int mt = 0; mt < 3; mt++)
      acc[mt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(frag[(kt * 3 + mt) * 64 + lane], bf,
                                                      acc[mt], 0, 0, 0);
  }
#pragma unroll
  for (int mt = 0; mt < 3; mt++)
#pragma unroll
    for (int v = 0; v < 4; v++) red[wave][lane][4 * mt + v] = acc[mt][v];
  __syncthreads();
  if (wave != 0) return;
  // Lane (g, n) holds slices 16 mt + 4 g + v of plane c (scaled); the sum
  // is below 2^256 * 2^log2(row_len) < 2^288, kept in five 64-bit limbs.
  const int sc = cp == 0 ? 0 : (int)cp - 1;
  uint64_t V[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int mt = 0; mt < 3; mt++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const uint32_t m = 16 * mt + 4 * g + v;
      if (m >= 37) continue;
      const uint64_t t = (uint64_t)(uint32_t)(red[0][lane][4 * mt + v] + red[1][lane][4 * mt + v] +
                                              red[2][lane][4 * mt + v] + red[3][lane][4 * mt + v]) >> sc;
      const uint32_t b0 = 7 * m, wi = b0 / 64, sh = b0 % 64;
      // V += t * 2^b0 (t < 2^31, so it spans at most two limbs).
      uint64_t add[5] = {0, 0, 0, 0, 0};
      add[wi] = t << sh;
      if (wi + 1 < 5) add[wi + 1] = sh ? (t >> (64 - sh)) : 0;
      uint64_t cy = 0;
#pragma unroll
      for (int i = 0; i < 5; i++) {
        const uint64_t s1 = V[i] + add[i];
        const uint64_t c1 = s1 < add[i];
        const uint64_t s2 = s1 + cy;
        cy = c1 | (s2 < cy);
        V[i] = s2;
      }
    }
#pragma unroll
  for (int msk = 16; msk <= 32; msk *= 2) {
    uint64_t o[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)V[i], msk);
      const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(V[i] >> 32), msk);
      o[i] = ((uint64_t)hi << 32) | lo;
    }
    uint64_t cy = 0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const uint64_t s1 = V[i] + o[i];
      const uint64_t c1 = s1 < o[i];
      const uint64_t s2 = s1 + cy;
      cy = c1 | (s2 < cy);
      V[i] = s2;
    }
  }
  if (g != 0 || c >= D) return;
  uint32_t* o = part + (((uint64_t)row * nblk + blk) * D + c) * 9;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    o[2 * i] = (uint32_t)V[i];
    o[2 * i + 1] = (uint32_t)(V[i] >> 32);
  }
  o[8] = (uint32_t)V[4];
}

// One thread per (poly row, plane): adds the blocks' partial sums and
// reduces mod p.  The total is below row_len * p <= 2^red_bits * p: p * 2^k
// is subtracted where it fits, k = red_bits-1 .. 0.
__global__ void __launch_bounds__(256)
plane_reduce_kernel(const uint32_t* __restrict__ part, Fp* __restrict__ out, uint64_t rows,
                    uint32_t nblk, uint32_t D, uint32_t red_bits, FpParams P) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * D) return;
  const uint64_t row = i / D;
  const uint32_t c = (uint32_t)(i % D);
  uint32_t s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t b = 0; b < nblk; b++) {
    const uint32_t* q = part + ((row * nblk + b) * D + c) * 9;
    uint64_t carry = 0;
    for (int k = 0; k < 9; k++) {
      const uint64_t v = (uint64_t)s[k] + q[k] + carry;
      s[k] = (uint32_t)v;
      carry = v >> 32;
    }
  }
  for (int k = (int)red_bits - 1; k >= 0; k--) {
    uint32_t pk[9];
    const uint32_t* pw = (const uint32_t*)P.p;
    for (int w = 0; w < 9; w++) {
      const uint32_t lo = w < 8 ? pw[w] : 0;
      const uint32_t lo_prev = w > 0 ? pw[w - 1] : 0;
      pk[w] = k == 0 ? lo : (lo << k) | (w > 0 ? lo_prev >> (32 - k) : 0);
    }
    bool ge = true;
    for (int w = 8; w >= 0; w--) {
      if (s[w] != pk[w]) {
        ge = s[w] > pk[w];
        break;
      }
    }
    if (ge) {
      uint64_t br = 0;
      for (int w = 0; w < 9; w++) {
        const uint64_t d = (uint64_t)s[w] - pk[w] - br;
        s[w] = (uint32_t)d;
        br = (d >> 32) & 1;
      }
    }
  }
  Fp r;
  for (int w = 0; w < 4; w++) r.w[w] = (uint64_t)s[2 * w] | ((uint64_t)s[2 * w + 1] << 32);
  out[i] = r;
}

// b[j] = sum_b sum_c alpha_{b,c} * P[b][j][c] (all Montgomery form).  One
// wave per row: lane l takes the (poly, plane) pairs l, l + 64, ..., then
// the wave adds its 64 partial sums in a butterfly.
__global__ void __launch_bounds__(256)
b_kernel(const Fp* __restrict__ planes, const Fp* __restrict__ alphas, Fp* __restrict__ b,
         uint64_t num_rows, uint32_t D, uint32_t batch, FpParams P) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t j = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (j >= num_rows) return;  // whole waves leave together
  Fp acc = {{0, 0, 0, 0}};
  for (uint32_t t = lane; t < batch * D; t += 64) {
    const uint32_t p = t / D, c = t % D;
    acc = fp_add(P, acc, fp_mul(P, alphas[p * D + c], planes[((uint64_t)p * num_rows + j) * D + c]));
  }
#pragma unroll
  for (int msk = 1; msk < 64; msk *= 2) {
    Fp o;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)acc.w[i], msk);
      const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(acc.w[i] >> 32), msk);
      o.w[i] = ((uint64_t)hi << 32) | lo;
    }
    acc = fp_add(P, acc, o);
  }
  if (lane == 0) b[j] = acc;
}

// i32 evals as u64 lanes (bit pattern zero-extended) for the combined row.
__global__ void widen_i32_kernel(const uint32_t* __restrict__ in, uint64_t* __restrict__ out,
                                 uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// Column openings: out[(k*rows + r)*words + w] = cw[(r*N + col_k)*words + w].
__global__ void gather_columns_kernel(const uint32_t* __restrict__ cw,
                                      const uint64_t* __restrict__ cols, uint32_t* __restrict__ out,
                                      uint64_t rows, uint64_t N, uint32_t words, uint32_t n_cols) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)n_cols * rows * words;
  if (idx >= total) return;
  const uint32_t w = (uint32_t)(idx % words);
  const uint64_t r = (idx / words) % rows;
  const uint64_t k = idx / words / rows;
  out[idx] = cw[(r * N + cols[k]) * words + w];
}

// Merkle paths: out[(k*levels + l)*8 + w] = word w of the sibling of column
// k's ancestor at level l (layers: leaves first, 32 bytes per node).
__global__ void gather_paths_kernel(const uint32_t* __restrict__ layers,
                                    const uint64_t* __restrict__ cols, uint32_t* __restrict__ out,
                                    uint64_t N, uint32_t levels, uint32_t n_cols) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)n_cols * levels * 8;
  if (idx >= total) return;
  const uint32_t w = (uint32_t)(idx % 8);
  const uint32_t l = (uint32_t)((idx / 8) % levels);
  const uint64_t k = idx / 8 / levels;
  const uint64_t off = 2 * N - 2 * (N >> l);  // first node of level l
  const uint64_t node = (cols[k] >> l) ^ 1;
  out[idx] = layers[(off + node) * 8 + w];
}

// The openings in proof-stream order (phase_prove.rs:282-304 with
// write_merkle_proof, pcs_transcript.rs:210-221): per opening k, the
// column of every poly's rows, then leaf index, leaf count and path length
// as u64, then the sibling path bottom-up.  One thread per output word.
__global__ void pack_openings_kernel(const uint32_t* __restrict__ cw,
                                     const uint32_t* __restrict__ layers,
                                     const uint64_t* __restrict__ cols, uint32_t* __restrict__ out,
                                     uint64_t rows, uint64_t N, uint32_t words, uint32_t levels,
                                     uint32_t n_cols) {
  const uint64_t col_words = rows * words;
  const uint64_t per = col_words + 6 + (uint64_t)levels * 8;
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per * n_cols) return;
  const uint64_t k = idx / per, w = idx % per;
  const uint64_t col = cols[k];
  uint32_t v;
  if (w < col_words) {
    v = cw[((w / words) * N + col) * words + w % words];
  } else if (w < col_words + 6) {
    const uint32_t h = (uint32_t)(w - col_words);
    const uint64_t f = h < 2 ? col : (h < 4 ? N : (uint64_t)levels);
    v = (h & 1) ? (uint32_t)(f >> 32) : (uint32_t)f;
  } else {
    const uint64_t pw = w - col_words - 6;
    const uint32_t l = (uint32_t)(pw / 8);
    const uint64_t off = 2 * N - 2 * (N >> l);  // first node of level l
    v = layers[(off + ((col >> l) ^ 1)) * 8 + pw % 8];
  }
  out[idx] = v;
}

uint32_t ceil_log2(uint64_t x) {
  uint32_t r = 0;
  while ((1ull << r) < x) r++;
  return r;
}

// Device buffers of one proof, carved from one allocation.
struct Bufs {
  std::vector<std::pair<void**, uint64_t>> parts;
  void* base = nullptr;
  void add(void** p, uint64_t bytes) { parts.push_back({p, (bytes + 255) & ~255ull}); }
  uint64_t total() const {
    uint64_t t = 0;
    for (auto& p : parts) t += p.second;
    return t;
  }
  int alloc(hipStream_t s) {
    ZP_TRY(dev_alloc(&base, std::max<uint64_t>(total(), 256), s));
    uint8_t* q = (uint8_t*)base;
    for (auto& p : parts) {
      *p.first = q;
      q += p.second;
    }
    return ZP_OK;
  }
};

}  // namespace

int launch_gather_openings(const zp_hint* h, const uint64_t* d_cols, uint32_t n_cols,
                           void* d_columns, void* d_paths, hipStream_t s) {
  if (!n_cols) return ZP_OK;
  const uint64_t rows = h->num_rows * h->batch, words = h->elem / 4;
  const uint32_t levels = h->height - 1;
  const uint64_t tc = n_cols * rows * words;
  hipLaunchKernelGGL(gather_columns_kernel, dim3((uint32_t)((tc + 255) / 256)), dim3(256), 0, s,
                     (const uint32_t*)h->d_cw, d_cols, (uint32_t*)d_columns, rows, h->N,
                     (uint32_t)words, n_cols);
  ZP_HIP(hipGetLastError());
  if (levels) {
    const uint64_t tp = (uint64_t)n_cols * levels * 8;
    hipLaunchKernelGGL(gather_paths_kernel, dim3((uint32_t)((tp + 255) / 256)), dim3(256), 0, s,
                       (const uint32_t*)h->d_layers, d_cols, (uint32_t*)d_paths, h->N, levels,
                       n_cols);
    ZP_HIP(hipGetLastError());
  }
  return ZP_OK;
}

}  // namespace zpi

using namespace zpi;

extern "C" int zp_prove(const zp_code* code, int eval_type, uint32_t D, uint32_t L,
                        const void* d_evals, uint64_t num_rows, uint32_t batch,
                        const zp_hint* hint, const int64_t* h_point, uint32_t num_vars,
                        const uint8_t h_modulus[32], uint32_t num_openings, zp_transcript* t,
                        uint8_t h_eval[32], zp_stream stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!code || !hint || !t || !h_modulus || !h_eval || (num_vars && !h_point) || !d_evals)
    return fail(ZP_ERR_INVALID_PARAM, "null argument");
  ZP_TRY(check_types(code, eval_type, D));
  if (eval_type != ZP_EVAL_BINPOLY && eval_type != ZP_EVAL_I32)
    return fail(ZP_ERR_INVALID_PARAM, "prove supports BinaryPoly and i32 evals");
  if (L < 3 || L > 8) return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 3..8");
  const uint64_t row_len = code->row_len, N = code->N;
  if (batch == 0) return fail(ZP_ERR_SHAPE, "Batch must contain at least one polynomial");
  if (!pow2(num_rows)) return fail(ZP_ERR_SHAPE, "num_rows must be a power of two");
  if (hint->N != N || hint->num_rows != num_rows || hint->batch != batch ||
      hint->elem != cw_elem_bytes_of(code, eval_type, D))
    return fail(ZP_ERR_SHAPE, "hint does not match the code, rows, batch or eval type");
  // validate_input (pcs/utils.rs:16-78): the combined row's bit budget, and
  // the point has one coordinate per variable of the polys.
  ZP_TRY(lc_budget_check(code->rep, row_len, eval_type, D, L, batch));
  {
    const uint64_t d = (uint64_t)__builtin_ctzll(row_len);
    const uint64_t nv = __builtin_ctzll(num_rows) + d;
    if (num_vars != nv)
      return fail(ZP_ERR_INVALID_PARAM, "Invalid point (expect point to have " +
                                            std::to_string(nv) + " variates but got " +
                                            std::to_string(num_vars) + ")");
  }
  uint64_t mod[4];
  std::memcpy(mod, h_modulus, 32);
  if (!(mod[0] & 1) || (mod[2] | mod[3]) == 0)
    return fail(ZP_ERR_INVALID_PARAM, "field modulus must be odd and above 2^128");
  const FpParams P = fp_params(mod);
  const bool binpoly = eval_type == ZP_EVAL_BINPOLY;
  const uint32_t Dk = binpoly ? D : 32;  // planes of one eval
  const uint32_t row_vars = __builtin_ctzll(num_rows), col_vars = num_vars - row_vars;

  // Point in the field: hi = point[..col_vars] -> q_1, lo -> q_0.
  std::vector<Fp> rr(2 * (size_t)num_vars);
  for (uint32_t k = 0; k < num_vars; k++) {
    const Fp r = fp_from_signed_limbs(P, (const uint64_t*)(h_point + 2 * k), 2);
    rr[2 * k] = r;
    rr[2 * k + 1] = fp_sub(P, fp_one(P), r);
  }
  // Per-poly alphas (get_challenges::<i128>(degree_bound + 1)), or [1].
  std::vector<int64_t> al((size_t)batch * Dk * 2);
  std::vector<Fp> alm((size_t)batch * Dk);
  for (uint32_t b = 0; b < batch; b++) {
    for (uint32_t c = 0; c < Dk; c++) {
      int64_t* a = &al[((size_t)b * Dk + c) * 2];
      if (binpoly && D > 1) {
        ts_challenge(t, (uint8_t*)a, 16);
      } else if (binpoly) {
        a[0] = 1, a[1] = 0;
      } else {  // i32 eval as bit planes: bit c weighs 2^c, bit 31 weighs -2^31
        a[0] = c < 31 ? (int64_t)(1ull << c) : INT64_MIN >> 32;
        a[1] = c < 31 ? 0 : -1;
      }
      alm[(size_t)b * Dk + c] = fp_from_signed_limbs(P, (const uint64_t*)a, 2);
    }
  }

  const uint64_t rows = num_rows * batch, elem = hint->elem;
  const uint32_t levels = hint->height - 1;
  Fp *d_rr, *d_q1, *d_q0, *d_planes, *d_b, *d_alm;
  int64_t *d_al, *d_co;
  uint64_t *d_cols, *d_comb, *d_wide = nullptr;
  void *d_scr, *d_open;
  Bufs bufs;
  bufs.add((void**)&d_rr, rr.size() * sizeof(Fp));
  bufs.add((void**)&d_q1, row_len * sizeof(Fp));
  bufs.add((void**)&d_q0, num_rows * sizeof(Fp));
  bufs.add((void**)&d_planes, rows * Dk * sizeof(Fp));
  // Plane sums on the matrix cores (default): column blocks of whole
  // K-steps (64 columns), enough for ~2048 workgroups over rows and plane
  // chunks; or, with zp_set_option("combined_row", 1), the VALU kernel with
  // blocks of at least 256 columns.
  const bool mfma_sums = g_force_comb != 1;
  const uint32_t ksteps = (uint32_t)((row_len + 63) / 64);
  uint32_t nblk, cpb = 0, spb = 0;
  if (mfma_sums) {
    const uint64_t units = rows * ((Dk + 15) / 16);
    nblk = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ksteps, (2048 + units - 1) / units));
    spb = (ksteps + nblk - 1) / nblk;
    if (spb > kMaxBlockSteps) spb = kMaxBlockSteps;
    nblk = (ksteps + spb - 1) / spb;
  } else {
    nblk = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(row_len / 256, (2048 + rows - 1) / rows));
    cpb = (uint32_t)((row_len + nblk - 1) / nblk);
  }
  uint32_t* d_part;
  bufs.add((void**)&d_part, rows * nblk * Dk * 9 * sizeof(uint32_t));
  v4i* d_qfrag = nullptr;
  if (mfma_sums) bufs.add((void**)&d_qfrag, (uint64_t)ksteps * 192 * sizeof(v4i));
  bufs.add((void**)&d_b, num_rows * sizeof(Fp));
  bufs.add((void**)&d_alm, alm.size() * sizeof(Fp));
  bufs.add((void**)&d_al, al.size() * 8);
  bufs.add((void**)&d_co, num_rows * 16);
  bufs.add((void**)&d_cols, (uint64_t)num_openings * 8);
  bufs.add((void**)&d_comb, row_len * L * 8);
  bufs.add(&d_scr, combined_row_scratch_bytes(Dk, row_len, batch, num_rows));
  const uint64_t open_bytes = rows * elem + 24 + (uint64_t)levels * 32;  // one opening
  bufs.add(&d_open, (uint64_t)num_openings * open_bytes);
  if (!binpoly) bufs.add((void**)&d_wide, rows * row_len * 8);
  ZP_TRY(bufs.alloc(s));
  int rc = ZP_OK;
  // Every exit goes through here.  On a failure, copies from the per-thread
  // pinned buffer may still be in flight: they must finish before the
  // buffer is reused or regrown by the next call on this thread.
  auto done = [&](int r) {
    if (r != ZP_OK) (void)hipStreamSynchronize(s);
    dev_free(bufs.base, s);
    return r;
  };
  const uint64_t up1 = rr.size() * sizeof(Fp) + alm.size() * sizeof(Fp);
  const uint64_t down1 = 2 * num_rows * sizeof(Fp);
  const uint64_t up2 = al.size() * 8 + num_rows * 16 + (uint64_t)num_openings * 8;
  const uint64_t down2 = row_len * L * 8 + (uint64_t)num_openings * open_bytes;
  uint8_t* pin = pinned_scratch(std::max(up1 + down1, up2 + down2));
  if (!pin) return done(fail(ZP_ERR_OOM, "pinned staging allocation failed"));

  // First prover message: q tables, plane sums, b.
  std::memcpy(pin, rr.data(), rr.size() * sizeof(Fp));
  std::memcpy(pin + rr.size() * sizeof(Fp), alm.data(), alm.size() * sizeof(Fp));
  if (hipMemcpyAsync(d_rr, pin, rr.size() * sizeof(Fp), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d_alm, pin + rr.size() * sizeof(Fp), alm.size() * sizeof(Fp),
                     hipMemcpyHostToDevice, s) != hipSuccess)
    return done(fail(ZP_ERR_DEVICE, "upload failed"));
  hipLaunchKernelGGL(eq_table_kernel, dim3((uint32_t)((row_len + 255) / 256)), dim3(256), 0, s,
                     d_rr, col_vars, d_q1, P);
  hipLaunchKernelGGL(eq_table_kernel, dim3((uint32_t)((num_rows + 255) / 256)), dim3(256), 0, s,
                     d_rr + 2 * col_vars, row_vars, d_q0, P);
  const uint32_t dp_log2 = ceil_log2(Dk);
  const uint32_t red_bits = ceil_log2(row_len) + 1;
  if (mfma_sums) {
    hipLaunchKernelGGL(q1_afrag_kernel, dim3((ksteps * 192 + 255) / 256), dim3(256), 0, s, d_q1,
                       d_qfrag, row_len, ksteps);
    const uint32_t grid = (uint32_t)(rows * ((Dk + 15) / 16) * nblk);
    if (binpoly)
      hipLaunchKernelGGL(plane_sums_mfma_kernel<uint64_t>, dim3(grid), dim3(256), 0, s,
                         (const uint64_t*)d_evals, (const v4i*)d_qfrag, d_part, row_len, ksteps,
                         spb, nblk, Dk);
    else
      hipLaunchKernelGGL(plane_sums_mfma_kernel<uint32_t>, dim3(grid), dim3(256), 0, s,
                         (const uint32_t*)d_evals, (const v4i*)d_qfrag, d_part, row_len, ksteps,
                         spb, nblk, Dk);
  } else if (binpoly) {
    hipLaunchKernelGGL(plane_sums_kernel<uint64_t>, dim3((uint32_t)(rows * nblk)),
                       dim3(64 * kSumWaves), 0, s, (const uint64_t*)d_evals, d_q1, d_part, row_len,
                       cpb, nblk, Dk, dp_log2);
  } else {
    hipLaunchKernelGGL(plane_sums_kernel<uint32_t>, dim3((uint32_t)(rows * nblk)),
                       dim3(64 * kSumWaves), 0, s, (const uint32_t*)d_evals, d_q1, d_part, row_len,
                       cpb, nblk, Dk, dp_log2);
  }
  hipLaunchKernelGGL(plane_reduce_kernel, dim3((uint32_t)((rows * Dk + 255) / 256)), dim3(256), 0,
                     s, d_part, d_planes, rows, nblk, Dk, red_bits, P);
  hipLaunchKernelGGL(b_kernel, dim3((uint32_t)((num_rows + 3) / 4)), dim3(256), 0, s, d_planes,
                     d_alm, d_b, num_rows, Dk, batch, P);
  if (hipGetLastError() != hipSuccess) return done(fail(ZP_ERR_DEVICE, "launch failed"));
  uint8_t* dn = pin + up1;
  if (hipMemcpyAsync(dn, d_b, num_rows * sizeof(Fp), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(dn + num_rows * sizeof(Fp), d_q0, num_rows * sizeof(Fp),
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(ZP_ERR_DEVICE, "b copy failed"));
  std::vector<Fp> b(num_rows), q0(num_rows);
  std::memcpy(b.data(), dn, num_rows * sizeof(Fp));
  std::memcpy(q0.data(), dn + num_rows * sizeof(Fp), num_rows * sizeof(Fp));

  // Transcript: b, eval = <q_0, b>, s, column indices.
  ts_write_field_elements(t, mod, b.data(), b.size());
  Fp ev = {{0, 0, 0, 0}};
  for (uint64_t j = 0; j < num_rows; j++) ev = fp_add(P, ev, fp_mul(P, q0[j], b[j]));
  const Fp evc = fp_from_mont(P, ev);
  std::memcpy(h_eval, evc.w, 32);
  std::vector<int64_t> co(num_rows * 2);
  if (num_rows == 1)
    co[0] = 1, co[1] = 0;
  else
    for (uint64_t j = 0; j < num_rows; j++) ts_challenge(t, (uint8_t*)&co[2 * j], 16);
  // The combined row and the columns are written without being absorbed, so
  // every index follows from the state after s.
  std::vector<uint64_t> cols(num_openings);
  for (uint32_t k = 0; k < num_openings; k++) cols[k] = ts_squeeze_idx(t, N);

  // Second prover message: combined row, columns, paths.
  uint8_t* up = pin;
  std::memcpy(up, al.data(), al.size() * 8);
  std::memcpy(up + al.size() * 8, co.data(), num_rows * 16);
  std::memcpy(up + al.size() * 8 + num_rows * 16, cols.data(), (uint64_t)num_openings * 8);
  if (hipMemcpyAsync(d_al, up, al.size() * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d_co, up + al.size() * 8, num_rows * 16, hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      (num_openings && hipMemcpyAsync(d_cols, up + al.size() * 8 + num_rows * 16,
                                      (uint64_t)num_openings * 8, hipMemcpyHostToDevice,
                                      s) != hipSuccess))
    return done(fail(ZP_ERR_DEVICE, "upload failed"));
  const uint64_t* ev64 = (const uint64_t*)d_evals;
  if (!binpoly) {
    const uint64_t n = rows * row_len;
    hipLaunchKernelGGL(widen_i32_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s,
                       (const uint32_t*)d_evals, d_wide, n);
    ev64 = d_wide;
  }
  rc = launch_combined_row(Dk, L, num_rows, row_len, batch, ev64, d_al, d_co, d_comb, d_scr, s);
  if (rc != ZP_OK) return done(rc);
  if (num_openings) {
    const uint64_t tw = (uint64_t)num_openings * open_bytes / 4;
    hipLaunchKernelGGL(pack_openings_kernel, dim3((uint32_t)((tw + 255) / 256)), dim3(256), 0, s,
                       (const uint32_t*)hint->d_cw, (const uint32_t*)hint->d_layers, d_cols,
                       (uint32_t*)d_open, rows, N, (uint32_t)(elem / 4), levels, num_openings);
    if (hipGetLastError() != hipSuccess) return done(fail(ZP_ERR_DEVICE, "launch failed"));
  }
  // Proof bytes: the combined row (write_const_many), then the openings.
  const uint64_t comb_bytes = row_len * L * 8;
  const uint64_t open_total = (uint64_t)num_openings * open_bytes;
  // With pinned stream storage the copies land in the stream itself.
  uint8_t* direct = nullptr;
  try {
    direct = ts_grow_pinned(t, comb_bytes + open_total);
  } catch (const std::bad_alloc&) {
    return done(fail(ZP_ERR_OOM, "proof stream allocation failed"));
  }
  if (direct) {
    if (hipMemcpyAsync(direct, d_comb, comb_bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
        (open_total && hipMemcpyAsync(direct + comb_bytes, d_open, open_total,
                                      hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess) {
      ts_unput(t, comb_bytes + open_total);
      return done(fail(ZP_ERR_DEVICE, "proof copy failed"));
    }
    return done(ZP_OK);
  }
  // Otherwise the openings come back in pieces through the pinned staging
  // buffer, each followed by an event, so the host appends piece k to the
  // stream while later pieces are in flight.
  uint8_t* o_comb = pin + up2;
  uint8_t* o_open = o_comb + row_len * L * 8;
  constexpr uint64_t kPiece = 16ull << 20;
  const uint64_t npieces = (open_total + kPiece - 1) / kPiece;
  std::vector<hipEvent_t> pev(npieces, nullptr);
  rc = hipMemcpyAsync(o_comb, d_comb, comb_bytes, hipMemcpyDeviceToHost, s) == hipSuccess
           ? ZP_OK
           : fail(ZP_ERR_DEVICE, "proof copy failed");
  for (uint64_t k = 0; rc == ZP_OK && k < npieces; k++) {
    const uint64_t off = k * kPiece, len = std::min(kPiece, open_total - off);
    if (hipEventCreateWithFlags(&pev[k], hipEventDisableTiming) != hipSuccess ||
        hipMemcpyAsync(o_open + off, (const uint8_t*)d_open + off, len, hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipEventRecord(pev[k], s) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "proof copy failed");
  }
  if (rc == ZP_OK && npieces == 0 && hipStreamSynchronize(s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "proof copy failed");
  if (rc == ZP_OK) ts_reserve(t, down2);
  for (uint64_t k = 0; rc == ZP_OK && k < npieces; k++) {
    if (hipEventSynchronize(pev[k]) != hipSuccess) {
      rc = fail(ZP_ERR_DEVICE, "proof copy failed");
      break;
    }
    if (k == 0) ts_put(t, o_comb, comb_bytes);
    const uint64_t off = k * kPiece;
    ts_put(t, o_open + off, std::min(kPiece, open_total - off));
  }
  if (rc == ZP_OK && npieces == 0) ts_put(t, o_comb, comb_bytes);
  if (rc != ZP_OK) (void)hipStreamSynchronize(s);  // before the events go
  for (hipEvent_t e : pev)
    if (e) (void)hipEventDestroy(e);
  return done(rc);
}
