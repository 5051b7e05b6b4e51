// raa.hip -- RAA (repeat-accumulate-accumulate) row encoder for gfx950.
//
// Reference: zip-plus/src/code/raa.rs:93-126 encode_inner (repeat :185-195,
// accumulate :208-236, permutation by index :239-244) and the sign-flipping
// variant zip-plus/src/code/raa_sign_flip.rs:30-71 (odd indices negated
// after the repeat and after the first accumulate, :135-165).
//
// Per coefficient plane p of a row with codeword length N:
//   y1[i]  = s(perm1[i]) * bit_p(x[perm1[i] mod row_len])
//   a[i]   = s(i) * sum_{l <= i} y1[l]
//   out[i] = sum_{l <= i} a[perm2[l]]
// where s(i) = -1 for odd i under the sign-flipping code and +1 otherwise.
//
// A workgroup owns one row and a group of up to 8 planes.  It keeps a[] in
// LDS as 16-bit values (|a| <= N for the plain code, <= N/2 + 1 with sign
// flips, so u16 / i16 hold them for N <= 2^15) and never materialises y1 or
// the gathered sequence: each scan is a per-thread segment sum, a block-wide
// exclusive scan of the segment sums, and a second pass that emits values.
// Output values stay below 2^26 in magnitude (raa.rs:52-69) and wrap like
// the reference's i32 accumulate_unchecked.  Each position of the output
// row is 4*D contiguous bytes; a workgroup writes its planes' 4*PG bytes,
// and the plane groups of one row run on one XCD so the pieces of each line
// meet in the same L2.
#include "common.h"

namespace zpi {
namespace {

constexpr int kThreads = 512;

}  // namespace


namespace {

// Lane-wise 16-bit add / negate on packed pairs (no carry between lanes).
__device__ __forceinline__ uint32_t add16x2(uint32_t a, uint32_t b) {
  return ((a & 0xffff0000u) + (b & 0xffff0000u)) | ((a + b) & 0xffffu);
}
__device__ __forceinline__ uint32_t neg16x2(uint32_t a) { return add16x2(~a, 0x00010001u); }

// Exclusive scan over the block of K per-thread words; kPacked words hold
// two independent 16-bit lanes.  part has kThreads * K words.
template <int K, bool kPacked>
__device__ void block_exscan(uint32_t v[K], uint32_t* part) {
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < K; k++) part[k * kThreads + t] = v[k];
  __syncthreads();
  for (uint32_t off = 1; off < kThreads; off <<= 1) {
    uint32_t add[K];
#pragma unroll
    for (int k = 0; k < K; k++) add[k] = t >= off ? part[k * kThreads + t - off] : 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++)
      part[k * kThreads + t] = kPacked ? add16x2(part[k * kThreads + t], add[k])
                                       : part[k * kThreads + t] + add[k];
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    v[k] = kPacked ? add16x2(part[k * kThreads + t], neg16x2(v[k])) : part[k * kThreads + t] - v[k];
  __syncthreads();
}

// 2*NW plane bits of `b` as NW pairs of 16-bit lanes.
template <int NW>
__device__ __forceinline__ void spread_pairs(uint32_t b, uint32_t out[NW]) {
#pragma unroll
  for (int m = 0; m < NW; m++)
    out[m] = ((b >> (2 * m)) & 1u) | (((b >> (2 * m + 1)) & 1u) << 16);
}

template <int NW> struct Vec;
template <> struct Vec<4> { typedef uint4 T; };
template <> struct Vec<2> { typedef uint2 T; };
template <> struct Vec<1> { typedef uint32_t T; };
template <int NW> __device__ __forceinline__ void unpack(const typename Vec<NW>::T& v, uint32_t w[NW]);
template <> __device__ __forceinline__ void unpack<4>(const uint4& v, uint32_t w[4]) {
  w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
}
template <> __device__ __forceinline__ void unpack<2>(const uint2& v, uint32_t w[2]) {
  w[0] = v.x; w[1] = v.y;
}
template <> __device__ __forceinline__ void unpack<1>(const uint32_t& v, uint32_t w[1]) { w[0] = v; }
template <int NW> __device__ __forceinline__ typename Vec<NW>::T pack(const uint32_t w[NW]);
template <> __device__ __forceinline__ uint4 pack<4>(const uint32_t w[4]) {
  return make_uint4(w[0], w[1], w[2], w[3]);
}
template <> __device__ __forceinline__ uint2 pack<2>(const uint32_t w[2]) {
  return make_uint2(w[0], w[1]);
}
template <> __device__ __forceinline__ uint32_t pack<1>(const uint32_t w[1]) { return w[0]; }

template <bool kSigned>
__device__ __forceinline__ uint32_t lo16(uint32_t v) {
  return kSigned ? (uint32_t)(int32_t)(int16_t)(v & 0xffffu) : (v & 0xffffu);
}
template <bool kSigned>
__device__ __forceinline__ uint32_t hi16(uint32_t v) {
  return kSigned ? (uint32_t)((int32_t)v >> 16) : (v >> 16);
}

// S = positions per thread (N / kThreads, at least 1); all per-position
// inputs of a thread are loaded into registers once and reused by both the
// summing and the emitting pass of each scan.
template <int NW, int S, bool kFlip>
__global__ void __launch_bounds__(kThreads)
raa_binary_kernel(const uint64_t* __restrict__ evals, int32_t* __restrict__ out,
                  const uint32_t* __restrict__ perm1, const uint32_t* __restrict__ perm2,
                  uint32_t row_len, uint32_t N, uint32_t D, uint32_t npg, uint64_t rows,
                  uint64_t row0) {
  constexpr uint32_t PG = 2 * NW;  // planes per workgroup
  typedef typename Vec<NW>::T V;
  extern __shared__ uint32_t lds[];
  V* A = (V*)lds;                           // N x (PG planes x 16 bit)
  uint32_t* part = lds + NW * (size_t)N;    // kThreads x PG scan lanes
  const uint32_t t = threadIdx.x;

  // Block -> (row, plane group); the groups of one row share an XCD.
  const uint64_t b = blockIdx.x;
  const uint64_t q = b >> 3;
  const uint64_t row = row0 + (b & 7) + 8 * (q / npg);
  const uint32_t pg = (uint32_t)(q % npg);
  if (row >= rows) return;
  const uint32_t p0 = PG * pg;
  const uint32_t pmask = (1u << PG) - 1u;
  const uint64_t* x = evals + row * row_len;
  const uint32_t i0 = t * S;
  const bool active = i0 < N;  // N < kThreads leaves some threads idle
  const uint32_t rmask = row_len - 1;

  // Prefetch the plane bits of every position (4 per word), the parity of
  // each perm1 entry (one mask) and the perm2 entries (two 16-bit per word).
  constexpr int SB = (S + 3) / 4, SP = (S + 1) / 2;
  uint32_t bitw[SB], p2w[SP];
  uint64_t odd = 0;
  {
    uint32_t src[S];
#pragma unroll
    for (int k = 0; k < S; k++) src[k] = active ? perm1[i0 + k] : 0u;
#pragma unroll
    for (int k = 0; k < SB; k++) bitw[k] = 0;
#pragma unroll
    for (int k = 0; k < S; k++) {
      const uint32_t bv = active ? (uint32_t)(x[src[k] & rmask] >> p0) & pmask : 0u;
      bitw[k / 4] |= bv << (8 * (k % 4));
      odd |= (uint64_t)(src[k] & 1u) << k;
    }
#pragma unroll
    for (int k = 0; k < SP; k++) {
      const uint32_t a = active ? perm2[i0 + 2 * k] : 0u;
      const uint32_t c = (active && 2 * k + 1 < S) ? perm2[i0 + 2 * k + 1] : 0u;
      p2w[k] = a | (c << 16);
    }
  }
#define BITS(k) ((bitw[(k) / 4] >> (8 * ((k) % 4))) & 0xffu)
#define ODD(k) ((odd >> (k)) & 1u)
#define P2(k) ((p2w[(k) / 2] >> (16 * ((k) % 2))) & 0xffffu)

  // First accumulate over y1 (NW packed 16-bit pairs = PG planes).
  uint32_t sum[PG];
#pragma unroll
  for (int k = 0; k < (int)PG; k++) sum[k] = 0;
#pragma unroll
  for (int k = 0; k < S; k++) {
    uint32_t v[NW];
    spread_pairs<NW>(BITS(k), v);
#pragma unroll
    for (int m = 0; m < NW; m++) {
      if (kFlip && ODD(k)) v[m] = neg16x2(v[m]);
      sum[m] = add16x2(sum[m], v[m]);
    }
  }
  block_exscan<NW, true>(sum, part);
  if (active) {
#pragma unroll
    for (int k = 0; k < S; k++) {
      uint32_t v[NW];
      spread_pairs<NW>(BITS(k), v);
      uint32_t a[NW];
#pragma unroll
      for (int m = 0; m < NW; m++) {
        if (kFlip && ODD(k)) v[m] = neg16x2(v[m]);
        sum[m] = add16x2(sum[m], v[m]);
        a[m] = (kFlip && ((i0 + k) & 1u)) ? neg16x2(sum[m]) : sum[m];
      }
      A[i0 + k] = pack<NW>(a);
    }
  }
  __syncthreads();

  // Second accumulate over the perm2 gather, in 32-bit lanes.
  uint32_t g[S][NW];
#pragma unroll
  for (int k = 0; k < S; k++) {
    if (active) {
      unpack<NW>(A[P2(k)], g[k]);
    } else {
#pragma unroll
      for (int m = 0; m < NW; m++) g[k][m] = 0;
    }
  }
#pragma unroll
  for (int k = 0; k < (int)PG; k++) sum[k] = 0;
#pragma unroll
  for (int k = 0; k < S; k++)
#pragma unroll
    for (int m = 0; m < NW; m++) {
      sum[2 * m] += lo16<kFlip>(g[k][m]);
      sum[2 * m + 1] += hi16<kFlip>(g[k][m]);
    }
  block_exscan<PG, false>(sum, part);
  if (!active) return;
  const uint32_t np = min(PG, D - p0);
  int32_t* o = out + (row * N + i0) * D + p0;
  const bool vec = np == PG && PG >= 4 && (D & 3) == 0;
#pragma unroll
  for (int k = 0; k < S; k++) {
#pragma unroll
    for (int m = 0; m < NW; m++) {
      sum[2 * m] += lo16<kFlip>(g[k][m]);
      sum[2 * m + 1] += hi16<kFlip>(g[k][m]);
    }
    int32_t* dst = o + (uint64_t)k * D;
    if (vec) {
#pragma unroll
      for (int c = 0; c < (int)PG; c += 4)
        *(uint4*)(dst + c) = make_uint4(sum[c], sum[c + 1], sum[c + 2], sum[c + 3]);
    } else {
#pragma unroll
      for (uint32_t c = 0; c < PG; c++)
        if (c < np) dst[c] = (int32_t)sum[c];
    }
  }
#undef BITS
#undef ODD
#undef P2
}

template <int NW, int S, bool kFlip>
int launch_raa(const zp_code* cd, const uint64_t* evals, uint64_t rows, uint32_t D, int32_t* out,
               hipStream_t s) {
  const uint32_t N = (uint32_t)cd->N;
  const size_t lds = (size_t)N * 4 * NW + (size_t)kThreads * 2 * NW * 4;
  ZP_TRY((grant_lds<raa_binary_kernel<NW, S, kFlip>>(lds)));
  const uint32_t npg = (D + 2 * NW - 1) / (2 * NW);
  // Blocks are grouped 8 rows at a time: block = 8*(rowgroup*npg + pg) + r8.
  const uint64_t max_groups = std::max<uint64_t>(1, 0x7fffffffull / (8ull * npg));
  for (uint64_t r0 = 0; r0 < rows; r0 += 8 * max_groups) {
    const uint64_t nr = std::min<uint64_t>(8 * max_groups, rows - r0);
    const uint64_t groups = (nr + 7) / 8;
    const uint32_t blocks = (uint32_t)(groups * npg * 8);
    hipLaunchKernelGGL((raa_binary_kernel<NW, S, kFlip>), dim3(blocks), dim3(kThreads), lds, s,
                       evals, out, cd->d_perm1, cd->d_perm2, (uint32_t)cd->row_len, N, D, npg,
                       r0 + nr, r0);
    ZP_HIP(hipGetLastError());
  }
  return ZP_OK;
}

// Full-row kernel for the plain code with D = 64 or 32 planes.
//
// The first prefix sum of plane p is a count of set bits of y1, so it is
// stored compactly: the y1 bit-vector of every plane (N bits, in 32-bit
// words) plus a 16-bit checkpoint per word (the count before that word):
//   a[m][p] = cp[p][m / 32] + popcount(bits[p][m / 32] << (31 - m % 32)).
// The shift keeps bits 0 .. m % 32 of the word, so a rank is one xor, one
// shift and one bit count (which adds the checkpoint).  Plane rows are an
// odd number of 32-bit words apart (bits) and of bank words apart
// (checkpoints), so the 32 lanes that look up one position (one lane per
// plane) hit different banks.  A workgroup holds 32 planes of one row (N = 8192:
// 32 KiB of bits + 16 KiB of checkpoints + 16 KiB of perm2 as 16-bit
// indices), so two workgroups share a CU and one's stores overlap the
// other's scans; at D = 64 the two halves of a row are separate
// workgroups.  Phase 1 puts positions in lanes and turns each plane's bits
// into 64-bit words with one ballot per plane; phase 2 puts planes in
// lanes, two positions per wave, so each store instruction writes two
// 128-byte runs.  perm2 is copied to LDS before any store: the memory
// counter also counts stores, so a global load of perm2 between the stores
// would wait for all of them.
constexpr uint32_t kRowThreads = 512;
constexpr uint32_t kRowWaves = kRowThreads / 64;
constexpr uint32_t kRowPlanes = 32;  // planes per workgroup

// Word strides of one plane: bits (odd, in 32-bit words) and checkpoints
// (an odd number of 32-bit bank words, i.e. 2 mod 4 in 16-bit entries).
__host__ __device__ inline uint32_t raa_bits_stride(uint32_t N) { return N / 32 + 1; }
__host__ __device__ inline uint32_t raa_cp_stride(uint32_t N) {
  const uint32_t w = N / 32;
  return (w % 4 == 0) ? w + 2 : ((w + 3) / 4) * 4 + 2;
}

// Counts the bits of `word` at positions 0 .. (m mod 32), plus `base`.
__device__ __forceinline__ uint32_t rank_word(uint32_t word, uint32_t m, uint32_t base) {
  uint32_t t;
  // The shift uses the low 5 bits of m ^ 31, i.e. 31 - m % 32.
  asm("v_lshlrev_b32 %0, %1, %2" : "=v"(t) : "v"(m ^ 31u), "v"(word));
  return __popc(t) + base;
}

// D planes per position in the output row, PG of them (planes PG*pg ..
// PG*pg + PG - 1) per workgroup: blockIdx.x = row * ceil(D / PG) + pg.
// kD = 32 or 64 fixes D at compile time; kD = 0 takes D (16..64) at run
// time, and lanes of planes at or above D compute but do not store.
//
// kFlip (the sign-flipping code): the first accumulate adds -bit where the
// perm1 entry of the position is odd, and the accumulated value of an odd
// position is negated.  With O the bit-vector of odd perm1 entries (one
// ballot per 64 positions, shared by all planes), the prefix up to m is
// rank(bits, m) - 2 rank(bits & O, m); the checkpoints hold that signed
// prefix at word boundaries (|prefix| <= N <= 2^14 fits 16 bits).
template <int kD, bool kFlip>
__device__ __forceinline__ void raa_row_body(const uint64_t* __restrict__ evals,
                                             int32_t* __restrict__ out,
                                             const uint32_t* __restrict__ perm1,
                                             const uint32_t* __restrict__ perm2, uint32_t row_len,
                                             uint32_t N, uint64_t row0, uint32_t Drt) {
  constexpr uint32_t PG = kRowPlanes;
  const uint32_t D = kD ? (uint32_t)kD : Drt;
  constexpr uint32_t SUB = 64 / PG;          // position streams per wave
  constexpr uint32_t NS = kRowWaves * SUB;   // position streams per workgroup
  const uint32_t NPG = kD ? (uint32_t)kD / PG : (D + PG - 1) / PG;
  extern __shared__ uint32_t lds32[];
  const uint32_t W = N / 64;                 // 64-bit words (ballots) per plane
  const uint32_t W32 = N / 32;               // 32-bit words per plane
  const uint32_t WP = raa_bits_stride(N);
  const uint32_t CP = raa_cp_stride(N);
  uint32_t* bits = lds32;                                        // [PG][WP]
  uint16_t* cp = (uint16_t*)(lds32 + (size_t)PG * WP);           // [PG][CP]
  uint32_t* part = lds32 + (size_t)PG * WP + ((size_t)PG * CP) / 2;
  uint16_t* p2 = (uint16_t*)(part + 2 * kRowThreads);            // [N] perm2 (N <= 2^15)
  uint32_t* oddm = (uint32_t*)(p2 + N);                          // [W32] odd perm1 entries
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t row = row0 + blockIdx.x / NPG;
  const uint32_t poff = PG * (blockIdx.x % NPG);
  const uint64_t* x = evals + row * row_len;
  const uint32_t rmask = row_len - 1;

  for (uint32_t i = tid; i < N; i += kRowThreads) p2[i] = (uint16_t)perm2[i];
  // Phase 1a: y1 bit-vectors, 64 positions per wave step; the permutation
  // and gather loads of up to kPre steps are issued together.
  constexpr uint32_t kPre = 4;
  for (uint32_t b0 = wave; b0 < W; b0 += kRowWaves * kPre) {
    uint32_t src[kPre];
    uint64_t xv[kPre];
#pragma unroll
    for (uint32_t k = 0; k < kPre; k++) {
      const uint32_t blk = b0 + k * kRowWaves;
      src[k] = blk < W ? perm1[blk * 64 + lane] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kPre; k++) xv[k] = x[src[k] & rmask];
#pragma unroll
    for (uint32_t k = 0; k < kPre; k++) {
      const uint32_t blk = b0 + k * kRowWaves;
      if (blk >= W) break;
      if constexpr (kFlip) {
        const uint64_t om = __ballot((src[k] & 1u) != 0u);
        if (lane == 0) {
          oddm[2 * blk] = (uint32_t)om;
          oddm[2 * blk + 1] = (uint32_t)(om >> 32);
        }
      }
      // Lane q collects plane poff + q's word from the q-th ballot.
      const uint64_t xq = xv[k] >> poff;
      const uint32_t lo = (uint32_t)xq, hi = (uint32_t)(xq >> 32);
      uint32_t wlo = 0, whi = 0;
#pragma unroll
      for (int q = 0; q < PG; q++) {
        const uint64_t m = __ballot((((q < 32) ? (lo >> q) : (hi >> (q - 32))) & 1u) != 0u);
        if (lane == (uint32_t)q) {
          wlo = (uint32_t)m;
          whi = (uint32_t)(m >> 32);
        }
      }
      if (lane < (uint32_t)PG) {
        bits[lane * WP + 2 * blk] = wlo;
        bits[lane * WP + 2 * blk + 1] = whi;
      }
    }
  }
  __syncthreads();

  // Phase 1b: checkpoints (exclusive counts per word), one thread per
  // (plane, segment of words) plus a scan over the segments.
  {
    constexpr uint32_t SEGS = kRowThreads / PG;
    const uint32_t p = tid % PG, seg = tid / PG;
    const uint32_t ws = (W32 + SEGS - 1) / SEGS;
    const uint32_t w0 = min(W32, seg * ws), w1 = min(W32, w0 + ws);
    // Signed word counts of the flipped code wrap mod 2^32 like the
    // plain ones; every prefix fits 16 bits.
    auto count = [&](uint32_t w) -> uint32_t {
      const uint32_t b = bits[p * WP + w];
      if constexpr (kFlip) return __popc(b) - 2u * __popc(b & oddm[w]);
      return __popc(b);
    };
    uint32_t cnt = 0;
    for (uint32_t w = w0; w < w1; w++) cnt += count(w);
    part[tid] = cnt;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t sg = 0; sg < seg; sg++) before += part[sg * PG + p];
    for (uint32_t w = w0; w < w1; w++) {
      cp[p * CP + w] = (uint16_t)before;
      before += count(w);
    }
  }
  __syncthreads();

  // Phase 2: second accumulate, planes in lanes, in one pass of N/(NS*K)
  // steps.  In each step stream s takes K consecutive positions, prefixes
  // them in registers, and the streams exchange their totals through LDS
  // (double-buffered, one barrier per step) to get their carries.
  const uint32_t sub = lane / PG, p = lane % PG;
  const uint32_t s = wave * SUB + sub;
  constexpr uint32_t K = 64 / SUB;
  const uint32_t steps = N / (NS * K);
  const uint32_t* bp = bits + p * WP;
  const uint16_t* cpp = cp + p * CP;
  auto rank = [&](uint32_t m) -> uint32_t {
    const uint32_t w = m >> 5;
    if constexpr (kFlip) {
      const uint32_t word = bp[w];
      const uint32_t all = rank_word(word, m, (uint32_t)(int32_t)(int16_t)cpp[w]);
      const uint32_t odd = rank_word(word & oddm[w], m, 0u);
      // v = all - 2 odd in one 24-bit multiply-add (odd <= N <= 2^14), then
      // negated at an odd position as (v ^ mask) - mask, mask = -(m & 1).
      uint32_t v;
      asm("v_mad_i32_i24 %0, %1, -2, %2" : "=v"(v) : "v"(odd), "v"(all));
      const uint32_t mask = (uint32_t)((int32_t)(m << 31) >> 31);
      return (v ^ mask) - mask;
    }
    return rank_word(bp[w], m, cpp[w]);
  };
  int32_t* orow = out + row * N * D + poff + p;
  const bool stores = kD != 0 || poff + p < D;
  uint32_t run = 0;  // every position of the earlier steps
  for (uint32_t st = 0; st < steps; st++) {
    const uint32_t base = (st * NS + s) * K;
    uint32_t v[K];
    // The lookups of 16 positions are independent: all are issued first.
#pragma unroll
    for (uint32_t k0 = 0; k0 < K; k0 += 16) {
      uint32_t m[16];
#pragma unroll
      for (uint32_t k = 0; k < 16; k++) m[k] = p2[base + k0 + k];
#pragma unroll
      for (uint32_t k = 0; k < 16; k++) v[k0 + k] = rank(m[k]);
    }
#pragma unroll
    for (uint32_t k = 1; k < K; k++) v[k] += v[k - 1];
    uint32_t* tot = part + (st & 1) * NS * PG;
    tot[s * PG + p] = v[K - 1];
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t q = 0; q < NS; q++) {
      const uint32_t t = tot[q * PG + p];
      all += t;
      before += q < s ? t : 0u;
    }
    const uint32_t c = run + before;
    if (stores) {
      if constexpr (kFlip) {
        // One base address per step, the K stores at constant offsets
        // (k * D * 4 bytes, within the store's immediate field): this frees
        // the 64-bit address arithmetic the flipped lookups' registers
        // need (0.95 against 1.06 ms at 2^22, D = 64).  The plain kernel is
        // faster with per-store addresses (0.77 against 0.83 ms).
        int32_t* ob = orow + (uint64_t)base * D;
#pragma unroll
        for (uint32_t k = 0; k < K; k++) ob[k * D] = (int32_t)(c + v[k]);
      } else {
#pragma unroll
        for (uint32_t k = 0; k < K; k++) orow[(uint64_t)(base + k) * D] = (int32_t)(c + v[k]);
      }
    }
    run += all;
  }
}

template <int kD>
__global__ void __launch_bounds__(kRowThreads)
raa_row_kernel(const uint64_t* __restrict__ evals, int32_t* __restrict__ out,
               const uint32_t* __restrict__ perm1, const uint32_t* __restrict__ perm2,
               uint32_t row_len, uint32_t N, uint64_t row0, uint32_t D) {
  raa_row_body<kD, false>(evals, out, perm1, perm2, row_len, N, row0, D);
}

// The flipped lookups need more registers; 4 waves per SIMD keep two
// workgroups of 8 waves on a CU.
template <int kD>
__global__ void __launch_bounds__(kRowThreads, 4)
raa_row_flip_kernel(const uint64_t* __restrict__ evals, int32_t* __restrict__ out,
                    const uint32_t* __restrict__ perm1, const uint32_t* __restrict__ perm2,
                    uint32_t row_len, uint32_t N, uint64_t row0, uint32_t D) {
  raa_row_body<kD, true>(evals, out, perm1, perm2, row_len, N, row0, D);
}

// bits, checkpoints, the exchange area (phase-1b partial counts, then two
// buffers of per-stream totals: 2 * 8 * 64 words), and perm2.
constexpr uint32_t kRowPartWords = 2 * kRowThreads;

size_t raa_row_lds(uint32_t N, bool flip) {
  return (size_t)kRowPlanes * raa_bits_stride(N) * 4 +
         (size_t)kRowPlanes * raa_cp_stride(N) * 2 + (size_t)kRowPartWords * 4 + (size_t)N * 2 +
         (flip ? (size_t)N / 8 : 0);
}

template <int kD, bool kFlip>
int launch_raa_row(const zp_code* cd, const uint64_t* evals, uint64_t rows, uint32_t D,
                   int32_t* out, hipStream_t s) {
  const uint32_t N = (uint32_t)cd->N;
  const size_t lds = raa_row_lds(N, kFlip);
  constexpr auto kernel = kFlip ? raa_row_flip_kernel<kD> : raa_row_kernel<kD>;
  ZP_TRY(grant_lds<kernel>(lds));
  const uint32_t NPG = (D + kRowPlanes - 1) / kRowPlanes;
  const uint64_t max_rows = 0x7fffffffull / NPG;
  for (uint64_t r0 = 0; r0 < rows; r0 += max_rows) {
    const uint32_t nr = (uint32_t)std::min<uint64_t>(max_rows, rows - r0);
    hipLaunchKernelGGL(kernel, dim3(nr * NPG), dim3(kRowThreads), lds, s, evals, out,
                       cd->d_perm1, cd->d_perm2, (uint32_t)cd->row_len, N, r0, D);
    ZP_HIP(hipGetLastError());
  }
  return ZP_OK;
}

template <bool kFlip>
int dispatch_raa(const zp_code* cd, const uint64_t* evals, uint64_t rows, uint32_t D,
                 int32_t* out, hipStream_t s, bool wide) {
  // Planes per workgroup: the 16-bit first-scan array (N x planes x 2 B)
  // and the scan scratch must fit in 160 KiB of LDS.
  switch (cd->N) {
    case 32768: return launch_raa<1, 64, kFlip>(cd, evals, rows, D, out, s);
    case 16384: return launch_raa<2, 32, kFlip>(cd, evals, rows, D, out, s);
    case 8192:
      return wide ? launch_raa<4, 16, kFlip>(cd, evals, rows, D, out, s)
                  : launch_raa<2, 16, kFlip>(cd, evals, rows, D, out, s);
    case 4096: return launch_raa<4, 8, kFlip>(cd, evals, rows, D, out, s);
    case 2048: return launch_raa<4, 4, kFlip>(cd, evals, rows, D, out, s);
    case 1024: return launch_raa<4, 2, kFlip>(cd, evals, rows, D, out, s);
    default: return launch_raa<4, 1, kFlip>(cd, evals, rows, D, out, s);  // N <= 512
  }
}

}  // namespace

bool raa_full_row_eligible(const zp_code* cd, uint32_t D) {
  // Either code, 16 to 64 planes, every stream a whole number of positions,
  // and the bit-vectors within LDS (N <= 2^14 here, so a flipped prefix fits
  // the 16-bit checkpoints).  Below 16 planes most lanes of a 32-plane
  // workgroup would idle; the plane-group kernel is as fast or faster there
  // (2^22 evals, D = 1: 0.17 against 0.31 ms; D = 16: 0.74 against 0.40 ms).
  const bool flip = cd->kind == ZP_CODE_RAA_SIGNFLIP;
  return (cd->kind == ZP_CODE_RAA || flip) && D >= 16 && D <= 64 && cd->N >= 1024 &&
         cd->N <= (1u << 14) && raa_row_lds((uint32_t)cd->N, flip) <= 150 * 1024;
}

// variant: 0 full-row kernel where eligible, 1 plane groups of 8 (at
// N = 8192), 2 plane groups of 4 (at N = 8192).
int launch_raa_binary(const zp_code* cd, const uint64_t* evals, uint64_t rows, uint32_t D,
                      int32_t* out, hipStream_t s, int variant) {
  if (rows == 0) return ZP_OK;
  if (cd->N > (1u << 15))
    return fail(ZP_ERR_INVALID_PARAM, "RAA codeword length above 32768 is not supported");
  if (variant == 0 && raa_full_row_eligible(cd, D)) {
    if (cd->kind == ZP_CODE_RAA_SIGNFLIP) {
      if (D == 64) return launch_raa_row<64, true>(cd, evals, rows, D, out, s);
      if (D == 32) return launch_raa_row<32, true>(cd, evals, rows, D, out, s);
      return launch_raa_row<0, true>(cd, evals, rows, D, out, s);
    }
    if (D == 64) return launch_raa_row<64, false>(cd, evals, rows, D, out, s);
    if (D == 32) return launch_raa_row<32, false>(cd, evals, rows, D, out, s);
    return launch_raa_row<0, false>(cd, evals, rows, D, out, s);
  }
  const bool wide = variant != 2;
  return cd->kind == ZP_CODE_RAA_SIGNFLIP ? dispatch_raa<true>(cd, evals, rows, D, out, s, wide)
                                          : dispatch_raa<false>(cd, evals, rows, D, out, s, wide);
}

}  // namespace zpi
