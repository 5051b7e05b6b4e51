// iprs_mfma.hip -- IPRS encoder for binary polynomials with the base layer
// on the int8 matrix cores of gfx950.
//
// Reference: zip-plus/src/code/iprs/pntt/radix8.rs:117-161 (base layer) and
// :58-113 with butterfly.rs:24-43 (radix-8 stages).
//
// The base layer of one row is an exact integer product
//   B[r][(c,p)] = sum_k M[r][k] * bit_p(x[octrev(c) + k*8^d]),
// M[r][k] = f(r*k mod bd) in [-32769, 32767].  Each f is split into three
// int8 slices f = s0 + 128*s1 + 16384*s2 (s0, s1 in [0,127], s2 in [-3,1]),
// and every slice product is one v_mfma_i32_16x16x64_i8 whose partial sums
// stay below 2^15: the recombined i32 sum is exact (|B| < 2^23).
//
// The radix-8 stages run in FP64 FMA.  With |B| < 2^23 and |tw| <= 32769,
// stage-0 products and sums stay below 2^42, so they are exact.  Stage-1
// inputs (< 2^42) are split at 2^32 so every product and partial sum stays
// below 2^51, again exact; the two halves are recombined in int64.  Each
// butterfly is evaluated as even and odd half-sums (the base twiddles
// satisfy B[u+4] = -B[u]), with an exact correction where a coefficient's
// normalisation differs from its negation by p.  The results equal the
// reference's wrapping-i64 arithmetic bit for bit because no value on this
// path exceeds 2^59.
//
// Work split: a pre-pass writes, per row and chunk c, the words P[c][w][p]
// holding the bits of plane p for k = 16w .. 16w+15.  The main kernel owns
// (row, 16 base rows r, 16 planes p) and all 8^d chunks: MFMA tiles are
// (16 r) x (16 planes) for one chunk, the summed base values go to LDS, and
// each thread then runs every stage of one (r, p) column and writes its
// outputs; the 16 planes of one (c, r) are 128 contiguous bytes.  The 16
// base rows of r-block rb are r = rb + (bd/16) m, m < 16: any 16 rows make
// an MFMA tile, and with rows bd/16 apart (16 KiB of codeword at D = 32)
// the four lines of one store instruction fall in four different 4 KiB
// spans instead of one, which spreads them over the memory channels
// (0.245 -> 0.237 ms per 2^20-eval encode on one box, 0.255 -> 0.252 on
// another; rows 4 bd/16 apart were slower).  The 16
// threads of one r form one 16-lane DPP row, which lets the butterflies read
// r's twiddles from each other's registers instead of from LDS.
#include "common.h"

namespace zpi {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t octrev_d(uint32_t c, uint32_t d) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < d; i++) {
    r = (r << 3) | (c & 7u);
    c >>= 3;
  }
  return r;
}

// Exact conversion of an integer-valued double with |x| < 2^51.
__device__ __forceinline__ int64_t dbl_to_i64(double x) {
  const double magic = 6755399441055744.0;  // 1.5 * 2^52
  return __double_as_longlong(x + magic) - __double_as_longlong(magic);
}

// Stage-1 inputs y (|y| < 2^41) are split as y = 2^32 yh + yl with
// yl in [0, 2^32), so every product and partial sum of both halves is an
// integer below 2^50 in magnitude.  The accumulators are biased so that
// their bit patterns are the values they hold: a double in [2^52, 2^53) has
// ulp 1, so M + v (M = 1.5 * 2^52, |v| < 2^51) is exact and
// bits(M + v) = bits(M) + v.  The low half starts at M, the high half at
// M - K with K = bits(M) >> 32 (bits(M) has no set bit below 51).  Then,
// modulo 2^64,
//   bits(lo) + (bits(hi) << 32) = bits(M) + sl + ((sh - K) << 32) = sl + 2^32 sh,
// and since the low word of bits(M) is zero, (bits(hi) << 32) only needs the
// low word of bits(hi): the output is one 32-bit add on the high word.
constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
constexpr uint64_t kMagicBits = 0x4338000000000000ull;
constexpr double kMagicHi = kMagic - (double)(kMagicBits >> 32);
static_assert((kMagicBits & 0xffffffffull) == 0, "bias must be a multiple of 2^32");

// Plane words: planes[((((row*NC + c)*(D/16) + p/16)*4 + g)*16 + p%16)*S2 + s2]
// (S2 = ceil(S/2)) holds, for plane p, the bits the MFMA lane group g takes
// in k-steps 2 s2 and 2 s2 + 1 (k = 64 s + 16 g + i, i < 16), in "spread"
// order: bit i = 4v + b of k-step 2 s2 + h sits at position 8b + v + 4h, so
// (word >> (4h + v)) & 0x01010101 is operand register v of that k-step's
// MFMA B fragment (bytes b = 0..3) in two instructions.  Two k-steps share a
// word and a lane's words of one tile are adjacent, so the main kernel reads
// a tile with one 8-byte load: 32 KiB of plane words per workgroup instead
// of 64 KiB in 4-byte loads (0.245-0.256 -> 0.238-0.241 ms per 2^20-eval
// encode).  Each group of 16 planes has its own run of words, so the loads
// read whole cache lines.  One thread per (row, c, g, s2, 16 planes; D is a
// multiple of 16 on this path), and the chunks whose octet reversals are
// adjacent follow on neighbouring lanes, so one load instruction reads
// adjacent evals.
constexpr uint32_t kPlanesPerLane = 16;
template <int NC>
__global__ void __launch_bounds__(256)
bitplane_kernel(const uint64_t* __restrict__ evals, uint32_t* __restrict__ planes,
                uint32_t row_len, uint32_t bl, uint32_t S, uint32_t D, uint32_t depth,
                uint64_t rows) {
  constexpr uint32_t PPL = kPlanesPerLane;
  const uint32_t PQ = D / PPL, S2 = (S + 1) / 2;
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * NC * 4 * S2 * PQ) return;
  const uint32_t pq = (uint32_t)(idx % PQ);
  uint64_t rest = idx / PQ;
  const uint32_t h = (uint32_t)(rest % 8);
  rest /= 8;
  const uint32_t a = (uint32_t)(rest % (NC / 8));
  rest /= NC / 8;
  const uint32_t s2 = (uint32_t)(rest % S2);
  rest /= S2;
  const uint32_t g = (uint32_t)(rest % 4);
  const uint64_t row = rest / 4;
  // Depth 2: chunk c = a + 8h has octet reversal 8a + h.  Depth 1: c = h.
  const uint32_t c = NC == 8 ? h : a + 8 * h;
  const uint64_t* x = evals + row * row_len + octrev_d(c, depth);
  uint32_t v[PPL];
#pragma unroll
  for (int q = 0; q < (int)PPL; q++) v[q] = 0;
#pragma unroll
  for (int hh = 0; hh < 2; hh++) {
    // k-step 2 s2 + hh, lane group g: k = 64 s + 16 g + i.
    const uint32_t k0 = 64 * (2 * s2 + hh) + 16 * g;
    uint64_t xs[16];
#pragma unroll
    for (int i = 0; i < 16; i++) xs[i] = x[(uint64_t)NC * min(k0 + i, bl - 1)];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t k = k0 + i;
      const uint32_t bits = k < bl ? (uint32_t)(xs[i] >> (PPL * pq)) : 0u;
      const uint32_t pos = 8 * (i & 3) + (i >> 2) + 4 * hh;
#pragma unroll
      for (int q = 0; q < (int)PPL; q++) v[q] |= ((bits >> q) & 1u) << pos;
    }
  }
  uint32_t* dst = planes + ((((row * NC + c) * PQ + pq) * 4 + g) * PPL) * S2 + s2;
#pragma unroll
  for (int q = 0; q < (int)PPL; q++) dst[q * S2] = v[q];
}

// Radix-8 butterflies by symmetry.  Output t of a butterfly is
// sum_j x_j T(j, t) with T(j, t+4) = T(j, t) for even j and -T(j, t) for odd
// j (up to a rare -p, corrected separately), so with
//   E_t = x_0 + sum_{j even} x_j T(j, t),  O_t = sum_{j odd} x_j T(j, t)
// the outputs are E_t + O_t and E_t - O_t for t < 4: 28 products and 8 adds
// instead of 56 products.  The 28 coefficients T(j, t), t < 4, of one
// position are spread over the 16 lanes of a wave row (lanes of a row share
// the position): entry e = (j-1)*4 + t sits on lane e % 16 of register
// e / 16, and each FMA reads it with a row_newbcast DPP source
// (v_fmac_f64_dpp).  Each asm block starts with the two wait states a DPP
// read needs after a VALU write of its source.
template <int J, int T>
constexpr int sym_lane() {
  return ((J - 1) * 4 + T) & 15;
}

// a[t] += xa * T(JA, t) and b[t] += xb * T(JB, t), t < 4.
template <int JA, int JB>
__device__ __forceinline__ void pair4(double (&a)[4], double (&b)[4], double tw, double xa,
                                      double xb) {
  asm("s_nop 1\n"
      "v_fmac_f64_dpp %0, %8, %9 row_newbcast:%11 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %4, %8, %10 row_newbcast:%15 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %8, %9 row_newbcast:%12 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %5, %8, %10 row_newbcast:%16 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %2, %8, %9 row_newbcast:%13 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %6, %8, %10 row_newbcast:%17 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %3, %8, %9 row_newbcast:%14 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %7, %8, %10 row_newbcast:%18 row_mask:0xf bank_mask:0xf\n"
      : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(b[0]), "+v"(b[1]), "+v"(b[2]),
        "+v"(b[3])
      : "v"(tw), "v"(xa), "v"(xb), "i"(sym_lane<JA, 0>()), "i"(sym_lane<JA, 1>()),
        "i"(sym_lane<JA, 2>()), "i"(sym_lane<JA, 3>()), "i"(sym_lane<JB, 0>()),
        "i"(sym_lane<JB, 1>()), "i"(sym_lane<JB, 2>()), "i"(sym_lane<JB, 3>()));
}

// a[t] += x * T(J, t), t < 4.
template <int J>
__device__ __forceinline__ void single4(double (&a)[4], double tw, double x) {
  asm("s_nop 1\n"
      "v_fmac_f64_dpp %0, %4, %5 row_newbcast:%6 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %4, %5 row_newbcast:%7 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %2, %4, %5 row_newbcast:%8 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %3, %4, %5 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n"
      : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3])
      : "v"(tw), "v"(x), "i"(sym_lane<J, 0>()), "i"(sym_lane<J, 1>()), "i"(sym_lane<J, 2>()),
        "i"(sym_lane<J, 3>()));
}

// The 28 coefficients of one position (a table row of 32 at `row`): entry
// 16m + q on lane q of this lane's row.
__device__ __forceinline__ void load_spread(double (&tw)[2], const double* row, uint32_t q) {
  tw[0] = row[q];
  tw[1] = row[16 + q];
}

// out[t + 4] -= p * x_j for the odd j and t < 4 marked in `mask`
// (bit ((j-1)/2)*4 + t): the coefficient of output t + 4 there is
// -T(j, t) - p.
__device__ __forceinline__ void fix_odd(double (&o)[8], uint32_t mask, const double (&x)[8]) {
#pragma unroll
  for (int jo = 0; jo < 4; jo++)
#pragma unroll
    for (int t = 0; t < 4; t++)
      if ((mask >> (jo * 4 + t)) & 1u) o[t + 4] = __builtin_fma(-65537.0, x[2 * jo + 1], o[t + 4]);
}

// D is the plane-word layout's plane count (a multiple of 16) and Dout the
// codeword's; with kPart (Dout not a multiple of 16) the lanes of planes at
// or above Dout compute but do not store.
template <int NC, int S, bool kPart>
__global__ void __launch_bounds__(256, 2)
iprs_mfma_kernel(const uint32_t* __restrict__ planes, const v4i* __restrict__ afrag,
                 const double* __restrict__ tw0, const double* __restrict__ tw1,
                 const uint16_t* __restrict__ q0, const uint16_t* __restrict__ q1,
                 int64_t* __restrict__ out, uint32_t bd, uint32_t D, uint32_t Dout, uint32_t nrb,
                 uint32_t npg, uint64_t row0, uint32_t rmul, uint32_t rstep) {
  // Base sums [c][r_local][p], from the matrix cores to the stage threads.
  constexpr uint32_t kBsWords = NC * 16 * 16;
  __shared__ int32_t Bs[kBsWords];
  constexpr uint32_t kTiles = NC / 4;  // chunks per wave
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t g = lane >> 4, pcol = lane & 15;

  // Block -> (row, r-block, plane group).  With nrb a multiple of 8, each
  // XCD (blockIdx mod 8) keeps a fixed 1/8 of the r-blocks, so its L2 holds
  // only that part of the matrix slices and twiddles.
  uint32_t rb, pg;
  uint64_t row;
  {
    const uint64_t b = blockIdx.x;
    if ((nrb & 7) == 0) {
      const uint32_t xcd = (uint32_t)(b & 7), per = nrb >> 3;
      const uint64_t q = b >> 3;
      rb = xcd + 8 * (uint32_t)(q % per);
      const uint64_t rest = q / per;
      pg = (uint32_t)(rest % npg);
      row = rest / npg;
    } else {
      rb = (uint32_t)(b % nrb);
      const uint64_t rest = b / nrb;
      pg = (uint32_t)(rest % npg);
      row = rest / npg;
    }
  }

  // Issue every global load up front: the plane words of this wave's chunks,
  // the matrix slices and this thread's share of its r's stage-0 twiddles
  // (phase 2 has rl = tid / 16 and p = tid % 16).
  // One row's plane words span NC*W*D*4 <= 256 KiB: 32-bit offsets from the
  // (uniform) row base.
  constexpr uint32_t S2 = (S + 1) / 2;
  const char* prow = (const char*)(planes + row * NC * 4 * S2 * D);
  const uint32_t poff = ((((wave * npg + pg) * 4 + g) * 16 + pcol) * S2) * 4;
  uint32_t pw[kTiles][S2];
#pragma unroll
  for (uint32_t i = 0; i < kTiles; i++) {
    const char* src = prow + (poff + 4 * i * npg * 4 * 16 * S2 * 4);
    if constexpr (S2 == 2) {
      const uint2 w2 = *(const uint2*)src;
      pw[i][0] = w2.x;
      pw[i][1] = w2.y;
    } else {
      pw[i][0] = *(const uint32_t*)src;
    }
  }
  v4i A[S][3];
#pragma unroll
  for (int s = 0; s < S; s++)
#pragma unroll
    for (int q = 0; q < 3; q++) A[s][q] = afrag[(((uint64_t)rb * S + s) * 3 + q) * 64 + lane];
  double t0r[2];
  // Row m of r-block rb is r = rb rmul + m rstep: (1, bd/16) spreads the
  // block's rows, (16, 1) takes 16 consecutive rows (the A table matches).
  load_spread(t0r, tw0 + (uint64_t)(rb * rmul + rstep * (tid >> 4)) * 32, tid & 15);
  const uint32_t m0 = q0[rb * rmul + rstep * (tid >> 4)];
  // Keep the loads here: left to the scheduler they sink to their first
  // uses and their latency is exposed there.
  __builtin_amdgcn_sched_barrier(0);

  // Phase 1: base layer on the matrix cores.
#pragma unroll
  for (uint32_t i = 0; i < kTiles; i++) {
    const uint32_t ci = wave + 4 * i;
    v4i acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0}, acc2 = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t w = pw[i][s / 2] >> (4 * (s & 1));
      v4i bf;
      bf[0] = (int)(w & 0x01010101u);
      bf[1] = (int)((w >> 1) & 0x01010101u);
      bf[2] = (int)((w >> 2) & 0x01010101u);
      bf[3] = (int)((w >> 3) & 0x01010101u);
      acc0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[s][0], bf, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[s][1], bf, acc1, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[s][2], bf, acc2, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      // Two shift-adds per value: the empty asm keeps the compiler from
      // re-associating into two shifts and a three-way add.
      int32_t t = (acc2[v] << 7) + acc1[v];
      asm("" : "+v"(t));
      Bs[(ci * 16 + 4 * g + v) * 16 + pcol] = (t << 7) + acc0[v];
    }
  }
  __syncthreads();

  // Phase 2: every radix-8 stage of one (r, p) column per thread.
  const uint32_t rl = tid >> 4, p = tid & 15;
  const uint32_t r = rb * rmul + rstep * rl;
  const uint64_t N = (uint64_t)bd * NC;
  // One row's codewords span N*D*8 <= 32 MiB, so 32-bit element offsets from
  // the row base address every output.
  char* orow = (char*)(out + (row0 + row) * N * Dout);
  const uint32_t obase = (r * Dout + 16 * pg + p) * 8;  // byte offset of chunk 0
  const uint32_t cstride = bd * Dout * 8;               // byte stride between chunks
  const bool stores = !kPart || 16 * pg + p < Dout;

  // Stage 1's first twiddles are loaded before stage 0, which hides them.
  double t1n[2];
  uint32_t m1n = 0;
  if constexpr (NC == 64) {
    load_spread(t1n, tw1 + (uint64_t)r * 32, p);
    m1n = q1[r];
    __builtin_amdgcn_sched_barrier(0);
  }

  // Stage 0: one radix-8 butterfly per block of 8 chunks.
  double y0[NC];
#pragma unroll
  for (int j = 0; j < NC / 8; j++) {
    double b[8];
#pragma unroll
    for (int jj = 0; jj < 8; jj++) b[jj] = (double)Bs[((8 * j + jj) * 16 + rl) * 16 + p];
    double e[4], o[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      e[t] = b[0];
      o[t] = 0.0;
    }
    pair4<2, 1>(e, o, t0r[0], b[2], b[1]);
    pair4<4, 3>(e, o, t0r[0], b[4], b[3]);
    pair4<6, 5>(e, o, t0r[1], b[6], b[5]);
    single4<7>(o, t0r[1], b[7]);
    double y[8];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      y[t] = e[t] + o[t];
      y[t + 4] = e[t] - o[t];
    }
    if (__builtin_expect(m0 != 0, 0)) fix_odd(y, m0, b);
#pragma unroll
    for (int t = 0; t < 8; t++) y0[8 * j + t] = y[t];
  }

  if constexpr (NC == 8) {
    if (stores) {
#pragma unroll
      for (int c = 0; c < 8; c++)
        *(int64_t*)(orow + (obase + (uint32_t)c * cstride)) = dbl_to_i64(y0[c]);
    }
  } else {
    // Stage 1: output chunk 8t + cl from the inputs 8j + cl.  The next
    // step's twiddles are loaded before this step's stores: the memory
    // counter covers stores too, so a load issued after them would wait for
    // them as well.
    const double two32 = 4294967296.0, inv32 = 1.0 / 4294967296.0;
#pragma unroll
    for (int cl = 0; cl < 8; cl++) {
      double t1r[2];
      t1r[0] = t1n[0];
      t1r[1] = t1n[1];
      const uint32_t m1 = m1n;
      if (cl < 7) {
        load_spread(t1n, tw1 + ((uint64_t)(cl + 1) * bd + r) * 32, p);
        m1n = q1[(cl + 1) * bd + r];
        __builtin_amdgcn_sched_barrier(0);
      }
      double yh[8], yl[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const double v = y0[8 * j + cl];
        yh[j] = __builtin_floor(v * inv32);
        yl[j] = __builtin_fma(yh[j], -two32, v);
      }
      // Even and odd half-sums of the high and low halves.
      double eh[4], el[4], oh[4], ol[4];
      const double h0 = yh[0] + kMagicHi, l0 = yl[0] + kMagic;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        eh[t] = h0;
        el[t] = l0;
        oh[t] = 0.0;
        ol[t] = 0.0;
      }
      pair4<2, 2>(eh, el, t1r[0], yh[2], yl[2]);
      pair4<1, 1>(oh, ol, t1r[0], yh[1], yl[1]);
      pair4<4, 4>(eh, el, t1r[0], yh[4], yl[4]);
      pair4<3, 3>(oh, ol, t1r[0], yh[3], yl[3]);
      pair4<6, 6>(eh, el, t1r[1], yh[6], yl[6]);
      pair4<5, 5>(oh, ol, t1r[1], yh[5], yl[5]);
      pair4<7, 7>(oh, ol, t1r[1], yh[7], yl[7]);
      double sh[8], sl[8];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        sh[t] = eh[t] + oh[t];
        sh[t + 4] = eh[t] - oh[t];
        sl[t] = el[t] + ol[t];
        sl[t + 4] = el[t] - ol[t];
      }
      if (__builtin_expect(m1 != 0, 0)) {
        fix_odd(sh, m1, yh);
        fix_odd(sl, m1, yl);
      }
      uint32_t ooff = obase + (uint32_t)cl * cstride;  // output chunk 8t + cl
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const uint32_t hlo = (uint32_t)__double_as_longlong(sh[t]);
        const uint64_t lb = (uint64_t)__double_as_longlong(sl[t]);
        if (stores) *(uint2*)(orow + ooff) = make_uint2((uint32_t)lb, (uint32_t)(lb >> 32) + hlo);
        ooff += 8 * cstride;
      }
    }
  }
}

template <int NC, int S>
int launch_main(const zp_code* cd, const uint32_t* planes, uint64_t rows, uint32_t D,
                int64_t* out, hipStream_t s, bool spread) {
  const uint32_t Dp = (D + 15) / 16 * 16;  // plane-word layout
  const uint32_t nrb = cd->bd / 16, npg = Dp / 16;
  const v4i* afrag = (const v4i*)(spread ? cd->d_afrag : cd->d_afrag_c);
  const uint32_t rmul = spread ? 1u : 16u, rstep = spread ? nrb : 1u;
  const uint64_t per_row = (uint64_t)nrb * npg;
  const uint64_t max_rows = std::max<uint64_t>(1, 0x7fffffffull / per_row);
  for (uint64_t r0 = 0; r0 < rows; r0 += max_rows) {
    const uint64_t nr = std::min(max_rows, rows - r0);
    const uint32_t* pl = planes + r0 * NC * 4 * ((S + 1) / 2) * Dp;
    if (D % 16 == 0)
      hipLaunchKernelGGL((iprs_mfma_kernel<NC, S, false>), dim3((uint32_t)(nr * per_row)),
                         dim3(256), 0, s, pl, afrag, cd->d_tw0f, cd->d_tw1f, cd->d_q0f,
                         cd->d_q1f, out, cd->bd, Dp, D, nrb, npg, r0, rmul, rstep);
    else
      hipLaunchKernelGGL((iprs_mfma_kernel<NC, S, true>), dim3((uint32_t)(nr * per_row)),
                         dim3(256), 0, s, pl, afrag, cd->d_tw0f, cd->d_tw1f, cd->d_q0f,
                         cd->d_q1f, out, cd->bd, Dp, D, nrb, npg, r0, rmul, rstep);
    ZP_HIP(hipGetLastError());
  }
  return ZP_OK;
}

template <int NC>
int launch_nc(const zp_code* cd, const uint64_t* evals, uint64_t rows, uint32_t D, int64_t* out,
              hipStream_t s, bool spread, hipEvent_t mid) {
  const uint32_t S = cd->mfma_ksteps, S2 = (S + 1) / 2;
  const uint32_t Dp = (D + 15) / 16 * 16;  // planes above D are computed, not stored
  const uint64_t bytes = rows * NC * 4 * S2 * Dp * sizeof(uint32_t);
  uint32_t* planes = nullptr;
  ZP_TRY(dev_alloc((void**)&planes, bytes, s));
  const uint64_t threads = rows * NC * 4 * S2 * (Dp / kPlanesPerLane);
  hipLaunchKernelGGL((bitplane_kernel<NC>), dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0,
                     s, evals, planes, (uint32_t)cd->row_len, cd->bl, S, Dp, cd->depth, rows);
  int rc = hipGetLastError() == hipSuccess ? ZP_OK : fail(ZP_ERR_DEVICE, "bitplane launch");
  if (rc == ZP_OK && mid && hipEventRecord(mid, s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "event record failed");
  if (rc == ZP_OK) {
    switch (S) {
      case 1: rc = launch_main<NC, 1>(cd, planes, rows, D, out, s, spread); break;
      case 2: rc = launch_main<NC, 2>(cd, planes, rows, D, out, s, spread); break;
      case 3: rc = launch_main<NC, 3>(cd, planes, rows, D, out, s, spread); break;
      case 4: rc = launch_main<NC, 4>(cd, planes, rows, D, out, s, spread); break;
      default: rc = fail(ZP_ERR_INVALID_PARAM, "unsupported base length for the MFMA encoder");
    }
  }
  dev_free(planes, s);
  return rc;
}

}  // namespace

uint64_t iprs_mfma_scratch_bytes(const zp_code* cd, uint64_t rows, uint32_t D) {
  const uint64_t NC = cd->depth == 1 ? 8 : 64;
  return rows * NC * 4 * ((cd->mfma_ksteps + 1) / 2) * ((D + 15) / 16 * 16) * sizeof(uint32_t);
}

bool iprs_mfma_eligible(const zp_code* cd, uint32_t D) {
  return cd->d_afrag != nullptr && D >= 1 && D <= 64 && (cd->depth == 1 || cd->depth == 2) &&
         cd->row_len < (1ull << 31);
}

int launch_iprs_binary_mfma(const zp_code* cd, const uint64_t* evals, uint64_t rows, uint32_t D,
                            int64_t* out, hipStream_t s, bool spread, hipEvent_t mid) {
  if (rows == 0) {
    if (mid && hipEventRecord(mid, s) != hipSuccess)
      return fail(ZP_ERR_DEVICE, "event record failed");
    return ZP_OK;
  }
  if (cd->depth == 1) return launch_nc<8>(cd, evals, rows, D, out, s, spread, mid);
  return launch_nc<64>(cd, evals, rows, D, out, s, spread, mid);
}

}  // namespace zpi
