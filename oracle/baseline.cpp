// baseline.cpp -- TEST INFRASTRUCTURE ONLY: the timed CPU baseline.
//
// The reference's CPU commit of BinaryPoly<D> under IPRS, organised the way
// its benches run it (README.md:35-37: --features "simd parallel unchecked",
// RUSTFLAGS="-C target-cpu=native"):
//   * base layer: one fold per codeword index, parallel over the indices
//     (zip-plus/src/code/iprs/pntt/radix8.rs:117-161, cfg_into_iter over
//     0..codeword_len), each term the AVX-512 widening of the eval bits
//     times the matrix entry (poly/src/univariate/binary_u64.rs:445-483,
//     605-661: _mm512_maskz_set1_epi64 per 8 coefficients) added to the
//     DensePolynomial<i64, D> accumulator;
//   * combine stages: parallel over blocks of 8 sub-chunks (radix8.rs:
//     58-113), i64 products per coefficient (vpmullq);
//   * rows of a poly in parallel (pcs/phase_commit.rs:159-176);
//   * leaves: parallel over columns, each column's bytes copied into one
//     buffer and hashed in one update (merkle.rs:128-168), the full chunks 16
//     at a time across AVX-512 lanes as the blake3 crate's hash_many does;
//   * Merkle layers parallel over pairs (merkle.rs:172-221).
// The same .so is built here and run on other hosts, so the AVX-512 code is
// selected at run time (cpuid) instead of by -march=native; without AVX-512
// the portable loops run.  Results equal or_commit's bit for bit
// (tests/test_baseline.py); only tests/ and bench.py's cpu_baseline call it.
#include <immintrin.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

extern "C" int or_iprs_params(uint64_t row_len, int64_t depth, uint64_t rep, uint64_t* out6);
extern "C" int or_iprs_base_matrix(uint64_t row_len, uint64_t depth, uint64_t rep, int64_t* out);
extern "C" int or_iprs_twiddles(uint64_t row_len, uint64_t depth, uint64_t rep, uint64_t stage,
                                int64_t* out);
extern "C" int or_blake3(const uint8_t* data, uint64_t len, uint8_t* out32);
extern "C" int or_blake3_parent(const uint8_t* l, const uint8_t* r, int root, uint8_t* out32);
extern "C" int or_blake3_chunk_cv(const uint8_t* data, uint64_t len, uint64_t chunk_idx,
                                  uint8_t* out32);

namespace {

#define AVX512 __attribute__((target("avx512f,avx512dq,avx512bw,avx512vl")))

template <class F>
void pfor(int T, uint64_t n, F fn) {
  if (T <= 1 || n < 2) {
    for (uint64_t i = 0; i < n; i++) fn(i);
    return;
  }
  if ((uint64_t)T > n) T = (int)n;
  std::atomic<uint64_t> next(0);
  const uint64_t grain = std::max<uint64_t>(1, n / ((uint64_t)T * 16));
  auto worker = [&]() {
    for (;;) {
      const uint64_t s = next.fetch_add(grain);
      if (s >= n) break;
      const uint64_t e = std::min(n, s + grain);
      for (uint64_t i = s; i < e; i++) fn(i);
    }
  };
  std::vector<std::thread> ts;
  for (int t = 1; t < T; t++) ts.emplace_back(worker);
  worker();
  for (auto& t : ts) t.join();
}

uint64_t octrev(uint64_t c, uint64_t d) {
  uint64_t r = 0;
  for (uint64_t i = 0; i < d; i++) {
    r = (r << 3) | (c & 7);
    c >>= 3;
  }
  return r;
}

// butterfly.rs:8-17: the twiddle column of output r for input j (j >= 1).
const int kButterfly[8][7] = {{0, 0, 0, 0, 0, 0, 0}, {1, 2, 3, 4, 5, 6, 7},
                              {2, 4, 6, 0, 2, 4, 6}, {3, 6, 1, 4, 7, 2, 5},
                              {4, 0, 4, 0, 4, 0, 4}, {5, 2, 7, 4, 1, 6, 3},
                              {6, 4, 2, 0, 6, 4, 2}, {7, 6, 5, 4, 3, 2, 1}};

struct Params {
  uint64_t row_len, N, d, bl, bd, bd_log2;
  std::vector<int64_t> M;                // bd x bl
  std::vector<std::vector<int64_t>> tw;  // per stage: (bd 8^k) x 7 x 8
};

int load_params(uint64_t row_len, int64_t depth, uint64_t rep, Params& p) {
  uint64_t v[6];
  int rc = or_iprs_params(row_len, depth, rep, v);
  if (rc) return rc;
  p.row_len = v[0], p.N = v[1], p.d = v[2], p.bl = v[3], p.bd = v[4], p.bd_log2 = v[5];
  p.M.resize(p.bd * p.bl);
  rc = or_iprs_base_matrix(row_len, p.d, rep, p.M.data());
  if (rc) return rc;
  p.tw.resize(p.d);
  for (uint64_t k = 0; k < p.d; k++) {
    p.tw[k].resize((p.bd << (3 * k)) * 56);
    rc = or_iprs_twiddles(row_len, p.d, rep, k, p.tw[k].data());
    if (rc) return rc;
  }
  return 0;
}

// ---- base layer -----------------------------------------------------------

void base_scalar(const Params& p, uint64_t D, const uint64_t* in, int64_t* out, uint64_t i) {
  const uint64_t o = octrev(i >> p.bd_log2, p.d), row = i & (p.bd - 1), sh = 3 * p.d;
  int64_t* acc = out + i * D;
  const uint64_t x0 = in[o];
  for (uint64_t c = 0; c < D; c++) acc[c] = (int64_t)((x0 >> c) & 1);
  const int64_t* mrow = &p.M[row * p.bl];
  for (uint64_t col = 1; col < p.bl; col++) {
    const uint64_t x = in[o | (col << sh)];
    const uint64_t m = (uint64_t)mrow[col];
    for (uint64_t c = 0; c < D; c++) acc[c] = (int64_t)((uint64_t)acc[c] + (((x >> c) & 1) ? m : 0));
  }
}

template <int NV>
AVX512 void base_avx512(const Params& p, const uint64_t* in, int64_t* out, uint64_t i) {
  const uint64_t o = octrev(i >> p.bd_log2, p.d), row = i & (p.bd - 1), sh = 3 * p.d;
  __m512i acc[NV];
  const uint64_t x0 = in[o];
  for (int v = 0; v < NV; v++) acc[v] = _mm512_maskz_set1_epi64((__mmask8)(x0 >> (8 * v)), 1);
  const int64_t* mrow = &p.M[row * p.bl];
  // widen(x, m) + acc: the widening's masked broadcast folded into a masked
  // add (the same sum, one instruction per 8 coefficients).
  for (uint64_t col = 1; col < p.bl; col++) {
    const uint64_t x = in[o | (col << sh)];
    const __m512i mb = _mm512_set1_epi64(mrow[col]);
    for (int v = 0; v < NV; v++)
      acc[v] = _mm512_mask_add_epi64(acc[v], (__mmask8)(x >> (8 * v)), acc[v], mb);
  }
  for (int v = 0; v < NV; v++) _mm512_storeu_si512((void*)(out + i * 8 * NV + 8 * v), acc[v]);
}

// ---- combine stages -------------------------------------------------------

void stage_scalar(const Params& p, uint64_t D, int64_t* out, uint64_t k, uint64_t blk) {
  const uint64_t sub = p.bd << (3 * k);
  int64_t* chunk = out + blk * 8 * sub * D;
  std::vector<int64_t> xs(8 * D);
  for (uint64_t i = 0; i < sub; i++) {
    for (uint64_t j = 0; j < 8; j++) std::memcpy(&xs[j * D], chunk + (j * sub + i) * D, D * 8);
    const int64_t* twi = &p.tw[k][i * 56];
    for (int r = 0; r < 8; r++) {
      int64_t* y = chunk + (r * sub + i) * D;
      for (uint64_t c = 0; c < D; c++) {
        uint64_t a = (uint64_t)xs[c];
        for (int j = 0; j < 7; j++)
          a += (uint64_t)xs[(j + 1) * D + c] * (uint64_t)twi[j * 8 + kButterfly[r][j]];
        y[c] = (int64_t)a;
      }
    }
  }
}

template <int NV>
AVX512 void stage_avx512(const Params& p, int64_t* out, uint64_t k, uint64_t blk) {
  constexpr uint64_t D = 8 * NV;
  const uint64_t sub = p.bd << (3 * k);
  int64_t* chunk = out + blk * 8 * sub * D;
  for (uint64_t i = 0; i < sub; i++) {
    __m512i xs[8][NV];
    for (int j = 0; j < 8; j++)
      for (int v = 0; v < NV; v++)
        xs[j][v] = _mm512_loadu_si512((const void*)(chunk + (j * sub + i) * D + 8 * v));
    const int64_t* twi = &p.tw[k][i * 56];
    for (int r = 0; r < 8; r++) {
      for (int v = 0; v < NV; v++) {
        __m512i a = xs[0][v];
        for (int j = 0; j < 7; j++)
          a = _mm512_add_epi64(
              a, _mm512_mullo_epi64(xs[j + 1][v], _mm512_set1_epi64(twi[j * 8 + kButterfly[r][j]])));
        _mm512_storeu_si512((void*)(chunk + (r * sub + i) * D + 8 * v), a);
      }
    }
  }
}

// ---- BLAKE3, 16 chunks across the AVX-512 lanes -----------------------------

const uint32_t kIV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                         0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
const int kPerm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};

AVX512 inline void g16(__m512i& a, __m512i& b, __m512i& c, __m512i& d, __m512i x, __m512i y) {
  a = _mm512_add_epi32(_mm512_add_epi32(a, b), x);
  d = _mm512_ror_epi32(_mm512_xor_si512(d, a), 16);
  c = _mm512_add_epi32(c, d);
  b = _mm512_ror_epi32(_mm512_xor_si512(b, c), 12);
  a = _mm512_add_epi32(_mm512_add_epi32(a, b), y);
  d = _mm512_ror_epi32(_mm512_xor_si512(d, a), 8);
  c = _mm512_add_epi32(c, d);
  b = _mm512_ror_epi32(_mm512_xor_si512(b, c), 7);
}

// Chaining values of the 16 full chunks at data + 1024 l (chunk counters
// c0 + l), out[l * 8 + i].
AVX512 void chunks16(const uint8_t* data, uint64_t c0, uint32_t* out) {
  const __m512i off = _mm512_mullo_epi32(
      _mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15), _mm512_set1_epi32(1024));
  __m512i cv[8];
  for (int i = 0; i < 8; i++) cv[i] = _mm512_set1_epi32((int)kIV[i]);
  alignas(64) uint32_t clo[16], chi[16];
  for (int l = 0; l < 16; l++) {
    clo[l] = (uint32_t)(c0 + l);
    chi[l] = (uint32_t)((c0 + l) >> 32);
  }
  const __m512i ctr_lo = _mm512_load_si512(clo), ctr_hi = _mm512_load_si512(chi);
  for (int blk = 0; blk < 16; blk++) {
    __m512i m[16];
    for (int w = 0; w < 16; w++)
      m[w] = _mm512_i32gather_epi32(off, (const void*)(data + 64 * blk + 4 * w), 1);
    const uint32_t flags = (blk == 0 ? 1u : 0u) | (blk == 15 ? 2u : 0u);
    __m512i v[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                     _mm512_set1_epi32((int)kIV[0]), _mm512_set1_epi32((int)kIV[1]),
                     _mm512_set1_epi32((int)kIV[2]), _mm512_set1_epi32((int)kIV[3]),
                     ctr_lo, ctr_hi, _mm512_set1_epi32(64), _mm512_set1_epi32((int)flags)};
    for (int r = 0; r < 7; r++) {
      g16(v[0], v[4], v[8], v[12], m[0], m[1]);
      g16(v[1], v[5], v[9], v[13], m[2], m[3]);
      g16(v[2], v[6], v[10], v[14], m[4], m[5]);
      g16(v[3], v[7], v[11], v[15], m[6], m[7]);
      g16(v[0], v[5], v[10], v[15], m[8], m[9]);
      g16(v[1], v[6], v[11], v[12], m[10], m[11]);
      g16(v[2], v[7], v[8], v[13], m[12], m[13]);
      g16(v[3], v[4], v[9], v[14], m[14], m[15]);
      if (r < 6) {
        __m512i t[16];
        for (int i = 0; i < 16; i++) t[i] = m[kPerm[i]];
        for (int i = 0; i < 16; i++) m[i] = t[i];
      }
    }
    for (int i = 0; i < 8; i++) cv[i] = _mm512_xor_si512(v[i], v[i + 8]);
  }
  alignas(64) uint32_t tmp[8][16];
  for (int i = 0; i < 8; i++) _mm512_store_si512(tmp[i], cv[i]);
  for (int l = 0; l < 16; l++)
    for (int i = 0; i < 8; i++) out[l * 8 + i] = tmp[i][l];
}

// BLAKE3 tree over n chunk chaining values (the left subtree holds the
// largest power of two of chunks below n; ROOT on the top merge).
void merge_tree(const uint32_t* cvs, uint64_t n, bool root, uint32_t* out) {
  if (n == 1) {
    std::memcpy(out, cvs, 32);
    return;
  }
  uint64_t left = 1;
  while (left * 2 < n) left *= 2;
  uint32_t l[8], r[8];
  merge_tree(cvs, left, false, l);
  merge_tree(cvs + 8 * left, n - left, false, r);
  or_blake3_parent((const uint8_t*)l, (const uint8_t*)r, root ? 1 : 0, (uint8_t*)out);
}

// BLAKE3 of `len` bytes (one Hasher::update).
void hash_buf(const uint8_t* data, uint64_t len, bool simd, std::vector<uint32_t>& cvs,
              uint8_t* out) {
  const uint64_t nchunks = len == 0 ? 1 : (len + 1023) / 1024;
  if (nchunks == 1) {
    or_blake3(data, len, out);
    return;
  }
  cvs.resize(nchunks * 8);
  uint64_t c = 0;
  if (simd)
    for (; c + 16 <= nchunks && (c + 16) * 1024 <= len; c += 16) chunks16(data + c * 1024, c, &cvs[c * 8]);
  for (; c < nchunks; c++)
    or_blake3_chunk_cv(data + c * 1024, std::min<uint64_t>(1024, len - c * 1024), c,
                       (uint8_t*)&cvs[c * 8]);
  merge_tree(cvs.data(), nchunks, true, (uint32_t*)out);
}

}  // namespace

// ZipPlus::commit of one BinaryPoly<D> poly under IPRS (phase_commit.rs:
// 61-101): encode_rows, then the column Merkle tree.  threads: worker count;
// simd: 1 = AVX-512 where the CPU has it (D a multiple of 8), 0 = portable
// loops.  cw_out (num_rows x N x D i64) may be null.  Returns 0, or the
// or_iprs_params error.
extern "C" int or_baseline_commit_iprs_binary(uint64_t row_len, int64_t depth, uint64_t rep,
                                              uint64_t D, const uint64_t* evals,
                                              uint64_t num_rows, int threads, int simd,
                                              int64_t* cw_out, uint8_t* root_out) {
  if (D == 0 || D > 64 || num_rows == 0) return 3;
  Params p;
  int rc = load_params(row_len, depth, rep, p);
  if (rc) return rc;
  const bool avx = simd && (D % 8 == 0) && __builtin_cpu_supports("avx512f") &&
                   __builtin_cpu_supports("avx512dq");
  const uint64_t N = p.N, elem = 8 * D;
  std::vector<int64_t> own;
  int64_t* cw = cw_out;
  if (!cw) {
    own.resize(num_rows * N * D);
    cw = own.data();
  }
  // Base layer, parallel over (row, codeword index).
  pfor(threads, num_rows * N, [&](uint64_t t) {
    const uint64_t r = t / N, i = t % N;
    const uint64_t* in = evals + r * row_len;
    int64_t* out = cw + r * N * D;
    if (avx) {
      if (D == 64) base_avx512<8>(p, in, out, i);
      else if (D == 32) base_avx512<4>(p, in, out, i);
      else if (D == 16) base_avx512<2>(p, in, out, i);
      else if (D == 8) base_avx512<1>(p, in, out, i);
      else base_scalar(p, D, in, out, i);
    } else {
      base_scalar(p, D, in, out, i);
    }
  });
  // Combine stages, parallel over (row, block of 8 sub-chunks).
  for (uint64_t k = 0; k < p.d; k++) {
    const uint64_t nblk = N / (8 * (p.bd << (3 * k)));
    pfor(threads, num_rows * nblk, [&](uint64_t t) {
      int64_t* out = cw + (t / nblk) * N * D;
      const uint64_t blk = t % nblk;
      if (avx && D == 32) stage_avx512<4>(p, out, k, blk);
      else if (avx && D == 64) stage_avx512<8>(p, out, k, blk);
      else stage_scalar(p, D, out, k, blk);
    });
  }
  // Leaves: each column copied into one buffer and hashed.
  std::vector<uint8_t> layers((2 * N - 1) * 32);
  const uint64_t col_bytes = num_rows * elem;
  pfor(threads, N, [&](uint64_t j) {
    thread_local std::vector<uint8_t> buf;
    thread_local std::vector<uint32_t> cvs;
    buf.resize(col_bytes);
    for (uint64_t r = 0; r < num_rows; r++)
      std::memcpy(&buf[r * elem], (const uint8_t*)(cw + (r * N + j) * D), elem);
    hash_buf(buf.data(), col_bytes, avx, cvs, &layers[32 * j]);
  });
  // Layers, parallel over pairs.
  uint64_t off = 0, cnt = N;
  while (cnt > 1) {
    const bool root = cnt == 2;
    const uint8_t* prev = &layers[32 * off];
    uint8_t* cur = &layers[32 * (off + cnt)];
    pfor(threads, cnt / 2, [&](uint64_t m) {
      or_blake3_parent(prev + 64 * m, prev + 64 * m + 32, root ? 1 : 0, cur + 32 * m);
    });
    off += cnt;
    cnt /= 2;
  }
  std::memcpy(root_out, &layers[32 * off], 32);
  return 0;
}

// 1 when the AVX-512 code paths run on this CPU.
extern "C" int or_baseline_has_avx512(void) {
  return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq") ? 1 : 0;
}
