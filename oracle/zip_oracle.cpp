// zip_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's Zip/Zip+ commitment path
// (NethermindEth/zinc-plus, Rust).  It is the parity checker for the HIP
// kernels and the timed CPU baseline ("port").  Only tests/, the smoke check
// and bench.py's cpu_baseline leg load it; the product library never does.
//
// Every function cites the reference file:line it restates.  The reference
// cannot be compiled here (Rust, unvendored crates), so third-party pieces are
// restated from their public algorithms:
//   * ark-ff/ark-poly 0.5 F65537 domain: omega_n = 3^(65536/n) mod 65537
//     (MontConfig generator 3, two-adicity 16), elements in natural order.
//   * blake3 1.8: BLAKE3 spec (compression, chunking, hazmat parent merges).
//   * rand 0.9 StdRng::seed_from_u64 + SliceRandom::shuffle: PCG32 seed fill,
//     ChaCha12 keystream, IncreasingUniform + Canon's bounded u32 sampling.
//   * crypto-bigint Int<L>: two's complement over L little-endian u64 limbs.
//
// Arithmetic is wrapping two's complement (the reference's `unchecked`
// feature); no value on the path can overflow at supported sizes, so the
// CHECKED and unchecked reference builds agree with it.

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

typedef __int128 i128;
typedef unsigned __int128 u128;

enum {
  OR_OK = 0,
  OR_ERR_INVALID_PARAM = 1,
  OR_ERR_OVERFLOW = 2,
  OR_ERR_SHAPE = 3,
};

// ---------------------------------------------------------------------------
// Threading: a plain static split of [0, n) over T threads (mirrors the rayon
// par_iter axes of the reference; utils/src/parallel.rs:13-172).
// ---------------------------------------------------------------------------
static int g_threads = 1;

extern "C" int or_set_threads(int n) {
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  if (n <= 0) n = 1;
  g_threads = n;
  return g_threads;
}

template <class F>
static void parallel_for(uint64_t n, F fn) {
  int T = g_threads;
  if (T <= 1 || n < 2) {
    for (uint64_t i = 0; i < n; i++) fn(i);
    return;
  }
  if ((uint64_t)T > n) T = (int)n;
  std::atomic<uint64_t> next(0);
  const uint64_t grain = std::max<uint64_t>(1, n / ((uint64_t)T * 8));
  auto worker = [&]() {
    for (;;) {
      uint64_t s = next.fetch_add(grain);
      if (s >= n) break;
      uint64_t e = std::min(n, s + grain);
      for (uint64_t i = s; i < e; i++) fn(i);
    }
  };
  std::vector<std::thread> ts;
  for (int t = 1; t < T; t++) ts.emplace_back(worker);
  worker();
  for (auto& t : ts) t.join();
}

// ---------------------------------------------------------------------------
// F65537 parameters of the radix-8 pseudo NTT.
// zip-plus/src/code/iprs/pntt/radix8/params.rs
// ---------------------------------------------------------------------------
static const uint64_t P = 65537;
// params.rs:220 -- the 8th roots of unity lifted to integers.
static const int64_t BASE_TWIDDLES[8] = {1, 4096, -256, 16, -1, -4096, 256, -16};

static uint64_t powmod(uint64_t b, uint64_t e) {
  uint64_t r = 1;
  b %= P;
  while (e) {
    if (e & 1) r = r * b % P;
    b = b * b % P;
    e >>= 1;
  }
  return r;
}

// params.rs:176-181 normalize_field_element: x >= (p-1)/2 ? x - p : x.
// Note the asymmetry: 32768 -> -32769 but 32769 -> -32768.
static inline int64_t norm(uint64_t x) {
  int64_t v = (int64_t)x;
  return v >= (int64_t)((P - 1) / 2) ? v - (int64_t)P : v;
}

// params.rs:193-195 to_positive_mod_repr.
static inline uint64_t to_pos(int64_t v) {
  int64_t m = v % (int64_t)P;
  if (m < 0) m += (int64_t)P;
  return (uint64_t)m;
}

// ark-poly Radix2EvaluationDomain::new(n).group_gen for F65537 with generator
// 3 and two-adicity 16: omega_n = 3^(2^16 / n).
static uint64_t omega(uint64_t n) { return powmod(3, 65536 / n); }

struct PnttParams {
  uint64_t row_len, codeword_len, depth, base_len, base_dim, base_dim_log2;
  std::vector<int64_t> base_matrix;            // [base_dim][base_len]
  std::vector<std::vector<int64_t>> twiddles;  // stage k: [sub][7][8]
};

static bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

// params.rs:59-103 Radix8PnttParams::new, with precompute at :114-171.
static int pntt_params(uint64_t row_len, uint64_t depth, uint64_t rep, PnttParams& p,
                       bool with_tables = true) {
  const uint64_t codeword_len = row_len * rep;
  if (codeword_len >= P) return OR_ERR_INVALID_PARAM;           // :60-65
  if (3 * depth >= 64) return OR_ERR_INVALID_PARAM;
  const uint64_t coeff = 1ull << (3 * depth);
  if (row_len % coeff != 0) return OR_ERR_INVALID_PARAM;        // :67-73
  const uint64_t base_len = row_len / coeff;
  const uint64_t base_dim = base_len * rep;
  if (!is_pow2(base_dim)) return OR_ERR_INVALID_PARAM;          // :75-80
  p.row_len = row_len;
  p.codeword_len = codeword_len;
  p.depth = depth;
  p.base_len = base_len;
  p.base_dim = base_dim;
  p.base_dim_log2 = (uint64_t)__builtin_ctzll(base_dim);
  if (!with_tables) return OR_OK;
  // :145-160 precompute_base_matrix: row r is domain element w^r raised to
  // the column index, normalized.
  p.base_matrix.assign(base_dim * base_len, 0);
  const uint64_t wb = omega(base_dim);
  for (uint64_t r = 0; r < base_dim; r++) {
    const uint64_t root = powmod(wb, r);
    uint64_t acc = 1;
    for (uint64_t k = 0; k < base_len; k++) {
      p.base_matrix[r * base_len + k] = norm(acc);
      acc = acc * root % P;
    }
  }
  // :163-171 precompute_roots_of_unity(codeword_len), normalized.
  std::vector<int64_t> roots(codeword_len);
  {
    const uint64_t w = omega(codeword_len);
    uint64_t acc = 1;
    for (uint64_t i = 0; i < codeword_len; i++) {
      roots[i] = norm(acc);
      acc = acc * w % P;
    }
  }
  // :114-143 precompute_butterfly_twiddles with :184-190
  // mul_and_normalize_twiddle.
  p.twiddles.assign(depth, {});
  for (uint64_t k = 0; k < depth; k++) {
    const uint64_t sub = base_dim << (3 * k);
    const uint64_t cpp = 1ull << (3 * (depth - 1 - k));
    auto& tw = p.twiddles[k];
    tw.assign(sub * 56, 0);
    for (uint64_t i = 0; i < sub; i++) {
      for (uint64_t jm1 = 0; jm1 < 7; jm1++) {
        const int64_t root = roots[cpp * i * (jm1 + 1)];
        for (uint64_t t = 0; t < 8; t++) {
          const uint64_t prod = to_pos(BASE_TWIDDLES[t]) * to_pos(root) % P;
          tw[(i * 7 + jm1) * 8 + t] = norm(prod);
        }
      }
    }
  }
  return OR_OK;
}

// zip-plus/src/code/iprs.rs:43-51 new_with_optimal_depth.
static uint64_t optimal_depth(uint64_t row_len) {
  uint64_t q = row_len / 256;
  if (q < 1) q = 1;
  uint64_t lg = 63 - (uint64_t)__builtin_clzll(q);
  uint64_t d = (lg + 2) / 3;
  return d < 1 ? 1 : d;
}

// octet_reversal.rs:6-29.
static int octet_reversal(uint64_t x, uint64_t num_octets, uint64_t* out) {
  if (num_octets == 0) {
    if (x != 0) return OR_ERR_INVALID_PARAM;  // :8-11 assert
    *out = 0;
    return OR_OK;
  }
  uint64_t result = 0;
  uint64_t left_mask = 7, right_mask = 7ull << (3 * (num_octets - 1));
  for (uint64_t i = 0; i < num_octets / 2 + num_octets % 2; i++) {
    const uint64_t lo = x & left_mask, hi = x & right_mask;
    result |= lo << (3 * (num_octets - 1 - 2 * i));
    result |= hi >> (3 * (num_octets - 1 - 2 * i));
    left_mask <<= 3;
    right_mask >>= 3;
  }
  *out = result;
  return OR_OK;
}

static uint64_t octrev(uint64_t x, uint64_t d) {
  uint64_t o = 0;
  octet_reversal(x, d, &o);
  return o;
}

extern "C" int or_octet_reversal(uint64_t x, uint64_t num_octets, uint64_t* out) {
  return octet_reversal(x, num_octets, out);
}

extern "C" uint64_t or_iprs_optimal_depth(uint64_t row_len) { return optimal_depth(row_len); }

// out6 = {row_len, codeword_len, depth, base_len, base_dim, base_dim_log2}
extern "C" int or_iprs_params(uint64_t row_len, int64_t depth, uint64_t rep, uint64_t* out6) {
  PnttParams p;
  const uint64_t d = depth < 0 ? optimal_depth(row_len) : (uint64_t)depth;
  int rc = pntt_params(row_len, d, rep, p, false);
  if (rc) return rc;
  out6[0] = p.row_len;
  out6[1] = p.codeword_len;
  out6[2] = p.depth;
  out6[3] = p.base_len;
  out6[4] = p.base_dim;
  out6[5] = p.base_dim_log2;
  return OR_OK;
}

extern "C" int or_iprs_base_matrix(uint64_t row_len, uint64_t depth, uint64_t rep, int64_t* out) {
  PnttParams p;
  int rc = pntt_params(row_len, depth, rep, p);
  if (rc) return rc;
  std::memcpy(out, p.base_matrix.data(), p.base_matrix.size() * 8);
  return OR_OK;
}

extern "C" int or_iprs_twiddles(uint64_t row_len, uint64_t depth, uint64_t rep, uint64_t stage,
                                int64_t* out) {
  PnttParams p;
  int rc = pntt_params(row_len, depth, rep, p);
  if (rc) return rc;
  if (stage >= depth) return OR_ERR_INVALID_PARAM;
  std::memcpy(out, p.twiddles[stage].data(), p.twiddles[stage].size() * 8);
  return OR_OK;
}

extern "C" int or_roots_of_unity(uint64_t n, int64_t* out) {
  if (!is_pow2(n) || n > 65536) return OR_ERR_INVALID_PARAM;
  const uint64_t w = omega(n);
  uint64_t acc = 1;
  for (uint64_t i = 0; i < n; i++) {
    out[i] = norm(acc);
    acc = acc * w % P;
  }
  return OR_OK;
}

// butterfly.rs:8-17.
static const int BUTTERFLY_TABLE[8][7] = {
    {0, 0, 0, 0, 0, 0, 0}, {1, 2, 3, 4, 5, 6, 7}, {2, 4, 6, 0, 2, 4, 6}, {3, 6, 1, 4, 7, 2, 5},
    {4, 0, 4, 0, 4, 0, 4}, {5, 2, 7, 4, 1, 6, 3}, {6, 4, 2, 0, 6, 4, 2}, {7, 6, 5, 4, 3, 2, 1},
};

// ---------------------------------------------------------------------------
// pNTT over D coefficient planes of bit-packed binary polynomials:
// In = BinaryU64Poly<D> (u64, bit c = coefficient c), Out = [i64; D].
// radix8.rs:21-52 pntt, :117-161 base_multiply_into_output (widen:
// poly/src/univariate/binary_u64.rs:445-483, bit ? scalar : 0 per
// coefficient; from_ref: poly/src/univariate/dense.rs:508-525), :58-113
// combine_stages with butterfly.rs:24-43 and the per-coefficient i64 product
// of DensePolynomial<i64>::mul_by_scalar (utils/src/mul_by_scalar.rs:12-29).
// ---------------------------------------------------------------------------
static void pntt_binary_row(const PnttParams& p, uint64_t D, const uint64_t* in, int64_t* out) {
  const uint64_t N = p.codeword_len, bd = p.base_dim, bl = p.base_len, d = p.depth;
  const uint64_t shift = 3 * d;
  // Base layer (:133-160).
  for (uint64_t i = 0; i < N; i++) {
    const uint64_t chunk = i >> p.base_dim_log2, row = i & (bd - 1);
    const uint64_t o = octrev(chunk, d);
    int64_t* acc = out + i * D;
    const uint64_t x0 = in[o];
    for (uint64_t c = 0; c < D; c++) acc[c] = (int64_t)((x0 >> c) & 1);  // from_ref
    const int64_t* mrow = &p.base_matrix[row * bl];
    for (uint64_t col = 1; col < bl; col++) {
      const uint64_t x = in[o | (col << shift)];
      const int64_t m = mrow[col];
      for (uint64_t c = 0; c < D; c++)
        acc[c] = (int64_t)((uint64_t)acc[c] + (((x >> c) & 1) ? (uint64_t)m : 0));
    }
  }
  // Combine stages (:66-112).
  std::vector<int64_t> xs(8 * D);
  for (uint64_t k = 0; k < d; k++) {
    const uint64_t sub = bd << (3 * k);
    const int64_t* tw = p.twiddles[k].data();
    for (uint64_t blk = 0; blk < N; blk += 8 * sub) {
      int64_t* chunk = out + blk * D;
      for (uint64_t i = 0; i < sub; i++) {
        for (uint64_t j = 0; j < 8; j++)
          std::memcpy(&xs[j * D], chunk + (j * sub + i) * D, D * 8);
        const int64_t* twi = tw + i * 56;
        for (uint64_t r = 0; r < 8; r++) {
          int64_t* y = chunk + (r * sub + i) * D;
          for (uint64_t c = 0; c < D; c++) {
            uint64_t a = (uint64_t)xs[c];
            for (uint64_t jm1 = 0; jm1 < 7; jm1++) {
              const int64_t t = twi[jm1 * 8 + BUTTERFLY_TABLE[r][jm1]];
              a += (uint64_t)xs[(jm1 + 1) * D + c] * (uint64_t)t;
            }
            y[c] = (int64_t)a;
          }
        }
      }
    }
  }
}

// pNTT for Eval = i32, Cw = i128 (the plain Zip types of
// zip-plus/benches/zip_benches.rs:25-40): i32 x i64 -> i128
// (utils/src/mul_by_scalar.rs:89-95), i128 x i64 (:97-108), from_ref
// i32 -> i128 (utils/src/from_ref.rs:35).
static void pntt_i32_row(const PnttParams& p, const int32_t* in, i128* out) {
  const uint64_t N = p.codeword_len, bd = p.base_dim, bl = p.base_len, d = p.depth;
  const uint64_t shift = 3 * d;
  for (uint64_t i = 0; i < N; i++) {
    const uint64_t chunk = i >> p.base_dim_log2, row = i & (bd - 1);
    const uint64_t o = octrev(chunk, d);
    i128 acc = (i128)in[o];
    const int64_t* mrow = &p.base_matrix[row * bl];
    for (uint64_t col = 1; col < bl; col++)
      acc = (i128)((u128)acc + (u128)((i128)in[o | (col << shift)] * (i128)mrow[col]));
    out[i] = acc;
  }
  for (uint64_t k = 0; k < d; k++) {
    const uint64_t sub = bd << (3 * k);
    const int64_t* tw = p.twiddles[k].data();
    for (uint64_t blk = 0; blk < N; blk += 8 * sub) {
      i128* chunk = out + blk;
      for (uint64_t i = 0; i < sub; i++) {
        i128 xs[8];
        for (uint64_t j = 0; j < 8; j++) xs[j] = chunk[j * sub + i];
        const int64_t* twi = tw + i * 56;
        for (uint64_t r = 0; r < 8; r++) {
          u128 a = (u128)xs[0];
          for (uint64_t jm1 = 0; jm1 < 7; jm1++)
            a += (u128)xs[jm1 + 1] * (u128)(i128)twi[jm1 * 8 + BUTTERFLY_TABLE[r][jm1]];
          chunk[r * sub + i] = (i128)a;
        }
      }
    }
  }
}

static int resolve_params(uint64_t row_len, int64_t depth, uint64_t rep, PnttParams& p) {
  const uint64_t d = depth < 0 ? optimal_depth(row_len) : (uint64_t)depth;
  return pntt_params(row_len, d, rep, p);
}

// zip-plus/src/pcs/phase_commit.rs:159-176 encode_rows (rayon over rows).
extern "C" int or_iprs_encode_rows_binary(uint64_t row_len, int64_t depth, uint64_t rep,
                                          uint64_t D, const uint64_t* evals, uint64_t num_rows,
                                          int64_t* out) {
  if (D == 0 || D > 64) return OR_ERR_INVALID_PARAM;
  PnttParams p;
  int rc = resolve_params(row_len, depth, rep, p);
  if (rc) return rc;
  parallel_for(num_rows, [&](uint64_t r) {
    pntt_binary_row(p, D, evals + r * row_len, out + r * p.codeword_len * D);
  });
  return OR_OK;
}

// out: i128 as 2 little-endian i64 words per element.
extern "C" int or_iprs_encode_rows_i32(uint64_t row_len, int64_t depth, uint64_t rep,
                                       const int32_t* evals, uint64_t num_rows, int64_t* out) {
  PnttParams p;
  int rc = resolve_params(row_len, depth, rep, p);
  if (rc) return rc;
  parallel_for(num_rows, [&](uint64_t r) {
    pntt_i32_row(p, evals + r * row_len, (i128*)(out + 2 * r * p.codeword_len));
  });
  return OR_OK;
}

// ---------------------------------------------------------------------------
// rand 0.9: StdRng (ChaCha12) seeded by seed_from_u64, and
// SliceRandom::shuffle.  Used by zip-plus/src/utils.rs:6-9 shuffle_seeded.
// ---------------------------------------------------------------------------
struct ChaCha12 {
  uint32_t key[8];
  uint64_t counter = 0;
  uint32_t buf[16];
  int idx = 16;

  static inline uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

  void refill() {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                      key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                      (uint32_t)counter, (uint32_t)(counter >> 32), 0u, 0u};
    uint32_t x[16];
    std::memcpy(x, s, sizeof x);
#define QR(a, b, c, d)                 \
  x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16); \
  x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12); \
  x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);  \
  x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
    for (int i = 0; i < 6; i++) {  // 12 rounds = 6 double rounds
      QR(0, 4, 8, 12) QR(1, 5, 9, 13) QR(2, 6, 10, 14) QR(3, 7, 11, 15)
      QR(0, 5, 10, 15) QR(1, 6, 11, 12) QR(2, 7, 8, 13) QR(3, 4, 9, 14)
    }
#undef QR
    for (int i = 0; i < 16; i++) buf[i] = x[i] + s[i];
    counter++;
    idx = 0;
  }
  uint32_t next_u32() {
    if (idx >= 16) refill();
    return buf[idx++];
  }
};

// rand_core 0.9 SeedableRng::seed_from_u64: PCG32 output words fill the seed.
static void seed_from_u64(ChaCha12& rng, uint64_t state) {
  for (int i = 0; i < 8; i++) {
    state = state * 6364136223846793005ull + 11634580027462260723ull;
    const uint64_t s = state;
    const uint32_t xorshifted = (uint32_t)(((s >> 18) ^ s) >> 27);
    const uint32_t rot = (uint32_t)(s >> 59);
    rng.key[i] = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
  }
  rng.counter = 0;
  rng.idx = 16;
}

// rand 0.9 UniformInt<u32>::sample_single_inclusive(0, bound-1) (Canon's
// method: one widening multiply, one extra draw to reduce bias).
static uint32_t random_range_u32(ChaCha12& rng, uint32_t bound) {
  const uint32_t range = bound;
  uint64_t m = (uint64_t)rng.next_u32() * range;
  uint32_t result = (uint32_t)(m >> 32), lo = (uint32_t)m;
  if (lo > (uint32_t)(0u - range)) {
    const uint32_t new_hi = (uint32_t)(((uint64_t)rng.next_u32() * range) >> 32);
    if ((uint64_t)lo + new_hi > 0xffffffffull) result += 1;
  }
  return result;
}

// rand 0.9 seq::increasing_uniform::calculate_bound_u32.
static void calculate_bound_u32(uint32_t m, uint32_t* bound, uint32_t* count) {
  uint64_t product = m, current = (uint64_t)m + 1;
  for (;;) {
    uint64_t pr = product * current;
    if (pr > 0xffffffffull) break;
    product = pr;
    current++;
  }
  *bound = (uint32_t)product;
  *count = (uint32_t)(current - m);
}

// rand 0.9 SliceRandom::shuffle -> partial_shuffle(len) with
// IncreasingUniform(rng, 0): for i in 0..len: swap(i, index in [0, i]).
template <class T>
static void shuffle_seeded(T* data, uint64_t len, uint64_t seed) {
  if (len <= 1) return;
  ChaCha12 rng;
  seed_from_u64(rng, seed);
  uint32_t n = 0, chunk = 0, chunk_remaining = 1;  // new(rng, 0): n == 0 -> 1
  for (uint64_t i = 0; i < len; i++) {
    const uint32_t next_n = n + 1;
    uint32_t next_remaining;
    if (chunk_remaining == 0) {
      uint32_t bound, count;
      calculate_bound_u32(next_n, &bound, &count);
      chunk = random_range_u32(rng, bound);
      next_remaining = count - 1;
    } else {
      next_remaining = chunk_remaining - 1;
    }
    uint32_t index;
    if (next_remaining == 0) {
      index = chunk;
    } else {
      index = chunk % next_n;
      chunk /= next_n;
    }
    chunk_remaining = next_remaining;
    n = next_n;
    std::swap(data[i], data[index]);
  }
}

extern "C" int or_shuffle_u64(uint64_t* data, uint64_t len, uint64_t seed) {
  shuffle_seeded(data, len, seed);
  return OR_OK;
}

// zip-plus/src/code/raa.rs:72-80: perm_s = shuffle(0..N, seed s).
extern "C" int or_raa_perm(uint64_t n, uint64_t seed, uint64_t* out) {
  for (uint64_t i = 0; i < n; i++) out[i] = i;
  shuffle_seeded(out, n, seed);
  return OR_OK;
}

extern "C" int or_chacha12_words(uint64_t seed, uint32_t* out, uint64_t n) {
  ChaCha12 rng;
  seed_from_u64(rng, seed);
  for (uint64_t i = 0; i < n; i++) out[i] = rng.next_u32();
  return OR_OK;
}

// raa.rs:52-69 width assert: 1*Eval bits + 2*log2(row_len) + 2*log2(REP) must
// fit Cw::COEFF_BIT_WIDTH.
static int raa_check(uint64_t row_len, uint64_t rep, uint64_t eval_bits, uint64_t cw_bits) {
  if (!is_pow2(rep) || !is_pow2(row_len)) return OR_ERR_INVALID_PARAM;  // :39-46
  const uint64_t w = eval_bits + 2 * (uint64_t)__builtin_ctzll(row_len) +
                     2 * (uint64_t)__builtin_ctzll(rep);
  return w <= cw_bits ? OR_OK : OR_ERR_INVALID_PARAM;
}

// raa.rs:93-126 encode_inner (and raa_sign_flip.rs:30-71 with sign_flip=1,
// flip_even_signs at :135-165 negating odd indices) for one coefficient
// plane of T values.  repeat (:185-195), accumulate (:208-236),
// clone_shuffled (:239-244; the in-place shuffle gives the same order).
template <class T>
static void raa_encode_plane(const T* row, uint64_t row_len, uint64_t rep, const uint64_t* perm1,
                             const uint64_t* perm2, int sign_flip, T* tmp, T* out) {
  const uint64_t N = row_len * rep;
  for (uint64_t i = 0; i < N; i++) tmp[i] = row[i % row_len];
  if (sign_flip)
    for (uint64_t i = 1; i < N; i += 2) tmp[i] = (T)(0 - (typename std::make_unsigned<T>::type)tmp[i]);
  for (uint64_t i = 0; i < N; i++) out[i] = tmp[perm1[i]];
  typedef typename std::make_unsigned<T>::type U;
  U acc = 0;
  for (uint64_t i = 0; i < N; i++) { acc += (U)out[i]; out[i] = (T)acc; }
  if (sign_flip)
    for (uint64_t i = 1; i < N; i += 2) out[i] = (T)(0 - (U)out[i]);
  for (uint64_t i = 0; i < N; i++) tmp[i] = out[perm2[i]];
  acc = 0;
  for (uint64_t i = 0; i < N; i++) { acc += (U)tmp[i]; out[i] = (T)acc; }
}

// Integer rows (the reference's TestZipTypes: Int<1> -> Int<4>; values here
// stay inside i64).  Used for the raa.rs:396-413 known answer.
extern "C" int or_raa_encode_i64(uint64_t row_len, uint64_t rep, int sign_flip, const int64_t* row,
                                 int64_t* out) {
  if (!is_pow2(rep) || !is_pow2(row_len)) return OR_ERR_INVALID_PARAM;
  const uint64_t N = row_len * rep;
  std::vector<uint64_t> p1(N), p2(N);
  or_raa_perm(N, 1, p1.data());
  or_raa_perm(N, 2, p2.data());
  std::vector<int64_t> tmp(N);
  raa_encode_plane<int64_t>(row, row_len, rep, p1.data(), p2.data(), sign_flip, tmp.data(), out);
  return OR_OK;
}

// BinaryPoly<D> -> [i32; D] (zip-plus/benches/zip_plus_benches.rs:74-75);
// from_ref per coefficient (dense.rs:508-525).  Rows in parallel.
extern "C" int or_raa_encode_rows_binary(uint64_t row_len, uint64_t rep, uint64_t D, int sign_flip,
                                         const uint64_t* evals, uint64_t num_rows, int32_t* out) {
  if (D == 0 || D > 64) return OR_ERR_INVALID_PARAM;
  int rc = raa_check(row_len, rep, 1, 32);
  if (rc) return rc;
  const uint64_t N = row_len * rep;
  std::vector<uint64_t> p1(N), p2(N);
  or_raa_perm(N, 1, p1.data());
  or_raa_perm(N, 2, p2.data());
  parallel_for(num_rows, [&](uint64_t r) {
    std::vector<int32_t> plane_in(row_len), tmp(N), o(N);
    const uint64_t* x = evals + r * row_len;
    int32_t* dst = out + r * N * D;
    for (uint64_t c = 0; c < D; c++) {
      for (uint64_t i = 0; i < row_len; i++) plane_in[i] = (int32_t)((x[i] >> c) & 1);
      raa_encode_plane<int32_t>(plane_in.data(), row_len, rep, p1.data(), p2.data(), sign_flip,
                                tmp.data(), o.data());
      for (uint64_t i = 0; i < N; i++) dst[i * D + c] = o[i];
    }
  });
  return OR_OK;
}

// ---------------------------------------------------------------------------
// BLAKE3 (public spec) -- the `blake3` crate 1.8 used by
// zip-plus/src/merkle.rs:128-221 (Hasher::update/finalize and
// hazmat::merge_subtrees_{non_root,root} with Mode::Hash).
// ---------------------------------------------------------------------------
static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const int B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

static inline uint32_t rotr32(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }

static void b3_compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out[16]) {
  uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                    B3_IV[0], B3_IV[1], B3_IV[2], B3_IV[3],
                    (uint32_t)counter, (uint32_t)(counter >> 32), block_len, flags};
  uint32_t m[16];
  std::memcpy(m, block, 64);
#define G(a, b, c, d, x, y)                          \
  s[a] = s[a] + s[b] + (x); s[d] = rotr32(s[d] ^ s[a], 16); \
  s[c] = s[c] + s[d]; s[b] = rotr32(s[b] ^ s[c], 12);       \
  s[a] = s[a] + s[b] + (y); s[d] = rotr32(s[d] ^ s[a], 8);  \
  s[c] = s[c] + s[d]; s[b] = rotr32(s[b] ^ s[c], 7);
  for (int round = 0; round < 7; round++) {
    G(0, 4, 8, 12, m[0], m[1]) G(1, 5, 9, 13, m[2], m[3])
    G(2, 6, 10, 14, m[4], m[5]) G(3, 7, 11, 15, m[6], m[7])
    G(0, 5, 10, 15, m[8], m[9]) G(1, 6, 11, 12, m[10], m[11])
    G(2, 7, 8, 13, m[12], m[13]) G(3, 4, 9, 14, m[14], m[15])
    if (round < 6) {
      uint32_t t[16];
      for (int i = 0; i < 16; i++) t[i] = m[B3_PERM[i]];
      std::memcpy(m, t, 64);
    }
  }
#undef G
  for (int i = 0; i < 8; i++) {
    out[i] = s[i] ^ s[i + 8];
    out[i + 8] = s[i + 8] ^ cv[i];
  }
}

static void load_block(const uint8_t* p, uint64_t len, uint32_t m[16]) {
  uint8_t b[64] = {0};
  if (len) std::memcpy(b, p, len);  // p may be null for the empty input
  for (int i = 0; i < 16; i++)
    m[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
           ((uint32_t)b[4 * i + 3] << 24);
}

// Chaining value of one chunk (<= 1024 bytes); if `is_root`, the last block
// carries ROOT and the output is the hash.
static void b3_chunk(const uint8_t* data, uint64_t len, uint64_t chunk_idx, bool is_root,
                     uint32_t cv_out[8]) {
  uint32_t cv[8];
  std::memcpy(cv, B3_IV, 32);
  const uint64_t nblocks = len == 0 ? 1 : (len + 63) / 64;
  for (uint64_t b = 0; b < nblocks; b++) {
    const uint64_t off = b * 64;
    const uint64_t blen = std::min<uint64_t>(64, len - off);
    uint32_t m[16], o[16];
    load_block(data + off, blen, m);
    uint32_t flags = 0;
    if (b == 0) flags |= CHUNK_START;
    if (b == nblocks - 1) flags |= CHUNK_END | (is_root ? ROOT : 0);
    b3_compress(cv, m, chunk_idx, (uint32_t)blen, flags, o);
    std::memcpy(cv, o, 32);
  }
  std::memcpy(cv_out, cv, 32);
}

static void b3_parent(const uint32_t l[8], const uint32_t r[8], bool root, uint32_t out[8]) {
  uint32_t m[16], o[16];
  std::memcpy(m, l, 32);
  std::memcpy(m + 8, r, 32);
  b3_compress(B3_IV, m, 0, 64, PARENT | (root ? ROOT : 0), o);
  std::memcpy(out, o, 32);
}

// Subtree over chunks [c0, c0 + n) of `data`; the left subtree holds the
// largest power of two number of chunks strictly less than n.
static void b3_subtree(const uint8_t* data, uint64_t total_len, uint64_t c0, uint64_t n, bool root,
                       uint32_t out[8]) {
  if (n == 1) {
    const uint64_t off = c0 * 1024;
    b3_chunk(data + off, std::min<uint64_t>(1024, total_len - off), c0, root, out);
    return;
  }
  uint64_t left = 1;
  while (left * 2 < n) left *= 2;
  uint32_t lcv[8], rcv[8];
  b3_subtree(data, total_len, c0, left, false, lcv);
  b3_subtree(data, total_len, c0 + left, n - left, false, rcv);
  b3_parent(lcv, rcv, root, out);
}

static void b3_hash(const uint8_t* data, uint64_t len, uint8_t out[32]) {
  const uint64_t nchunks = len == 0 ? 1 : (len + 1023) / 1024;
  uint32_t h[8];
  b3_subtree(data, len, 0, nchunks, true, h);
  for (int i = 0; i < 8; i++) std::memcpy(out + 4 * i, &h[i], 4);
}

static void words_from_bytes(const uint8_t* b, uint32_t w[8]) { std::memcpy(w, b, 32); }
static void bytes_from_words(const uint32_t w[8], uint8_t* b) { std::memcpy(b, w, 32); }

extern "C" int or_blake3(const uint8_t* data, uint64_t len, uint8_t* out32) {
  b3_hash(data, len, out32);
  return OR_OK;
}

// hazmat::merge_subtrees_non_root / merge_subtrees_root (Mode::Hash).
extern "C" int or_blake3_parent(const uint8_t* l, const uint8_t* r, int root, uint8_t* out32) {
  uint32_t lw[8], rw[8], o[8];
  words_from_bytes(l, lw);
  words_from_bytes(r, rw);
  b3_parent(lw, rw, root != 0, o);
  bytes_from_words(o, out32);
  return OR_OK;
}

// Chaining value of chunk `chunk_idx` of a longer input (non-root).
extern "C" int or_blake3_chunk_cv(const uint8_t* data, uint64_t len, uint64_t chunk_idx,
                                  uint8_t* out32) {
  if (len > 1024) return OR_ERR_SHAPE;
  uint32_t o[8];
  b3_chunk(data, len, chunk_idx, false, o);
  bytes_from_words(o, out32);
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Merkle tree over columns.  merkle.rs:74-89 new, :148-168 hash_leaves
// (leaf j = BLAKE3 of the LE bytes of column j over all rows), :172-221
// build_merkle_tree_from_leaves, :224-257 build_sibling_path.
// rows: [num_rows][width][elem_bytes] row-major.
// ---------------------------------------------------------------------------
extern "C" int or_merkle_leaves(const uint8_t* rows, uint64_t num_rows, uint64_t width,
                                uint64_t elem_bytes, uint8_t* leaves) {
  if (num_rows == 0 || width == 0 || !is_pow2(width)) return OR_ERR_SHAPE;  // :78-85
  parallel_for(width, [&](uint64_t j) {
    std::vector<uint8_t> buf(num_rows * elem_bytes);
    for (uint64_t r = 0; r < num_rows; r++)
      std::memcpy(&buf[r * elem_bytes], rows + (r * width + j) * elem_bytes, elem_bytes);
    b3_hash(buf.data(), buf.size(), leaves + 32 * j);
  });
  return OR_OK;
}

// layers_out receives 2n-1 hashes: leaves, then each layer, root last.
extern "C" int or_merkle_layers(const uint8_t* leaves, uint64_t n, uint8_t* layers_out) {
  if (n == 0) {
    b3_hash(nullptr, 0, layers_out);  // :175-179
    return OR_OK;
  }
  if (!is_pow2(n)) return OR_ERR_SHAPE;
  std::memcpy(layers_out, leaves, 32 * n);
  uint64_t off = 0, cnt = n;
  while (cnt > 1) {
    const bool is_root = cnt == 2;
    const uint8_t* prev = layers_out + 32 * off;
    uint8_t* cur = layers_out + 32 * (off + cnt);
    parallel_for(cnt / 2, [&](uint64_t m) {
      or_blake3_parent(prev + 64 * m, prev + 64 * m + 32, is_root, cur + 32 * m);
    });
    off += cnt;
    cnt /= 2;
  }
  return OR_OK;
}

// merkle.rs:224-257: siblings bottom-up; returns the count in *len.
extern "C" int or_merkle_prove(const uint8_t* layers, uint64_t n, uint64_t leaf, uint8_t* siblings,
                               uint64_t* len) {
  if (leaf >= n || n == 0) return OR_ERR_INVALID_PARAM;
  uint64_t off = 0, cnt = n, idx = leaf, k = 0;
  while (cnt > 1) {
    const uint64_t sib = idx ^ 1;
    std::memcpy(siblings + 32 * k, layers + 32 * (off + sib), 32);
    k++;
    off += cnt;
    cnt /= 2;
    idx /= 2;
  }
  *len = k;
  return OR_OK;
}

// merkle.rs:282-359 MerkleProof::verify over a column's bytes (returns 0 when
// the proof checks out).
extern "C" int or_merkle_verify(const uint8_t* root, const uint8_t* column_bytes, uint64_t col_len,
                                uint64_t leaf, uint64_t leaf_count, const uint8_t* siblings,
                                uint64_t nsib) {
  uint8_t cur[32];
  b3_hash(column_bytes, col_len, cur);
  if (leaf_count == 1) {
    if (leaf == 0 && nsib == 0) return std::memcmp(cur, root, 32) == 0 ? OR_OK : 2;
    return 2;
  }
  // get_path_directions (:385-410) for a power-of-two leaf count reduces to
  // the bits of the index.
  uint64_t depth = 0;
  for (uint64_t c = leaf_count; c > 1; c = (c + 1) / 2) depth++;
  if (depth != nsib || !is_pow2(leaf_count)) return 2;
  for (uint64_t k = 0; k < nsib; k++) {
    const bool is_left = ((leaf >> k) & 1) == 0;
    const bool last = k + 1 == nsib;
    uint8_t nxt[32];
    if (is_left)
      or_blake3_parent(cur, siblings + 32 * k, last, nxt);
    else
      or_blake3_parent(siblings + 32 * k, cur, last, nxt);
    std::memcpy(cur, nxt, 32);
  }
  return std::memcmp(cur, root, 32) == 0 ? OR_OK : 2;
}

// ---------------------------------------------------------------------------
// ZipPlus::commit (zip-plus/src/pcs/phase_commit.rs:61-101) for a batch of
// `batch` same-size polys: encode every poly's rows, then one Merkle tree over
// the poly-major concatenation of all rows (:93).  root_out gets the root;
// cw_out (optional) the codewords; layers_out (optional) the 2N-1 hashes.
// kind: 0 IPRS, 1 RAA, 2 RAA sign-flip.  eval: 0 i32 (Cw i128), 1 binary
// (Cw [i64;D] for IPRS, [i32;D] for RAA).
// ---------------------------------------------------------------------------
extern "C" int or_commit(int kind, uint64_t row_len, int64_t depth, uint64_t rep, int eval,
                         uint64_t D, const void* evals, uint64_t num_rows, uint64_t batch,
                         uint8_t* cw_out, uint8_t* layers_out, uint8_t* root_out) {
  if (batch == 0) return OR_ERR_SHAPE;  // :65-68
  uint64_t N = row_len * rep;
  uint64_t elem = 0;
  if (kind == 0 && eval == 1) elem = 8 * D;
  else if (kind == 0 && eval == 0) elem = 16;
  else if (kind != 0 && eval == 1) elem = 4 * D;
  else return OR_ERR_INVALID_PARAM;
  const uint64_t poly_bytes = num_rows * N * elem;
  std::vector<uint8_t> own;
  uint8_t* cw = cw_out;
  if (!cw) {
    own.resize(poly_bytes * batch);
    cw = own.data();
  }
  for (uint64_t b = 0; b < batch; b++) {
    int rc;
    if (kind == 0 && eval == 1)
      rc = or_iprs_encode_rows_binary(row_len, depth, rep, D,
                                      (const uint64_t*)evals + b * num_rows * row_len, num_rows,
                                      (int64_t*)(cw + b * poly_bytes));
    else if (kind == 0)
      rc = or_iprs_encode_rows_i32(row_len, depth, rep,
                                   (const int32_t*)evals + b * num_rows * row_len, num_rows,
                                   (int64_t*)(cw + b * poly_bytes));
    else
      rc = or_raa_encode_rows_binary(row_len, rep, D, kind == 2,
                                     (const uint64_t*)evals + b * num_rows * row_len, num_rows,
                                     (int32_t*)(cw + b * poly_bytes));
    if (rc) return rc;
  }
  std::vector<uint8_t> leaves(32 * N);
  int rc = or_merkle_leaves(cw, batch * num_rows, N, elem, leaves.data());
  if (rc) return rc;
  std::vector<uint8_t> layers(32 * (2 * N - 1));
  rc = or_merkle_layers(leaves.data(), N, layers.data());
  if (rc) return rc;
  if (layers_out) std::memcpy(layers_out, layers.data(), layers.size());
  std::memcpy(root_out, layers.data() + 32 * (2 * N - 2), 32);
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Prover combined row (zip-plus/src/pcs/phase_prove.rs:153-174 alpha
// projection with BinaryU64PolyInnerProduct, poly/src/univariate/
// binary_u64.rs:393-432; :205-244 combined row with Int<5>::mul_by_scalar
// over i128 coefficients, utils/src/mul_by_scalar.rs:46-74).  CombR is
// Int<L> (two's complement over L little-endian u64 limbs, mod 2^(64L)).
// ---------------------------------------------------------------------------
struct BigL {
  uint64_t w[9];  // up to Int<8>, plus the guard limb of the checked folds
};

static void big_from_i128(BigL& a, i128 v, int L) {
  const uint64_t ext = v < 0 ? ~0ull : 0ull;
  a.w[0] = (uint64_t)v;
  a.w[1] = (uint64_t)((u128)v >> 64);
  for (int i = 2; i < L; i++) a.w[i] = ext;
}
static void big_add(BigL& a, const BigL& b, int L) {
  u128 c = 0;
  for (int i = 0; i < L; i++) {
    c += (u128)a.w[i] + b.w[i];
    a.w[i] = (uint64_t)c;
    c >>= 64;
  }
}
// a * (signed i128 s), mod 2^(64L): product of the L-limb two's complement
// value with the sign-extended L-limb form of s.
static void big_mul_i128(BigL& out, const BigL& a, i128 s, int L) {
  BigL b;
  big_from_i128(b, s, L);
  uint64_t r[9] = {0};
  for (int i = 0; i < L; i++) {
    u128 carry = 0;
    for (int j = 0; i + j < L; j++) {
      u128 t = (u128)a.w[i] * b.w[j] + r[i + j] + carry;
      r[i + j] = (uint64_t)t;
      carry = t >> 64;
    }
  }
  for (int i = 0; i < L; i++) out.w[i] = r[i];
}

static i128 load_i128(const int64_t* p) {
  return (i128)(((u128)(uint64_t)p[1] << 64) | (uint64_t)p[0]);
}

// evals: [batch][num_rows * row_len] u64 bit-packed; alphas: [batch][D] i128
// (2 LE words); coeffs: [num_rows] i128; out: [row_len][L] u64 limbs.
// Per the reference, coeffs is [1] when num_rows == 1 (:205-211); the caller
// passes that vector.
extern "C" int or_combined_row_binary(uint64_t D, uint64_t L, uint64_t num_rows, uint64_t row_len,
                                      uint64_t batch, const uint64_t* evals, const int64_t* alphas,
                                      const int64_t* coeffs, uint64_t* out) {
  if (D == 0 || D > 64 || L < 2 || L > 8) return OR_ERR_INVALID_PARAM;
  // Alpha projection per eval (:163-172): w' = sum of alpha_c over set bits.
  std::vector<BigL> proj(batch * num_rows * row_len);
  parallel_for(batch * num_rows, [&](uint64_t br) {
    const uint64_t b = br / num_rows;
    for (uint64_t col = 0; col < row_len; col++) {
      const uint64_t idx = br * row_len + col;
      BigL acc;
      big_from_i128(acc, 0, (int)L);
      uint64_t bits = evals[idx];
      if (D < 64) bits &= (1ull << D) - 1;  // only DEGREE_PLUS_ONE coefficients exist
      while (bits) {
        const int c = __builtin_ctzll(bits);
        BigL t;
        big_from_i128(t, load_i128(alphas + 2 * (b * D + c)), (int)L);
        big_add(acc, t, (int)L);
        bits &= bits - 1;
      }
      proj[idx] = acc;
    }
  });
  // Combined row (:213-244): comb[col] = sum_poly sum_j s_j * w'[j*row_len+col].
  parallel_for(row_len, [&](uint64_t col) {
    BigL acc;
    big_from_i128(acc, 0, (int)L);
    for (uint64_t b = 0; b < batch; b++)
      for (uint64_t j = 0; j < num_rows; j++) {
        BigL t;
        big_mul_i128(t, proj[(b * num_rows + j) * row_len + col], load_i128(coeffs + 2 * j), (int)L);
        big_add(acc, t, (int)L);
      }
    for (uint64_t i = 0; i < L; i++) out[col * L + i] = acc.w[i];
  });
  return OR_OK;
}

// The alpha projection alone (w' values, [n][L] limbs), for tests.
extern "C" int or_alpha_projection(uint64_t D, uint64_t L, uint64_t n, const uint64_t* evals,
                                   const int64_t* alphas, uint64_t* out) {
  if (D == 0 || D > 64 || L < 2 || L > 8) return OR_ERR_INVALID_PARAM;
  for (uint64_t k = 0; k < n; k++) {
    BigL acc;
    big_from_i128(acc, 0, (int)L);
    uint64_t bits = evals[k];
    if (D < 64) bits &= (1ull << D) - 1;
    while (bits) {
      const int c = __builtin_ctzll(bits);
      BigL t;
      big_from_i128(t, load_i128(alphas + 2 * c), (int)L);
      big_add(acc, t, (int)L);
      bits &= bits - 1;
    }
    for (uint64_t i = 0; i < L; i++) out[k * L + i] = acc.w[i];
  }
  return OR_OK;
}

// ---------------------------------------------------------------------------
// LinearCode::encode_wide (zip-plus/src/code.rs:55, IprsCode
// code/iprs.rs:148-150 -> encode_inner :54-87 with In = Out = CombR;
// RaaCode code/raa.rs:151-153 and RaaSignFlippingCode
// code/raa_sign_flip.rs:97-99 -> encode_inner).  CombR = Int<L>; the pNTT
// multiplies by the i64 table entries (Int<L>::mul_by_scalar(&i64),
// utils/src/mul_by_scalar.rs:31-74) and adds, all modulo 2^(64L) (the
// unchecked build; the checked build panics instead of wrapping, so the two
// agree on every input that does not overflow).
// ---------------------------------------------------------------------------
static void big_neg(BigL& a, int L) {
  u128 c = 1;
  for (int i = 0; i < L; i++) {
    c += (u128)(~a.w[i]);
    a.w[i] = (uint64_t)c;
    c >>= 64;
  }
}
static BigL big_load(const uint64_t* p, int L) {
  BigL a;
  for (int i = 0; i < L; i++) a.w[i] = p[i];
  return a;
}
static void big_store(uint64_t* p, const BigL& a, int L) {
  for (int i = 0; i < L; i++) p[i] = a.w[i];
}

// One row: in [row_len][L] limbs -> out [N][L] limbs (radix8.rs:21-161).
static void pntt_wide_row(const PnttParams& p, int L, const uint64_t* in, uint64_t* out) {
  const uint64_t N = p.codeword_len, bd = p.base_dim, bl = p.base_len, d = p.depth;
  const uint64_t shift = 3 * d;
  for (uint64_t i = 0; i < N; i++) {
    const uint64_t chunk = i >> p.base_dim_log2, row = i & (bd - 1);
    const uint64_t o = octrev(chunk, d);
    BigL acc = big_load(in + o * L, L);  // from_ref
    const int64_t* mrow = &p.base_matrix[row * bl];
    for (uint64_t col = 1; col < bl; col++) {
      BigL t;
      big_mul_i128(t, big_load(in + (o | (col << shift)) * L, L), (i128)mrow[col], L);
      big_add(acc, t, L);
    }
    big_store(out + i * L, acc, L);
  }
  for (uint64_t k = 0; k < d; k++) {
    const uint64_t sub = bd << (3 * k);
    const int64_t* tw = p.twiddles[k].data();
    for (uint64_t blk = 0; blk < N; blk += 8 * sub) {
      uint64_t* chunk = out + blk * L;
      for (uint64_t i = 0; i < sub; i++) {
        BigL xs[8];
        for (uint64_t j = 0; j < 8; j++) xs[j] = big_load(chunk + (j * sub + i) * L, L);
        const int64_t* twi = tw + i * 56;
        for (uint64_t r = 0; r < 8; r++) {
          BigL a = xs[0];
          for (uint64_t jm1 = 0; jm1 < 7; jm1++) {
            BigL t;
            big_mul_i128(t, xs[jm1 + 1], (i128)twi[jm1 * 8 + BUTTERFLY_TABLE[r][jm1]], L);
            big_add(a, t, L);
          }
          big_store(chunk + (r * sub + i) * L, a, L);
        }
      }
    }
  }
}

extern "C" int or_iprs_encode_wide(uint64_t row_len, int64_t depth, uint64_t rep, uint64_t L,
                                   const uint64_t* in, uint64_t num_rows, uint64_t* out) {
  if (L < 1 || L > 8) return OR_ERR_INVALID_PARAM;
  PnttParams p;
  int rc = resolve_params(row_len, depth, rep, p);
  if (rc) return rc;
  parallel_for(num_rows, [&](uint64_t r) {
    pntt_wide_row(p, (int)L, in + r * row_len * L, out + r * p.codeword_len * L);
  });
  return OR_OK;
}

// raa.rs:93-126 over Int<L> (sign flips: raa_sign_flip.rs:45-57).
extern "C" int or_raa_encode_wide(uint64_t row_len, uint64_t rep, int sign_flip, uint64_t L,
                                  const uint64_t* in, uint64_t num_rows, uint64_t* out) {
  if (L < 1 || L > 8 || !is_pow2(rep) || !is_pow2(row_len)) return OR_ERR_INVALID_PARAM;
  const uint64_t N = row_len * rep;
  std::vector<uint64_t> p1(N), p2(N);
  or_raa_perm(N, 1, p1.data());
  or_raa_perm(N, 2, p2.data());
  const int l = (int)L;
  parallel_for(num_rows, [&](uint64_t r) {
    const uint64_t* x = in + r * row_len * L;
    uint64_t* dst = out + r * N * L;
    std::vector<BigL> tmp(N), a(N);
    for (uint64_t i = 0; i < N; i++) {
      tmp[i] = big_load(x + (i % row_len) * L, l);  // repeat
      if (sign_flip && (i & 1)) big_neg(tmp[i], l);
    }
    BigL acc;
    big_from_i128(acc, 0, l);
    for (uint64_t i = 0; i < N; i++) {  // perm1, accumulate
      big_add(acc, tmp[p1[i]], l);
      a[i] = acc;
      if (sign_flip && (i & 1)) big_neg(a[i], l);
    }
    big_from_i128(acc, 0, l);
    for (uint64_t i = 0; i < N; i++) {  // perm2, accumulate
      big_add(acc, a[p2[i]], l);
      big_store(dst + i * L, acc, l);
    }
  });
  return OR_OK;
}

// ---------------------------------------------------------------------------
// The checked build of encode_wide: IprsCode<.., CHECKED> folds with
// mul_by_scalar::<true>(..).expect (code/iprs.rs:76-84: checked_mul of
// Int<L> by the i64 entry) and add! (pntt/radix8.rs:150-157 over
// col = 1..bl-1, butterfly.rs:34-42 over j = 1..7), and the checked RAA
// accumulate add! (raa.rs:205-218) with neg! sign flips
// (raa_sign_flip.rs:146-155); each panics at the first result outside
// Int<L>.  Here every value is carried exactly in L + 1 limbs (a
// sign-extended Int<L> times an entry below 2^31, and the sum of two
// Int<L> values, both fit) and range-checked after each step; the first
// failure returns OR_ERR_OVERFLOW (the panic).  Values that pass equal the
// unchecked ones.
// ---------------------------------------------------------------------------
static BigL big_widen(const uint64_t* p, int L) {
  BigL a;
  for (int i = 0; i < L; i++) a.w[i] = p[i];
  a.w[L] = (int64_t)p[L - 1] < 0 ? ~0ull : 0ull;
  return a;
}
// An (L+1)-limb value lies in Int<L> iff its top limb is the sign
// extension of limb L-1.
static bool big_in_range(const BigL& a, int L) {
  return a.w[L] == ((int64_t)a.w[L - 1] < 0 ? ~0ull : 0ull);
}
static bool big_is_min(const uint64_t* p, int L) {
  for (int i = 0; i < L - 1; i++)
    if (p[i]) return false;
  return p[L - 1] == (1ull << 63);
}

static int pntt_wide_row_checked(const PnttParams& p, int L, const uint64_t* in, uint64_t* out) {
  const uint64_t N = p.codeword_len, bd = p.base_dim, bl = p.base_len, d = p.depth;
  const uint64_t shift = 3 * d;
  const int W = L + 1;
  for (uint64_t i = 0; i < N; i++) {
    const uint64_t chunk = i >> p.base_dim_log2, row = i & (bd - 1);
    const uint64_t o = octrev(chunk, d);
    BigL acc = big_widen(in + o * L, L);  // from_ref
    const int64_t* mrow = &p.base_matrix[row * bl];
    for (uint64_t col = 1; col < bl; col++) {
      BigL t;
      big_mul_i128(t, big_widen(in + (o | (col << shift)) * L, L), (i128)mrow[col], W);
      if (!big_in_range(t, L)) return OR_ERR_OVERFLOW;
      big_add(acc, t, W);
      if (!big_in_range(acc, L)) return OR_ERR_OVERFLOW;
    }
    big_store(out + i * L, acc, L);
  }
  for (uint64_t k = 0; k < d; k++) {
    const uint64_t sub = bd << (3 * k);
    const int64_t* tw = p.twiddles[k].data();
    for (uint64_t blk = 0; blk < N; blk += 8 * sub) {
      uint64_t* chunk = out + blk * L;
      for (uint64_t i = 0; i < sub; i++) {
        BigL xs[8];
        for (uint64_t j = 0; j < 8; j++) xs[j] = big_widen(chunk + (j * sub + i) * L, L);
        const int64_t* twi = tw + i * 56;
        for (uint64_t r = 0; r < 8; r++) {
          BigL a = xs[0];
          for (uint64_t jm1 = 0; jm1 < 7; jm1++) {
            BigL t;
            big_mul_i128(t, xs[jm1 + 1], (i128)twi[jm1 * 8 + BUTTERFLY_TABLE[r][jm1]], W);
            if (!big_in_range(t, L)) return OR_ERR_OVERFLOW;
            big_add(a, t, W);
            if (!big_in_range(a, L)) return OR_ERR_OVERFLOW;
          }
          big_store(chunk + (r * sub + i) * L, a, L);
        }
      }
    }
  }
  return OR_OK;
}

extern "C" int or_iprs_encode_wide_checked(uint64_t row_len, int64_t depth, uint64_t rep,
                                           uint64_t L, const uint64_t* in, uint64_t num_rows,
                                           uint64_t* out) {
  if (L < 1 || L > 8) return OR_ERR_INVALID_PARAM;
  PnttParams p;
  int rc = resolve_params(row_len, depth, rep, p);
  if (rc) return rc;
  std::vector<int> rcs(num_rows, OR_OK);
  parallel_for(num_rows, [&](uint64_t r) {
    rcs[r] = pntt_wide_row_checked(p, (int)L, in + r * row_len * L, out + r * p.codeword_len * L);
  });
  for (int v : rcs)
    if (v) return v;
  return OR_OK;
}

extern "C" int or_raa_encode_wide_checked(uint64_t row_len, uint64_t rep, int sign_flip,
                                          uint64_t L, const uint64_t* in, uint64_t num_rows,
                                          uint64_t* out) {
  if (L < 1 || L > 8 || !is_pow2(rep) || !is_pow2(row_len)) return OR_ERR_INVALID_PARAM;
  const uint64_t N = row_len * rep;
  std::vector<uint64_t> p1(N), p2(N);
  or_raa_perm(N, 1, p1.data());
  or_raa_perm(N, 2, p2.data());
  const int l = (int)L, W = l + 1;
  std::vector<int> rcs(num_rows, OR_OK);
  parallel_for(num_rows, [&](uint64_t r) {
    const uint64_t* x = in + r * row_len * L;
    uint64_t* dst = out + r * N * L;
    std::vector<BigL> tmp(N), a(N);
    for (uint64_t i = 0; i < N; i++) {  // repeat, flip_even_signs
      const uint64_t* xi = x + (i % row_len) * L;
      if (sign_flip && (i & 1) && big_is_min(xi, l)) {
        rcs[r] = OR_ERR_OVERFLOW;
        return;
      }
      tmp[i] = big_widen(xi, l);
      if (sign_flip && (i & 1)) big_neg(tmp[i], W);
    }
    for (int pass = 0; pass < 2; pass++) {  // shuffle, accumulate (twice)
      const std::vector<uint64_t>& perm = pass == 0 ? p1 : p2;
      const std::vector<BigL>& src = pass == 0 ? tmp : a;
      std::vector<BigL> res(N);
      BigL acc = src[perm[0]];
      res[0] = acc;
      for (uint64_t i = 1; i < N; i++) {
        big_add(acc, src[perm[i]], W);
        if (!big_in_range(acc, l)) {
          rcs[r] = OR_ERR_OVERFLOW;
          return;
        }
        res[i] = acc;
      }
      if (pass == 0) {
        for (uint64_t i = 0; i < N; i++) {
          if (sign_flip && (i & 1)) {  // flip_even_signs after the first accumulate
            if (big_is_min(res[i].w, l)) {
              rcs[r] = OR_ERR_OVERFLOW;
              return;
            }
            big_neg(res[i], W);
          }
          a[i] = res[i];
        }
      } else {
        for (uint64_t i = 0; i < N; i++) big_store(dst + i * L, res[i], l);
      }
    }
  });
  for (int v : rcs)
    if (v) return v;
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Opening path.
//
// BLAKE3 extended output (finalize_xof) of a whole input: the root node's
// compression repeated with output block counters 0, 1, ... (BLAKE3 spec
// section 2.6).  The Fiat-Shamir transcript (transcript/src/lib.rs:10-72)
// reads its challenges this way from the hash of everything absorbed.
// ---------------------------------------------------------------------------
static void b3_root_block(const uint8_t* data, uint64_t len, uint64_t counter, uint32_t o[16]) {
  const uint64_t nchunks = len == 0 ? 1 : (len + 1023) / 1024;
  if (nchunks == 1) {
    uint32_t cv[8];
    std::memcpy(cv, B3_IV, 32);
    const uint64_t nblocks = len == 0 ? 1 : (len + 63) / 64;
    for (uint64_t b = 0; b < nblocks; b++) {
      const uint64_t off = b * 64, blen = std::min<uint64_t>(64, len - off);
      uint32_t m[16];
      load_block(data + off, blen, m);
      uint32_t flags = b == 0 ? CHUNK_START : 0;
      if (b + 1 < nblocks) {
        uint32_t t[16];
        b3_compress(cv, m, 0, (uint32_t)blen, flags, t);
        std::memcpy(cv, t, 32);
      } else {
        b3_compress(cv, m, counter, (uint32_t)blen, flags | CHUNK_END | ROOT, o);
      }
    }
    return;
  }
  uint64_t left = 1;
  while (left * 2 < nchunks) left *= 2;
  uint32_t m[16];
  b3_subtree(data, len, 0, left, false, m);
  b3_subtree(data, len, left, nchunks - left, false, m + 8);
  b3_compress(B3_IV, m, counter, 64, PARENT | ROOT, o);
}

extern "C" int or_blake3_xof(const uint8_t* data, uint64_t len, uint8_t* out, uint64_t n) {
  for (uint64_t blk = 0; blk * 64 < n; blk++) {
    uint32_t o[16];
    b3_root_block(data, len, blk, o);
    const uint64_t take = std::min<uint64_t>(64, n - blk * 64);
    std::memcpy(out + blk * 64, o, take);  // little-endian words
  }
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Prime field F_p, p < 2^256, by definition: residues kept canonical,
// products and sums reduced by binary long division (no Montgomery
// arithmetic).  The proof stream carries an element a as a * 2^256 mod p,
// which is also computed by division here.
// ---------------------------------------------------------------------------
// r = x mod p for an n-limb unsigned x.
static void big_mod(const uint64_t* x, int n, const uint64_t p[4], uint64_t r[4]) {
  uint64_t acc[5] = {0, 0, 0, 0, 0};  // < 2p < 2^257
  for (int bit = 64 * n - 1; bit >= 0; bit--) {
    for (int i = 4; i > 0; i--) acc[i] = (acc[i] << 1) | (acc[i - 1] >> 63);
    acc[0] = (acc[0] << 1) | ((x[bit / 64] >> (bit % 64)) & 1);
    bool ge = acc[4] != 0;
    if (!ge) {
      ge = true;
      for (int i = 3; i >= 0; i--)
        if (acc[i] != p[i]) {
          ge = acc[i] > p[i];
          break;
        }
    }
    if (ge) {
      uint64_t br = 0;
      for (int i = 0; i < 5; i++) {
        const uint64_t pi = i < 4 ? p[i] : 0;
        const u128 d = (u128)acc[i] - pi - br;
        acc[i] = (uint64_t)d;
        br = (uint64_t)(d >> 64) & 1;
      }
    }
  }
  std::memcpy(r, acc, 32);
}

// out[na+nb] = a * b (unsigned).
static void big_mul_u(const uint64_t* a, int na, const uint64_t* b, int nb, uint64_t* out) {
  for (int i = 0; i < na + nb; i++) out[i] = 0;
  for (int i = 0; i < na; i++) {
    u128 c = 0;
    for (int j = 0; j < nb; j++) {
      const u128 t = (u128)a[i] * b[j] + out[i + j] + c;
      out[i + j] = (uint64_t)t;
      c = t >> 64;
    }
    out[i + nb] = (uint64_t)c;
  }
}

static void fmod_add(const uint64_t a[4], const uint64_t b[4], const uint64_t p[4], uint64_t r[4]) {
  uint64_t s[5];
  u128 c = 0;
  for (int i = 0; i < 4; i++) {
    c += (u128)a[i] + b[i];
    s[i] = (uint64_t)c;
    c >>= 64;
  }
  s[4] = (uint64_t)c;
  big_mod(s, 5, p, r);
}

static void fmod_sub(const uint64_t a[4], const uint64_t b[4], const uint64_t p[4], uint64_t r[4]) {
  uint64_t nb[4], z[4] = {0, 0, 0, 0};
  // p - b (b < p), then a + (p - b)
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    const u128 d = (u128)p[i] - b[i] - br;
    nb[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  if (std::memcmp(b, z, 32) == 0) std::memcpy(nb, z, 32);
  fmod_add(a, nb, p, r);
}

extern "C" int or_field_mul(const uint64_t* a, const uint64_t* b, const uint64_t* p, uint64_t* r) {
  uint64_t t[8];
  big_mul_u(a, 4, b, 4, t);
  big_mod(t, 8, p, r);
  return OR_OK;
}

// a * 2^256 mod p: the element's bytes in the proof stream.
extern "C" int or_field_inner(const uint64_t* a, const uint64_t* p, uint64_t* r) {
  uint64_t t[8] = {0, 0, 0, 0, a[0], a[1], a[2], a[3]};
  big_mod(t, 8, p, r);
  return OR_OK;
}

// A two's complement integer of n limbs mod p.
static void signed_mod(const uint64_t* x, int n, const uint64_t p[4], uint64_t r[4]) {
  const bool neg = (int64_t)x[n - 1] < 0;
  std::vector<uint64_t> m(x, x + n);
  if (neg) {
    u128 c = 1;
    for (int i = 0; i < n; i++) {
      c += (u128)(~m[i]);
      m[i] = (uint64_t)c;
      c >>= 64;
    }
  }
  uint64_t v[4];
  big_mod(m.data(), n, p, v);
  if (neg) {
    const uint64_t z[4] = {0, 0, 0, 0};
    fmod_sub(z, v, p, r);
  } else {
    std::memcpy(r, v, 32);
  }
}

extern "C" int or_field_from_signed(const uint64_t* x, uint64_t n, const uint64_t* p, uint64_t* r) {
  if (n == 0 || n > 16) return OR_ERR_INVALID_PARAM;
  signed_mod(x, (int)n, p, r);
  return OR_OK;
}

// b of prove_f (zip-plus/src/pcs/phase_prove.rs:176-198): for every row j,
//   b_j = sum_poly sum_col F(w'_poly[j*row_len + col]) * q1[col]  (mod p),
// with w' the alpha projection (binary evals, :153-174) or the eval itself
// (i32 evals, ScalarProduct with alphas = [1]).  The sum of signed products
// is formed exactly (positive and negative parts) and reduced once per row.
// q1: [row_len][4] canonical; out: [num_rows][4] canonical.
extern "C" int or_prove_b(int binpoly, uint64_t D, uint64_t num_rows, uint64_t row_len,
                          uint64_t batch, const void* evals, const int64_t* alphas,
                          const uint64_t* q1, const uint64_t* p, uint64_t* out) {
  if (binpoly && (D == 0 || D > 64)) return OR_ERR_INVALID_PARAM;
  parallel_for(num_rows, [&](uint64_t j) {
    uint64_t b[4] = {0, 0, 0, 0};
    for (uint64_t pb = 0; pb < batch; pb++) {
      uint64_t pos[9] = {0}, neg[9] = {0};
      for (uint64_t col = 0; col < row_len; col++) {
        const uint64_t idx = (pb * num_rows + j) * row_len + col;
        i128 hi = 0;  // w' = hi * 2^128 + lo, exact in 192 bits
        u128 lo = 0;
        if (binpoly) {
          uint64_t bits = ((const uint64_t*)evals)[idx];
          if (D < 64) bits &= (1ull << D) - 1;
          while (bits) {
            const int c = __builtin_ctzll(bits);
            const i128 a = load_i128(alphas + 2 * (pb * D + c));
            const u128 s = lo + (u128)a;
            hi += (s < lo ? 1 : 0) + (a < 0 ? -1 : 0);
            lo = s;
            bits &= bits - 1;
          }
        } else {
          const i128 a = ((const int32_t*)evals)[idx];
          lo = (u128)a;
          hi = a < 0 ? -1 : 0;
        }
        uint64_t w[3] = {(uint64_t)lo, (uint64_t)(lo >> 64), (uint64_t)hi};
        const bool negative = (int64_t)w[2] < 0;
        if (negative) {
          u128 c = 1;
          for (int i = 0; i < 3; i++) {
            c += (u128)(~w[i]);
            w[i] = (uint64_t)c;
            c >>= 64;
          }
        }
        uint64_t prod[7];
        big_mul_u(w, 3, q1 + 4 * col, 4, prod);
        uint64_t* acc = negative ? neg : pos;
        u128 c = 0;
        for (int i = 0; i < 9; i++) {
          c += (u128)acc[i] + (i < 7 ? prod[i] : 0);
          acc[i] = (uint64_t)c;
          c >>= 64;
        }
      }
      uint64_t rp[4], rn[4], d[4];
      big_mod(pos, 9, p, rp);
      big_mod(neg, 9, p, rn);
      fmod_sub(rp, rn, p, d);
      fmod_add(b, d, p, b);
    }
    std::memcpy(out + 4 * j, b, 32);
  });
  return OR_OK;
}

// Combined row of i32 evals (phase_prove.rs:205-244 with ScalarProduct
// alpha projection: w' = eval): comb[col] = sum_poly sum_j s_j * x_j[col]
// in Int<L>.
extern "C" int or_combined_row_i32(uint64_t L, uint64_t num_rows, uint64_t row_len,
                                   uint64_t batch, const int32_t* evals, const int64_t* coeffs,
                                   uint64_t* out) {
  if (L < 2 || L > 8) return OR_ERR_INVALID_PARAM;
  parallel_for(row_len, [&](uint64_t col) {
    BigL acc;
    big_from_i128(acc, 0, (int)L);
    for (uint64_t b = 0; b < batch; b++)
      for (uint64_t j = 0; j < num_rows; j++) {
        BigL x, t;
        big_from_i128(x, evals[(b * num_rows + j) * row_len + col], (int)L);
        big_mul_i128(t, x, load_i128(coeffs + 2 * j), (int)L);
        big_add(acc, t, (int)L);
      }
    for (uint64_t i = 0; i < L; i++) out[col * L + i] = acc.w[i];
  });
  return OR_OK;
}
