// api.hip -- the C ABI declared in include/zipgpu.h.
#include <algorithm>
#include <array>
#include <map>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace zpi {

static thread_local std::string g_last_error;

int fail(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

}  // namespace zpi

using namespace zpi;

namespace zpi {

// Library-wide overrides of the kernel variants; -1 leaves the choice to
// the code's tuned variant (zp_autotune) or the default.
int g_force_iprs = -1, g_force_raa = -1, g_force_overlap = -1, g_force_wide = -1,
    g_force_tree = -1, g_force_ctree = -1, g_force_comb = -1;


uint64_t cw_elem_bytes_of(const zp_code* c, int eval_type, uint32_t D) {
  if (eval_type == ZP_EVAL_I64) return 16;  // i128 under either code
  if (c->kind == ZP_CODE_IPRS) {
    if (eval_type == ZP_EVAL_I32) return 16;
    if (eval_type == ZP_EVAL_BINPOLY) return 8ull * D;
    if (eval_type == ZP_EVAL_DENSE_I64) return 8ull * D;
  } else if (eval_type == ZP_EVAL_BINPOLY) {
    return 4ull * D;
  } else if (eval_type == ZP_EVAL_DENSE_I64) {
    return 32ull * D;  // Int<4> coefficients
  }
  return 0;
}

uint64_t eval_bytes_of(int eval_type, uint32_t D) {
  switch (eval_type) {
    case ZP_EVAL_I32: return 4;
    case ZP_EVAL_DENSE_I64: return 8ull * D;
    default: return 8;
  }
}

namespace {

template <class T>
int upload(const std::vector<T>& host, T** dev) {
  ZP_HIP(hipMalloc((void**)dev, host.size() * sizeof(T)));
  ZP_HIP(hipMemcpy(*dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  return ZP_OK;
}

// Butterfly coefficients T(j, t) = tw[(j-1)*8 + (t*j mod 8)] of each
// position (butterfly.rs:24-43).  The base twiddles satisfy B[u+4] = -B[u],
// so T(j, t+4) = T(j, t) for even j and, for odd j, the normalised negation
// of T(j, t): -T(j, t), or -T(j, t) - p where the negation leaves
// [-32769, 32767].  Keeps T(j, t) for t < 4 and marks the -p cases.
int symmetric_twiddles(const std::vector<int32_t>& tw, std::vector<double>& out,
                       std::vector<uint16_t>& mask) {
  const int64_t p = 65537;
  const uint64_t npos = tw.size() / 56;
  out.assign(npos * 32, 0.0);
  mask.assign(npos, 0);
  for (uint64_t i = 0; i < npos; i++) {
    const int32_t* row = &tw[i * 56];
    auto T = [&](int j, int t) { return (int64_t)row[(j - 1) * 8 + ((t * j) & 7)]; };
    for (int j = 1; j <= 7; j++)
      for (int t = 0; t < 4; t++) {
        out[i * 32 + (j - 1) * 4 + t] = (double)T(j, t);
        const int64_t hi = T(j, t + 4);
        if (j % 2 == 0) {
          if (hi != T(j, t)) return fail(ZP_ERR_INVALID_PARAM, "twiddle table is not symmetric");
        } else if (hi == -T(j, t) - p) {
          mask[i] |= (uint16_t)(1u << (((j - 1) / 2) * 4 + t));
        } else if (hi != -T(j, t)) {
          return fail(ZP_ERR_INVALID_PARAM, "twiddle table is not symmetric");
        }
      }
  }
  return ZP_OK;
}

void free_code(zp_code* c) {
  if (!c) return;
  (void)hipFree(c->d_ftab);
  for (auto* p : c->d_tw) (void)hipFree(p);
  (void)hipFree(c->d_afrag);
  (void)hipFree(c->d_afrag_c);
  (void)hipFree(c->d_tw0f);
  (void)hipFree(c->d_tw1f);
  (void)hipFree(c->d_q0f);
  (void)hipFree(c->d_q1f);
  (void)hipFree(c->d_perm1);
  (void)hipFree(c->d_perm2);
  delete c;
}

int pick(int forced, int tuned, int dflt) { return forced >= 0 ? forced : tuned >= 0 ? tuned : dflt; }

}  // namespace

int check_types(const zp_code* c, int eval_type, uint32_t D) {
  if (eval_type == ZP_EVAL_BINPOLY && (D == 0 || D > 64))
    return fail(ZP_ERR_INVALID_PARAM, "BinaryPoly degree bound must be within 1..64");
  if (eval_type == ZP_EVAL_DENSE_I64 && (D == 0 || D > 64))
    return fail(ZP_ERR_INVALID_PARAM, "DensePolynomial degree bound must be within 1..64");
  if (cw_elem_bytes_of(c, eval_type, D) == 0)
    return fail(ZP_ERR_INVALID_PARAM, "unsupported eval type for this code");
  if (c->kind != ZP_CODE_IPRS && eval_type == ZP_EVAL_BINPOLY) {
    // raa.rs:52-69: 1 + 2*log2(row_len) + 2*log2(REP) bits must fit i32.
    const uint64_t w = 1 + 2 * (uint64_t)__builtin_ctzll(c->row_len) +
                       2 * (uint64_t)__builtin_ctzll(c->rep);
    if (w > 32)
      return fail(ZP_ERR_SHAPE, "Cannot fit " + std::to_string(w) +
                                    "-bit wide codeword entries in 32 bits entries");
  }
  return ZP_OK;
}

// validate_input's bit budget of the linear combinations
// (zip-plus/src/pcs/utils.rs:24-54): with d = log2(row_len),
//   2*ceil_log2(max(REP*d, 1)) + ceil_log2(max(d, 1)) + challenge_bits
//     + (COEFF_BIT_WIDTH - 1) + ceil_log2(batch)  <=  CombR::NUM_BITS,
// where Chal = i128 (every ZipTypes of the reference: protocol/src/lib.rs:
// 475-593, zip-plus/benches) and challenge_bits doubles when Comb has a
// degree bound above zero (alphas are drawn too), i.e. for BinaryPoly<D>
// and DensePolynomial<i64, D> with D > 1.  COEFF_BIT_WIDTH is 1 for binary
// polys (binary_u64.rs:330), 32 for i32, 64 for i64 and DensePolynomial<i64,
// D> (dense.rs:431).  Returns ZP_ERR_SHAPE where the reference's assert
// fires, with its message.
int lc_budget_check(uint32_t rep, uint64_t row_len, int eval_type, uint32_t D, uint32_t L,
                    uint64_t batch) {
  auto clog2 = [](uint64_t x) {
    uint32_t r = 0;
    while ((1ull << r) < x) r++;
    return (uint64_t)r;
  };
  if (row_len == 0) return fail(ZP_ERR_SHAPE, "row length must be positive");
  const uint64_t d = (uint64_t)(63 - __builtin_clzll(row_len));  // usize::ilog2
  const uint64_t cw_bits = clog2(std::max<uint64_t>((uint64_t)rep * d, 1));
  const bool poly = eval_type == ZP_EVAL_BINPOLY || eval_type == ZP_EVAL_DENSE_I64;
  const uint64_t chal_bits = (poly && D > 1) ? 256 : 128;
  const uint64_t coeff_bits = eval_type == ZP_EVAL_BINPOLY ? 1 : eval_type == ZP_EVAL_I32 ? 32 : 64;
  const uint64_t lc = 2 * cw_bits + clog2(std::max<uint64_t>(d, 1)) + chal_bits +
                      (coeff_bits - 1) + clog2(batch);
  if (lc > 64ull * L)
    return fail(ZP_ERR_SHAPE, "The number of bits used for linear combinations is too large: " +
                                  std::to_string(lc) + " bits, max allowed is " +
                                  std::to_string(64 * L) + " bits");
  return ZP_OK;
}

bool iprs_uses_mfma(const zp_code* c, uint32_t D) {
  return iprs_mfma_eligible(c, D) && pick(g_force_iprs, c->tune_iprs, 0) != 1;
}

// Encodes with the variant in force for this code (types already checked).
int encode_rows_variant(const zp_code* c, int eval_type, uint32_t D, const void* evals,
                        uint64_t num_rows, void* out, hipStream_t s, hipEvent_t mid) {
  const bool mfma_binary = !c->checked && c->kind == ZP_CODE_IPRS &&
                           eval_type == ZP_EVAL_BINPOLY && iprs_uses_mfma(c, D);
  if (mfma_binary)
    return launch_iprs_binary_mfma(c, (const uint64_t*)evals, num_rows, D, (int64_t*)out, s,
                                   pick(g_force_iprs, c->tune_iprs, 0) != 2, mid);
  if (mid && hipEventRecord(mid, s) != hipSuccess)
    return fail(ZP_ERR_DEVICE, "event record failed");
  if (c->checked) {
    bool done = false;
    ZP_TRY(encode_checked(c, eval_type, D, evals, num_rows, out, s, &done));
    if (done) return ZP_OK;
  }
  // i64 and DensePolynomial<i64, D> evals: the Int<L> encoders per
  // coefficient plane (wrapping in the Cw type, as the unchecked build).
  if (eval_type == ZP_EVAL_I64)
    return launch_encode_i64_planes(c, 1, 2, (const int64_t*)evals, num_rows, out, s);
  if (eval_type == ZP_EVAL_DENSE_I64)
    return launch_encode_i64_planes(c, D, c->kind == ZP_CODE_IPRS ? 1 : 4,
                                    (const int64_t*)evals, num_rows, out, s);
  if (c->kind == ZP_CODE_IPRS) {
    if (eval_type == ZP_EVAL_BINPOLY) {
      if (iprs_uses_mfma(c, D))
        return launch_iprs_binary_mfma(c, (const uint64_t*)evals, num_rows, D, (int64_t*)out, s,
                                       pick(g_force_iprs, c->tune_iprs, 0) != 2);
      return launch_iprs_binary(c, (const uint64_t*)evals, num_rows, D, (int64_t*)out, s);
    }
    if (wide_mfma_eligible(c))
      return launch_encode_i32_mfma(c, (const int32_t*)evals, num_rows, out, s);
    return launch_iprs_i32(c, (const int32_t*)evals, num_rows, out, s);
  }
  return launch_raa_binary(c, (const uint64_t*)evals, num_rows, D, (int32_t*)out, s,
                           pick(g_force_raa, c->tune_raa, 0));
}

namespace {

// Encoder timing inside zp_commit: while it is on, a timed commit takes the
// next three events of a pool of timing events (made by
// zp_encode_timing_start, so a timed commit creates nothing) and records
// them on the commit stream before its encoder, before the encoder's last
// kernel and after it.  The events carry no system-scope fence: recording
// them writes back and invalidates no cache.  zp_encode_timing_stop reads
// and destroys the pool.  The store is never destroyed, so a commit running
// during process exit still finds it.
struct EncodeTiming {
  std::mutex mu;
  bool on = false;
  std::vector<hipEvent_t> pool;  // triples: pool[3i .. 3i + 2]
  size_t used = 0;               // triples handed out
  uint32_t every = 1;            // one commit in `every` is timed
  uint64_t seen = 0;             // commits that asked while on
  std::vector<uint8_t> ok;       // per triple handed out
};
EncodeTiming& encode_timing() {
  static auto* t = new EncodeTiming;
  return *t;
}

// The next triple of the pool (its index in *slot), or false when timing is
// off, this commit is not sampled or the pool is used up.
bool timing_begin(hipEvent_t (&e)[3], size_t* slot) {
  e[0] = e[1] = e[2] = nullptr;
  EncodeTiming& t = encode_timing();
  std::lock_guard<std::mutex> g(t.mu);
  if (!t.on || 3 * (t.used + 1) > t.pool.size()) return false;
  if (t.seen++ % t.every != 0) return false;
  *slot = t.used++;
  for (int i = 0; i < 3; i++) e[i] = t.pool[3 * *slot + i];
  t.ok.push_back(0);
  return true;
}

// Marks the triple of a successful encode as one to read.
void timing_end(size_t slot, bool ok) {
  EncodeTiming& t = encode_timing();
  std::lock_guard<std::mutex> g(t.mu);
  if (ok && slot < t.ok.size()) t.ok[slot] = 1;
}

// Encode all rows of all polys, then hash the leaves.  For a single poly
// whose rows split into whole 1 KiB BLAKE3 chunks, the rows are encoded in
// groups on `s` while the chunks of finished groups are hashed on the
// auxiliary stream; the streams join before the chunk trees are merged.
int commit_encode_and_hash(const zp_code* code, int eval_type, uint32_t D, const void* d_evals,
                           uint64_t num_rows, uint32_t batch, zp_hint* h, hipStream_t s,
                           bool allow_overlap, bool timed) {
  const uint64_t N = h->N, elem = h->elem, rows_all = num_rows * batch;
  const uint64_t in_stride = num_rows * code->row_len * eval_bytes_of(eval_type, D);
  const uint64_t rpc = (elem <= 1024 && 1024 % elem == 0) ? 1024 / elem : 0;  // rows per chunk
  const uint64_t nchunks = leaf_chunks(rows_all, (uint32_t)elem);
  const uint32_t kGroups = 4;
  const bool pipelined = allow_overlap && pick(g_force_overlap, code->tune_overlap, 0) == 1 &&
                         batch == 1 && rpc && num_rows % (rpc * kGroups) == 0 &&
                         nchunks >= kGroups;
  if (!pipelined) {
    hipEvent_t te[3] = {nullptr, nullptr, nullptr};
    size_t slot = 0;
    const bool tmd = timed && timing_begin(te, &slot);
    int rc = ZP_OK;
    if (tmd && hipEventRecord(te[0], s) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "event record failed");
    for (uint32_t b = 0; rc == ZP_OK && b < batch; b++)  // the last poly's last kernel
      rc = encode_rows_variant(code, eval_type, D, (const uint8_t*)d_evals + b * in_stride,
                               num_rows, h->d_cw + b * h->per_poly, s,
                               b + 1 == batch ? te[1] : nullptr);
    if (tmd && rc == ZP_OK && hipEventRecord(te[2], s) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "event record failed");
    if (tmd) timing_end(slot, rc == ZP_OK);
    ZP_TRY(rc);
    return launch_leaf_hash(h->d_cw, rows_all, N, (uint32_t)elem, h->d_layers, s,
                            code->tune_ctree_m);
  }
  hipStream_t s2;
  ZP_TRY(aux_stream(&s2));
  uint32_t* cv[2] = {nullptr, nullptr};
  ZP_TRY(dev_alloc((void**)&cv[0], nchunks * N * 32, s));
  int rc = dev_alloc((void**)&cv[1], ((nchunks + 1) / 2) * N * 32, s);
  hipEvent_t ev[kGroups + 1] = {};
  for (uint32_t q = 0; rc == ZP_OK && q <= kGroups; q++)
    if (hipEventCreateWithFlags(&ev[q], hipEventDisableTiming) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "event creation failed");
  const uint64_t rows_q = num_rows / kGroups, chunks_q = nchunks / kGroups;
  for (uint32_t q = 0; rc == ZP_OK && q < kGroups; q++) {
    rc = encode_rows_variant(
        code, eval_type, D,
        (const uint8_t*)d_evals + q * rows_q * code->row_len * eval_bytes_of(eval_type, D), rows_q,
        h->d_cw + q * rows_q * N * elem, s);
    if (rc == ZP_OK && (hipEventRecord(ev[q], s) != hipSuccess ||
                        hipStreamWaitEvent(s2, ev[q], 0) != hipSuccess))
      rc = fail(ZP_ERR_DEVICE, "event record/wait failed");
    if (rc == ZP_OK)
      rc = launch_chunk_cvs(h->d_cw, rows_all, N, (uint32_t)elem, q * chunks_q,
                            (q + 1) * chunks_q, cv[0], h->d_layers, s2);
  }
  // Join: later work on `s` (and the reuse of cv on `s`) waits for s2.
  if (hipEventRecord(ev[kGroups], s2) != hipSuccess ||
      hipStreamWaitEvent(s, ev[kGroups], 0) != hipSuccess)
    rc = rc == ZP_OK ? fail(ZP_ERR_DEVICE, "stream join failed") : rc;
  if (rc == ZP_OK) rc = launch_cv_merge(cv[0], cv[1], nchunks, N, h->d_layers, s, code->tune_ctree_m);
  for (auto e : ev)
    if (e) (void)hipEventDestroy(e);
  dev_free(cv[0], s);
  dev_free(cv[1], s);
  return rc;
}

}  // namespace

int hint_alloc(const zp_code* code, int eval_type, uint32_t D, uint64_t num_rows, uint32_t batch,
               hipStream_t s, zp_hint** out) {
  *out = nullptr;
  if (batch == 0) return fail(ZP_ERR_SHAPE, "Batch must contain at least one polynomial");
  if (num_rows == 0) return fail(ZP_ERR_SHAPE, "num_rows must be positive");
  ZP_TRY(check_types(code, eval_type, D));
  const uint64_t N = code->N;
  if (!pow2(N)) return fail(ZP_ERR_SHAPE, "codeword length must be a power of two");
  zp_hint* h = new zp_hint();
  h->device = code->device;
  h->N = N;
  h->num_rows = num_rows;
  h->elem = cw_elem_bytes_of(code, eval_type, D);
  h->per_poly = num_rows * N * h->elem;
  h->batch = batch;
  h->height = (uint32_t)__builtin_ctzll(N) + 1;
  h->stream = s;
  int rc = dev_alloc((void**)&h->d_cw, h->per_poly * batch, s);
  if (rc == ZP_OK) rc = dev_alloc((void**)&h->d_layers, (2 * N - 1) * 32, s);
  if (rc != ZP_OK) {
    zp_hint_free(h);
    return rc;
  }
  *out = h;
  return ZP_OK;
}

int commit_enqueue(const zp_code* code, int eval_type, uint32_t D, const void* d_evals,
                   uint64_t num_rows, uint32_t batch, zp_hint* h, hipStream_t s,
                   bool allow_overlap, bool timed) {
  ZP_TRY(commit_encode_and_hash(code, eval_type, D, d_evals, num_rows, batch, h, s,
                                allow_overlap, timed));
  return launch_merkle_layers(h->d_layers, h->N, s);
}

}  // namespace zpi

extern "C" {

int zp_init(int device) {
  ZP_HIP(hipSetDevice(device));
  return noalloc_init_device();
}

const char* zp_status_string(int status) {
  switch (status) {
    case ZP_OK: return "ok";
    case ZP_ERR_INVALID_PARAM: return "invalid parameter";
    case ZP_ERR_OVERFLOW: return "overflow";
    case ZP_ERR_SHAPE: return "shape";
    case ZP_ERR_DEVICE: return "device";
    case ZP_ERR_OOM: return "out of memory";
    case ZP_ERR_TIMEOUT: return "timeout";
  }
  return "unknown";
}

const char* zp_last_error(void) { return g_last_error.c_str(); }

int zp_set_option(const char* name, int64_t value) {
  if (!name) return fail(ZP_ERR_INVALID_PARAM, "null option name");
  struct Opt {
    const char* name;
    int* var;
    int64_t max;
  };
  const Opt opts[] = {{"iprs_encoder", &g_force_iprs, 2},
                      {"raa_kernel", &g_force_raa, 2},
                      {"commit_overlap", &g_force_overlap, 1},
                      {"wide_base", &g_force_wide, 1},
                      {"merkle_tree", &g_force_tree, 1},
                      {"chunk_tree", &g_force_ctree, 1},
                      {"combined_row", &g_force_comb, 1}};
  for (const auto& o : opts) {
    if (std::strcmp(name, o.name) != 0) continue;
    if (value < -1 || value > o.max)
      return fail(ZP_ERR_INVALID_PARAM, std::string(name) + " is -1 (automatic) or 0.." +
                                            std::to_string(o.max));
    *o.var = (int)value;
    return ZP_OK;
  }
  return fail(ZP_ERR_INVALID_PARAM, std::string("unknown option ") + name);
}

int zp_encode_timing_start(uint32_t max_commits, uint32_t every) {
  auto& t = zpi::encode_timing();
  std::lock_guard<std::mutex> g(t.mu);
  if (t.on) return fail(ZP_ERR_INVALID_PARAM, "encoder timing is already on");
  if (max_commits == 0 || max_commits > (1u << 20))
    return fail(ZP_ERR_INVALID_PARAM, "max_commits must be within 1..2^20");
  if (every == 0) return fail(ZP_ERR_INVALID_PARAM, "every must be positive");
  std::vector<hipEvent_t> pool(3 * (size_t)max_commits, nullptr);
  for (auto& e : pool)
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
      (void)hipGetLastError();
      for (auto& x : pool)
        if (x) (void)hipEventDestroy(x);
      return fail(ZP_ERR_DEVICE, "timing event creation failed");
    }
  t.pool.swap(pool);
  t.used = 0;
  t.every = every;
  t.seen = 0;
  t.ok.clear();
  t.on = true;
  return ZP_OK;
}

int zp_encode_timing_stop(double* total_ms, double* last_ms, uint64_t* launches) {
  auto& t = zpi::encode_timing();
  std::vector<hipEvent_t> pool;
  std::vector<uint8_t> ok;
  size_t used = 0;
  {
    std::lock_guard<std::mutex> g(t.mu);
    if (!t.on) return fail(ZP_ERR_INVALID_PARAM, "encoder timing is not on");
    t.on = false;
    pool.swap(t.pool);
    ok.swap(t.ok);
    used = t.used;
    t.used = 0;
  }
  double sum = 0.0, sum_last = 0.0;
  uint64_t n = 0;
  int rc = ZP_OK;
  for (size_t i = 0; i < used; i++) {
    hipEvent_t* e = &pool[3 * i];
    // Every handed-out triple is waited for: a failed encode may have
    // recorded some of its events.
    for (int k = 0; k < 3; k++)
      if (hipEventSynchronize(e[k]) != hipSuccess)
        rc = rc == ZP_OK ? fail(ZP_ERR_DEVICE, "encoder timing events failed") : rc;
    if (!ok[i] || rc != ZP_OK) continue;
    float all = 0.0f, last = 0.0f;
    if (hipEventElapsedTime(&all, e[0], e[2]) != hipSuccess ||
        hipEventElapsedTime(&last, e[1], e[2]) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "encoder timing events failed");
    sum += all;
    sum_last += last;
    n++;
  }
  for (auto e : pool) (void)hipEventDestroy(e);
  if (total_ms) *total_ms = rc == ZP_OK ? sum : 0.0;
  if (last_ms) *last_ms = rc == ZP_OK ? sum_last : 0.0;
  if (launches) *launches = rc == ZP_OK ? n : 0;
  return rc;
}

int zp_code_create(int kind, uint32_t rep, uint64_t row_len, int32_t depth, zp_code** out) {
  if (!out) return fail(ZP_ERR_INVALID_PARAM, "null output");
  *out = nullptr;
  zp_code* c = new zp_code();
  c->kind = kind;
  c->rep = rep;
  c->row_len = row_len;
  c->N = row_len * rep;
  if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
  int rc = ZP_OK;
  if (kind == ZP_CODE_IPRS) {
    const uint64_t d = depth < 0 ? iprs_optimal_depth(row_len) : (uint64_t)depth;
    uint32_t bl = 0, bd = 0;
    rc = iprs_params(row_len, d, rep, &bl, &bd);
    if (rc == ZP_OK && d > kMaxIprsDepth)
      rc = fail(ZP_ERR_INVALID_PARAM, "IPRS depth > 3 is not supported on the device");
    if (rc == ZP_OK) {
      c->depth = (uint32_t)d;
      c->bl = bl;
      c->bd = bd;
      c->bd_log2 = (uint32_t)__builtin_ctz(bd);
      std::vector<int32_t> f(bd);
      iprs_ftab(bd, f.data());
      rc = upload(f, &c->d_ftab);
      std::vector<std::vector<int32_t>> tws(c->depth);
      for (uint32_t k = 0; rc == ZP_OK && k < c->depth; k++) {
        tws[k].resize(((uint64_t)bd << (3 * k)) * 56);
        iprs_twiddles(bd, c->N, c->depth, k, tws[k].data());
        rc = upload(tws[k], &c->d_tw[k]);
      }
      // Growth bound of the checked encodings: the base layer sums bl
      // products with |f| <= max|f|; stage k multiplies by at most
      // max over positions and outputs of 1 + sum_j |twiddle|.
      int64_t fmax = 0;
      for (int32_t v : f) fmax = std::max<int64_t>(fmax, v < 0 ? -(int64_t)v : v);
      c->growth_log2 = std::log2((double)bl * (double)std::max<int64_t>(fmax, 1));
      for (uint32_t k = 0; k < c->depth; k++) {
        int64_t smax = 1;
        const uint64_t npos = (uint64_t)bd << (3 * k);
        for (uint64_t i = 0; i < npos; i++)
          for (int r = 0; r < 8; r++) {
            int64_t sum = 1;
            for (int jm1 = 0; jm1 < 7; jm1++) {
              const int64_t t = tws[k][i * 56 + jm1 * 8 + ((r * (jm1 + 1)) & 7)];
              sum += t < 0 ? -t : t;
            }
            smax = std::max(smax, sum);
          }
        c->growth_log2 += std::log2((double)smax);
      }
      // Tables of the matrix-core encoder (depth 1 or 2, K up to 256).
      const uint32_t ksteps = (bl + 63) / 64;
      if (rc == ZP_OK && (c->depth == 1 || c->depth == 2) && bd >= 16 && ksteps <= 4) {
        std::vector<int8_t> af((uint64_t)(bd / 16) * ksteps * 3 * 64 * 16);
        iprs_afrag(bd, bl, ksteps, f.data(), af.data(), true);
        rc = upload(af, (int8_t**)&c->d_afrag);
        if (rc == ZP_OK) {
          iprs_afrag(bd, bl, ksteps, f.data(), af.data(), false);
          rc = upload(af, (int8_t**)&c->d_afrag_c);
        }
        for (uint32_t k = 0; rc == ZP_OK && k < c->depth; k++) {
          std::vector<double> tf;
          std::vector<uint16_t> qm;
          rc = symmetric_twiddles(tws[k], tf, qm);
          if (rc == ZP_OK) rc = upload(tf, k == 0 ? &c->d_tw0f : &c->d_tw1f);
          if (rc == ZP_OK) rc = upload(qm, k == 0 ? &c->d_q0f : &c->d_q1f);
        }
        if (rc == ZP_OK) c->mfma_ksteps = ksteps;
      }
    }
  } else if (kind == ZP_CODE_RAA || kind == ZP_CODE_RAA_SIGNFLIP) {
    if (!pow2(rep)) rc = fail(ZP_ERR_SHAPE, "Repetition factor must be a power of two");
    else if (!pow2(row_len)) rc = fail(ZP_ERR_SHAPE, "Row length must be a power of two");
    else if (c->N > (1ull << 32)) rc = fail(ZP_ERR_INVALID_PARAM, "codeword too long");
    if (rc == ZP_OK) {
      c->growth_log2 = 2 * std::log2((double)c->N);  // two accumulates of N terms
      std::vector<uint32_t> p1(c->N), p2(c->N);
      raa_perm(c->N, 1, p1.data());
      raa_perm(c->N, 2, p2.data());
      rc = upload(p1, &c->d_perm1);
      if (rc == ZP_OK) rc = upload(p2, &c->d_perm2);
    }
  } else {
    rc = fail(ZP_ERR_INVALID_PARAM, "unknown code kind");
  }
  if (rc != ZP_OK) {
    free_code(c);
    return rc;
  }
  *out = c;
  return ZP_OK;
}

void zp_code_destroy(zp_code* code) { free_code(code); }

int zp_validate_input(uint32_t rep, uint64_t row_len, int eval_type, uint32_t D,
                      uint32_t comb_limbs, uint32_t batch) {
  if (batch == 0) return fail(ZP_ERR_SHAPE, "Batch must contain at least one polynomial");
  if (comb_limbs < 1 || comb_limbs > 8)
    return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 1..8");
  if (eval_type < ZP_EVAL_I32 || eval_type > ZP_EVAL_DENSE_I64)
    return fail(ZP_ERR_INVALID_PARAM, "unknown eval type");
  if ((eval_type == ZP_EVAL_BINPOLY || eval_type == ZP_EVAL_DENSE_I64) && (D == 0 || D > 64))
    return fail(ZP_ERR_INVALID_PARAM, "degree bound must be within 1..64");
  return lc_budget_check(rep, row_len, eval_type, D, comb_limbs, batch);
}
int zp_code_set_checked(zp_code* c, int checked) {
  if (!c) return fail(ZP_ERR_INVALID_PARAM, "null code");
  c->checked = checked != 0;
  return ZP_OK;
}

int zp_code_is_checked(const zp_code* c) { return c && c->checked ? 1 : 0; }

uint64_t zp_code_row_len(const zp_code* c) { return c ? c->row_len : 0; }
uint64_t zp_code_codeword_len(const zp_code* c) { return c ? c->N : 0; }
uint32_t zp_code_depth(const zp_code* c) { return c ? c->depth : 0; }
int zp_code_kind_of(const zp_code* c) { return c ? c->kind : -1; }

uint64_t zp_cw_elem_bytes(const zp_code* c, int eval_type, uint32_t D) {
  return c ? cw_elem_bytes_of(c, eval_type, D) : 0;
}

int zp_encode_rows(const zp_code* code, int eval_type, uint32_t D, const void* d_evals,
                   uint64_t num_rows, void* d_out, zp_stream stream) {
  if (!code) return fail(ZP_ERR_INVALID_PARAM, "null code");
  if (num_rows && (!d_evals || !d_out)) return fail(ZP_ERR_INVALID_PARAM, "null buffer");
  ZP_TRY(check_types(code, eval_type, D));
  return encode_rows_variant(code, eval_type, D, d_evals, num_rows, d_out, (hipStream_t)stream);
}

int zp_encode_wide(const zp_code* code, uint32_t L, const uint64_t* d_in, uint64_t num_rows,
                   uint64_t* d_out, zp_stream stream) {
  if (!code) return fail(ZP_ERR_INVALID_PARAM, "null code");
  if (L < 1 || L > 8) return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 1..8");
  if (num_rows && (!d_in || !d_out)) return fail(ZP_ERR_INVALID_PARAM, "null buffer");
  return launch_encode_wide(code, L, d_in, num_rows, d_out, (hipStream_t)stream);
}

int zp_merkle_root(const void* d_rows, uint64_t num_rows, uint64_t width, uint32_t elem,
                   uint8_t h_root[32], zp_stream stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!d_rows || !h_root) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  if (num_rows == 0 || width == 0 || elem == 0) return fail(ZP_ERR_SHAPE, "empty matrix");
  if (!pow2(width)) return fail(ZP_ERR_SHAPE, "row width must be a power of two");
  uint8_t* layers = nullptr;
  ZP_TRY(dev_alloc((void**)&layers, (2 * width - 1) * 32, s));
  uint8_t* pin = pinned_scratch(32);
  if (!pin) {
    dev_free(layers, s);
    return fail(ZP_ERR_OOM, "pinned staging allocation failed");
  }
  int rc = launch_leaf_hash((const uint8_t*)d_rows, num_rows, width, elem, layers, s);
  if (rc == ZP_OK) rc = launch_merkle_layers(layers, width, s);
  if (rc == ZP_OK && hipMemcpyAsync(pin, layers + (2 * width - 2) * 32, 32,
                                    hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "root copy failed");
  if (rc == ZP_OK && hipStreamSynchronize(s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "stream synchronize failed");
  if (rc == ZP_OK) std::memcpy(h_root, pin, 32);
  dev_free(layers, s);
  return rc;
}

int zp_commit(const zp_code* code, int eval_type, uint32_t D, const void* d_evals,
              uint64_t num_rows, uint32_t batch, zp_hint** hint_out, uint8_t h_root[32],
              zp_stream stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!code || !hint_out || !d_evals || !h_root)
    return fail(ZP_ERR_INVALID_PARAM, "null argument");
  *hint_out = nullptr;
  zp_hint* h = nullptr;
  ZP_TRY(hint_alloc(code, eval_type, D, num_rows, batch, s, &h));
  uint8_t* pin = pinned_scratch(32);
  int rc = pin ? ZP_OK : fail(ZP_ERR_OOM, "pinned staging allocation failed");
  if (rc == ZP_OK)
    rc = commit_enqueue(code, eval_type, D, d_evals, num_rows, batch, h, s, true, true);
  if (rc == ZP_OK && hipMemcpyAsync(pin, h->d_layers + (2 * h->N - 2) * 32, 32,
                                    hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "root copy failed");
  if (rc == ZP_OK && hipStreamSynchronize(s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "stream synchronize failed");
  if (rc == ZP_OK) std::memcpy(h_root, pin, 32);
  if (rc != ZP_OK) {
    zp_hint_free(h);
    return rc;
  }
  *hint_out = h;
  return ZP_OK;
}

namespace zpi {
namespace {

// The five ordering events of zp_commit_host, made once per host thread and
// device and reused by every later call (creating and destroying them took
// host time ahead of the first copy).
int host_commit_events(hipEvent_t (&ev)[5]) {
  static thread_local std::map<int, std::array<hipEvent_t, 5>> cache;
  int dev = 0;
  ZP_HIP(hipGetDevice(&dev));
  auto it = cache.find(dev);
  if (it == cache.end()) {
    std::array<hipEvent_t, 5> e{};
    for (auto& x : e)
      if (hipEventCreateWithFlags(&x, hipEventDisableTiming) != hipSuccess) {
        for (auto y : e)
          if (y) (void)hipEventDestroy(y);
        return fail(ZP_ERR_DEVICE, "event creation failed");
      }
    it = cache.emplace(dev, e).first;
  }
  for (int i = 0; i < 5; i++) ev[i] = it->second[i];
  return ZP_OK;
}

}  // namespace
}  // namespace zpi

int zp_commit_host(const zp_code* code, int eval_type, uint32_t D, const void* h_evals,
                   uint64_t num_rows, uint32_t batch, zp_hint** hint_out, uint8_t h_root[32],
                   zp_stream stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!code || !hint_out || !h_evals || !h_root)
    return fail(ZP_ERR_INVALID_PARAM, "null argument");
  *hint_out = nullptr;
  zp_hint* h = nullptr;
  ZP_TRY(hint_alloc(code, eval_type, D, num_rows, batch, s, &h));
  const uint64_t row_bytes = code->row_len * eval_bytes_of(eval_type, D);
  const uint64_t per_poly_in = num_rows * row_bytes, elem = h->elem, N = h->N;
  // Pageable input: rows in four groups per poly, group g copied on a
  // second stream while the earlier groups encode on `s` (0.86 against
  // 0.92 ms for copy-then-commit at 2^20 x 64 rows).  Pinned input: one copy
  // on `s`, which measured as fast as the split or faster (the split's
  // cross-stream waits and four smaller encoder launches cost what the
  // overlap saves).
  hipPointerAttribute_t attr{};
  const bool pinned = hipPointerGetAttributes(&attr, h_evals) == hipSuccess &&
                      attr.type == hipMemoryTypeHost;
  (void)hipGetLastError();  // pageable memory is reported as an error
  const bool split = !pinned && num_rows >= 4;
  const uint64_t groups = split ? 4 : 1;
  const uint64_t rows_g = (num_rows + groups - 1) / groups;
  uint8_t* d_in = nullptr;
  hipStream_t cs = s;
  hipEvent_t ev[5] = {};  // ev[0]: the input buffer is ready for the copy stream
  int rc = dev_alloc((void**)&d_in, batch * per_poly_in, s);
  if (rc == ZP_OK && split) rc = aux_stream(&cs);
  if (rc == ZP_OK && split) rc = zpi::host_commit_events(ev);
  if (rc == ZP_OK && split &&
      (hipEventRecord(ev[0], s) != hipSuccess || hipStreamWaitEvent(cs, ev[0], 0) != hipSuccess))
    rc = fail(ZP_ERR_DEVICE, "stream ordering failed");
  // Each event is recorded again only after `s` has enqueued its wait on
  // the previous record, so the wait sees the record it follows.
  for (uint32_t b = 0; rc == ZP_OK && b < batch; b++)
    for (uint64_t g = 0; rc == ZP_OK && g < groups; g++) {
      const uint64_t r0 = g * rows_g, nr = std::min(rows_g, num_rows - std::min(r0, num_rows));
      if (nr == 0) break;
      const uint64_t off = b * per_poly_in + r0 * row_bytes;
      hipEvent_t e = ev[1 + g];
      if (hipMemcpyAsync(d_in + off, (const uint8_t*)h_evals + off, nr * row_bytes,
                         hipMemcpyHostToDevice, cs) != hipSuccess ||
          (split && (hipEventRecord(e, cs) != hipSuccess ||
                     hipStreamWaitEvent(s, e, 0) != hipSuccess)))
        rc = fail(ZP_ERR_DEVICE, "input copy failed");
      if (rc == ZP_OK)
        rc = encode_rows_variant(code, eval_type, D, d_in + off, nr,
                                 h->d_cw + b * h->per_poly + r0 * N * elem, s);
    }
  if (rc == ZP_OK)
    rc = launch_leaf_hash(h->d_cw, num_rows * batch, N, (uint32_t)elem, h->d_layers, s,
                          code->tune_ctree_m);
  if (rc == ZP_OK) rc = launch_merkle_layers(h->d_layers, N, s);
  uint8_t* pin = rc == ZP_OK ? pinned_scratch(32) : nullptr;
  if (rc == ZP_OK && !pin) rc = fail(ZP_ERR_OOM, "pinned staging allocation failed");
  if (rc == ZP_OK && hipMemcpyAsync(pin, h->d_layers + (2 * N - 2) * 32, 32,
                                    hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "root copy failed");
  if (rc == ZP_OK && hipStreamSynchronize(s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "stream synchronize failed");
  if (rc != ZP_OK) {  // copies may still be in flight on either stream
    if (cs != s) (void)hipStreamSynchronize(cs);
    (void)hipStreamSynchronize(s);
  }
  if (rc == ZP_OK) std::memcpy(h_root, pin, 32);
  if (d_in) dev_free(d_in, s);
  if (rc != ZP_OK) {
    zp_hint_free(h);
    return rc;
  }
  *hint_out = h;
  return ZP_OK;
}

int zp_hint_info(const zp_hint* h, uint64_t* cw_bytes_per_poly, uint32_t* batch,
                 uint64_t* num_leaves, uint32_t* height) {
  if (!h) return fail(ZP_ERR_INVALID_PARAM, "null hint");
  if (cw_bytes_per_poly) *cw_bytes_per_poly = h->per_poly;
  if (batch) *batch = h->batch;
  if (num_leaves) *num_leaves = h->N;
  if (height) *height = h->height;
  return ZP_OK;
}

const void* zp_hint_codewords(const zp_hint* h) { return h ? h->d_cw : nullptr; }

int zp_hint_copy_codewords(const zp_hint* h, uint32_t poly, void* h_out, zp_stream stream) {
  if (!h || poly >= h->batch) return fail(ZP_ERR_INVALID_PARAM, "bad hint or poly index");
  if (!h_out) return fail(ZP_ERR_INVALID_PARAM, "null output");
  hipStream_t s = (hipStream_t)stream;
  ZP_HIP(hipMemcpyAsync(h_out, h->d_cw + poly * h->per_poly, h->per_poly, hipMemcpyDeviceToHost,
                        s));
  ZP_HIP(hipStreamSynchronize(s));
  return ZP_OK;
}

int zp_hint_copy_layer(const zp_hint* h, uint32_t level, uint8_t* h_out, zp_stream stream) {
  if (!h || level >= h->height) return fail(ZP_ERR_INVALID_PARAM, "bad hint or layer");
  if (!h_out) return fail(ZP_ERR_INVALID_PARAM, "null output");
  uint64_t off = 0, cnt = h->N;
  for (uint32_t l = 0; l < level; l++) {
    off += cnt;
    cnt /= 2;
  }
  hipStream_t s = (hipStream_t)stream;
  ZP_HIP(hipMemcpyAsync(h_out, h->d_layers + off * 32, cnt * 32, hipMemcpyDeviceToHost, s));
  ZP_HIP(hipStreamSynchronize(s));
  return ZP_OK;
}

int zp_merkle_prove(const zp_hint* h, uint64_t leaf, uint8_t* h_siblings, uint32_t* h_len,
                    zp_stream stream) {
  if (!h || leaf >= h->N) return fail(ZP_ERR_INVALID_PARAM, "Invalid leaf index");
  if (!h_siblings || !h_len) return fail(ZP_ERR_INVALID_PARAM, "null output");
  hipStream_t s = (hipStream_t)stream;
  uint64_t off = 0, cnt = h->N, idx = leaf;
  uint32_t k = 0;
  while (cnt > 1) {
    ZP_HIP(hipMemcpyAsync(h_siblings + 32 * k, h->d_layers + (off + (idx ^ 1)) * 32, 32,
                          hipMemcpyDeviceToHost, s));
    k++;
    off += cnt;
    cnt /= 2;
    idx /= 2;
  }
  ZP_HIP(hipStreamSynchronize(s));
  if (h_len) *h_len = k;
  return ZP_OK;
}

int zp_hint_open_column(const zp_hint* h, uint64_t col, void* h_out, zp_stream stream) {
  if (!h || col >= h->N) return fail(ZP_ERR_INVALID_PARAM, "bad hint or column");
  if (!h_out) return fail(ZP_ERR_INVALID_PARAM, "null output");
  hipStream_t s = (hipStream_t)stream;
  const uint64_t rows = h->num_rows * h->batch;
  ZP_HIP(hipMemcpy2DAsync(h_out, h->elem, h->d_cw + col * h->elem, h->N * h->elem, h->elem, rows,
                          hipMemcpyDeviceToHost, s));
  ZP_HIP(hipStreamSynchronize(s));
  return ZP_OK;
}

int zp_hint_record_stream(zp_hint* h, zp_stream stream) {
  if (!h) return fail(ZP_ERR_INVALID_PARAM, "null hint");
  hipStream_t s = (hipStream_t)stream;
  if (s == h->stream) return ZP_OK;
  for (hipStream_t u : h->used_on)
    if (u == s) return ZP_OK;
  h->used_on.push_back(s);
  return ZP_OK;
}

void zp_hint_free(zp_hint* h) {
  if (!h || !h->owned) return;
  dev_free_after(h->d_cw, h->stream, h->used_on.data(), h->used_on.size());
  dev_free_after(h->d_layers, h->stream, h->used_on.data(), h->used_on.size());
  delete h;
}

int zp_combined_row(uint32_t D, uint32_t L, uint64_t num_rows, uint64_t row_len, uint32_t batch,
                    const uint64_t* d_evals, const int64_t* d_alphas, const int64_t* d_coeffs,
                    uint64_t* d_out, zp_stream stream) {
  hipStream_t s = (hipStream_t)stream;
  if (D == 0 || D > 64) return fail(ZP_ERR_INVALID_PARAM, "degree bound must be within 1..64");
  if (L < 3 || L > 8) return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 3..8");
  if (batch == 0 || num_rows == 0 || row_len == 0) return fail(ZP_ERR_SHAPE, "empty input");
  if (!d_evals || !d_alphas || !d_coeffs || !d_out)
    return fail(ZP_ERR_INVALID_PARAM, "null buffer");
  void* scratch = nullptr;
  ZP_TRY(dev_alloc(&scratch, combined_row_scratch_bytes(D, row_len, batch, num_rows), s));
  // Stream-ordered: the scratch goes back to this stream's cache, where only
  // later work on the same stream can take it.
  const int rc = launch_combined_row(D, L, num_rows, row_len, batch, d_evals, d_alphas, d_coeffs,
                                     d_out, scratch, s);
  dev_free(scratch, s);
  return rc;
}

}  // extern "C"
