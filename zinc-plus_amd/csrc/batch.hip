// batch.hip -- batched commitment plans and the kernel-variant autotuner.
//
// A plan holds a list of independent commitments (each one the reference's
// ZipPlus::commit over its own polys, phase_commit.rs:61-140, with its own
// code and shape), their hints in one device pool, and a fixed scratch
// arena.  Commitments are spread over a few branch streams (largest first,
// each to the least loaded branch) so that small items overlap on the
// device.  The leaf hashing either follows each item's encoder on its
// branch, or (grouped, the default) runs for all items at once after the
// branches join: one persistent launch for every item's BLAKE3 chunks, one
// launch per chunk-tree shape and one per band of Merkle levels.  The first
// commit of a plan runs eagerly and measures the scratch it needs (with
// ZP_BATCH_AUTOTUNE it first times the schedules: grouped or per item, 2, 4
// or 8 branches, and keeps the fastest); later commits either run eagerly
// from the arena or, with ZP_BATCH_GRAPH, replay one HIP graph of the whole
// batch (captured again only when the input addresses change).  The roots
// are gathered into one device buffer and copied to the host once.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"

using namespace zpi;

struct zp_batch {
  int device = 0;
  uint32_t flags = 0;
  uint32_t n = 0;
  std::vector<zp_commit_desc> desc;
  std::vector<zp_hint> hints;
  std::vector<uint32_t> order;      // enqueue order (largest first)
  std::vector<uint32_t> branch_of;  // item -> branch stream
  std::vector<uint64_t> cost;       // codeword bytes per item
  uint32_t active = 0;              // branch streams in use (<= branch.size())
  bool grouped = true;              // leaf hashing in grouped launches
  std::vector<hipStream_t> branch;
  std::vector<hipEvent_t> ev_join;
  hipEvent_t ev_fork = nullptr;
  hipStream_t cap = nullptr;  // stream the graph is captured on
  uint8_t* d_pool = nullptr;  // every hint's codewords and layers
  uint64_t pool_bytes = 0;
  const uint8_t** d_rootptr = nullptr;  // item -> its root in the pool
  uint8_t* d_roots = nullptr;           // n x 32 gathered roots
  uint8_t* h_roots = nullptr;           // pinned copy of d_roots
  Arena arena;
  bool measured = false;
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> inputs;  // the input addresses the graph was captured with
  // A checked code whose overflow check reads a verdict on the host cannot
  // be captured: such a plan replays its commits eagerly.
  bool eager_only = false;
  bool schedule_set = false;  // tuned, or set by the caller
};

namespace zpi {
namespace {

__global__ void fill_kernel(uint64_t* out, uint64_t n_words, uint64_t mask, uint64_t seed) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words; i += stride) {
    uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    out[i] = (z ^ (z >> 31)) & mask;
  }
}

// One thread per 4-byte word of the n roots.
__global__ void gather_roots_kernel(const uint8_t* const* roots, uint32_t n, uint8_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * 8) return;
  ((uint32_t*)out)[i] = ((const uint32_t*)roots[i / 8])[i % 8];
}

void destroy_plan(zp_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  for (auto s : b->branch) (void)hipStreamDestroy(s);
  for (auto e : b->ev_join) (void)hipEventDestroy(e);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  (void)hipFree(b->d_pool);
  (void)hipFree(b->d_rootptr);
  (void)hipFree(b->d_roots);
  (void)hipFree(b->arena.base);
  if (b->h_roots) (void)hipHostFree(b->h_roots);
  delete b;
}

// Largest first, each to the least loaded of the first `k` branches.
void assign_branches(zp_batch* b, uint32_t k) {
  b->active = std::max(1u, std::min<uint32_t>(k, (uint32_t)b->branch.size()));
  std::vector<uint64_t> load(b->active, 0);
  b->branch_of.assign(b->n, 0);
  for (uint32_t i : b->order) {
    const uint32_t k2 = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
    b->branch_of[i] = k2;
    load[k2] += b->cost[i] + 1;
  }
}

// Encodes every poly of item i on stream st (the encoding half of
// commit_enqueue).
int encode_item(zp_batch* b, uint32_t i, const void* d_evals, hipStream_t st) {
  const zp_commit_desc& d = b->desc[i];
  const zp_hint& h = b->hints[i];
  const uint64_t in_stride =
      d.num_rows * d.code->row_len * eval_bytes_of(d.eval_type, d.degree_plus_one);
  for (uint32_t p = 0; p < d.batch; p++)
    ZP_TRY(encode_rows_variant(d.code, d.eval_type, d.degree_plus_one,
                               (const uint8_t*)d_evals + p * in_stride, d.num_rows,
                               h.d_cw + p * h.per_poly, st));
  return ZP_OK;
}

// Forks the branches from `s`, enqueues every commitment, joins, gathers.
int enqueue_all(zp_batch* b, const void* const* d_evals, hipStream_t s) {
  ZP_HIP(hipEventRecord(b->ev_fork, s));
  for (uint32_t k = 0; k < b->active; k++) ZP_HIP(hipStreamWaitEvent(b->branch[k], b->ev_fork, 0));
  int rc = ZP_OK;
  for (uint32_t i : b->order) {
    const zp_commit_desc& d = b->desc[i];
    hipStream_t st = b->branch[b->branch_of[i]];
    rc = b->grouped ? encode_item(b, i, d_evals[i], st)
                    : commit_enqueue(d.code, d.eval_type, d.degree_plus_one, d_evals[i],
                                     d.num_rows, d.batch, &b->hints[i], st, false);
    if (rc != ZP_OK) break;
  }
  // Join every branch even after an error, so that a capture can end.
  for (uint32_t k = 0; k < b->active; k++) {
    if (hipEventRecord(b->ev_join[k], b->branch[k]) != hipSuccess ||
        hipStreamWaitEvent(s, b->ev_join[k], 0) != hipSuccess)
      rc = rc == ZP_OK ? fail(ZP_ERR_DEVICE, "branch join failed") : rc;
  }
  if (rc != ZP_OK) return rc;
  if (b->grouped) {
    std::vector<LeafJob> jobs(b->n);
    for (uint32_t i = 0; i < b->n; i++) {
      const zp_commit_desc& d = b->desc[i];
      const zp_hint& h = b->hints[i];
      jobs[i] = {h.d_cw, d.num_rows * d.batch, h.N, (uint32_t)h.elem, h.d_layers,
                 d.code->tune_ctree_m};
    }
    ZP_TRY(launch_grouped_hash(jobs, s));
  }
  hipLaunchKernelGGL(gather_roots_kernel, dim3((b->n * 8 + 255) / 256), dim3(256), 0, s,
                     b->d_rootptr, b->n, b->d_roots);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

int capture(zp_batch* b, const void* const* d_evals) {
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  b->arena.used = 0;
  set_arena(&b->arena);
  ZP_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_all(b, d_evals, b->cap);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->cap, &g);
  set_arena(nullptr);
  if (rc == ZP_OK && e != hipSuccess) rc = fail(ZP_ERR_DEVICE, "graph capture failed");
  if (rc == ZP_OK && hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    b->exec = nullptr;
    rc = fail(ZP_ERR_DEVICE, "graph instantiation failed");
  }
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  if (rc == ZP_OK) b->inputs.assign(d_evals, d_evals + b->n);
  return rc;
}

template <class F>
int time_ms(hipStream_t s, int reps, F&& f, float* ms) {
  hipEvent_t a = nullptr, z = nullptr;
  ZP_HIP(hipEventCreate(&a));
  int rc = hipEventCreate(&z) == hipSuccess ? ZP_OK : fail(ZP_ERR_DEVICE, "event creation failed");
  if (rc == ZP_OK) rc = f();  // warm-up
  if (rc == ZP_OK && hipEventRecord(a, s) != hipSuccess) rc = fail(ZP_ERR_DEVICE, "event record");
  for (int r = 0; rc == ZP_OK && r < reps; r++) rc = f();
  if (rc == ZP_OK && (hipEventRecord(z, s) != hipSuccess || hipEventSynchronize(z) != hipSuccess ||
                      hipEventElapsedTime(ms, a, z) != hipSuccess))
    rc = fail(ZP_ERR_DEVICE, "event timing failed");
  if (rc == ZP_OK) *ms /= (float)reps;
  (void)hipEventDestroy(a);
  if (z) (void)hipEventDestroy(z);
  return rc;
}

// Times every schedule (grouped or per-item hashing, 2, 4 or 8 branches)
// eagerly on the plan's own inputs and keeps the fastest.  The runs write
// the plan's hints like any commit.
int tune_schedule(zp_batch* b, const void* const* d_evals, hipStream_t s) {
  float best = 0.f;
  bool best_grouped = b->grouped;
  uint32_t best_k = b->active;
  bool first = true;
  for (int g = 1; g >= 0; g--)
    for (uint32_t k : {2u, 4u, 8u}) {
      if (k > b->branch.size() && k != 2) continue;
      b->grouped = g == 1;
      assign_branches(b, k);
      float ms = 0.f;
      ZP_TRY(time_ms(s, 2, [&] { return enqueue_all(b, d_evals, s); }, &ms));
      if (first || ms < best) {
        best = ms;
        best_grouped = b->grouped;
        best_k = b->active;
        first = false;
      }
    }
  b->grouped = best_grouped;
  assign_branches(b, best_k);
  b->schedule_set = true;
  return ZP_OK;
}

}  // namespace
}  // namespace zpi

extern "C" {

int zp_autotune(zp_code* code, int eval_type, uint32_t D, uint64_t num_rows, uint32_t batch,
                zp_stream stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!code) return fail(ZP_ERR_INVALID_PARAM, "null code");
  // The variants are timed on synthetic rows with the ordinary kernels: a
  // checked code's overflow check does not depend on the variant, and the
  // synthetic values may well overflow.
  struct Unchecked {
    zp_code* c;
    bool was;
    ~Unchecked() { c->checked = was; }
  } unchecked{code, code->checked};
  code->checked = false;
  zp_hint* h = nullptr;
  ZP_TRY(hint_alloc(code, eval_type, D, num_rows, batch, s, &h));
  const uint64_t words = num_rows * batch * code->row_len * eval_bytes_of(eval_type, D) / 8;
  uint64_t* evals = nullptr;
  int rc = dev_alloc((void**)&evals, words * 8 + 8, s);
  if (rc == ZP_OK) {
    const uint64_t mask = eval_type == ZP_EVAL_I32 ? ~0ull : (D >= 64 ? ~0ull : (1ull << D) - 1);
    hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, s, evals, words + 1, mask,
                       0x5eedull);
    if (hipGetLastError() != hipSuccess) rc = fail(ZP_ERR_DEVICE, "fill launch failed");
  }
  const int kReps = 3;
  auto encode_all = [&] {
    return encode_rows_variant(code, eval_type, D, evals, num_rows * batch, h->d_cw, s);
  };
  auto tune = [&](int* field, const std::vector<int>& cands, auto&& run) {
    if (rc != ZP_OK || cands.size() < 2) {
      if (rc == ZP_OK && cands.size() == 1) *field = cands[0];
      return;
    }
    float best = 0.f;
    int best_v = cands[0];
    for (int v : cands) {
      *field = v;
      float ms = 0.f;
      rc = time_ms(s, kReps, run, &ms);
      if (rc != ZP_OK) return;
      if (v == cands[0] || ms < best) {
        best = ms;
        best_v = v;
      }
    }
    *field = best_v;
  };
  if (eval_type == ZP_EVAL_I64 || eval_type == ZP_EVAL_DENSE_I64) {
    // One encoder (the Int<L> kernels): nothing to tune.
  } else if (code->kind == ZP_CODE_IPRS) {
    // The matrix-core encoder wherever it applies (the VALU encoder is 14x
    // slower at every such shape), its r-blocks' rows spread (0) or
    // consecutive (2) as timed; otherwise the VALU encoder.
    if (eval_type == ZP_EVAL_BINPOLY && iprs_mfma_eligible(code, D))
      tune(&code->tune_iprs, {0, 2}, encode_all);
    else
      code->tune_iprs = 1;
  } else {
    std::vector<int> c;
    if (raa_full_row_eligible(code, D)) c.push_back(0);
    c.push_back(1);
    if (code->N == 8192) c.push_back(2);
    tune(&code->tune_raa, c, encode_all);
  }
  auto commit_all = [&] {
    return commit_enqueue(code, eval_type, D, evals, num_rows, batch, h, s, true);
  };
  // The chunk-tree fold of the leaf hashing (chunks per lane), then the
  // encode/hash overlap with that fold.
  const std::vector<int> folds = ctree_fold_candidates(
      leaf_chunks(num_rows * batch, (uint32_t)cw_elem_bytes_of(code, eval_type, D)));
  tune(&code->tune_ctree_m, folds, commit_all);
  tune(&code->tune_overlap, {0, 1}, commit_all);
  if (rc == ZP_OK && hipStreamSynchronize(s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "stream synchronize failed");
  dev_free(evals, s);
  zp_hint_free(h);
  return rc;
}

int zp_code_tuning(const zp_code* code, int32_t* iprs_encoder, int32_t* raa_kernel,
                   int32_t* commit_overlap, int32_t* chunk_tree_fold) {
  if (!code) return fail(ZP_ERR_INVALID_PARAM, "null code");
  if (iprs_encoder) *iprs_encoder = code->tune_iprs;
  if (raa_kernel) *raa_kernel = code->tune_raa;
  if (commit_overlap) *commit_overlap = code->tune_overlap;
  if (chunk_tree_fold) *chunk_tree_fold = code->tune_ctree_m;
  return ZP_OK;
}

int zp_batch_create(const zp_commit_desc* items, uint32_t n_items, uint32_t flags,
                    uint32_t n_streams, zp_batch** out) {
  if (!out) return fail(ZP_ERR_INVALID_PARAM, "null output");
  *out = nullptr;
  if (!items || n_items == 0) return fail(ZP_ERR_SHAPE, "empty batch");
  if (flags & ~(uint32_t)(ZP_BATCH_GRAPH | ZP_BATCH_AUTOTUNE))
    return fail(ZP_ERR_INVALID_PARAM, "unknown batch flags");
  zp_batch* b = new zp_batch();
  b->flags = flags;
  b->n = n_items;
  b->desc.assign(items, items + n_items);
  b->hints.resize(n_items);
  int rc = hipGetDevice(&b->device) == hipSuccess ? ZP_OK : fail(ZP_ERR_DEVICE, "no device");
  // Hint shapes (validated like zp_commit) and their places in the pool.
  std::vector<uint64_t>& cost = b->cost;
  cost.assign(n_items, 0);
  for (uint32_t i = 0; rc == ZP_OK && i < n_items; i++) {
    const zp_commit_desc& d = items[i];
    if (!d.code) {
      rc = fail(ZP_ERR_INVALID_PARAM, "null code in batch item " + std::to_string(i));
      break;
    }
    zp_hint& h = b->hints[i];
    if (d.batch == 0) rc = fail(ZP_ERR_SHAPE, "Batch must contain at least one polynomial");
    else if (d.num_rows == 0) rc = fail(ZP_ERR_SHAPE, "num_rows must be positive");
    if (rc == ZP_OK) rc = check_types(d.code, d.eval_type, d.degree_plus_one);
    if (rc != ZP_OK) break;
    if (checked_may_sync(d.code, d.eval_type, d.degree_plus_one)) b->eager_only = true;
    h.device = b->device;
    h.owned = false;
    h.N = d.code->N;
    h.num_rows = d.num_rows;
    h.elem = cw_elem_bytes_of(d.code, d.eval_type, d.degree_plus_one);
    h.per_poly = d.num_rows * h.N * h.elem;
    h.batch = d.batch;
    h.height = (uint32_t)__builtin_ctzll(h.N) + 1;
    cost[i] = h.per_poly * d.batch;
    b->pool_bytes += (cost[i] + 255) / 256 * 256 + ((2 * h.N - 1) * 32 + 255) / 256 * 256;
  }
  if (rc == ZP_OK && hipMalloc((void**)&b->d_pool, b->pool_bytes) != hipSuccess) {
    (void)hipGetLastError();
    b->d_pool = nullptr;
    rc = fail(ZP_ERR_OOM, "hint pool allocation of " + std::to_string(b->pool_bytes) +
                              " bytes failed");
  }
  std::vector<const uint8_t*> rootptr(n_items);
  if (rc == ZP_OK) {
    uint64_t off = 0;
    for (uint32_t i = 0; i < n_items; i++) {
      zp_hint& h = b->hints[i];
      h.d_cw = b->d_pool + off;
      off += (cost[i] + 255) / 256 * 256;
      h.d_layers = b->d_pool + off;
      off += ((2 * h.N - 1) * 32 + 255) / 256 * 256;
      rootptr[i] = h.d_layers + (2 * h.N - 2) * 32;
    }
  }
  // Largest first, each to the least loaded branch.  The autotuner may pick
  // up to 8 branches, so a tuned plan creates 8.
  const uint32_t want = n_streams ? n_streams : 4u;
  const uint32_t nb = std::max(1u, std::min(std::max(want, (flags & ZP_BATCH_AUTOTUNE) ? 8u : 1u),
                                            n_items));
  if (rc == ZP_OK) {
    b->order.resize(n_items);
    for (uint32_t i = 0; i < n_items; i++) b->order[i] = i;
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
  }
  for (uint32_t k = 0; rc == ZP_OK && k < nb; k++) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "stream creation failed");
    else
      b->branch.push_back(st);
    if (rc == ZP_OK && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "event creation failed");
    else if (rc == ZP_OK)
      b->ev_join.push_back(ev);
  }
  if (rc == ZP_OK) assign_branches(b, std::min(want, nb));
  if (rc == ZP_OK && (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
                      hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess))
    rc = fail(ZP_ERR_DEVICE, "stream/event creation failed");
  if (rc == ZP_OK && (hipMalloc((void**)&b->d_rootptr, n_items * sizeof(void*)) != hipSuccess ||
                      hipMalloc((void**)&b->d_roots, n_items * 32) != hipSuccess ||
                      hipHostMalloc((void**)&b->h_roots, n_items * 32, hipHostMallocDefault) !=
                          hipSuccess)) {
    (void)hipGetLastError();
    rc = fail(ZP_ERR_OOM, "root buffers allocation failed");
  }
  if (rc == ZP_OK && hipMemcpy(b->d_rootptr, rootptr.data(), n_items * sizeof(void*),
                               hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "root pointer upload failed");
  if (rc == ZP_OK && (flags & ZP_BATCH_AUTOTUNE)) {
    // One tuning run per distinct code, on its first item's shape.
    std::vector<const zp_code*> seen;
    for (uint32_t i = 0; rc == ZP_OK && i < n_items; i++) {
      const zp_commit_desc& d = items[i];
      if (std::find(seen.begin(), seen.end(), d.code) != seen.end()) continue;
      seen.push_back(d.code);
      rc = zp_autotune(const_cast<zp_code*>(d.code), d.eval_type, d.degree_plus_one, d.num_rows, d.batch,
                       (zp_stream)b->cap);
    }
  }
  if (rc != ZP_OK) {
    destroy_plan(b);
    return rc;
  }
  *out = b;
  return ZP_OK;
}

int zp_batch_commit(zp_batch* b, const void* const* d_evals, uint8_t* h_roots, zp_stream stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!b || !d_evals || !h_roots) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  int dev = -1;
  ZP_HIP(hipGetDevice(&dev));
  if (dev != b->device) return fail(ZP_ERR_INVALID_PARAM, "plan belongs to another device");
  for (uint32_t i = 0; i < b->n; i++)
    if (!d_evals[i]) return fail(ZP_ERR_INVALID_PARAM, "null input of batch item " +
                                                           std::to_string(i));
  int rc = ZP_OK;
  if (!b->measured) {
    if ((b->flags & ZP_BATCH_AUTOTUNE) && !b->schedule_set) rc = tune_schedule(b, d_evals, s);
    // First commit: eager, scratch from the cache, sizes counted.
    Arena m;
    m.measuring = true;
    set_arena(&m);
    if (rc == ZP_OK) rc = enqueue_all(b, d_evals, s);
    set_arena(nullptr);
    if (rc == ZP_OK && m.used) {  // an arena of size 0 refuses every request
      if (hipMalloc((void**)&b->arena.base, m.used) != hipSuccess) {
        (void)hipGetLastError();
        b->arena.base = nullptr;
        rc = fail(ZP_ERR_OOM, "batch scratch allocation failed");
      }
    }
    if (rc == ZP_OK) {
      b->arena.size = m.used;
      b->measured = true;
    }
  } else if ((b->flags & ZP_BATCH_GRAPH) && !b->eager_only) {
    if (!b->exec || !std::equal(b->inputs.begin(), b->inputs.end(), d_evals))
      rc = capture(b, d_evals);
    if (rc == ZP_OK && hipGraphLaunch(b->exec, s) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "graph launch failed");
  } else {
    // An eager-only plan takes its scratch from the cache: a checked
    // encoding's scratch depends on the inputs (the matrix-core fast path
    // or the checked kernels), so the first commit's measure need not hold.
    if (!b->eager_only) {
      b->arena.used = 0;
      set_arena(&b->arena);
    }
    rc = enqueue_all(b, d_evals, s);
    set_arena(nullptr);
  }
  if (rc == ZP_OK &&
      hipMemcpyAsync(b->h_roots, b->d_roots, b->n * 32, hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "root copy failed");
  if (rc == ZP_OK && hipStreamSynchronize(s) != hipSuccess)
    rc = fail(ZP_ERR_DEVICE, "stream synchronize failed");
  if (rc == ZP_OK) std::memcpy(h_roots, b->h_roots, b->n * 32);
  return rc;
}

zp_hint* zp_batch_hint(zp_batch* b, uint32_t item) {
  if (!b || item >= b->n) {
    fail(ZP_ERR_INVALID_PARAM, "bad plan or item");
    return nullptr;
  }
  return &b->hints[item];
}

int zp_batch_info(const zp_batch* b, uint64_t* hint_bytes, uint64_t* scratch_bytes,
                  uint32_t* graph_ready) {
  if (!b) return fail(ZP_ERR_INVALID_PARAM, "null plan");
  if (hint_bytes) *hint_bytes = b->pool_bytes;
  if (scratch_bytes) *scratch_bytes = b->arena.size;
  if (graph_ready) *graph_ready = b->exec ? 1u : 0u;
  return ZP_OK;
}

int zp_batch_schedule(const zp_batch* b, uint32_t* grouped, uint32_t* streams) {
  if (!b) return fail(ZP_ERR_INVALID_PARAM, "null plan");
  if (grouped) *grouped = b->grouped ? 1u : 0u;
  if (streams) *streams = b->active;
  return ZP_OK;
}

int zp_batch_set_schedule(zp_batch* b, uint32_t grouped, uint32_t streams) {
  if (!b) return fail(ZP_ERR_INVALID_PARAM, "null plan");
  if (grouped > 1 || streams == 0 || streams > b->branch.size())
    return fail(ZP_ERR_INVALID_PARAM, "grouped is 0 or 1 and streams 1.." +
                                          std::to_string(b->branch.size()));
  // The next commit measures the scratch again and captures a new graph.
  if (hipDeviceSynchronize() != hipSuccess) return fail(ZP_ERR_DEVICE, "synchronize failed");
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  (void)hipFree(b->arena.base);
  b->arena = Arena{};
  b->measured = false;
  b->grouped = grouped == 1;
  assign_branches(b, streams);
  b->schedule_set = true;
  return ZP_OK;
}

void zp_batch_destroy(zp_batch* b) {
  if (!b) return;
  (void)hipDeviceSynchronize();
  destroy_plan(b);
}

}  // extern "C"
