// common.h -- shared declarations of the zipgpu library (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <string>
#include <vector>

#include "../../include/zipgpu.h"

namespace zpi {

// Records a message for zp_last_error() and returns `status`.
int fail(int status, const std::string& msg);

#define ZP_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::zpi::fail(ZP_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define ZP_TRY(expr)        \
  do {                      \
    int rc_ = (expr);       \
    if (rc_ != ZP_OK) return rc_; \
  } while (0)

// The deepest radix-8 recursion the fused IPRS kernel holds on chip.
constexpr uint32_t kMaxIprsDepth = 3;

}  // namespace zpi

// A linear code with its device-resident tables.
struct zp_code {
  int kind = 0;
  uint32_t rep = 0;
  uint64_t row_len = 0;
  uint64_t N = 0;  // codeword_len
  int device = 0;
  // IPRS (Radix8PnttParams)
  uint32_t depth = 0, bl = 0, bd = 0, bd_log2 = 0;
  int32_t* d_ftab = nullptr;                         // f(e) = norm(w_bd^e), e < bd
  int32_t* d_tw[zpi::kMaxIprsDepth] = {nullptr, nullptr, nullptr};  // [bd*8^k][7][8]
  // MFMA encoder tables (depth 1 or 2): int8 slices of the base matrix in
  // the v_mfma_i32_16x16x64_i8 operand layout, and stage twiddles as double.
  void* d_afrag = nullptr;    // [bd/16][ksteps][3 slices][64 lanes][16 bytes]; r-block rb
                              // holds rows r = rb + (bd/16) m, m = lane % 16 (binary rows)
  void* d_afrag_c = nullptr;  // the same with rows r = 16 rb + m (encode_wide, step 0)
  // Stage twiddles as double in butterfly order: entry (j-1)*4 + t of a
  // position holds the coefficient of input j in output t < 4 (outputs
  // t + 4 follow by symmetry); 32 entries per position, 28 used.
  double* d_tw0f = nullptr;   // stage 0, [bd][32]
  double* d_tw1f = nullptr;   // stage 1, [8*bd][32] (depth 2)
  // Per position, bit ((j-1)/2)*4 + t set where the odd-j coefficient of
  // output t + 4 is -(coefficient of t) - p instead of its negation.
  uint16_t* d_q0f = nullptr;  // stage 0, [bd]
  uint16_t* d_q1f = nullptr;  // stage 1, [8*bd]
  uint32_t mfma_ksteps = 0;   // ceil(bl / 64)
  // RAA
  uint32_t* d_perm1 = nullptr;
  uint32_t* d_perm2 = nullptr;
  // Kernel variants picked by the autotuner (-1 = not tuned: defaults).
  mutable int tune_iprs = -1;     // 0 matrix cores (rows spread), 1 VALU, 2 matrix cores (rows consecutive)
  mutable int tune_raa = -1;      // 0 full-row kernel, 1 plane groups (8), 2 plane groups (4)
  mutable int tune_overlap = -1;  // commit: 1 overlap encode and hash on two streams, 0 not
  // Chunks folded per lane by the chunk-tree kernel of this code's commits
  // (1, 2, 4 or 8; -1 = the shape heuristic).
  mutable int tune_ctree_m = -1;
  // The CHECKED build of the code (IprsCode<.., CHECKED>, code/iprs.rs:20;
  // RaaConfig::CHECK_FOR_OVERFLOWS, code/raa.rs:14): encodings whose exact
  // values leave the Cw type fail with ZP_ERR_OVERFLOW instead of wrapping.
  bool checked = false;
  // log2 of how much larger than the largest input any intermediate value
  // of an encoding can be (IPRS: bl * max|f| times, per stage, the largest
  // 1 + sum_j |twiddle|; RAA: N^2 over the two accumulates).
  double growth_log2 = 0;
};

// A commitment hint kept on the device (ZipPlusHint).
struct zp_hint {
  int device = 0;
  hipStream_t stream = nullptr;  // buffers return to the cache of this stream
  bool owned = true;             // false: buffers belong to a batch plan
  uint8_t* d_cw = nullptr;       // batch x num_rows x N elements, poly-major
  uint8_t* d_layers = nullptr;   // (2N-1) x 32 bytes, leaves first, root last
  uint64_t N = 0, num_rows = 0, elem = 0, per_poly = 0;
  uint32_t batch = 0, height = 0;
  // Other streams that used the buffers (zp_hint_record_stream): the free
  // keeps the blocks from the next user until their work has finished.
  std::vector<hipStream_t> used_on;
};

namespace zpi {

// Device memory cache (alloc.hip): blocks are reused only on the stream
// that freed them.  A thread-local arena, when set, serves requests from
// its fixed block instead (fails when it is exhausted), or, when it is
// measuring, lets the cache serve them and counts their rounded sizes.
struct Arena {
  uint8_t* base = nullptr;
  size_t size = 0, used = 0;
  bool measuring = false;
};
void set_arena(Arena* a);
int dev_alloc(void** out, size_t bytes, hipStream_t s);
void dev_free(void* p, hipStream_t s);
// dev_free of a block that was also used on the n_also streams `also`: the
// next allocation that takes it first waits for the work those streams held
// at this call.
void dev_free_after(void* p, hipStream_t s, const hipStream_t* also, size_t n_also);
void dev_release_cache();
uint8_t* pinned_scratch(size_t bytes);

// Raises kernel K's dynamic-LDS limit to `lds` bytes the first time a launch
// needs more than the 64 KiB default (not again afterwards, so that launches
// recorded into a graph make no attribute calls).
template <auto K>
int grant_lds(size_t lds) {
  static std::atomic<size_t> granted{64 * 1024};
  if (lds <= granted.load()) return ZP_OK;
  ZP_HIP(hipFuncSetAttribute((const void*)K, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  size_t g = granted.load();
  while (g < lds && !granted.compare_exchange_weak(g, lds)) {
  }
  return ZP_OK;
}

// Library-wide overrides of the kernel variants (zp_set_option): -1 = use
// the code's tuned or default choice.
extern int g_force_iprs, g_force_raa, g_force_overlap, g_force_wide, g_force_tree, g_force_ctree,
    g_force_comb;

// Host-side tables (params.cpp).
int iprs_params(uint64_t row_len, uint64_t depth, uint64_t rep, uint32_t* bl, uint32_t* bd);
uint64_t iprs_optimal_depth(uint64_t row_len);
void iprs_ftab(uint32_t bd, int32_t* out);
void iprs_twiddles(uint32_t bd, uint64_t N, uint32_t depth, uint32_t stage, int32_t* out);
void raa_perm(uint64_t n, uint64_t seed, uint32_t* out);
void iprs_afrag(uint32_t bd, uint32_t bl, uint32_t ksteps, const int32_t* ftab, int8_t* out,
                bool spread);

// Kernel launchers (asynchronous on `s`).
int launch_iprs_binary(const zp_code* c, const uint64_t* evals, uint64_t rows, uint32_t D,
                       int64_t* out, hipStream_t s);
bool iprs_mfma_eligible(const zp_code* c, uint32_t D);
// Bytes of bit-plane words the matrix-core encoder takes from the cache for
// `rows` rows (one allocation per launch).
uint64_t iprs_mfma_scratch_bytes(const zp_code* c, uint64_t rows, uint32_t D);
// spread: r-block rows bd/16 apart (d_afrag), else 16 consecutive rows
// (d_afrag_c); iprs_encoder variants 0 and 2.
// `mid` (or null) is recorded between the bit-plane pre-pass and the
// encoder proper.
int launch_iprs_binary_mfma(const zp_code* c, const uint64_t* evals, uint64_t rows, uint32_t D,
                            int64_t* out, hipStream_t s, bool spread = true,
                            hipEvent_t mid = nullptr);
int launch_iprs_i32(const zp_code* c, const int32_t* evals, uint64_t rows, void* out,
                    hipStream_t s);
bool wide_mfma_eligible(const zp_code* c);
int launch_encode_i32_mfma(const zp_code* c, const int32_t* evals, uint64_t rows, void* out,
                           hipStream_t s);
int launch_raa_binary(const zp_code* c, const uint64_t* evals, uint64_t rows, uint32_t D,
                      int32_t* out, hipStream_t s, int variant);
bool raa_full_row_eligible(const zp_code* c, uint32_t D);

// Commit core (api.hip): allocate a hint, and enqueue encode + Merkle on a
// stream without synchronising (the root stays in the hint's last layer).
int hint_alloc(const zp_code* code, int eval_type, uint32_t D, uint64_t num_rows, uint32_t batch,
               hipStream_t s, zp_hint** out);
// `timed` (zp_commit only): the encoder launches are bracketed by timing
// events when zp_encode_timing_start is in effect.
int commit_enqueue(const zp_code* code, int eval_type, uint32_t D, const void* d_evals,
                   uint64_t num_rows, uint32_t batch, zp_hint* h, hipStream_t s,
                   bool allow_overlap, bool timed = false);
// `mid` (or null) is recorded before the encoder's last kernel: after the
// bit-plane pre-pass on the IPRS matrix-core path, before the first launch
// on the others.
int encode_rows_variant(const zp_code* c, int eval_type, uint32_t D, const void* evals,
                        uint64_t num_rows, void* out, hipStream_t s, hipEvent_t mid = nullptr);
// The variant encode_rows_variant runs for an IPRS binary-poly code: true
// for the matrix-core encoder.
bool iprs_uses_mfma(const zp_code* c, uint32_t D);
// noalloc.hip: resolves the device address of the static scratch array of
// the current device (once per device).
int noalloc_init_device();
// wide.hip: LinearCode::encode_wide over Int<L>, L in 1..8.
int launch_encode_wide(const zp_code* c, uint32_t L, const uint64_t* in, uint64_t rows,
                       uint64_t* out, hipStream_t s);
int launch_encode_i64_planes(const zp_code* c, uint32_t D, uint32_t L, const int64_t* in,
                             uint64_t rows, void* out, hipStream_t s);
// The checked encoding of a checked code's rows (eval_type as
// encode_rows_variant): *done false means these inputs cannot overflow and
// the ordinary encoder is to run; true means `out` holds the encoding.
// ZP_ERR_OVERFLOW where the reference's checked build panics.
int encode_checked(const zp_code* c, int eval_type, uint32_t D, const void* in, uint64_t rows,
                   void* out, hipStream_t s, bool* done);
inline double code_growth_log2(const zp_code* c) { return c->growth_log2; }
// True when a checked code's encoding of these rows may need the checks
// that read a verdict on the host (so it cannot run under stream capture).
bool checked_may_sync(const zp_code* c, int eval_type, uint32_t D);
// prove.hip: column values and Merkle paths of n_cols columns of a hint
// (columns: n_cols x rows x elem bytes; paths: n_cols x (height-1) x 32).
int launch_gather_openings(const zp_hint* h, const uint64_t* d_cols, uint32_t n_cols,
                           void* d_columns, void* d_paths, hipStream_t s);
inline bool pow2(uint64_t x) { return x && !(x & (x - 1)); }
int check_types(const zp_code* c, int eval_type, uint32_t D);
int lc_budget_check(uint32_t rep, uint64_t row_len, int eval_type, uint32_t D, uint32_t L,
                    uint64_t batch);
uint64_t cw_elem_bytes_of(const zp_code* c, int eval_type, uint32_t D);
uint64_t eval_bytes_of(int eval_type, uint32_t D);
// ctree_m: chunks folded per lane in the chunk trees (1, 2, 4, 8), or -1
// for the shape heuristic.
int launch_leaf_hash(const uint8_t* rows, uint64_t nrows, uint64_t width, uint32_t elem,
                     uint8_t* leaves, hipStream_t s, int ctree_m = -1);
uint64_t leaf_chunks(uint64_t nrows, uint32_t elem);
int launch_chunk_cvs(const uint8_t* rows, uint64_t nrows, uint64_t width, uint32_t elem,
                     uint64_t ci_begin, uint64_t ci_end, uint32_t* cvs, uint8_t* leaves,
                     hipStream_t s);
int launch_cv_merge(uint32_t* cvs, uint32_t* tmp, uint64_t nchunks, uint64_t width,
                    uint8_t* leaves, hipStream_t s, int ctree_m = -1);
// The chunk-tree folds the quad kernel can take for a shape (empty when it
// does not apply): the tuner's candidates.
std::vector<int> ctree_fold_candidates(uint64_t nchunks);
// Chaining values (8 words each) of n full chunks of the stream that frames
// `entry_bytes`-byte entries of `data` as 0x6 || entry || 0x7, starting at
// stream offset q0 with chunk counter c0.
int launch_framed_chunk_cvs(const uint8_t* data, uint64_t entry_bytes, uint64_t q0, uint64_t c0,
                            uint64_t n, uint32_t* cvs, hipStream_t s);
// A second stream (with events) per device for overlapping commit stages.
int aux_stream(hipStream_t* out);
int launch_merkle_layers(uint8_t* layers, uint64_t n, hipStream_t s);
// One commitment's leaf hashing for launch_grouped_hash: nrows rows of
// `width` elements of `elem` bytes; leaves and layers at `layers`.
struct LeafJob {
  const uint8_t* rows;
  uint64_t nrows, width;
  uint32_t elem;
  uint8_t* layers;
  int ctree_m;
};
// Leaves and Merkle layers of every job in grouped launches.
int launch_grouped_hash(const std::vector<LeafJob>& jobs, hipStream_t s);
uint32_t ctree_fold(uint64_t nchunks, uint64_t width, int ctree_m);
int launch_combined_row(uint32_t D, uint32_t L, uint64_t num_rows, uint64_t row_len,
                        uint32_t batch, const uint64_t* evals, const int64_t* alphas,
                        const int64_t* coeffs, uint64_t* out, void* scratch, hipStream_t s);
uint64_t combined_row_scratch_bytes(uint32_t D, uint64_t row_len, uint32_t batch,
                                    uint64_t num_rows);

}  // namespace zpi
