/*
 * zipgpu.h -- C ABI of the MI355X (gfx950) Zip/Zip+ commitment library.
 *
 * This is the drop-in boundary for the reference's commitment hot path
 * (NethermindEth/zinc-plus, crate zip-plus).  Every entry point names the
 * reference interface it replaces.  All signatures use plain pointers and
 * sizes; "d_" pointers are device (HBM) pointers, "h_" pointers are host
 * pointers.  A stream argument of NULL means the default HIP stream.
 *
 * Element layouts are the reference's in-memory layouts with the `simd`
 * feature (poly/src/univariate.rs:41-50):
 *   BinaryPoly<D>        u64, bit c = coefficient c   (binary_u64.rs:26-28)
 *   i32 / i128           little-endian two's complement
 *   DensePolynomial<T,D> [T; D], coefficient 0 first  (dense.rs:30-34)
 *   Int<L>               L little-endian u64 limbs, two's complement
 * so codeword buffers are byte-identical to the reference's DenseRowMatrix
 * (row-major num_rows x codeword_len) and to the Merkle leaf bytes
 * (transcript/src/traits.rs:432-517).
 *
 * Functions return a zp_status.  Where the reference returns
 * Err(ZipError::InvalidPcsParam) the ABI returns ZP_ERR_INVALID_PARAM; where
 * it panics on an assert it returns ZP_ERR_SHAPE.  zp_last_error() gives the
 * message of the calling thread's last failure.
 *
 * By default arithmetic follows the reference's unchecked build (the
 * `unchecked` feature of its benches, CHECK = false): codeword and
 * combined-row values are exact modulo their types, i.e. they wrap where a
 * checked build would panic.  A code marked with zp_code_set_checked is the
 * checked build (IprsCode<.., CHECKED>, RaaConfig::CHECK_FOR_OVERFLOWS, as
 * the protocol's codes are, protocol/src/lib.rs:574-583): its encodings
 * return ZP_ERR_OVERFLOW exactly where the reference panics, and equal the
 * unchecked ones everywhere else.  The path's binary-poly, i32 and i64
 * codes cannot overflow at depth <= 2 (every new_with_optimal_depth code);
 * an explicit depth-3 IPRS code over binary polys, IPRS over
 * DensePolynomial<i64, D> with large coefficients, and encode_wide of large
 * Int<L> rows can.
 *
 * Thread safety: codes are immutable after creation and may be shared by any
 * number of threads; hints are owned by one caller at a time.
 */
#ifndef ZIPGPU_H
#define ZIPGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  ZP_OK = 0,
  ZP_ERR_INVALID_PARAM = 1, /* ZipError::InvalidPcsParam (zip-plus/src/lib.rs:13-29) */
  ZP_ERR_OVERFLOW = 2,      /* the CHECK "should not overflow" panics */
  ZP_ERR_SHAPE = 3,         /* assert! panics on shapes */
  ZP_ERR_DEVICE = 4,        /* HIP runtime failure */
  ZP_ERR_OOM = 5,           /* device allocation failed / admission refused */
  ZP_ERR_TIMEOUT = 6        /* zp_server_poll: no event within the timeout */
} zp_status;

typedef enum { ZP_CODE_IPRS = 0, ZP_CODE_RAA = 1, ZP_CODE_RAA_SIGNFLIP = 2 } zp_code_kind;

typedef enum {
  ZP_EVAL_I32 = 0,     /* Eval = i32, Cw = i128 (zip-plus/benches/zip_benches.rs:25-40) */
  ZP_EVAL_BINPOLY = 1, /* Eval = BinaryPoly<D>, Cw = [i64;D] (IPRS) or [i32;D] (RAA),
                          zip-plus/benches/zip_plus_benches.rs:37-78 */
  ZP_EVAL_I64 = 2,     /* Eval = i64, Cw = i128 (IntZipTypes, protocol/src/lib.rs:546-559) */
  ZP_EVAL_DENSE_I64 = 3 /* Eval = DensePolynomial<i64, D>, Cw = [i64;D] (IPRS) or
                           [Int<4>;D] (RAA) (ArbitraryPolyZipTypesIprs/Raa,
                           protocol/src/lib.rs:490-541) */
} zp_eval_type;

typedef struct zp_code zp_code;
typedef struct zp_hint zp_hint;
typedef void* zp_stream; /* hipStream_t */

/* Library set-up on a device (selects it for the calling thread). */
int zp_init(int device);
const char* zp_status_string(int status);
const char* zp_last_error(void);

/* Library-wide overrides of the kernel variants (no reference counterpart).
 * -1 (the initial value) leaves the choice to each code's tuned variant
 * (zp_autotune) or, for an untuned code, the default (variant 0 / 1 below).
 *   "iprs_encoder": 0 = matrix-core encoder where eligible, an r-block's
 *                   16 base rows bd/16 apart (default), 1 = VALU encoder
 *                   for every shape, 2 = matrix-core encoder with 16
 *                   consecutive base rows per r-block.
 *   "raa_kernel":   0 = full-row kernel where eligible (default),
 *                   1 = plane groups (8 planes per workgroup at codeword
 *                       length 8192), 2 = plane groups (4 planes at 8192).
 *   "commit_overlap": 1 = hash finished row groups on a second stream while
 *                   later rows encode, 0 = encode, then hash (default).
 *   "wide_base":    0 = IPRS base layer of Int<L>, i64 and i32 rows on the
 *                   matrix cores where the code has its int8 tables
 *                   (default), 1 = the per-element kernels.
 *   "merkle_tree":  0 = Merkle layers with one 4-lane quad per parent
 *                   (default), 1 = one thread per parent.
 *   "chunk_tree":   0 = a column's chunk tree (4 to 2048 chunks, a power
 *                   of two) folded in registers by up to 256 lanes and
 *                   merged by 4-lane quads (default), 1 = the whole tree
 *                   in LDS, or one launch per level, for every chunk count.
 *   "combined_row": 0 = the combined row's slice sums on the int8 matrix
 *                   cores (default, up to 2^24 rows), 1 = 64-bit VALU sums.
 * Returns ZP_ERR_INVALID_PARAM for an unknown name or value. */
int zp_set_option(const char* name, int64_t value);

/* Encoder timing inside commits.  No counterpart in the reference: for
 * profiling callers that want the encoder's device time as it runs inside
 * whole commits rather than in a loop of its own.
 * zp_encode_timing_start creates 3 * max_commits HIP timing events (no
 * system-scope fence) up front.  Then one zp_commit in `every` (the 1st,
 * the every+1-th, ...) among those that encode their rows in one sequence
 * on their stream takes the next three, while any are left, and records
 * them before its encoder, before the encoder's last kernel and after it.
 * For an IPRS binary-poly code on the matrix cores the last kernel is the
 * encoder proper, after the bit-plane pre-pass; other encoders record the
 * middle event at the start.  A timed commit creates nothing; it adds three
 * event records to its stream.
 * zp_encode_timing_stop ends the recording, waits for the recorded events,
 * and destroys them.  It returns the number of timed commits and two sums
 * in ms: of the whole encoder spans, and of the last-kernel spans.  The
 * last-kernel span is the tighter measure of a kernel: the stream is busy
 * with the pre-pass while the host submits the last kernel.  The first event
 * of a whole span is stamped as soon as the idle stream reaches it, so the
 * span also holds the host's submission of the first kernel.
 * Batch plans and the server are never timed.
 * ZP_ERR_INVALID_PARAM: start while on, stop while off, max_commits outside
 * 1..2^20, every = 0. */
int zp_encode_timing_start(uint32_t max_commits, uint32_t every);
int zp_encode_timing_stop(double* total_ms, double* last_ms, uint64_t* launches);

/* ---- LinearCode (zip-plus/src/code.rs:10-72) -------------------------- */

/* IprsCode::new(row_len, depth) (zip-plus/src/code/iprs.rs:30-37; errors as
 * Radix8PnttParams::new, code/iprs/pntt/radix8/params.rs:59-103), or
 * IprsCode::new_with_optimal_depth (iprs.rs:43-51) when depth < 0.
 * RaaCode::new(row_len) (zip-plus/src/code/raa.rs:38-90, perms from seeds 1
 * and 2) and RaaSignFlippingCode::new (code/raa_sign_flip.rs) ignore depth.
 * rep is the const REP (REPETITION_FACTOR). */
int zp_code_create(int kind, uint32_t rep, uint64_t row_len, int32_t depth, zp_code** out);
void zp_code_destroy(zp_code* code);
/* Marks a code checked (checked != 0) or unchecked (0, the default); the
 * CHECK const of IprsCode<Zt, Config, REP, CHECK> (code/iprs.rs:20) and
 * RaaConfig::CHECK_FOR_OVERFLOWS (code/raa.rs:14).  A checked code's
 * zp_encode_rows, zp_encode_wide, zp_commit and served commits return
 * ZP_ERR_OVERFLOW where the reference's mul_by_scalar::<true>(..).expect,
 * add! or neg! panics (iprs.rs:76-84, pntt/radix8.rs:150-157,
 * butterfly.rs:34-42, raa.rs:205-218, raa_sign_flip.rs:146-155); the
 * output buffer's contents are then unspecified.  Inputs that provably
 * cannot overflow (from the code's tables and, for integer rows, a scan of
 * the inputs' magnitudes) take the ordinary kernels; others take checked
 * kernels whose verdict the host reads, so such calls synchronize the
 * stream and fail under stream capture.  Set it before the code is shared
 * between threads. */
int zp_code_set_checked(zp_code* code, int checked);
int zp_code_is_checked(const zp_code* code);
/* LinearCode::row_len / codeword_len; depth is Radix8PnttParams::depth. */
uint64_t zp_code_row_len(const zp_code* code);
uint64_t zp_code_codeword_len(const zp_code* code);
uint32_t zp_code_depth(const zp_code* code);
int zp_code_kind_of(const zp_code* code);

/* validate_input's bit budget of the linear combinations, the assert of
 * zip-plus/src/pcs/utils.rs:24-54 that ZipPlus::commit (phase_commit.rs:
 * 70-77) and prove run before any work: ZP_ERR_SHAPE exactly where that
 * assert panics, for a code of repetition factor `rep` and row length
 * `row_len` (Lc::REPETITION_FACTOR, row_len), CombR = Int<comb_limbs>,
 * Chal = i128 and `batch` polys.  zp_commit carries no CombR, so a caller
 * that mirrors ZipPlus::commit calls this first; zp_prove applies it
 * itself.  Host only: needs no device. */
int zp_validate_input(uint32_t rep, uint64_t row_len, int eval_type, uint32_t degree_plus_one,
                      uint32_t comb_limbs, uint32_t batch);

/* Bytes of one codeword element for an eval type (16 for i128, 8*D or 4*D
 * for binary polys under IPRS / RAA). */
uint64_t zp_cw_elem_bytes(const zp_code* code, int eval_type, uint32_t degree_plus_one);

/* ZipPlus::encode_rows (zip-plus/src/pcs/phase_commit.rs:159-176), i.e.
 * LinearCode::encode (iprs.rs:120-130, raa.rs:147-149) over num_rows
 * consecutive rows of row_len evals.  d_out receives num_rows x
 * codeword_len elements, row-major.  Asynchronous on `stream`. */
int zp_encode_rows(const zp_code* code, int eval_type, uint32_t degree_plus_one,
                   const void* d_evals, uint64_t num_rows, void* d_out, zp_stream stream);

/* LinearCode::encode_wide (zip-plus/src/code.rs:55; iprs.rs:148-150,
 * raa.rs:151-153, raa_sign_flip.rs:97-99): the code applied to num_rows
 * rows of row_len CombR = Int<comb_limbs> values (comb_limbs little-endian
 * u64 limbs each, two's complement), arithmetic mod 2^(64*comb_limbs).
 * d_in: num_rows x row_len x comb_limbs u64; d_out: num_rows x
 * codeword_len x comb_limbs u64.  comb_limbs in 1..8.  Asynchronous on
 * `stream`. */
int zp_encode_wide(const zp_code* code, uint32_t comb_limbs, const uint64_t* d_in,
                   uint64_t num_rows, uint64_t* d_out, zp_stream stream);

/* ---- Merkle tree (zip-plus/src/merkle.rs) ----------------------------- */

/* MerkleTree::new(rows).root() (merkle.rs:74-102) over a device row matrix
 * of num_rows x width elements of elem_bytes each. */
int zp_merkle_root(const void* d_rows, uint64_t num_rows, uint64_t width, uint32_t elem_bytes,
                   uint8_t h_root[32], zp_stream stream);

/* ---- Commit (zip-plus/src/pcs/phase_commit.rs:61-101) ----------------- */

/* ZipPlus::commit(pp, polys) for `batch` polys of num_rows*row_len evals,
 * stored contiguously (poly-major) at d_evals.  The hint (ZipPlusHint,
 * pcs/structs.rs:114-129) stays on the device: codewords and every Merkle
 * layer.  h_root receives ZipPlusCommitment::root. */
int zp_commit(const zp_code* code, int eval_type, uint32_t degree_plus_one, const void* d_evals,
              uint64_t num_rows, uint32_t batch, zp_hint** hint_out, uint8_t h_root[32],
              zp_stream stream);

/* ZipPlus::commit(pp, polys) (phase_commit.rs:61-101) with the evaluations
 * in host memory, as the reference's callers hold them: the same commit as
 * zp_commit of the same bytes.  Pageable h_evals is copied to the device
 * in four row groups per poly (polys of 4 rows or more) on a second
 * stream, and each group is encoded on `stream` as soon as its rows have
 * landed, so the copy of one group overlaps the encoding of the groups
 * before it.  Pinned h_evals (hipHostMalloc) is copied once, on `stream`.
 * Returns when the root is back; h_evals may be reused then. */
int zp_commit_host(const zp_code* code, int eval_type, uint32_t degree_plus_one,
                   const void* h_evals, uint64_t num_rows, uint32_t batch, zp_hint** hint_out,
                   uint8_t h_root[32], zp_stream stream);

/* zp_commit without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device or pinned memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same commit as zp_commit for `batch` BinaryPoly<degree_plus_one>
 * polys (eval_type must be ZP_EVAL_BINPOLY), written into caller buffers:
 * d_codewords (batch x num_rows x codeword_len elements, as
 * zp_hint_codewords) and d_layers ((2 codeword_len - 1) x 32 bytes, leaves
 * first; the root is the last 32 bytes).  Fully asynchronous on `stream`;
 * no host synchronisation.  Transient scratch comes from one static device
 * array of 64 MiB per device (held until the process ends); a commit that
 * needs more (bit-plane words of the IPRS matrix-core encoder plus 1.5 x
 * 32 bytes per chunk and column of multi-chunk leaves) returns ZP_ERR_OOM
 * before any work.  Every call on a device shares that array: calls that
 * overlap on one device race and must be serialised by the caller.  A
 * checked IPRS code whose binary rows can overflow (depth 3) is refused
 * with ZP_ERR_INVALID_PARAM: its check reads a verdict on the host. */
int zp_commit_noalloc(const zp_code* code, int eval_type, uint32_t degree_plus_one,
                      const void* d_evals, uint64_t num_rows, uint32_t batch, void* d_codewords,
                      uint8_t* d_layers, zp_stream stream);

/* Hint accessors (the reference keeps these in host memory). */
int zp_hint_info(const zp_hint* hint, uint64_t* cw_bytes_per_poly, uint32_t* batch,
                 uint64_t* num_leaves, uint32_t* height);
const void* zp_hint_codewords(const zp_hint* hint); /* device pointer */
int zp_hint_copy_codewords(const zp_hint* hint, uint32_t poly, void* h_out, zp_stream stream);
/* Layer `level` of the tree (0 = leaves ... height-1 = root), 32 B each. */
int zp_hint_copy_layer(const zp_hint* hint, uint32_t level, uint8_t* h_out, zp_stream stream);
/* MerkleTree::prove(leaf) (merkle.rs:105-120, 224-257): siblings bottom-up. */
int zp_merkle_prove(const zp_hint* hint, uint64_t leaf, uint8_t* h_siblings, uint32_t* h_len,
                    zp_stream stream);
/* Column `col` over all rows of all polys (the bytes a column opening sends,
 * zip-plus/src/pcs/phase_prove.rs:247-304): batch*num_rows elements. */
int zp_hint_open_column(const zp_hint* hint, uint64_t col, void* h_out, zp_stream stream);
/* Records that work on `stream` uses the hint's device memory (reads through
 * zp_hint_codewords, or any call given a stream other than the commit's).
 * No reference counterpart: the reference's hint is host memory.  Without
 * it, zp_hint_free returns the memory to the commit stream's cache at once,
 * and the next commit on that stream may overwrite it while work on another
 * stream still reads it.  With it, the free records an event on each such
 * stream, and the next user of the memory waits for those events first
 * (the rule of torch's Tensor.record_stream). */
int zp_hint_record_stream(zp_hint* hint, zp_stream stream);
/* Releases the hint.  Its memory goes back to the library's cache of the
 * commit stream; work on other streams that reads it must be recorded with
 * zp_hint_record_stream, or have finished. */
void zp_hint_free(zp_hint* hint);

/* ---- Batched commitments ------------------------------------------------
 * A plan of independent commitments, each one ZipPlus::commit
 * (phase_commit.rs:61-140) with its own code, shape and inputs, as the
 * reference's callers issue them one by one (protocol/src/prover.rs:710-726,
 * zip-plus/benches/zip_common.rs:191-230).  The plan owns every hint (one
 * device pool) and a fixed scratch arena sized by its first commit; later
 * commits run from the arena, and with ZP_BATCH_GRAPH replay one HIP graph
 * of the whole batch (recaptured only when the input addresses change).
 * Results are those of zp_commit on each item. */

typedef struct zp_batch zp_batch;

typedef struct {
  const zp_code* code;
  int32_t eval_type;        /* zp_eval_type */
  uint32_t degree_plus_one; /* D of BinaryPoly<D> (ignored for i32) */
  uint64_t num_rows;
  uint32_t batch; /* polys under this one root (1 = commit_single) */
} zp_commit_desc;

enum { ZP_BATCH_GRAPH = 1, ZP_BATCH_AUTOTUNE = 2 };

/* flags: ZP_BATCH_GRAPH, ZP_BATCH_AUTOTUNE (run zp_autotune once per
 * distinct code at creation).  n_streams: branch streams (0 = 4).  A plan
 * holding a checked code whose overflow check reads a verdict on the host
 * (see zp_code_set_checked; mark codes before creating the plan) replays
 * its commits eagerly from the arena instead of as a graph. */
int zp_batch_create(const zp_commit_desc* items, uint32_t n_items, uint32_t flags,
                    uint32_t n_streams, zp_batch** out);
/* d_evals[i]: item i's evals (batch x num_rows x row_len, poly-major).
 * h_roots: n_items x 32 bytes.  Synchronous.  The hints of the previous
 * commit are overwritten; calls on one plan must not overlap. */
int zp_batch_commit(zp_batch* plan, const void* const* d_evals, uint8_t* h_roots,
                    zp_stream stream);
/* Item i's hint (valid until the next commit or destroy of the plan;
 * zp_hint_free on it does nothing). */
zp_hint* zp_batch_hint(zp_batch* plan, uint32_t item);
int zp_batch_info(const zp_batch* plan, uint64_t* hint_bytes, uint64_t* scratch_bytes,
                  uint32_t* graph_ready);
/* The plan's schedule (no reference counterpart): grouped = 1 when the
 * leaf hashing of all items runs in grouped launches after the encoders
 * (one persistent launch for every item's BLAKE3 chunks, one per
 * chunk-tree shape, one per band of 8 Merkle levels), 0 when each item's
 * hashing follows its encoder on its branch stream; streams = branch
 * streams in use.  With ZP_BATCH_AUTOTUNE the first commit times every
 * schedule (grouped or not, 2, 4 or 8 branches) on the plan's inputs and
 * keeps the fastest; otherwise it is grouped on min(n_streams, items). */
int zp_batch_schedule(const zp_batch* plan, uint32_t* grouped, uint32_t* streams);
/* Sets the schedule (streams 1..the plan's branch streams); the next commit
 * measures the scratch again and, with ZP_BATCH_GRAPH, captures anew.
 * Results do not depend on the schedule. */
int zp_batch_set_schedule(zp_batch* plan, uint32_t grouped, uint32_t streams);
void zp_batch_destroy(zp_batch* plan);

/* ---- Autotuner (no reference counterpart) --------------------------------
 * Times the kernel variants and tile parameters of `code` on synthetic
 * inputs of the given shape and stores the fastest in the code; later
 * encodes and commits use them: the RAA kernel (full row, 8- or 4-plane
 * groups), the chunk-tree fold of the leaf hashing (1, 2, 4 or 8 chunk
 * chaining values per lane before the quad merges), and the encode/hash
 * overlap.  IPRS binary codes take the matrix-core encoder wherever it
 * applies, with its r-block rows spread or consecutive as timed.  Must not
 * run concurrently with other uses of the code. */
int zp_autotune(zp_code* code, int eval_type, uint32_t degree_plus_one, uint64_t num_rows,
                uint32_t batch, zp_stream stream);
/* The tuned variants (-1 = not tuned), numbered as in zp_set_option. */
int zp_code_tuning(const zp_code* code, int32_t* iprs_encoder, int32_t* raa_kernel,
                   int32_t* commit_overlap, int32_t* chunk_tree_fold);

/* ---- Commit/open service ------------------------------------------------
 * Continuous batching on one device (no reference counterpart; it serves
 * the reference's per-request sequence ZipPlus::commit (phase_commit.rs:
 * 61-140), then on the challenge the combined row (phase_prove.rs:153-244)
 * and the column openings with Merkle paths (phase_prove.rs:247-304,
 * merkle.rs:105-120)).  `slots` worker streams run requests side by side; a
 * free slot takes the oldest pending opening, else the oldest pending commit
 * whose HBM footprint (codewords, Merkle layers, evaluations, scratch) fits
 * the budget.  Codewords, layers and evaluations stay in HBM from the commit
 * to the end of the opening.  Results arrive as events (zp_server_poll). */

typedef struct zp_server zp_server;

typedef struct {
  uint64_t hbm_budget; /* bytes the server may hold for requests (0 = 3/4 of free memory) */
  uint32_t slots;      /* requests on the device at once, 1..16 (0 = 4) */
} zp_server_config;

typedef struct {
  uint64_t request;
  int32_t kind;   /* 0 commit finished (root valid), 1 opening finished, 2 proof finished */
  int32_t status; /* zp_status of the job; a failed job ends the request */
  uint8_t root[32];
  double queued_us;  /* arrival (submission) to start on a slot */
  double service_us; /* start to results on the host */
  double latency_us; /* arrival to results on the host */
  char message[128]; /* zp_last_error of a failed job */
  uint64_t proof_bytes; /* kind 2: length of the proof (zp_server_take_proof) */
} zp_server_event;

int zp_server_create(const zp_server_config* cfg, zp_server** out);
/* Queue one commit of desc->batch polys.  Exactly one of h_evals (host,
 * copied to the device by the server) and d_evals (device; must stay valid
 * until the request ends) is given.  Returns at once with the request id. */
int zp_server_commit(zp_server* server, const zp_commit_desc* desc, const void* h_evals,
                     const void* d_evals, uint64_t* request);
/* Host buffer sizes of an opening (combined row, columns, Merkle paths). */
int zp_server_open_sizes(zp_server* server, uint64_t request, uint32_t comb_limbs,
                         uint32_t n_cols, uint64_t* comb_bytes, uint64_t* columns_bytes,
                         uint64_t* siblings_bytes);
/* Queue the opening of a committed request: the combined row in
 * Int<comb_limbs> (0 = none) for h_alphas (batch x D i128) and h_coeffs
 * (num_rows i128), and the n_cols (<= 256) columns h_cols with their paths.
 * h_comb: row_len x comb_limbs u64; h_columns: per column, batch*num_rows
 * elements; h_siblings: per column, log2(N) x 32 bytes bottom-up.  The
 * buffers must stay valid until the opening's event; the request's device
 * memory is released after it. */
int zp_server_open(zp_server* server, uint64_t request, uint32_t comb_limbs,
                   const int64_t* h_alphas, const int64_t* h_coeffs, const uint64_t* h_cols,
                   uint32_t n_cols, uint64_t* h_comb, uint8_t* h_columns, uint8_t* h_siblings);
/* Queue ZipPlus::prove::<MontyField<4>, _> of a committed request
 * (phase_prove.rs:85-304) instead of an opening: the server starts the
 * transcript from the request's root (PcsProverTranscript::
 * new_from_commitment), draws the field configuration from it
 * (get_random_field_cfg) and proves at the i128 point (num_vars x 2 words,
 * copied) with CombR = Int<comb_limbs> and num_openings (<= 256) column
 * openings, as zp_prove.  The event (kind 2) gives the proof's length; the
 * request's device memory is released after the proof. */
int zp_server_prove(zp_server* server, uint64_t request, const int64_t* h_point,
                    uint32_t num_vars, uint32_t comb_limbs, uint32_t num_openings);
/* The proof bytes, the field modulus and the canonical evaluation of a
 * finished proof; *len receives the proof length (ZP_ERR_SHAPE, with *len
 * set, when cap is smaller).  Ends the request. */
int zp_server_take_proof(zp_server* server, uint64_t request, uint8_t* h_proof, uint64_t cap,
                         uint64_t* len, uint8_t h_modulus[32], uint8_t h_eval[32]);
/* Next finished job; timeout_us < 0 waits without limit.  ZP_ERR_TIMEOUT
 * when none finished in time. */
int zp_server_poll(zp_server* server, int64_t timeout_us, zp_server_event* event);
int zp_server_stats(zp_server* server, uint64_t* hbm_in_use, uint64_t* hbm_peak,
                    uint64_t* budget, uint64_t* evals_committed, uint64_t* pending);
/* Stops the workers after their current job and releases every request. */
void zp_server_destroy(zp_server* server);

/* ---- Prover combined row (zip-plus/src/pcs/phase_prove.rs:153-244) ----- */

/* For `batch` binary polys (D = degree_plus_one) of num_rows x row_len evals:
 *   w'_b[k]  = sum of alpha_{b,c} over the set bits c of eval k
 *              (EvalDotChal, poly/src/univariate/binary_u64.rs:393-432)
 *   comb[col] = sum_b sum_j s_j * w'_b[j*row_len + col]
 * in CombR = Int<comb_limbs> (two's complement mod 2^(64*comb_limbs)).
 * d_alphas: batch x D i128; d_coeffs: num_rows i128 (the reference uses [1]
 * when num_rows == 1); d_out: row_len x comb_limbs u64.  Asynchronous on
 * `stream`: d_out is complete once the stream passes this call. */
int zp_combined_row(uint32_t degree_plus_one, uint32_t comb_limbs, uint64_t num_rows,
                    uint64_t row_len, uint32_t batch, const uint64_t* d_evals,
                    const int64_t* d_alphas, const int64_t* d_coeffs, uint64_t* d_out,
                    zp_stream stream);

/* ---- Fiat-Shamir transcript and proof stream ------------------------- */

/* PcsProverTranscript (zip-plus/src/pcs_transcript.rs:40-221): a
 * Blake3Transcript (transcript/src/lib.rs:10-72) plus the proof byte
 * stream.  Not thread-safe; one owner at a time. */
typedef struct zp_transcript zp_transcript;

/* Blake3Transcript::new with an empty stream.  PcsProverTranscript::
 * new_from_commitment(s) is this followed by zp_transcript_absorb_slice of
 * each root. */
int zp_transcript_create(zp_transcript** out);
int zp_transcript_clone(const zp_transcript* t, zp_transcript** out); /* Clone */
void zp_transcript_destroy(zp_transcript* t);

/* Transcript::absorb_slice (transcript/src/traits.rs:388-392). */
int zp_transcript_absorb_slice(zp_transcript* t, const uint8_t* bytes, uint64_t len);

/* absorb_slice of each of `count` consecutive entries of entry_bytes bytes:
 * absorb_public_columns of the protocol (protocol/src/lib.rs:305-316) for
 * one column of transcription bytes. */
int zp_transcript_absorb_entries(zp_transcript* t, const uint8_t* bytes, uint64_t entry_bytes,
                                 uint64_t count);

/* zp_transcript_absorb_entries with the entries in device memory (the same
 * interface, protocol/src/lib.rs:305-316, for columns that live in HBM):
 * the same bytes are absorbed and the transcript ends in the same state.
 * The full 1 KiB BLAKE3 chunks of the framed stream are hashed on the GPU
 * and only their chaining values come back; small streams are copied and
 * hashed on the host.  Enqueued on `stream` and synchronised before
 * returning. */
int zp_transcript_absorb_entries_device(zp_transcript* t, const void* d_bytes,
                                        uint64_t entry_bytes, uint64_t count, zp_stream stream);

/* Transcript::get_challenge::<T> with T::NUM_BYTES = num_bytes: the
 * little-endian bytes of T (lib.rs:48-55). */
int zp_transcript_challenge(zp_transcript* t, uint8_t* h_out, uint32_t num_bytes);

/* PcsProverTranscript::squeeze_challenge_idx (pcs_transcript.rs:31-36):
 * a u32 challenge mod cap; cap = 0 is ZP_ERR_SHAPE. */
int zp_transcript_squeeze_idx(zp_transcript* t, uint64_t cap, uint64_t* idx);

/* Transcript::get_random_field_cfg::<MontyField<4>, Uint<4>, MillerRabin>
 * (traits.rs:371-382, lib.rs:57-72): the 256-bit prime modulus,
 * little-endian. */
int zp_transcript_random_field_cfg(zp_transcript* t, uint8_t h_modulus[32]);

/* Appends raw bytes to the proof stream without absorbing them
 * (PcsProverTranscript::write_const_many of bytes). */
int zp_transcript_write(zp_transcript* t, const uint8_t* bytes, uint64_t len);

/* The proof bytes (PcsProverTranscript::stream). */
uint64_t zp_transcript_stream_len(const zp_transcript* t);
int zp_transcript_stream(const zp_transcript* t, uint8_t* h_out, uint64_t cap);

/* ---- Opening proof (zip-plus/src/pcs/phase_prove.rs:85-304) ----------- */

/* ZipPlus::prove::<MontyField<4>, _> of `batch` committed polys at an
 * i128 point (num_vars x 2 little-endian i64 words), with the hint of
 * their zp_commit, CombR = Int<comb_limbs> and num_openings column
 * openings (ZipTypes::NUM_COLUMN_OPENINGS).  Draws alphas, writes b,
 * draws s, writes the combined row and the openings into `t`, and returns
 * the evaluation <q_0, b> as its canonical value (32 bytes little-endian).
 * Errors: ZP_ERR_INVALID_PARAM for a point of the wrong length (the
 * reference's InvalidPcsParam) or a modulus that is even or below 2^128;
 * ZP_ERR_SHAPE for the linear-combination bit budget (validate_input's
 * assert) or a hint of another shape.  Synchronous. */
int zp_prove(const zp_code* code, int eval_type, uint32_t degree_plus_one, uint32_t comb_limbs,
             const void* d_evals, uint64_t num_rows, uint32_t batch, const zp_hint* hint,
             const int64_t* h_point, uint32_t num_vars, const uint8_t h_modulus[32],
             uint32_t num_openings, zp_transcript* t, uint8_t h_eval[32], zp_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* ZIPGPU_H */
