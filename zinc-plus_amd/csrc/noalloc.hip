// noalloc.hip -- ZipPlus::commit for callers that cannot allocate.
//
// zp_commit_noalloc runs the commit of zp_commit (encode, chunk chaining
// values, chunk trees, Merkle layers) into buffers the caller owns, taking
// its transient scratch (bit-plane words of the matrix-core encoder, the
// chunk chaining values of multi-chunk leaves) from `scratch`, a static
// device array of kScratchBytes.  Each device that loads the library holds
// its own copy until the process ends.  Host code never passes the array's
// name to a kernel: that name is a host-side stand-in.  Its device address
// is resolved once per device with hipGetSymbolAddress, at zp_init or at the
// entry point's first call on that device, and only that address is used.
//
// Every call on a device shares that device's array, so calls that overlap
// on one device (two host threads, or two streams whose work overlaps)
// race on it.  Nothing in it needs to start at zero: every word a kernel
// reads was written earlier in the same call.
#include <mutex>

#include "common.h"

namespace zpi {
namespace {

constexpr uint64_t kScratchBytes = 64ull << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

int scratch_ptr(void** out) {
  int dev = 0;
  ZP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ZP_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    ZP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(ZP_ERR_DEVICE, "scratch symbol has no device address");
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return ZP_OK;
}

uint64_t round256(uint64_t n) { return (n + 255) / 256 * 256; }

}  // namespace

int noalloc_init_device() {
  void* p = nullptr;
  return scratch_ptr(&p);
}

}  // namespace zpi

using namespace zpi;

extern "C" int zp_commit_noalloc(const zp_code* code, int eval_type, uint32_t D,
                                 const void* d_evals, uint64_t num_rows, uint32_t batch,
                                 void* d_codewords, uint8_t* d_layers, zp_stream stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!code || !d_evals || !d_codewords || !d_layers)
    return fail(ZP_ERR_INVALID_PARAM, "null argument");
  if (eval_type != ZP_EVAL_BINPOLY)
    return fail(ZP_ERR_INVALID_PARAM, "zp_commit_noalloc takes BinaryPoly evals");
  if (batch == 0) return fail(ZP_ERR_SHAPE, "Batch must contain at least one polynomial");
  if (num_rows == 0) return fail(ZP_ERR_SHAPE, "num_rows must be positive");
  ZP_TRY(check_types(code, eval_type, D));
  // A checked IPRS code whose binary rows can overflow i64 (depth 3) is
  // checked by kernels whose result the host reads back: not here.
  if (checked_may_sync(code, eval_type, D))
    return fail(ZP_ERR_INVALID_PARAM,
                "zp_commit_noalloc takes no checked code whose encoding needs an overflow check");
  const uint64_t N = code->N;
  if (!pow2(N)) return fail(ZP_ERR_SHAPE, "codeword length must be a power of two");

  // The scratch this commit takes, in the order it is taken: the encoder's
  // bit-plane words (all rows of all polys in one launch), then the chunk
  // chaining values and the merge buffer of multi-chunk leaves.
  const uint64_t rows_all = num_rows * batch;
  const uint64_t elem = cw_elem_bytes_of(code, eval_type, D);
  uint64_t need = 0;
  if (code->kind == ZP_CODE_IPRS && iprs_uses_mfma(code, D))
    need += round256(iprs_mfma_scratch_bytes(code, rows_all, D));
  const uint64_t nchunks = leaf_chunks(rows_all, (uint32_t)elem);
  if (nchunks > 1)
    need += round256(nchunks * N * 32) + round256(((nchunks + 1) / 2) * N * 32);
  if (need > kScratchBytes)
    return fail(ZP_ERR_OOM, "zp_commit_noalloc needs " + std::to_string(need) +
                                " bytes of scratch; its static buffer holds " +
                                std::to_string(kScratchBytes));

  void* base = nullptr;
  ZP_TRY(scratch_ptr(&base));
  // Hint view of the caller's buffers (nothing here is freed).
  zp_hint h;
  h.device = code->device;
  h.stream = s;
  h.owned = false;
  h.d_cw = (uint8_t*)d_codewords;
  h.d_layers = d_layers;
  h.N = N;
  h.num_rows = num_rows;
  h.elem = elem;
  h.per_poly = num_rows * N * elem;
  h.batch = batch;
  h.height = (uint32_t)__builtin_ctzll(N) + 1;

  // Every dev_alloc below is served from the scratch array in order.
  Arena arena;
  arena.base = (uint8_t*)base;
  arena.size = kScratchBytes;
  set_arena(&arena);
  int rc = encode_rows_variant(code, eval_type, D, d_evals, rows_all, d_codewords, s);
  if (rc == ZP_OK)
    rc = launch_leaf_hash(h.d_cw, rows_all, N, (uint32_t)elem, d_layers, s, code->tune_ctree_m);
  if (rc == ZP_OK) rc = launch_merkle_layers(d_layers, N, s);
  set_arena(nullptr);
  return rc;
}
