// common.h -- shared declarations of the zipgpu library (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

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
  void* d_afrag = nullptr;    // [bd/16][ksteps][3 slices][64 lanes][16 bytes]
  double* d_tw0f = nullptr;   // stage 0, [bd][7][8]
  double* d_tw1f = nullptr;   // stage 1, [8*bd][7][8] (depth 2)
  uint32_t mfma_ksteps = 0;   // ceil(bl / 64)
  // RAA
  uint32_t* d_perm1 = nullptr;
  uint32_t* d_perm2 = nullptr;
};

namespace zpi {

// Device memory cache (alloc.hip): blocks are reused only on the stream
// that freed them.
int dev_alloc(void** out, size_t bytes, hipStream_t s);
void dev_free(void* p, hipStream_t s);
void dev_release_cache();
uint8_t* pinned_scratch(size_t bytes);

// Tunables (set through zp_set_option / the autotuner).
extern int g_raa_wide;
extern int g_raa_planes;

// Host-side tables (params.cpp).
int iprs_params(uint64_t row_len, uint64_t depth, uint64_t rep, uint32_t* bl, uint32_t* bd);
uint64_t iprs_optimal_depth(uint64_t row_len);
void iprs_ftab(uint32_t bd, int32_t* out);
void iprs_twiddles(uint32_t bd, uint64_t N, uint32_t depth, uint32_t stage, int32_t* out);
void raa_perm(uint64_t n, uint64_t seed, uint32_t* out);
void iprs_afrag(uint32_t bd, uint32_t bl, uint32_t ksteps, const int32_t* ftab, int8_t* out);

// Kernel launchers (asynchronous on `s`).
int launch_iprs_binary(const zp_code* c, const uint64_t* evals, uint64_t rows, uint32_t D,
                       int64_t* out, hipStream_t s);
bool iprs_mfma_eligible(const zp_code* c, uint32_t D);
int launch_iprs_binary_mfma(const zp_code* c, const uint64_t* evals, uint64_t rows, uint32_t D,
                            int64_t* out, hipStream_t s);
int launch_iprs_i32(const zp_code* c, const int32_t* evals, uint64_t rows, void* out,
                    hipStream_t s);
int launch_raa_binary(const zp_code* c, const uint64_t* evals, uint64_t rows, uint32_t D,
                      int32_t* out, hipStream_t s);
int launch_leaf_hash(const uint8_t* rows, uint64_t nrows, uint64_t width, uint32_t elem,
                     uint8_t* leaves, hipStream_t s);
uint64_t leaf_chunks(uint64_t nrows, uint32_t elem);
int launch_chunk_cvs(const uint8_t* rows, uint64_t nrows, uint64_t width, uint32_t elem,
                     uint64_t ci_begin, uint64_t ci_end, uint32_t* cvs, uint8_t* leaves,
                     hipStream_t s);
int launch_cv_merge(uint32_t* cvs, uint32_t* tmp, uint64_t nchunks, uint64_t width,
                    uint8_t* leaves, hipStream_t s);
// A second stream (with events) per device for overlapping commit stages.
int aux_stream(hipStream_t* out);
int launch_merkle_layers(uint8_t* layers, uint64_t n, hipStream_t s);
int launch_combined_row(uint32_t D, uint32_t L, uint64_t num_rows, uint64_t row_len,
                        uint32_t batch, const uint64_t* evals, const int64_t* alphas,
                        const int64_t* coeffs, uint64_t* out, void* scratch, hipStream_t s);
uint64_t combined_row_scratch_bytes(uint32_t D, uint64_t row_len, uint32_t batch,
                                    uint64_t num_rows);

}  // namespace zpi
