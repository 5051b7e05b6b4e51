// server.hip -- continuous-batching commit/open service on one device.
//
// A request is one ZipPlus::commit (phase_commit.rs:61-140) followed, when
// its challenge arrives, by the prover's opening work on the kept hint: the
// combined row (phase_prove.rs:153-244) and the column openings with their
// Merkle paths (phase_prove.rs:247-304, merkle.rs:105-120).  Between the two
// the codewords, every Merkle layer and the evaluations stay in HBM.
//
// Worker slots each own a stream and a host thread.  A free slot takes the
// oldest pending opening first (openings release memory), else the oldest
// pending commit if its HBM footprint fits the budget; requests that arrive
// while others run are admitted as soon as a slot and memory are free, so
// their device work overlaps.  Every job ends with a synchronisation of its
// slot's stream, after which its result and timestamps are posted as an
// event.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"

using namespace zpi;

namespace zpi {
namespace {

using Clock = std::chrono::steady_clock;

double us_since(Clock::time_point t0, Clock::time_point t) {
  return std::chrono::duration<double, std::micro>(t - t0).count();
}

}  // namespace
}  // namespace zpi

struct zp_server {
  struct Request {
    zp_commit_desc desc{};
    const void* h_evals = nullptr;
    const void* d_evals = nullptr;  // caller's device input, or the server's copy
    bool input_owned = false;
    zp_hint* hint = nullptr;
    uint64_t footprint = 0;
    int state = 0;  // 0 queued, 1 committing, 2 committed, 3 open queued, 4 opening
    hipStream_t home = nullptr;  // stream whose cache holds the request's buffers
    Clock::time_point t_commit, t_open;
    // Opening.
    uint32_t comb_limbs = 0;
    std::vector<int64_t> alphas, coeffs;
    std::vector<uint64_t> cols;
    uint64_t* h_comb = nullptr;
    uint8_t* h_columns = nullptr;
    uint8_t* h_siblings = nullptr;
    // Proof (ZipPlus::prove) instead of an opening.
    bool prove = false;
    std::vector<int64_t> point;
    uint32_t num_openings = 0;
    uint8_t root[32] = {}, modulus[32] = {}, eval[32] = {};
    std::vector<uint8_t> proof;
  };
  int device = 0;
  uint64_t budget = 0, used = 0, peak = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_event;
  bool stop = false;
  uint64_t next_id = 1;
  std::map<uint64_t, Request> req;
  std::deque<uint64_t> commit_q, open_q;
  std::deque<zp_server_event> events;
  std::vector<std::thread> workers;
  std::vector<hipStream_t> streams;
  uint64_t jobs_done = 0, evals_committed = 0;
};

namespace {

uint64_t commit_scratch(const zp_commit_desc& d) {
  const zp_code* c = d.code;
  const uint64_t elem = cw_elem_bytes_of(c, d.eval_type, d.degree_plus_one);
  uint64_t planes = 0;
  if (c->kind == ZP_CODE_IPRS && d.eval_type == ZP_EVAL_BINPOLY &&
      iprs_mfma_eligible(c, d.degree_plus_one))
    planes = d.num_rows * (1ull << (3 * c->depth)) * 4 * c->mfma_ksteps * d.degree_plus_one * 4;
  const uint64_t nchunks = leaf_chunks(d.num_rows * d.batch, (uint32_t)elem);
  const uint64_t cvs = nchunks > 1 ? (nchunks + (nchunks + 1) / 2) * c->N * 32 : 0;
  return planes + cvs + (8u << 20);
}

uint64_t open_scratch(const zp_commit_desc& d, uint32_t n_cols) {
  const zp_code* c = d.code;
  const uint64_t elem = cw_elem_bytes_of(c, d.eval_type, d.degree_plus_one);
  const uint64_t comb = d.eval_type == ZP_EVAL_BINPOLY
                            ? combined_row_scratch_bytes(d.degree_plus_one, c->row_len, d.batch,
                                                         d.num_rows) +
                                  c->row_len * 8 * 8 + (uint64_t)d.batch * 64 * 16 +
                                  d.num_rows * 16
                            : 0;
  return comb + (uint64_t)n_cols * (d.num_rows * d.batch * elem + 64 * 32 + 8) + (8u << 20);
}

// Device scratch of zp_prove with up to n_open openings and CombR limbs.
uint64_t prove_scratch(const zp_commit_desc& d, uint32_t n_open, uint32_t limbs) {
  const zp_code* c = d.code;
  const uint64_t elem = cw_elem_bytes_of(c, d.eval_type, d.degree_plus_one);
  const uint32_t Dk = d.eval_type == ZP_EVAL_BINPOLY ? d.degree_plus_one : 32;
  const uint64_t rows = d.num_rows * d.batch;
  uint64_t b = rows * Dk * 32 + (c->row_len + 2 * d.num_rows) * 32 + (uint64_t)d.batch * Dk * 48 +
               d.num_rows * 16 + combined_row_scratch_bytes(Dk, c->row_len, d.batch, d.num_rows) +
               c->row_len * limbs * 8 + (uint64_t)n_open * (rows * elem + 24 + 64 * 32 + 8);
  if (d.eval_type != ZP_EVAL_BINPOLY) b += rows * c->row_len * 8;
  // Plane-sum partials (at most ~2048 workgroups' worth beyond one block
  // per row and plane chunk, 36 bytes per plane) and the q1 slice operand
  // (3 KiB per 64 columns).
  const uint64_t units = rows * ((Dk + 15) / 16);
  b += (units + 2048) * 16 * 36 + (c->row_len + 63) / 64 * 3072;
  return b + (8u << 20);
}

uint64_t resident_bytes(const zp_commit_desc& d) {
  const zp_code* c = d.code;
  const uint64_t elem = cw_elem_bytes_of(c, d.eval_type, d.degree_plus_one);
  const uint64_t in = (uint64_t)d.batch * d.num_rows * c->row_len * eval_bytes_of(d.eval_type, d.degree_plus_one);
  return d.batch * d.num_rows * c->N * elem + (2 * c->N - 1) * 32 + in;
}

void post(zp_server* sv, const zp_server_event& ev) {
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->events.push_back(ev);
  sv->jobs_done++;
  sv->cv_event.notify_all();
}

void release(zp_server* sv, zp_server::Request& r) {
  if (r.hint) {
    zp_hint_free(r.hint);
    r.hint = nullptr;
  }
  if (r.input_owned && r.d_evals) dev_free(const_cast<void*>(r.d_evals), r.home);
  r.d_evals = nullptr;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->used -= r.footprint;
  r.footprint = 0;
  sv->cv_work.notify_all();
}

int run_commit(zp_server* sv, zp_server::Request& r, hipStream_t s, uint8_t* pin) {
  const zp_commit_desc& d = r.desc;
  r.home = s;
  if (r.h_evals) {
    const uint64_t in = (uint64_t)d.batch * d.num_rows * d.code->row_len *
                        eval_bytes_of(d.eval_type, d.degree_plus_one);
    void* p = nullptr;
    ZP_TRY(dev_alloc(&p, in, s));
    r.d_evals = p;
    r.input_owned = true;
    ZP_HIP(hipMemcpyAsync(p, r.h_evals, in, hipMemcpyHostToDevice, s));
  }
  ZP_TRY(hint_alloc(d.code, d.eval_type, d.degree_plus_one, d.num_rows, d.batch, s, &r.hint));
  ZP_TRY(commit_enqueue(d.code, d.eval_type, d.degree_plus_one, r.d_evals, d.num_rows, d.batch,
                        r.hint, s, false));
  ZP_HIP(hipMemcpyAsync(pin, r.hint->d_layers + (2 * r.hint->N - 2) * 32, 32,
                        hipMemcpyDeviceToHost, s));
  ZP_HIP(hipStreamSynchronize(s));
  return ZP_OK;
}

int run_open(zp_server::Request& r, hipStream_t s) {
  const zp_commit_desc& d = r.desc;
  zp_hint* h = r.hint;
  const uint32_t n_cols = (uint32_t)r.cols.size();
  const uint64_t rows = h->num_rows * h->batch;
  const uint32_t levels = h->height - 1;
  std::vector<void*> tmp;
  auto alloc = [&](uint64_t bytes, void** p) {
    int rc = dev_alloc(p, bytes ? bytes : 1, s);
    if (rc == ZP_OK) tmp.push_back(*p);
    return rc;
  };
  int rc = ZP_OK;
  if (r.comb_limbs) {
    void *al = nullptr, *co = nullptr, *scr = nullptr, *out = nullptr;
    const uint64_t row_len = d.code->row_len;
    rc = alloc(r.alphas.size() * 8, &al);
    if (rc == ZP_OK) rc = alloc(r.coeffs.size() * 8, &co);
    if (rc == ZP_OK)
      rc = alloc(combined_row_scratch_bytes(d.degree_plus_one, row_len, d.batch, d.num_rows),
                 &scr);
    if (rc == ZP_OK) rc = alloc(row_len * r.comb_limbs * 8, &out);
    if (rc == ZP_OK &&
        (hipMemcpyAsync(al, r.alphas.data(), r.alphas.size() * 8, hipMemcpyHostToDevice, s) !=
             hipSuccess ||
         hipMemcpyAsync(co, r.coeffs.data(), r.coeffs.size() * 8, hipMemcpyHostToDevice, s) !=
             hipSuccess))
      rc = fail(ZP_ERR_DEVICE, "challenge upload failed");
    if (rc == ZP_OK)
      rc = launch_combined_row(d.degree_plus_one, r.comb_limbs, d.num_rows, row_len, d.batch,
                               (const uint64_t*)r.d_evals, (const int64_t*)al,
                               (const int64_t*)co, (uint64_t*)out, scr, s);
    if (rc == ZP_OK && hipMemcpyAsync(r.h_comb, out, row_len * r.comb_limbs * 8,
                                      hipMemcpyDeviceToHost, s) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "combined row copy failed");
  }
  if (rc == ZP_OK && n_cols) {
    void *dc = nullptr, *colout = nullptr, *pathout = nullptr;
    rc = alloc(n_cols * 8, &dc);
    if (rc == ZP_OK) rc = alloc(n_cols * rows * h->elem, &colout);
    if (rc == ZP_OK) rc = alloc((uint64_t)n_cols * levels * 32, &pathout);
    if (rc == ZP_OK &&
        hipMemcpyAsync(dc, r.cols.data(), n_cols * 8, hipMemcpyHostToDevice, s) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "column index upload failed");
    if (rc == ZP_OK)
      rc = launch_gather_openings(h, (const uint64_t*)dc, n_cols, colout, pathout, s);
    if (rc == ZP_OK &&
        (hipMemcpyAsync(r.h_columns, colout, n_cols * rows * h->elem, hipMemcpyDeviceToHost,
                        s) != hipSuccess ||
         (levels && hipMemcpyAsync(r.h_siblings, pathout, (uint64_t)n_cols * levels * 32,
                                   hipMemcpyDeviceToHost, s) != hipSuccess)))
      rc = fail(ZP_ERR_DEVICE, "opening copy failed");
  }
  if (hipStreamSynchronize(s) != hipSuccess && rc == ZP_OK)
    rc = fail(ZP_ERR_DEVICE, "stream synchronize failed");
  for (void* p : tmp) dev_free(p, s);
  return rc;
}

// ZipPlus::prove of a committed request on the worker's stream: transcript
// from the root (PcsProverTranscript::new_from_commitment), field config
// drawn from it, then zp_prove; the proof bytes stay with the request.
int run_prove(zp_server::Request& r, hipStream_t s) {
  const zp_commit_desc& d = r.desc;
  zp_transcript* t = nullptr;
  ZP_TRY(zp_transcript_create(&t));
  int rc = zp_transcript_absorb_slice(t, r.root, 32);
  if (rc == ZP_OK) rc = zp_transcript_random_field_cfg(t, r.modulus);
  if (rc == ZP_OK)
    rc = zp_prove(d.code, d.eval_type, d.degree_plus_one, r.comb_limbs, r.d_evals, d.num_rows,
                  d.batch, r.hint, r.point.data(), (uint32_t)(r.point.size() / 2), r.modulus,
                  r.num_openings, t, r.eval, s);
  if (rc == ZP_OK) {
    r.proof.resize(zp_transcript_stream_len(t));
    rc = zp_transcript_stream(t, r.proof.data(), r.proof.size());
  }
  zp_transcript_destroy(t);
  return rc;
}

void worker(zp_server* sv, uint32_t slot) {
  (void)hipSetDevice(sv->device);
  hipStream_t s = sv->streams[slot];
  uint8_t* pin = nullptr;
  if (hipHostMalloc((void**)&pin, 32, hipHostMallocDefault) != hipSuccess) pin = nullptr;
  for (;;) {
    uint64_t id = 0;
    bool is_open = false;
    zp_server::Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      for (;;) {
        if (sv->stop) break;
        if (!sv->open_q.empty()) {
          id = sv->open_q.front();
          sv->open_q.pop_front();
          is_open = true;
          break;
        }
        if (!sv->commit_q.empty()) {
          auto& head = sv->req[sv->commit_q.front()];
          // Admit the oldest commit when it fits; a request larger than the
          // whole budget runs alone once nothing else holds memory.
          if (sv->used + head.footprint <= sv->budget || sv->used == 0) {
            id = sv->commit_q.front();
            sv->commit_q.pop_front();
            sv->used += head.footprint;
            if (sv->used > sv->peak) sv->peak = sv->used;
            break;
          }
        }
        sv->cv_work.wait(lk);
      }
      if (sv->stop) break;
      r = &sv->req[id];
      r->state = is_open ? 4 : 1;
    }
    zp_server_event ev{};
    ev.request = id;
    ev.kind = is_open ? (r->prove ? 2 : 1) : 0;
    const Clock::time_point t_arrive = is_open ? r->t_open : r->t_commit;
    const Clock::time_point t_start = Clock::now();
    int rc = ZP_OK;
    if (!is_open) {
      rc = pin ? run_commit(sv, *r, s, pin) : fail(ZP_ERR_OOM, "pinned staging failed");
      if (rc == ZP_OK) {
        std::memcpy(ev.root, pin, 32);
        std::memcpy(r->root, pin, 32);
      }
    } else if (r->prove) {
      rc = run_prove(*r, s);
      ev.proof_bytes = r->proof.size();
    } else {
      rc = run_open(*r, s);
    }
    const Clock::time_point t_end = Clock::now();
    ev.status = rc;
    ev.queued_us = us_since(t_arrive, t_start);
    ev.service_us = us_since(t_start, t_end);
    ev.latency_us = us_since(t_arrive, t_end);
    if (rc != ZP_OK) {
      const char* msg = zp_last_error();
      std::strncpy(ev.message, msg, sizeof(ev.message) - 1);
    }
    // A failed commit or any opening ends the request: its memory returns.
    // A finished proof keeps only its bytes until zp_server_take_proof.
    if (is_open && r->prove && rc == ZP_OK) {
      release(sv, *r);
      std::lock_guard<std::mutex> lk(sv->mu);
      r->state = 5;
    } else if (is_open || rc != ZP_OK) {
      release(sv, *r);
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->req.erase(id);
    } else {
      std::lock_guard<std::mutex> lk(sv->mu);
      r->state = 2;
      sv->evals_committed += (uint64_t)r->desc.batch * r->desc.num_rows * r->desc.code->row_len;
    }
    post(sv, ev);
  }
  if (pin) (void)hipHostFree(pin);
}

}  // namespace

extern "C" {

int zp_server_create(const zp_server_config* cfg, zp_server** out) {
  if (!out) return fail(ZP_ERR_INVALID_PARAM, "null output");
  *out = nullptr;
  const uint32_t slots = cfg && cfg->slots ? cfg->slots : 4;
  if (slots > 16) return fail(ZP_ERR_INVALID_PARAM, "at most 16 slots");
  zp_server* sv = new zp_server();
  int rc = hipGetDevice(&sv->device) == hipSuccess ? ZP_OK : fail(ZP_ERR_DEVICE, "no device");
  if (rc == ZP_OK) {
    sv->budget = cfg ? cfg->hbm_budget : 0;
    if (sv->budget == 0) {
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
        rc = fail(ZP_ERR_DEVICE, "memory query failed");
      sv->budget = free_b / 4 * 3;
    }
  }
  for (uint32_t k = 0; rc == ZP_OK && k < slots; k++) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
      rc = fail(ZP_ERR_DEVICE, "stream creation failed");
    else
      sv->streams.push_back(st);
  }
  if (rc != ZP_OK) {
    for (auto s : sv->streams) (void)hipStreamDestroy(s);
    delete sv;
    return rc;
  }
  for (uint32_t k = 0; k < slots; k++) sv->workers.emplace_back(worker, sv, k);
  *out = sv;
  return ZP_OK;
}

int zp_server_commit(zp_server* sv, const zp_commit_desc* desc, const void* h_evals,
                     const void* d_evals, uint64_t* request) {
  if (!sv || !desc || !desc->code || !request) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  if (!h_evals == !d_evals) return fail(ZP_ERR_INVALID_PARAM, "give exactly one of h/d evals");
  if (desc->batch == 0) return fail(ZP_ERR_SHAPE, "Batch must contain at least one polynomial");
  if (desc->num_rows == 0) return fail(ZP_ERR_SHAPE, "num_rows must be positive");
  ZP_TRY(check_types(desc->code, desc->eval_type, desc->degree_plus_one));
  zp_server::Request r;
  r.desc = *desc;
  r.h_evals = h_evals;
  r.d_evals = d_evals;
  r.footprint = resident_bytes(*desc) +
                std::max({commit_scratch(*desc), open_scratch(*desc, 256),
                          prove_scratch(*desc, 256, 8)});
  r.t_commit = Clock::now();
  std::lock_guard<std::mutex> lk(sv->mu);
  const uint64_t id = sv->next_id++;
  sv->req.emplace(id, std::move(r));
  sv->commit_q.push_back(id);
  *request = id;
  sv->cv_work.notify_one();
  return ZP_OK;
}

int zp_server_open_sizes(zp_server* sv, uint64_t request, uint32_t comb_limbs, uint32_t n_cols,
                         uint64_t* comb_bytes, uint64_t* columns_bytes,
                         uint64_t* siblings_bytes) {
  if (!sv) return fail(ZP_ERR_INVALID_PARAM, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->req.find(request);
  if (it == sv->req.end()) return fail(ZP_ERR_INVALID_PARAM, "unknown request");
  const zp_commit_desc& d = it->second.desc;
  const uint64_t elem = cw_elem_bytes_of(d.code, d.eval_type, d.degree_plus_one);
  if (comb_bytes) *comb_bytes = d.code->row_len * comb_limbs * 8;
  if (columns_bytes) *columns_bytes = (uint64_t)n_cols * d.num_rows * d.batch * elem;
  if (siblings_bytes)
    *siblings_bytes = (uint64_t)n_cols * (uint64_t)__builtin_ctzll(d.code->N) * 32;
  return ZP_OK;
}

int zp_server_open(zp_server* sv, uint64_t request, uint32_t comb_limbs, const int64_t* h_alphas,
                   const int64_t* h_coeffs, const uint64_t* h_cols, uint32_t n_cols,
                   uint64_t* h_comb, uint8_t* h_columns, uint8_t* h_siblings) {
  if (!sv) return fail(ZP_ERR_INVALID_PARAM, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->req.find(request);
  if (it == sv->req.end()) return fail(ZP_ERR_INVALID_PARAM, "unknown request");
  zp_server::Request& r = it->second;
  if (r.state != 2) return fail(ZP_ERR_INVALID_PARAM, "request is not committed yet");
  const zp_commit_desc& d = r.desc;
  if (comb_limbs) {
    if (d.eval_type != ZP_EVAL_BINPOLY)
      return fail(ZP_ERR_INVALID_PARAM, "combined row needs binary polys");
    if (comb_limbs < 3 || comb_limbs > 8)
      return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 3..8");
    if (!h_alphas || !h_coeffs || !h_comb) return fail(ZP_ERR_INVALID_PARAM, "null buffer");
  }
  if (n_cols > 256) return fail(ZP_ERR_INVALID_PARAM, "at most 256 column openings");
  if (n_cols && (!h_cols || !h_columns || !h_siblings))
    return fail(ZP_ERR_INVALID_PARAM, "null buffer");
  for (uint32_t k = 0; k < n_cols; k++)
    if (h_cols[k] >= d.code->N) return fail(ZP_ERR_INVALID_PARAM, "Invalid leaf index");
  r.comb_limbs = comb_limbs;
  if (comb_limbs) {
    r.alphas.assign(h_alphas, h_alphas + 2ull * d.batch * d.degree_plus_one);
    r.coeffs.assign(h_coeffs, h_coeffs + 2 * d.num_rows);
  }
  r.cols.assign(h_cols, h_cols + n_cols);
  r.h_comb = h_comb;
  r.h_columns = h_columns;
  r.h_siblings = h_siblings;
  r.t_open = Clock::now();
  r.state = 3;
  sv->open_q.push_back(request);
  sv->cv_work.notify_one();
  return ZP_OK;
}

int zp_server_prove(zp_server* sv, uint64_t request, const int64_t* h_point, uint32_t num_vars,
                    uint32_t comb_limbs, uint32_t num_openings) {
  if (!sv) return fail(ZP_ERR_INVALID_PARAM, "null server");
  if (num_vars && !h_point) return fail(ZP_ERR_INVALID_PARAM, "null point");
  if (num_openings > 256) return fail(ZP_ERR_INVALID_PARAM, "at most 256 column openings");
  if (comb_limbs < 3 || comb_limbs > 8)
    return fail(ZP_ERR_INVALID_PARAM, "CombR limbs must be within 3..8");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->req.find(request);
  if (it == sv->req.end()) return fail(ZP_ERR_INVALID_PARAM, "unknown request");
  zp_server::Request& r = it->second;
  if (r.state != 2) return fail(ZP_ERR_INVALID_PARAM, "request is not committed yet");
  r.prove = true;
  r.point.assign(h_point, h_point + 2ull * num_vars);
  r.comb_limbs = comb_limbs;
  r.num_openings = num_openings;
  r.t_open = Clock::now();
  r.state = 3;
  sv->open_q.push_back(request);
  sv->cv_work.notify_one();
  return ZP_OK;
}

int zp_server_take_proof(zp_server* sv, uint64_t request, uint8_t* h_proof, uint64_t cap,
                         uint64_t* len, uint8_t h_modulus[32], uint8_t h_eval[32]) {
  if (!sv || !len) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->req.find(request);
  if (it == sv->req.end()) return fail(ZP_ERR_INVALID_PARAM, "unknown request");
  zp_server::Request& r = it->second;
  if (r.state != 5) return fail(ZP_ERR_INVALID_PARAM, "no finished proof for this request");
  *len = r.proof.size();
  if (cap < r.proof.size()) return fail(ZP_ERR_SHAPE, "buffer smaller than the proof");
  if (!h_proof && !r.proof.empty()) return fail(ZP_ERR_INVALID_PARAM, "null buffer");
  if (!r.proof.empty()) std::memcpy(h_proof, r.proof.data(), r.proof.size());
  if (h_modulus) std::memcpy(h_modulus, r.modulus, 32);
  if (h_eval) std::memcpy(h_eval, r.eval, 32);
  sv->req.erase(it);
  return ZP_OK;
}

int zp_server_poll(zp_server* sv, int64_t timeout_us, zp_server_event* ev) {
  if (!sv || !ev) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto ready = [&] { return !sv->events.empty(); };
  if (timeout_us < 0)
    sv->cv_event.wait(lk, ready);
  else if (!sv->cv_event.wait_for(lk, std::chrono::microseconds(timeout_us), ready))
    return ZP_ERR_TIMEOUT;
  *ev = sv->events.front();
  sv->events.pop_front();
  return ZP_OK;
}

int zp_server_stats(zp_server* sv, uint64_t* hbm_in_use, uint64_t* hbm_peak, uint64_t* budget,
                    uint64_t* evals_committed, uint64_t* pending) {
  if (!sv) return fail(ZP_ERR_INVALID_PARAM, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (hbm_in_use) *hbm_in_use = sv->used;
  if (hbm_peak) *hbm_peak = sv->peak;
  if (budget) *budget = sv->budget;
  if (evals_committed) *evals_committed = sv->evals_committed;
  if (pending) *pending = sv->commit_q.size() + sv->open_q.size();
  return ZP_OK;
}

void zp_server_destroy(zp_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->cv_work.notify_all();
  }
  for (auto& t : sv->workers) t.join();
  for (auto& kv : sv->req) release(sv, kv.second);
  for (auto s : sv->streams) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  delete sv;
}

}  // extern "C"
