// alloc.hip -- device memory cache for hints and transient scratch.
//
// Blocks are cached per (device, stream).  A block freed on stream s is only
// handed out again to work on the same stream s, so stream order makes the
// reuse safe without synchronisation.  A block that was also used on other
// streams (dev_free_after) carries one event per such stream, recorded at
// the free; the next allocation that takes it makes its stream wait for
// those events first, so work on the other streams that was enqueued before
// the free finishes before the block is written again.  Sizes are rounded up
// (2 MiB granules above 1 MiB) so that repeated commits of one shape hit the
// cache.  When hipMalloc fails, the cache of that device is released (after
// a device synchronisation) and the allocation is retried once.
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "common.h"

namespace zpi {
namespace {

struct Key {
  int dev;
  hipStream_t stream;
  bool operator<(const Key& o) const {
    return dev != o.dev ? dev < o.dev : (uintptr_t)stream < (uintptr_t)o.stream;
  }
};

// A cached block and the events (recorded on other streams at its free)
// that the next user's stream must wait for.
struct Block {
  void* p;
  std::vector<hipEvent_t> waits;
};

std::mutex g_mu;
std::map<Key, std::multimap<size_t, Block>> g_free;
std::unordered_map<void*, std::pair<int, size_t>> g_size;  // every cached or live block

size_t round_size(size_t n) {
  if (n == 0) n = 1;
  const size_t g = n >= (1u << 20) ? (2u << 20) : 512;
  return (n + g - 1) / g * g;
}

void release_device_locked(int dev) {
  (void)hipDeviceSynchronize();
  for (auto it = g_free.begin(); it != g_free.end();) {
    if (it->first.dev != dev) {
      ++it;
      continue;
    }
    for (auto& kv : it->second) {
      for (hipEvent_t e : kv.second.waits) (void)hipEventDestroy(e);
      (void)hipFree(kv.second.p);
      g_size.erase(kv.second.p);
    }
    it = g_free.erase(it);
  }
}

}  // namespace

// A bump arena (set by the caller for the current thread) replaces the
// cache while a graph is being captured: no device allocation may happen
// then, and the graph must keep using the same scratch addresses.  A
// measuring arena counts the requests the cache serves.
static thread_local Arena* t_arena = nullptr;

void set_arena(Arena* a) { t_arena = a; }

int dev_alloc(void** out, size_t bytes, hipStream_t s) {
  *out = nullptr;
  if (t_arena) {
    const size_t want = (bytes + 255) / 256 * 256;
    if (t_arena->measuring) {  // pass through to the cache, count bytes
      t_arena->used += want;
    } else {
      if (t_arena->used + want > t_arena->size)
        return fail(ZP_ERR_OOM, "graph scratch arena exhausted");
      *out = t_arena->base + t_arena->used;
      t_arena->used += want;
      return ZP_OK;
    }
  }
  int dev = 0;
  ZP_HIP(hipGetDevice(&dev));
  const size_t want = round_size(bytes);
  std::lock_guard<std::mutex> lk(g_mu);
  auto& fl = g_free[Key{dev, s}];
  auto it = fl.lower_bound(want);
  if (it != fl.end() && it->first <= 2 * want) {
    Block blk = std::move(it->second);
    fl.erase(it);
    int rc = ZP_OK;
    for (hipEvent_t e : blk.waits) {
      if (rc == ZP_OK && hipStreamWaitEvent(s, e, 0) != hipSuccess)
        rc = fail(ZP_ERR_DEVICE, "stream wait on a freed block's event failed");
      (void)hipEventDestroy(e);
    }
    if (rc != ZP_OK) {  // keep the block out of circulation: its users may still run
      g_size.erase(blk.p);
      (void)hipDeviceSynchronize();
      (void)hipFree(blk.p);
      return rc;
    }
    *out = blk.p;
    return ZP_OK;
  }
  void* p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    release_device_locked(dev);
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      return fail(ZP_ERR_OOM, "device allocation of " + std::to_string(want) + " bytes failed");
    }
  }
  g_size[p] = {dev, want};
  *out = p;
  return ZP_OK;
}

void dev_free(void* p, hipStream_t s) { dev_free_after(p, s, nullptr, 0); }

void dev_free_after(void* p, hipStream_t s, const hipStream_t* also, size_t n_also) {
  if (!p) return;
  if (t_arena && !t_arena->measuring) return;  // arena memory is released with the arena
  Block blk{p, {}};
  for (size_t i = 0; i < n_also; i++) {
    if (also[i] == s) continue;
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(e, also[i]) != hipSuccess) {
      (void)hipGetLastError();
      if (e) (void)hipEventDestroy(e);
      (void)hipStreamSynchronize(also[i]);  // no event: wait for that stream here instead
      continue;
    }
    blk.waits.push_back(e);
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_size.find(p);
  if (it == g_size.end()) {
    for (hipEvent_t e : blk.waits) (void)hipEventDestroy(e);
    return;
  }
  g_free[Key{it->second.first, s}].emplace(it->second.second, std::move(blk));
}

void dev_release_cache() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(g_mu);
  release_device_locked(dev);
}

// One auxiliary non-blocking stream per device (default priority), used to
// run a commit's leaf hashing beside its encoding.
int aux_stream(hipStream_t* out) {
  static std::mutex mu;
  static std::map<int, hipStream_t> streams;
  int dev = 0;
  ZP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto it = streams.find(dev);
  if (it == streams.end()) {
    hipStream_t st;
    ZP_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    it = streams.emplace(dev, st).first;
  }
  *out = it->second;
  return ZP_OK;
}

// A per-thread pinned staging buffer for small device-to-host results.
uint8_t* pinned_scratch(size_t bytes) {
  static thread_local uint8_t* buf = nullptr;
  static thread_local size_t cap = 0;
  if (cap < bytes) {
    if (buf) (void)hipHostFree(buf);
    buf = nullptr;
    cap = 0;
    if (hipHostMalloc((void**)&buf, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    cap = bytes;
  }
  return buf;
}

}  // namespace zpi
