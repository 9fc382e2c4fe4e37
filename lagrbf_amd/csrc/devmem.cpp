// Process-wide cache of large device blocks.  The factor, member and coefficient
// pools of a context run to tens or hundreds of GB (C4-sized workloads: ~158 GB of
// FP64 factors), and memory returned to the driver is cleared before the driver
// hands it out again: a context created after another one was destroyed waited
// seconds in hipMalloc for that clearing.  Released blocks of at least
// kCacheMin bytes are therefore kept and handed, cleared, to the next reservation
// on the same device that they fit (best fit, at most twice the request); everything
// cached is returned to the driver when hipMalloc fails, and
// LAGRBF_DEVICE_CACHE=0 in the environment turns the cache off.
#include "ctx.hpp"

#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace lrb {
namespace {

constexpr size_t kCacheMin = size_t(64) << 20;

struct Block {
  void* p;
  size_t bytes;
  int device;
};

// never destroyed: contexts may be released during process teardown
struct State {
  std::mutex mu;
  std::vector<Block> free;              // cached blocks
  std::map<void*, Block> live;          // large blocks handed out
};
State& st() {
  static State* s = new State;
  return *s;
}

bool cache_on() {
  static const bool on = [] {
    const char* e = std::getenv("LAGRBF_DEVICE_CACHE");
    return !(e && (!std::strcmp(e, "0") || !std::strcmp(e, "off") || !std::strcmp(e, "false")));
  }();
  return on;
}

// Free every cached block of `device` (all devices for -1); the lock held.
void drop_cached(int device) {
  std::vector<Block>& g_free = st().free;
  int cur = -1;
  (void)hipGetDevice(&cur);
  std::vector<Block> keep;
  for (const Block& b : g_free) {
    if (device >= 0 && b.device != device) { keep.push_back(b); continue; }
    if (b.device != cur) (void)hipSetDevice(b.device);
    (void)hipFree(b.p);
    if (b.device != cur) (void)hipSetDevice(cur);
  }
  g_free.swap(keep);
}

}  // namespace

hipError_t dev_malloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes < kCacheMin || !cache_on()) return hipMalloc(p, bytes);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(st().mu);
  std::vector<Block>& g_free = st().free;
  std::map<void*, Block>& g_live = st().live;
  int best = -1;
  for (int i = 0; i < (int)g_free.size(); i++) {
    const Block& b = g_free[i];
    if (b.device != dev || b.bytes < bytes || b.bytes > 2 * bytes) continue;
    if (best < 0 || b.bytes < g_free[best].bytes) best = i;
  }
  if (best >= 0) {
    // handed out as the driver hands out fresh memory: cleared (~30 ms for 150 GB,
    // against seconds of waiting on the driver's own clearing)
    const Block b = g_free[best];
    e = hipMemset(b.p, 0, bytes);
    // the clearing runs on the null stream, which the contexts' non-blocking streams
    // do not wait for: it must end before the block is handed out
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    g_free.erase(g_free.begin() + best);
    g_live[b.p] = b;
    *p = b.p;
    return hipSuccess;
  }
  e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    drop_cached(dev);
    e = hipMalloc(p, bytes);
    if (e != hipSuccess) return e;
  }
  g_live[*p] = Block{*p, bytes, dev};
  return hipSuccess;
}

void dev_free(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(st().mu);
    std::map<void*, Block>& g_live = st().live;
    auto it = g_live.find(p);
    if (it != g_live.end()) {
      const Block b = it->second;
      g_live.erase(it);
      // work still queued on the block must end before another context reuses it
      int cur = -1;
      (void)hipGetDevice(&cur);
      if (cur != b.device) (void)hipSetDevice(b.device);
      (void)hipDeviceSynchronize();
      if (cur != b.device) (void)hipSetDevice(cur);
      st().free.push_back(b);
      return;
    }
  }
  (void)hipFree(p);
}

size_t dev_cached_bytes() {
  std::lock_guard<std::mutex> lk(st().mu);
  size_t s = 0;
  for (const Block& b : st().free) s += b.bytes;
  return s;
}

void dev_cache_release() {
  std::lock_guard<std::mutex> lk(st().mu);
  drop_cached(-1);
}

}  // namespace lrb
