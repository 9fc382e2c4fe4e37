// Process-wide cache of device memory for the large pools.  The factor, member and
// coefficient pools of a context run to tens or hundreds of GB (C4-sized workloads:
// ~146 GB of FP64 factors), and memory returned to the driver is cleared before the
// driver hands it out again: a context created after another one was destroyed
// waited seconds in hipMalloc for that clearing.
//
// Reservations of at least kCacheMin bytes are therefore built from physical chunks
// (hipMemCreate) mapped back to back into a reserved virtual address range.  A
// released reservation is unmapped and its chunks stay with the process; the next
// reservation on the same device maps them again, whatever its size: the cached
// chunks that fit in what it still needs are taken largest first, then at most one
// that overshoots it by a bounded amount, then new chunks of at most kChunkMax bytes
// for the rest.  Reused chunks are cleared (the driver clears new ones) before the
// range is handed out.  When a new chunk cannot be created, the cached chunks not
// taken go back to the driver and the creation is tried once more.  A device without
// virtual memory management gets plain hipMalloc / hipFree, and so does every
// reservation when LAGRBF_DEVICE_CACHE=0 is set in the environment.
//
// Kernels and the runtime's linear copies and fills work across the chunks of a
// range; its pitched copies (hipMemcpy2D*) refuse a region that spans more than one
// chunk ("invalid argument"), so the library makes none on these ranges.
#include "ctx.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace lrb {
namespace {

constexpr size_t kCacheMin = size_t(64) << 20;
constexpr size_t kAlign = size_t(2) << 20;     // chunk sizes and range offsets
constexpr size_t kChunkMax = size_t(4) << 30;  // largest chunk created

struct Chunk {
  hipMemGenericAllocationHandle_t h;
  size_t bytes;
  int device;
};

struct Range {
  size_t bytes;                // mapped (sum of the chunks)
  int device;
  std::vector<Chunk> chunks;   // in address order
};

// never destroyed: contexts may be released during process teardown
struct State {
  std::mutex mu;
  std::vector<Chunk> free;           // cached chunks
  std::map<void*, Range> live;       // ranges handed out
  std::map<int, int> vmm;            // device -> 1 supported, 0 not
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

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// the lock held
bool vmm_supported(int dev) {
  auto it = st().vmm.find(dev);
  if (it != st().vmm.end()) return it->second != 0;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeVirtualMemoryManagementSupported, dev) != hipSuccess) v = 0;
  (void)hipGetLastError();
  st().vmm[dev] = v ? 1 : 0;
  return v != 0;
}

hipMemAllocationProp chunk_prop(int dev) {
  hipMemAllocationProp prop = {};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = dev;
  return prop;
}

// Return the cached chunks of `device` (all devices for -1) to the driver; the lock held.
void drop_cached(int device) {
  std::vector<Chunk> keep;
  for (const Chunk& c : st().free) {
    if (device >= 0 && c.device != device) { keep.push_back(c); continue; }
    (void)hipMemRelease(c.h);
  }
  st().free.swap(keep);
}

// Unmap and free a range; its chunks go to the cache.  The lock held.
void unmap_range(void* va, Range& r) {
  size_t off = 0;
  for (const Chunk& c : r.chunks) {
    (void)hipMemUnmap((char*)va + off, c.bytes);
    off += c.bytes;
  }
  (void)hipMemAddressFree(va, r.bytes);
  for (const Chunk& c : r.chunks) st().free.push_back(c);
  r.chunks.clear();
}

}  // namespace

hipError_t dev_malloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes < kCacheMin || !cache_on()) return hipMalloc(p, bytes);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(st().mu);
  if (!vmm_supported(dev)) return hipMalloc(p, bytes);
  std::vector<Chunk>& cache = st().free;
  const size_t need = round_up(bytes, kAlign);
  // cached chunks of this device, largest first
  std::vector<int> order;
  for (int i = 0; i < (int)cache.size(); i++)
    if (cache[i].device == dev) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return cache[a].bytes > cache[b].bytes; });
  std::vector<int> take;
  size_t got = 0;
  for (int i : order)
    if (got + cache[i].bytes <= need) { take.push_back(i); got += cache[i].bytes; }
  if (got < need) {
    // the smallest chunk left that covers the rest, if it overshoots by at most the rest
    // and kCacheMin
    const size_t rest = need - got;
    int best = -1;
    for (int i : order) {
      if (std::find(take.begin(), take.end(), i) != take.end()) continue;
      if (cache[i].bytes >= rest && cache[i].bytes <= 2 * rest + kCacheMin &&
          (best < 0 || cache[i].bytes < cache[best].bytes))
        best = i;
    }
    if (best >= 0) { take.push_back(best); got += cache[best].bytes; }
  }
  Range r;
  r.device = dev;
  for (int i : take) r.chunks.push_back(cache[i]);
  const size_t reused = got;
  std::sort(take.begin(), take.end(), [](int a, int b) { return a > b; });
  for (int i : take) cache.erase(cache.begin() + i);
  // new chunks for the rest
  const hipMemAllocationProp prop = chunk_prop(dev);
  bool dropped = false;
  while (got < need) {
    const size_t sz = std::min(kChunkMax, need - got);
    Chunk c{{}, sz, dev};
    e = hipMemCreate(&c.h, sz, &prop, 0);
    if (e != hipSuccess && !dropped) {
      (void)hipGetLastError();
      drop_cached(dev);
      dropped = true;
      e = hipMemCreate(&c.h, sz, &prop, 0);
    }
    if (e != hipSuccess) break;
    r.chunks.push_back(c);
    got += sz;
  }
  void* va = nullptr;
  if (e == hipSuccess) {
    r.bytes = got;
    e = hipMemAddressReserve(&va, got, kAlign, nullptr, 0);
  }
  size_t mapped = 0;
  if (e == hipSuccess) {
    for (const Chunk& c : r.chunks) {
      e = hipMemMap((char*)va + mapped, c.bytes, 0, c.h, 0);
      if (e != hipSuccess) break;
      mapped += c.bytes;
    }
  }
  if (e == hipSuccess) {
    hipMemAccessDesc ad = {};
    ad.location.type = hipMemLocationTypeDevice;
    ad.location.id = dev;
    ad.flags = hipMemAccessFlagsProtReadWrite;
    e = hipMemSetAccess(va, got, &ad, 1);
  }
  if (e == hipSuccess && reused > 0) {
    // reused chunks lead the range; handed out cleared, as the driver hands out new
    // memory.  The clearing runs on the null stream, which the contexts' non-blocking
    // streams do not wait for: it must end before the range is handed out.
    e = hipMemset(va, 0, reused);
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    size_t off = 0;
    for (const Chunk& c : r.chunks) {
      if (off < mapped) (void)hipMemUnmap((char*)va + off, c.bytes);
      off += c.bytes;
    }
    if (va) (void)hipMemAddressFree(va, got);
    for (const Chunk& c : r.chunks) cache.push_back(c);
    return e;
  }
  st().live[va] = std::move(r);
  *p = va;
  return hipSuccess;
}

void dev_free(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(st().mu);
    auto it = st().live.find(p);
    if (it != st().live.end()) {
      Range r = std::move(it->second);
      st().live.erase(it);
      // work still queued on the range must end before its chunks are mapped again
      int cur = -1;
      (void)hipGetDevice(&cur);
      if (cur != r.device) (void)hipSetDevice(r.device);
      (void)hipDeviceSynchronize();
      unmap_range(p, r);
      if (cur != r.device) (void)hipSetDevice(cur);
      return;
    }
  }
  (void)hipFree(p);
}

size_t dev_cached_bytes() {
  std::lock_guard<std::mutex> lk(st().mu);
  size_t s = 0;
  for (const Chunk& c : st().free) s += c.bytes;
  return s;
}

void dev_cache_release() {
  std::lock_guard<std::mutex> lk(st().mu);
  drop_cached(-1);
}

}  // namespace lrb
