// Memory of the contexts and of the calls' temporary buffers, cached by size for the next owner
// in the same process.
//
// A solve creates and destroys contexts of the same shapes over and over (one per level of the
// coarse-to-fine ladder, one per solve of a batch of fresh problems), and hipMalloc / hipFree and
// hipHostMalloc / hipHostFree of ~50 buffers each time cost milliseconds.  Released blocks are kept
// on a free list per (device, size) and handed to the next request of that size on that device.
// An owner releases a block only after the stream that used it has drained (the context's
// destructor and CallBuffers synchronise first), so the next owner, on any stream, finds it idle.
// The cache holds at most kDeviceCacheBytes per device and kPinnedCacheBytes of pinned host memory;
// a release beyond that frees the block.  When an allocation fails, the cached blocks of that
// kind are freed and the allocation is tried once more.  Cached blocks are not freed at process
// exit (the runtime may already be gone by then); the driver reclaims them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace cot {

namespace {

constexpr size_t kDeviceCacheBytes = size_t(8) << 30;   // of 288 GB per MI355X
constexpr size_t kPinnedCacheBytes = size_t(512) << 20;
constexpr size_t kGranule = 512;

size_t round_up(size_t bytes) { return (std::max<size_t>(bytes, 1) + kGranule - 1) / kGranule * kGranule; }

struct Pool {
  std::mutex mu;
  // key: (device << 48) ^ size for device memory, size for pinned memory
  std::unordered_map<unsigned long long, std::vector<void*>> free_dev, free_pin;
  std::unordered_map<void*, unsigned long long> live_dev, live_pin;  // block -> key
  std::unordered_map<int, size_t> cached_dev;                        // device -> cached bytes
  size_t cached_pin = 0;
};

Pool& pool() {
  static Pool* p = new Pool();  // never destroyed: see above
  return *p;
}

unsigned long long dev_key(int dev, size_t bytes) {
  return ((unsigned long long)(unsigned)dev << 48) ^ (unsigned long long)bytes;
}

// Frees every cached block of device dev (dev < 0: the pinned cache).  Caller holds the lock.
void drop_cached(Pool& P, int dev) {
  if (dev < 0) {
    for (auto& kv : P.free_pin)
      for (void* q : kv.second) (void)hipHostFree(q);
    P.free_pin.clear();
    P.cached_pin = 0;
    return;
  }
  for (auto it = P.free_dev.begin(); it != P.free_dev.end();) {
    if ((int)(it->first >> 48) == dev) {
      for (void* q : it->second) (void)hipFree(q);
      it = P.free_dev.erase(it);
    } else {
      ++it;
    }
  }
  P.cached_dev[dev] = 0;
}

}  // namespace

// Device memory of `bytes` on the current device, or nullptr.
void* pool_alloc_device(size_t bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const size_t sz = round_up(bytes);
  const unsigned long long key = dev_key(dev, sz);
  Pool& P = pool();
  std::lock_guard<std::mutex> lock(P.mu);
  auto it = P.free_dev.find(key);
  if (it != P.free_dev.end() && !it->second.empty()) {
    void* q = it->second.back();
    it->second.pop_back();
    P.cached_dev[dev] -= sz;
    P.live_dev[q] = key;
    return q;
  }
  void* q = nullptr;
  if (hipMalloc(&q, sz) != hipSuccess) {
    (void)hipGetLastError();
    drop_cached(P, dev);
    q = nullptr;
    if (hipMalloc(&q, sz) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  P.live_dev[q] = key;
  return q;
}

void pool_free_device(void* q) {
  if (!q) return;
  Pool& P = pool();
  std::lock_guard<std::mutex> lock(P.mu);
  auto it = P.live_dev.find(q);
  if (it == P.live_dev.end()) return;  // not ours
  const unsigned long long key = it->second;
  P.live_dev.erase(it);
  const int dev = (int)(key >> 48);
  const size_t sz = (size_t)(key & ((1ull << 48) - 1));
  if (P.cached_dev[dev] + sz > kDeviceCacheBytes) {
    (void)hipFree(q);
    return;
  }
  P.free_dev[key].push_back(q);
  P.cached_dev[dev] += sz;
}

// Pinned host memory of `bytes`, or nullptr.
void* pool_alloc_pinned(size_t bytes) {
  const size_t sz = round_up(bytes);
  Pool& P = pool();
  std::lock_guard<std::mutex> lock(P.mu);
  auto it = P.free_pin.find(sz);
  if (it != P.free_pin.end() && !it->second.empty()) {
    void* q = it->second.back();
    it->second.pop_back();
    P.cached_pin -= sz;
    P.live_pin[q] = sz;
    return q;
  }
  void* q = nullptr;
  if (hipHostMalloc(&q, sz) != hipSuccess) {
    (void)hipGetLastError();
    drop_cached(P, -1);
    q = nullptr;
    if (hipHostMalloc(&q, sz) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  P.live_pin[q] = sz;
  return q;
}

void pool_free_pinned(void* q) {
  if (!q) return;
  Pool& P = pool();
  std::lock_guard<std::mutex> lock(P.mu);
  auto it = P.live_pin.find(q);
  if (it == P.live_pin.end()) return;
  const size_t sz = (size_t)it->second;
  P.live_pin.erase(it);
  if (P.cached_pin + sz > kPinnedCacheBytes) {
    (void)hipHostFree(q);
    return;
  }
  P.free_pin[sz].push_back(q);
  P.cached_pin += sz;
}

}  // namespace cot
