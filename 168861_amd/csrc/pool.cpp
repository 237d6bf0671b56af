// Caching device allocator (see nrt_internal.hpp). Size classes: powers of two up
// to 2 MiB, then multiples of 2 MiB; one free list per (device, class).
#include <map>
#include <mutex>
#include <vector>
#include "nrt_internal.hpp"

namespace nrt {
namespace {

constexpr size_t kSmallLimit = 2u << 20;

size_t size_class(size_t bytes) {
  if (bytes <= 256) return 256;
  if (bytes <= kSmallLimit) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
  }
  return (bytes + kSmallLimit - 1) / kSmallLimit * kSmallLimit;
}

struct Pool {
  std::mutex mu;
  std::map<std::pair<int, size_t>, std::vector<void*>> free_;  // (device, class) -> blocks
  std::map<int, int64_t> cached;                               // device -> bytes
};

Pool& pool() {
  static Pool* p = new Pool;  // never destroyed: blocks outlive static teardown
  return *p;
}

int current_device() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}

}  // namespace

void* pool_alloc(size_t bytes) {
  const size_t c = size_class(bytes);
  const int dev = current_device();
  Pool& P = pool();
  {
    std::lock_guard<std::mutex> g(P.mu);
    auto it = P.free_.find({dev, c});
    if (it != P.free_.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      P.cached[dev] -= (int64_t)c;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pool_trim(dev);  // give the cached blocks back and try once more
    e = hipMalloc(&p, c);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(NRT_ERR_NOMEM, "hipMalloc of " + std::to_string(c) +
                                   " bytes failed: " + hipGetErrorString(e));
  }
  return p;
}

void pool_free(void* p, size_t bytes) {
  if (!p) return;
  const size_t c = size_class(bytes);
  hipPointerAttribute_t attr{};
  int dev = current_device();
  if (hipPointerGetAttributes(&attr, p) == hipSuccess) dev = attr.device;
  else (void)hipGetLastError();
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  P.free_[{dev, c}].push_back(p);
  P.cached[dev] += (int64_t)c;
}

int64_t pool_cached_bytes(int device) {
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.cached.find(device);
  return it == P.cached.end() ? 0 : it->second;
}

int64_t pool_trim(int device) {
  Pool& P = pool();
  std::vector<void*> blocks;
  int64_t n = 0;
  {
    std::lock_guard<std::mutex> g(P.mu);
    for (auto& kv : P.free_) {
      if (kv.first.first != device) continue;
      n += (int64_t)kv.first.second * (int64_t)kv.second.size();
      blocks.insert(blocks.end(), kv.second.begin(), kv.second.end());
      kv.second.clear();
    }
    P.cached[device] = 0;
  }
  for (void* p : blocks) (void)hipFree(p);
  return n;
}

}  // namespace nrt
