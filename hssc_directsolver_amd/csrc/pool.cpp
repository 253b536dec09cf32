// Device memory cache (see pool.h).
#include "pool.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace hssc {
namespace pool {
namespace {

struct Device {
  std::unordered_map<void*, size_t> live;   // handed out: size of each block
  std::multimap<size_t, void*> ready;       // reusable now
  std::vector<std::pair<void*, size_t>> released;  // reusable after the next device sync
  size_t cached = 0;                        // bytes in ready + released
};

std::mutex mu;
int no_sync = 0;  // live NoSync scopes (guarded by mu)
// never destroyed: blocks outlive static destruction order (the runtime may be gone by then)
std::map<int, Device>& devices() {
  static auto* d = new std::map<int, Device>();
  return *d;
}

size_t cap_bytes() {
  static const size_t cap = [] {
    const char* v = std::getenv("HSSC_POOL_CAP_GB");
    const double gb = v ? std::atof(v) : 96.0;
    return size_t(gb > 0 ? gb * (1ull << 30) : 0);
  }();
  return cap;
}

// sizes are rounded so that requests of nearly the same size share blocks
size_t round_size(size_t b) {
  if (b == 0) b = 1;
  const size_t g = b >= (size_t(1) << 20) ? (size_t(2) << 20) : 4096;
  return (b + g - 1) / g * g;
}

Device& current() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return devices()[dev];
}

void make_ready_locked(Device& d) {
  if (d.released.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto& b : d.released) d.ready.emplace(b.second, b.first);
  d.released.clear();
}

void free_all_locked(Device& d) {
  make_ready_locked(d);
  for (auto& b : d.ready) (void)hipFree(b.second);
  d.ready.clear();
  d.cached = 0;
}

// HSSC_POOL_POISON=<byte> (diagnostics): every block handed out is first filled with that
// byte, e.g. 0x4E (about 1.6e70 as a double, 8.7e8 as a float), so that a kernel which reads
// device memory it has not written gives a result that differs from a run without it.
int poison_byte() {
  static const int v = [] {
    const char* e = std::getenv("HSSC_POOL_POISON");
    return e ? int(std::strtol(e, nullptr, 0)) & 0xff : -1;
  }();
  return v;
}

void poison(void* p, size_t s) {
  const int v = poison_byte();
  if (v < 0) return;
  // complete before the block is handed out: work that the caller then queues on streams that
  // do not synchronise with the null stream must not race with the fill
  (void)hipMemsetAsync(p, v, s, nullptr);
  (void)hipStreamSynchronize(nullptr);
}

// HSSC_POOL_GUARD=<KiB> (diagnostics): every block is allocated on its own with guard zones
// of that size before and after the requested bytes, filled with a pattern; a release waits
// for the device and reports to stderr any guard byte that changed (a write out of bounds of
// the block), then frees the block.  The cache is bypassed.
size_t guard_bytes() {
  static const size_t v = [] {
    const char* e = std::getenv("HSSC_POOL_GUARD");
    return e ? size_t(std::strtoull(e, nullptr, 0)) * 1024 : size_t(0);
  }();
  return v;
}
constexpr unsigned char kGuard = 0xA5;
struct Guarded {
  void* raw;
  size_t bytes;
  uint64_t serial;
};
std::unordered_map<void*, Guarded>& guarded_blocks() {
  static auto* m = new std::unordered_map<void*, Guarded>();
  return *m;
}

hipError_t guarded_alloc(void** p, size_t bytes) {
  static uint64_t serial = 0;
  const size_t g = guard_bytes(), rb = (bytes + 255) / 256 * 256;
  void* raw = nullptr;
  hipError_t e = hipMalloc(&raw, rb + 2 * g);
  if (e != hipSuccess) return e;
  (void)hipMemsetAsync(raw, kGuard, rb + 2 * g, nullptr);
  (void)hipStreamSynchronize(nullptr);
  *p = static_cast<char*>(raw) + g;
  guarded_blocks()[*p] = Guarded{raw, bytes, ++serial};
  return hipSuccess;
}

void guarded_release(void* p, const Guarded& b) {
  const size_t g = guard_bytes(), rb = (b.bytes + 255) / 256 * 256;
  (void)hipDeviceSynchronize();
  std::vector<unsigned char> front(g), back(g + rb - b.bytes);
  (void)hipMemcpy(front.data(), b.raw, g, hipMemcpyDeviceToHost);
  (void)hipMemcpy(back.data(), static_cast<char*>(p) + b.bytes, back.size(),
                  hipMemcpyDeviceToHost);
  size_t nf = 0, nb = 0, last_f = 0, first_b = 0, last_b = 0;
  for (size_t i = 0; i < g; ++i)
    if (front[i] != kGuard) ++nf, last_f = g - i;
  for (size_t i = 0; i < back.size(); ++i)
    if (back[i] != kGuard) {
      if (!nb) first_b = i;
      ++nb, last_b = i;
    }
  if (nf || nb)
    std::fprintf(stderr,
                 "[hssc pool guard] block #%llu of %zu bytes: %zu bytes written before it (up to "
                 "%zu before), %zu after it (from +%zu to +%zu past the end)\n",
                 (unsigned long long)b.serial, b.bytes, nf, last_f, nb, first_b, last_b);
  (void)hipFree(b.raw);
}

bool take_locked(Device& d, size_t s, void** p) {
  auto it = d.ready.find(s);
  if (it == d.ready.end()) return false;
  *p = it->second;
  d.ready.erase(it);
  d.cached -= s;
  d.live[*p] = s;
  poison(*p, s);
  return true;
}

}  // namespace

hipError_t alloc(void** p, size_t bytes) {
  *p = nullptr;
  const size_t s = round_size(bytes);
  std::lock_guard<std::mutex> g(mu);
  if (guard_bytes()) return guarded_alloc(p, bytes);
  Device& d = current();
  if (take_locked(d, s, p)) return hipSuccess;
  for (auto& b : d.released)
    if (b.second == s && no_sync == 0) {
      make_ready_locked(d);
      if (take_locked(d, s, p)) return hipSuccess;
      break;
    }
  hipError_t e = hipMalloc(p, s);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    free_all_locked(d);
    e = hipMalloc(p, s);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    return e;
  }
  d.live[*p] = s;
  poison(*p, s);
  return hipSuccess;
}

void release(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu);
  if (guard_bytes()) {
    auto it = guarded_blocks().find(p);
    if (it != guarded_blocks().end()) {
      const Guarded b = it->second;
      guarded_blocks().erase(it);
      guarded_release(p, b);
      return;
    }
  }
  Device& d = current();
  auto it = d.live.find(p);
  if (it == d.live.end()) {
    // not from this device's cache (another device current): free it directly
    for (auto& kv : devices())
      if (kv.second.live.count(p)) {
        kv.second.released.emplace_back(p, kv.second.live[p]);
        kv.second.cached += kv.second.live[p];
        kv.second.live.erase(p);
        return;
      }
    (void)hipFree(p);
    return;
  }
  d.released.emplace_back(p, it->second);
  d.cached += it->second;
  d.live.erase(it);
  if (d.cached > cap_bytes()) {
    // over the cap: free the largest reusable blocks first
    make_ready_locked(d);
    while (d.cached > cap_bytes() && !d.ready.empty()) {
      auto last = std::prev(d.ready.end());
      (void)hipFree(last->second);
      d.cached -= last->first;
      d.ready.erase(last);
    }
  }
}

NoSync::NoSync() {
  std::lock_guard<std::mutex> g(mu);
  ++no_sync;
}
NoSync::~NoSync() {
  std::lock_guard<std::mutex> g(mu);
  --no_sync;
}

void reclaim() {
  std::lock_guard<std::mutex> g(mu);
  make_ready_locked(current());
}

void trim() {
  std::lock_guard<std::mutex> g(mu);
  free_all_locked(current());
}

size_t cached_bytes() {
  std::lock_guard<std::mutex> g(mu);
  return current().cached;
}

}  // namespace pool
}  // namespace hssc
