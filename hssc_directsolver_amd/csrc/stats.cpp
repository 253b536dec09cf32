// Per-kernel instrumentation (see stats.h and hssc_stats_* in include/hssc.h).
#include "stats.h"

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hssc.h"

namespace hssc {
namespace stats {

struct Slot {
  std::string name;
  hipEvent_t start = nullptr, stop = nullptr;
  double flops = 0, bytes = 0;
};

namespace {
struct Totals {
  int64_t launches = 0;
  double flops = 0, bytes = 0, ms = 0;
};
std::mutex mu;
bool on = false;
std::vector<Slot*> pending;
std::vector<Slot*> pool;
std::map<std::string, Totals> totals;

// Reads the timings of the recorded launches.  The event calls can fail (an event recorded
// on a stream that has since been destroyed); such a launch is not counted, and the error
// it leaves is cleared so that it does not surface as the library's own last error (an
// error that was already pending is left in place).
void drain_locked() {
  const bool clean = hipPeekAtLastError() == hipSuccess;
  bool failed = false;
  for (Slot* s : pending) {
    float ms = 0;
    if (hipEventSynchronize(s->stop) == hipSuccess &&
        hipEventElapsedTime(&ms, s->start, s->stop) == hipSuccess) {
      Totals& t = totals[s->name];
      t.launches += 1;
      t.flops += s->flops;
      t.bytes += s->bytes;
      t.ms += ms;
    } else {
      failed = true;
    }
    pool.push_back(s);
  }
  pending.clear();
  if (failed && clean) (void)hipGetLastError();
}
}  // namespace

bool enabled() { return on; }

Slot* begin(const char* name, hipStream_t st, double flops, double bytes) {
  std::lock_guard<std::mutex> g(mu);
  Slot* s;
  if (!pool.empty()) {
    s = pool.back();
    pool.pop_back();
  } else {
    s = new Slot();
    if (hipEventCreate(&s->start) != hipSuccess || hipEventCreate(&s->stop) != hipSuccess) {
      delete s;
      return nullptr;
    }
  }
  s->name = name;
  s->flops = flops;
  s->bytes = bytes;
  (void)hipEventRecord(s->start, st);
  return s;
}

void flush() {
  std::lock_guard<std::mutex> g(mu);
  if (!pending.empty()) drain_locked();
}

void end(Slot* s, hipStream_t st) {
  (void)hipEventRecord(s->stop, st);
  std::lock_guard<std::mutex> g(mu);
  pending.push_back(s);
  if (pending.size() > 4096) drain_locked();
}

}  // namespace stats
}  // namespace hssc

extern "C" {

int hssc_stats_enable(int32_t on) {
  std::lock_guard<std::mutex> g(hssc::stats::mu);
  hssc::stats::on = on != 0;
  return HSSC_OK;
}

int hssc_stats_reset(void) {
  std::lock_guard<std::mutex> g(hssc::stats::mu);
  hssc::stats::drain_locked();
  hssc::stats::totals.clear();
  return HSSC_OK;
}

int hssc_stats_get(const char* kernel, int64_t* launches, double* flops, double* bytes,
                   double* ms) {
  std::lock_guard<std::mutex> g(hssc::stats::mu);
  hssc::stats::drain_locked();
  auto it = hssc::stats::totals.find(kernel ? kernel : "");
  const hssc::stats::Totals t = it == hssc::stats::totals.end() ? hssc::stats::Totals{} : it->second;
  if (launches) *launches = t.launches;
  if (flops) *flops = t.flops;
  if (bytes) *bytes = t.bytes;
  if (ms) *ms = t.ms;
  return HSSC_OK;
}

}  // extern "C"
