// Optional per-kernel instrumentation: HIP events around each launch of a named kernel
// class, recorded on the launch stream, plus the algorithmic flops / bytes of that launch.
// Disabled by default (zero overhead beyond one branch per launch).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hssc {
namespace stats {

bool enabled();
// Returns an event pair slot to record around a launch (nullptr when disabled).
struct Slot;
Slot* begin(const char* name, hipStream_t st, double flops, double bytes);
void end(Slot* s, hipStream_t st);
// Reads the recorded launches now (call before destroying a stream they were recorded on).
void flush();

}  // namespace stats
}  // namespace hssc

// Brackets one kernel launch.
#define HSSC_TIMED(name, st, flops, bytes, launch)                        \
  do {                                                                    \
    ::hssc::stats::Slot* slot_ = ::hssc::stats::enabled()                 \
                                     ? ::hssc::stats::begin(name, st, flops, bytes) \
                                     : nullptr;                           \
    launch;                                                               \
    if (slot_) ::hssc::stats::end(slot_, st);                             \
  } while (0)
