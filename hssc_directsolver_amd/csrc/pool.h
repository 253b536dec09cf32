// Device memory cache of the library.  A tree build allocates and releases tens of GB of
// workspace (the least-squares blocks of every level); going to hipMalloc / hipFree for each
// build costs up to a second per call once the device's memory has been cycled.  Released
// blocks are kept per device and handed out again for requests of the same rounded size.
//
// A released block may still be read or written by work in flight, so it is reused only after
// a device-wide synchronisation that follows its release (hipFree gave the same guarantee by
// synchronising itself).  reclaim() performs that synchronisation at a point where the device
// is expected to be idle (the start of a tree build); alloc() falls back to it when only
// released blocks of the requested size are left.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace hssc {
namespace pool {

// hipMalloc semantics (hipErrorOutOfMemory after releasing the whole cache and retrying).
hipError_t alloc(void** p, size_t bytes);
// Returns a block obtained from alloc (nullptr is ignored).
void release(void* p);
// Synchronises the device and makes every released block of the current device reusable.
void reclaim();
// Synchronises the device and frees every cached block of the current device.
void trim();
// Bytes held in the cache (released, not yet freed) on the current device.
size_t cached_bytes();

// While a NoSync object lives, alloc() never synchronises the device to reuse released blocks
// (it takes a ready block or calls hipMalloc): code that queues work on its own streams and
// keeps allocating meanwhile (the pipelined tree build) is not stalled behind that work.
struct NoSync {
  NoSync();
  ~NoSync();
  NoSync(const NoSync&) = delete;
  NoSync& operator=(const NoSync&) = delete;
};

}  // namespace pool
}  // namespace hssc
