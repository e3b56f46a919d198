// Copies of host images into pinned staging buffers on several threads.
#pragma once

#include <cstdint>

namespace vqm::ss2 {

// std::memcpy(dst, src, bytes) for large host buffers, split into chunks of at
// least 1 MiB that the calling thread and up to three helper threads of a
// process-wide pool copy at once; returns when every chunk is copied. One
// thread's copy of a streamed 1080p pair (12.4 MB) is what limits a pool of
// pipelines fed from host memory, while the device upload of the same bytes
// takes a fraction of it. Buffers below 2 MiB, and processes whose CPU
// affinity allows one thread, copy on the calling thread alone. Safe to call
// from several threads at once.
void HostCopy(void* dst, const void* src, uint64_t bytes);

// Threads HostCopy may use for one call (the caller included).
unsigned HostCopyThreads();

}  // namespace vqm::ss2
