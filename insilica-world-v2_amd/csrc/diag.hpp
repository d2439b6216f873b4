// Per-kernel timing for measurement (bench.py): when switched on with isw_timing_start(), every launch through
// ISW_LAUNCH goes through hipExtLaunchKernelGGL with a start and a stop event, which the command processor stamps
// with the kernel's own begin and end (the durations a kernel trace reports), tagged with the kernel's name and its
// algorithmic bytes (each distinct array the kernel touches counted once at its stored width, twice if it is read
// and written; index tables included; neighbour re-reads free). isw_timing_stop() returns, per kernel, the launch
// count, the summed time and the summed bytes. Nothing is recorded while the stream is being captured.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

namespace isw_diag {
bool enabled();
// Registers a launch; on success returns true with the event pair to pass to hipExtLaunchKernelGGL.
bool begin(const char* name, double bytes, hipStream_t st, hipEvent_t* a, hipEvent_t* b);
}  // namespace isw_diag

#define ISW_LAUNCH(name, bytes, kern, grid, block, ...)                                       \
  do {                                                                                        \
    hipEvent_t isw_a_ = nullptr, isw_b_ = nullptr;                                            \
    if (isw_diag::enabled() && isw_diag::begin((name), (bytes), st, &isw_a_, &isw_b_))        \
      hipExtLaunchKernelGGL(kern, (grid), (block), 0, st, isw_a_, isw_b_, 0, __VA_ARGS__);    \
    else                                                                                      \
      kern<<<(grid), (block), 0, st>>>(__VA_ARGS__);                                          \
  } while (0)

// The same with `shm` bytes of dynamic LDS.
#define ISW_LAUNCH_SHM(name, bytes, kern, grid, block, shm, ...)                              \
  do {                                                                                        \
    hipEvent_t isw_a_ = nullptr, isw_b_ = nullptr;                                            \
    if (isw_diag::enabled() && isw_diag::begin((name), (bytes), st, &isw_a_, &isw_b_))        \
      hipExtLaunchKernelGGL(kern, (grid), (block), (shm), st, isw_a_, isw_b_, 0, __VA_ARGS__); \
    else                                                                                      \
      kern<<<(grid), (block), (shm), st>>>(__VA_ARGS__);                                      \
  } while (0)
