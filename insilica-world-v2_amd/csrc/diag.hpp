// Per-kernel timing with HIP events, for measurement (bench.py): when switched on with isw_timing_start(), every
// launch through ISW_LAUNCH records an event before and after itself on its stream, tagged with the kernel's name and
// its algorithmic bytes (each distinct array the kernel touches counted once at its stored width, twice if it is read
// and written; index tables included; neighbour re-reads free). isw_timing_stop() returns, per kernel, the launch
// count, the summed event time and the summed bytes. Nothing is recorded while the stream is being captured.
#pragma once
#include <hip/hip_runtime.h>

namespace isw_diag {
bool enabled();
int begin(const char* name, double bytes, hipStream_t st);
void end(int tok, hipStream_t st);
}  // namespace isw_diag

#define ISW_LAUNCH(name, bytes, ...)                                                        \
  do {                                                                                      \
    const int isw_tok_ = isw_diag::enabled() ? isw_diag::begin((name), (bytes), st) : -1; \
    __VA_ARGS__;                                                                            \
    if (isw_tok_ >= 0) isw_diag::end(isw_tok_, st);                                         \
  } while (0)
