// profiler.h -- HIP-event timing of named kernel launches (sedi_profile_*).
#pragma once

#include <hip/hip_runtime.h>
#include <string>

namespace sedi {

// True when launches of `name` are to be bracketed with events.
bool profile_wants(const char* name);
// True when any kernel is being timed (launches are then not graph-captured).
bool profile_any();
// Whether chunks of solver iterations are captured into HIP graphs and
// replayed (default on; sedi_graphs_enable).
bool graphs_enabled();
void set_graphs_enabled(bool on);
// While one of these lives, launches on this thread are not timed (timing
// creates HIP events; the allocation-free solve must not).
struct ProfileSuppress {
    ProfileSuppress();
    ~ProfileSuppress();
    bool prev;
};
// Records the start event (returns a token, -1 if not profiling this kernel).
int profile_begin(const char* name, hipStream_t stream);
void profile_end(int token, hipStream_t stream);
// Resolves all recorded pairs; call after the stream(s) synchronised.
void profile_flush();

// Launch helper: kernel<<<grid, block, 0, stream>>>(args...) with optional timing.
template <typename... KArgs, typename... Args>
inline hipError_t launch(const char* name, void (*kern)(KArgs...), dim3 grid, dim3 block,
                         hipStream_t stream, Args... args) {
    int tok = profile_begin(name, stream);
    hipLaunchKernelGGL(kern, grid, block, 0, stream, args...);
    hipError_t e = hipGetLastError();
    if (tok >= 0) profile_end(tok, stream);
    return e;
}

}  // namespace sedi
