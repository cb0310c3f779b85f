// rng.h -- device-resident mt19937 probe stream.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "runtime.h"
#include "sedi_internal.h"

namespace sedi {

struct MtState {
    uint32_t mt[624];
    int32_t idx;          // next unused output of the current block (624 = regenerate)
    int32_t carry_valid;  // an output waiting for its pair partner
    uint32_t carry;
};

class MtStream {
  public:
    int init();
    int seed(uint32_t seed, hipStream_t stream);
    // seed, then jump ahead by first_draw outputs on the host
    int seed_at(uint32_t seed, unsigned long long first_draw, hipStream_t stream);
    int skip(unsigned long long draws, hipStream_t stream);
    int fill(double* v, long long ncells, hipStream_t stream);

  private:
    DevBuf state_;
};

// Probes k0 .. k0+count-1 of the stream mt19937(seed) on a grid of ncells
// cells, generated in parallel: init() uploads the stream's start and the
// probes' jump polynomials and computes each probe's start (asynchronous on
// `stream` after a synchronous upload); emit() writes probes
// k0+first .. k0+first+nprobe-1 into slots 0 .. nprobe-1 of a batch layout
// [K/4][ncells][4] (float or double) and zeros into the other slots.
class ProbeSource {
  public:
    int init(uint32_t seed, long long ncells, int k0, int count, hipStream_t stream);
    int emit(int first, int nprobe, int K, bool fp32, void* out, hipStream_t stream) const;
    int count() const { return count_; }

  private:
    DevBuf up_, wk_;
    long long ncells_ = 0;
    int count_ = 0;
};

// Probes k0 .. k0+kact-1 of the stream mt19937(seed) on a grid of ncells
// cells, written as +-1 into the batch layout [K/4][ncells][4] (float or
// double); slots kact .. K-1 are zeroed.  K is a multiple of 4.  Uses
// `scratch` for the uploaded jump data; synchronises `stream`.
int fill_probe_batch(uint32_t seed, long long ncells, int k0, int K, int kact, bool fp32, void* out,
                     hipStream_t stream, DevBuf& scratch);

// [K/4][ncells][4] FP64 batch layout -> [count][ncells] rows (count <= K).
int unbatch_rows(const double* batch, long long ncells, int count, double* rows, hipStream_t stream);

}  // namespace sedi
