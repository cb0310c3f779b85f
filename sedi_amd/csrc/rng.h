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

}  // namespace sedi
