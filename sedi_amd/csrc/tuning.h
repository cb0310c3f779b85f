// tuning.h -- launch parameters measured by the offline autotuner
// (sedi_amd/autotune.py), read from a small text table at first use.
#pragma once

#include <string>

#include "sedi_internal.h"

namespace sedi {

// Value stored for `key` and grid s in the tuning table, or `fallback`.
// Keys: "zc:<kernel>" -- planes per tile of a march kernel (8, 16 or 32);
//       "sched:<kernel>" -- 1: workgroups take tiles in launch order, up to
//                           8 per CU; 0 (default): the resident set walks
//                           XCD-contiguous runs of tiles;
//       "k:<solver>:<precision>" -- probes per batch of the estimator;
//       "ty:batch_upd_kernel" -- rows per tile of the FP32 batched CG update
//                                pass (8 or 16; default 16 from 256 rows up);
//       "ty:batch_dir_kernel" -- the same for its direction pass (default 8).
// An entry matches the exact grid, or any grid when written for 0 0 0; the
// keys "zc:*" and "sched:*" stand for every kernel.  With `nearest`, an entry of the same
// key for the grid of nearest cell count also matches.
int tuned_value(const std::string& key, const Shape& s, int fallback, bool nearest = false);

// The table in use: SEDI_TUNING_FILE if set (empty = none), else tuning.txt
// next to the library.  "" when no table was found.
const std::string& tuning_path();

}  // namespace sedi
