// Kernels of the partitioned group-by update compiled at run time for one spec.
#pragma once

#include "vr_part_types.h"

namespace vr {

struct PartJit {
    hipModule_t module;
    hipFunction_t count;    // part_count_kernel<the spec's programs as constants>
    hipFunction_t scatter;  // part_scatter_kernel<RW, the same>
    hipFunction_t merge;    // part_merge_kernel<RW, the same>
};

// The kernels for these column programs, this record layout and key mode (fast_key: one 64-bit key
// lane) on device `ordinal`, compiled with hipRTC on first use and kept for the life of the process
// (one compile per spec and device); nullptr when the compile or the module load failed (the caller
// runs the generic kernels).
const PartJit* part_jit(int ordinal, const ColProg* prog, int nprog, const RecLayout& L, bool fast_key);

// After a compiled kernel failed to launch: part_jit returns nullptr from then on.
void part_jit_disable();

}  // namespace vr
