// Launch helpers of k_hybrid (defined in ttals_hybrid.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "ttals_engine.hpp"
#include "ttals_gang.hpp"

namespace ttals {

// Resident k_hybrid workgroups per CU with `dyn_lds` bytes of dynamic LDS.
hipError_t hybrid_occupancy(int* per_cu, size_t dyn_lds);

// Static LDS bytes of k_hybrid.
hipError_t hybrid_static_lds(size_t* bytes);

// Launches k_hybrid: `blocks` workgroups, the first `gang_blocks` of them the
// gangs of `gdescs` (block_sys: gang of each), the rest one system each of
// `descs`, in order.
hipError_t hybrid_launch(int blocks, size_t dyn_lds, hipStream_t stream, const GangDesc* gdescs,
                         const int* block_sys, int gang_blocks, const SysDesc* descs, long arena);

}  // namespace ttals
