// k_hybrid: one launch that mixes gangs and single-workgroup systems.
//
// A translation unit of its own: the engine's out-of-line device functions
// are compiled here for k_hybrid's occupancy (one workgroup per CU, set by
// the gang path's registers), while k_als keeps its own copies, compiled for
// two resident workgroups per CU.  With registers to spare, the blocked
// QR's streaming passes keep eight rows per lane in flight and the global
// GEMMs two K-steps (same arithmetic order as k_als's two rows and one step).
#define TTALS_HH_ROWS 8
#define TTALS_GEMM_PREFETCH 2
#include "ttals_hybrid.hpp"

namespace ttals {

// Mixed schedule of one batch in one launch: workgroups [0, gang_blocks)
// form gangs (the systems that would otherwise set the launch time), the
// rest run one system each, largest first.  Workgroups are dispatched in
// index order, so every gang is resident before any single-workgroup system
// takes a CU; the gang workgroups alone fit on the device (checked on the
// host), so every gang eventually runs whole.
__global__ __launch_bounds__(NT) void k_hybrid(const GangDesc* gdescs, const int* block_sys, int gang_blocks,
                                               const SysDesc* descs, long arena) {
  extern __shared__ double dyn_lds[];
  __shared__ SolveLDS S;
  const int blk = (int)blockIdx.x;
  if (blk < gang_blocks)
    gang_block(gdescs, block_sys, blk, dyn_lds, arena, S);
  else
    als_block(descs[blk - gang_blocks], dyn_lds, arena, S);
}

hipError_t hybrid_occupancy(int* per_cu, size_t dyn_lds) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, reinterpret_cast<const void*>(k_hybrid), NT, dyn_lds);
}

hipError_t hybrid_static_lds(size_t* bytes) {
  hipFuncAttributes fa;
  const hipError_t e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k_hybrid));
  if (e == hipSuccess) *bytes = fa.sharedSizeBytes;
  return e;
}

hipError_t hybrid_launch(int blocks, size_t dyn_lds, hipStream_t stream, const GangDesc* gdescs,
                         const int* block_sys, int gang_blocks, const SysDesc* descs, long arena) {
  hipLaunchKernelGGL(k_hybrid, dim3(blocks), dim3(NT), dyn_lds, stream, gdescs, block_sys, gang_blocks, descs, arena);
  return hipGetLastError();
}

}  // namespace ttals
