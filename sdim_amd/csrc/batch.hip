// Persistent batch kernels of one workgroup size.  This file is compiled once per
// size (-DSDIM_BATCH_NT=64, 128, 256, 512, 1024), so the sizes build in parallel;
// each object defines launch_batch_nt / prepare_batch for its size.
#include "tableau_dev.h"

#ifndef SDIM_BATCH_NT
#error "compile batch.hip with -DSDIM_BATCH_NT=<workgroup size>"
#endif

// occupancy target of the batch kernel: the register budget is cut for this many waves per
// SIMD (a build of the kernels for another target may define it empty)
#ifndef SDIM_WAVES_PER_EU
#define SDIM_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif
#ifndef SDIM_BATCH_WAVES
#define SDIM_BATCH_WAVES 4
#endif

namespace sdim {

// ------------------------------------------------------------------ persistent batch kernel
// Each workgroup pulls items (of any dimension) from P.queue until the queue is drained.
// One kernel per (workgroup size, placement, RB, OPB): the items of a launch keep their
// tableau in LDS (RES) or in HBM; RB rows in flight per thread in the rowsum sweep and OPB
// gate batches per op load (the HBM-resident classes pick them per launch, the LDS-resident
// classes use RB = 1, OPB = 2).
template <int NT, bool RES, int RB, int OPB>
__global__ __launch_bounds__(NT) SDIM_WAVES_PER_EU(SDIM_BATCH_WAVES) void batch_kernel(KParams P) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ uint32_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(P.queue, 1u);
    __syncthreads();
    const uint32_t idx = s_item;
    __syncthreads();
    if (idx >= P.n_items) break;
    const ItemDesc it = P.items[P.order[idx]];
    switch (it.d) {
      case 2: run_item_at<2, NT, RES, RB, OPB>(it, P, lds); break;
      case 3: run_item_at<3, NT, RES, RB, OPB>(it, P, lds); break;
      case 5: run_item_at<5, NT, RES, RB, OPB>(it, P, lds); break;
      case 7: run_item_at<7, NT, RES, RB, OPB>(it, P, lds); break;
      case 11: run_item_at<11, NT, RES, RB, OPB>(it, P, lds); break;
      default: break;
    }
    __syncthreads();
  }
}

template <int NT, bool RES, int RB, int OPB>
static bool set_lds(int lds) {
  return hipFuncSetAttribute((const void*)batch_kernel<NT, RES, RB, OPB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             lds) == hipSuccess;
}

template <int NT>
int prepare_batch() {
  const int lds = 163584;
  const bool ok = set_lds<NT, true, 1, 2>(lds) && set_lds<NT, false, 1, 1>(lds) && set_lds<NT, false, 1, 2>(lds) &&
                  set_lds<NT, false, 2, 1>(lds) && set_lds<NT, false, 2, 2>(lds) && set_lds<NT, false, 4, 1>(lds) &&
                  set_lds<NT, false, 4, 2>(lds);
  return ok ? 0 : -1;
}

template <int NT, int RB, int OPB>
static void launch_hbm(const KParams& P, int grid, hipStream_t st) {
  hipLaunchKernelGGL((batch_kernel<NT, false, RB, OPB>), dim3(grid), dim3(NT), (size_t)P.lds_bytes, st, P);
}

template <int NT>
int launch_batch_nt(const KParams& P, bool resident, int rb, int opb, int grid, hipStream_t st) {
  if (resident) {
    hipLaunchKernelGGL((batch_kernel<NT, true, 1, 2>), dim3(grid), dim3(NT), (size_t)P.lds_bytes, st, P);
  } else {
    switch (rb * 10 + opb) {
      case 11: launch_hbm<NT, 1, 1>(P, grid, st); break;
      case 12: launch_hbm<NT, 1, 2>(P, grid, st); break;
      case 21: launch_hbm<NT, 2, 1>(P, grid, st); break;
      case 22: launch_hbm<NT, 2, 2>(P, grid, st); break;
      case 41: launch_hbm<NT, 4, 1>(P, grid, st); break;
      case 42: launch_hbm<NT, 4, 2>(P, grid, st); break;
      default: return -2;
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template int prepare_batch<SDIM_BATCH_NT>();
template int launch_batch_nt<SDIM_BATCH_NT>(const KParams& P, bool resident, int rb, int opb, int grid,
                                             hipStream_t st);

}  // namespace sdim
