// Persistent batch kernel of one workgroup size.  This file is compiled once per
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

namespace sdim {

// ------------------------------------------------------------------ persistent batch kernel
// Each workgroup pulls items (of any dimension) from P.queue until the queue is drained.
template <int NT>
__global__ __launch_bounds__(NT) SDIM_WAVES_PER_EU(4) void batch_kernel(KParams P) {
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
      case 2: run_item<2, NT>(it, P, lds); break;
      case 3: run_item<3, NT>(it, P, lds); break;
      case 5: run_item<5, NT>(it, P, lds); break;
      case 7: run_item<7, NT>(it, P, lds); break;
      case 11: run_item<11, NT>(it, P, lds); break;
      default: break;
    }
    __syncthreads();
  }
}

template <int NT>
int prepare_batch() {
  return hipFuncSetAttribute((const void*)batch_kernel<NT>, hipFuncAttributeMaxDynamicSharedMemorySize, 163584) ==
                 hipSuccess ? 0 : -1;
}

template <int NT>
int launch_batch_nt(const KParams& P, int grid, hipStream_t st) {
  hipLaunchKernelGGL(batch_kernel<NT>, dim3(grid), dim3(NT), (size_t)P.lds_bytes, st, P);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template int prepare_batch<SDIM_BATCH_NT>();
template int launch_batch_nt<SDIM_BATCH_NT>(const KParams& P, int grid, hipStream_t st);

}  // namespace sdim
