// Positive and negative controls of the CPU build's checkers (CPU only: these
// kernels are compiled for the host emulation and never for a GPU).
//   0 clean: a write, __syncthreads, reads of it by every thread
//   1 race: thread 0 writes LDS, the other threads read it without a barrier
//           (ThreadSanitizer must report a data race)
//   2 out of bounds: one byte written past the launch's dynamic LDS
//           (AddressSanitizer must report it)
//   3 wave exchange: lane 0 writes LDS, a shuffle, lane 1 reads (ordered: clean)
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace sdim {

__global__ void selftest_kernel(int which, uint32_t shmem, int* out) {
  extern __shared__ uint8_t lds[];
  const uint32_t t = threadIdx.x;
  if (which == 0) {
    if (t == 0) lds[0] = 7;
    __syncthreads();
    atomicAdd(out, (int)lds[0]);
  } else if (which == 1) {
    if (t == 0) lds[0] = 7;
    atomicAdd(out, (int)lds[0]);
  } else if (which == 2) {
    if (t == 0) lds[shmem] = 1;
  } else if (which == 3) {
    if (t == 0) lds[8] = 5;
    const int v = __shfl((int)t, 0, 64);
    if (t == 1) atomicAdd(out, (int)lds[8] + v);
  }
}

}  // namespace sdim

extern "C" int sdim_sim_selftest(int which) {
  int out = 0;
  const uint32_t shmem = 256;
  hipLaunchKernelGGL(sdim::selftest_kernel, dim3(1), dim3(128), shmem, (hipStream_t) nullptr, which, shmem, &out);
  return out;
}
