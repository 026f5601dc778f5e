// Interface between the device engine (engine.hip) and the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "sdim_types.h"

namespace sdim {

// number of optional phase timers after the two byte counters
#define SDIM_NUM_TIMERS 12

struct KParams {
  const ItemDesc* items;
  const uint32_t* order;
  uint32_t n_items;
  uint32_t* queue;
  const DevOp* ops;
  const uint64_t* keep53;
  uint8_t* arena;
  unsigned long long* counters;   // [0] bytes on HBM-resident tableaux, [1] bytes on LDS-resident tableaux
  uint32_t lds_bytes;
};

// LDS scratch of one workgroup for an item with n qudits, row stride rs, S shots
// (must match carve() in engine.hip)
__host__ __device__ inline uint32_t scratch_bytes(uint32_t n, uint32_t rs, uint32_t S, uint32_t NT) {
  uint32_t b = 256 + 16 * 16;                 // counters, multiplication tables
  b += 2 * ((n + 15u) & ~15u);                // pivot column X, Z
  b += (2 * n + 15u) & ~15u;                  // touched rows
  b += ((rs / 2) + 15u) & ~15u;               // words with f != 0
  b += 6 * rs;                                // row q, f, dots (2 rs), segment delta (2 rs)
  b += NT * 16;                               // reductions
  b += ((2 * S * 4) + 15u) & ~15u;            // per-shot pivot phase / outcome
  return b;
}

// LDS bytes of the noise tables of a segment with nn noise ops
__host__ __device__ inline uint32_t noise_bytes(uint32_t nn, uint32_t rs, uint32_t S) {
  uint32_t mw = (S + 31) / 32;
  uint32_t b = nn * 2 * rs;
  b += nn * mw * 4;
  b += (nn * S + 15u) & ~15u;
  return b;
}

int prepare_kernels();
int launch_reset(uint32_t* queue, unsigned long long* counters, hipStream_t st);
int launch_batch(const KParams& P, int nt, int grid, hipStream_t st);
int launch_plain_op(const KParams& P, const DevOp& o, uint32_t n, uint32_t d, uint32_t n_items, uint32_t S,
                    hipStream_t st);

}  // namespace sdim
