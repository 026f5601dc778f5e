// Interface between the device engine (tableau_dev.h and the batch, gang, plain and
// frame kernels) and the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "sdim_types.h"

namespace sdim {

// number of optional phase timers after the two byte counters
#define SDIM_NUM_TIMERS 16
// device counters of a launch: [0] tableau bytes of HBM-resident items, [1] of LDS-resident
// items, [2, 2 + SDIM_NUM_TIMERS) phase timers, then the error word: the first capacity
// violation of the launch as code << 32 | item id (0: none), which fails the ticket
#define SDIM_ERR_SLOT (2 + SDIM_NUM_TIMERS)
#define SDIM_NUM_COUNTERS (SDIM_ERR_SLOT + 1)
enum KernelError : uint32_t {
  KERR_TOUCHED_ROWS = 1,    // more touched rows than the touched-row list holds
  KERR_FACTOR_GROUPS = 2,   // more 16-byte groups with f != 0 than the group list holds
};
// launch classes, and item queues (one per launch)
#define SDIM_MAX_CLASSES 32
#define SDIM_MAX_QUEUES 64

struct KParams {
  const ItemDesc* items;
  const uint32_t* order;
  uint32_t n_items;
  uint32_t* queue;
  const DevOp* ops;
  const uint64_t* keep53;
  uint8_t* arena;
  unsigned long long* counters;   // SDIM_NUM_COUNTERS device counters (below)
  uint32_t lds_bytes;
  uint32_t list_cap;              // nonzero: caps the LDS row / group lists (exercises the overflow check)
};

// LDS scratch of one workgroup for an item with n qudits, row stride rs, S shots
// (must match carve() in tableau_dev.h)
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

// ---- gangs: one item processed by K workgroups, each owning a range of column words
struct GangSync;
// bytes of one gang's barrier record (GangSync in gang.hip)
constexpr uint64_t GANG_SYNC_BYTES = 4096;
struct GangParams {
  GangSync* sync;       // one GANG_SYNC_BYTES record per gang
  uint8_t* scr;         // per gang: [n_max][Kp] + [S_max][Kp] bytes, acc[K], crs[K], tot[S_max] u32,
                        // two pivot-column staging buffers
  uint64_t scr_stride;  // bytes of one gang's scratch
  uint32_t K;           // members per gang
  uint32_t n_max, S_max;
  int* err;             // set when a gang barrier timed out
  uint64_t timeout;     // ticks of the 100 MHz constant clock a member waits at a barrier before giving up
};
// default barrier wait bound of a gang member: 10 s
constexpr uint64_t GANG_TIMEOUT_TICKS = 1000000000ull;

__host__ __device__ inline uint64_t gang_align(uint64_t x) { return (x + 255u) & ~255ull; }
// row stride of the per-member byte tables (16-byte multiples)
__host__ __device__ inline uint32_t gang_kp(uint32_t K) { return K < 16 ? 16u : (K + 15u) & ~15u; }
// column words owned by each member: ceil(words / K) rounded up to 16 bytes
__host__ __device__ inline uint32_t gang_words(uint32_t words, uint32_t K, uint32_t bytes_per_word) {
  const uint32_t unit = 16 / bytes_per_word;
  uint32_t c = (words + K - 1) / K;
  return (c + unit - 1) / unit * unit;
}
// staging of a pivot column: n bytes (x | z << 4 per row) then S pivot phases, two buffers
__host__ __device__ inline uint64_t gang_stage_bytes(uint32_t n_max, uint32_t S_max) {
  return gang_align((uint64_t)n_max + S_max);
}
__host__ __device__ inline uint64_t gang_scratch_bytes(uint32_t K, uint32_t n_max, uint32_t S_max) {
  const uint64_t kp = gang_kp(K);
  return gang_align(n_max * kp) + gang_align(S_max * kp) + gang_align(4ull * (2 * K + S_max)) +
         2 * gang_stage_bytes(n_max, S_max);
}

// ---- Pauli-frame sampler (frame.hip): one workgroup per item of up to FRAME_NT shots
#define FRAME_NT 256
struct FrameParams {
  const ItemDesc* items;
  const DevOp* ops;
  uint8_t* arena;      // records at it.rec_off, device-memory frames at it.tab_off
  uint32_t nt;
};
int launch_frame(const FrameParams& P, uint32_t n_items, uint32_t lds_bytes, hipStream_t st);
int prepare_frame_kernel();

int prepare_kernels();

// allocation-free path (noalloc.hip): device metadata of its submissions, per GPU
#define SDIM_NOALLOC_SCRATCH_BYTES (8u << 20)
int noalloc_init_device(int device);     // with `device` current; 0 on success
uint8_t* noalloc_scratch(int device);    // the device's scratch array (nullptr before init)
int launch_gang(const KParams& P, const GangParams& G, int d, int nt, int grid, hipStream_t st);
int launch_reset(uint32_t* queue, unsigned long long* counters, hipStream_t st);
// every item of a batch launch has the placement `resident` (tableau in LDS) or not; HBM-resident
// launches keep rb rows in flight per thread in the sweep (1, 2, 4) and load the ops of opb gate
// batches at once (1, 2)
int launch_batch(const KParams& P, int nt, bool resident, int rb, int opb, int grid, hipStream_t st);
int launch_plain_op(const KParams& P, const DevOp& o, uint32_t n, uint32_t d, uint32_t n_items, uint32_t S,
                    hipStream_t st);

// One object per workgroup size (batch.hip) and per dimension (gang.hip) defines these.
template <int NT> int prepare_batch();
template <int NT> int launch_batch_nt(const KParams& P, bool resident, int rb, int opb, int grid, hipStream_t st);
template <int D> int prepare_gang();
template <int D> int launch_gang_d(const KParams& P, const GangParams& G, int nt, int grid, hipStream_t st);
#define SDIM_EXTERN_BATCH(NT)                 \
  extern template int prepare_batch<NT>(); \
  extern template int launch_batch_nt<NT>(const KParams& P, bool resident, int rb, int opb, int grid, hipStream_t st);
SDIM_EXTERN_BATCH(64) SDIM_EXTERN_BATCH(128) SDIM_EXTERN_BATCH(256) SDIM_EXTERN_BATCH(512) SDIM_EXTERN_BATCH(1024)
#undef SDIM_EXTERN_BATCH
#define SDIM_EXTERN_GANG(D)                  \
  extern template int prepare_gang<D>(); \
  extern template int launch_gang_d<D>(const KParams& P, const GangParams& G, int nt, int grid, hipStream_t st);
SDIM_EXTERN_GANG(2) SDIM_EXTERN_GANG(3) SDIM_EXTERN_GANG(5) SDIM_EXTERN_GANG(7) SDIM_EXTERN_GANG(11)
#undef SDIM_EXTERN_GANG

}  // namespace sdim
