// Pauli-frame sampler on the device: the reference's default path for
// shots > 1 (sdim/program.py:45-165, dispatched at :244-265).
//
// One reference shot (noiseless tableau simulation) fixes every measurement's
// reference value and deterministic flag; each further shot only tracks a Pauli
// frame (x_q, z_q) per qudit, conjugated by the gates, and reports
// (reference value + x_q) mod d.  Shots are independent, so a shot is a lane:
// one thread walks the whole gate list for one shot, the gate list is read with
// wave-uniform (scalar) loads, and the frame of a shot lives in a byte column
// of LDS, x | z << 4 per qudit (frames too large for the LDS budget use the
// same column layout in device memory).  No barriers: a thread only touches its
// own column.
//
// Draws (Philox4x64-10 keyed (seed, stream), as everywhere in the engine):
//   initial z_q of shot s            counter (q, s, 1, 0), value mulhi64(out0, d)
//   z_q redrawn after M/M_X/RESET k  counter (k, s, 1, 1)
//   N1 at instruction k              counter (k, s, 0, 0), the tableau path's draw
#include <hip/hip_runtime.h>

#include "engine.h"
#include "sdim_arith.h"

namespace sdim {

namespace {

SDIM_DEV uint32_t frame_value(uint64_t seed, uint64_t stream, uint64_t c0, uint64_t shot, uint64_t kind, uint32_t d) {
  const U64x4 o = philox4x64_10(c0, shot, FRAME_DOMAIN, kind, seed, stream);
  return (uint32_t)__umul64hi(o.v[0], (uint64_t)d);
}

SDIM_DEV uint32_t addm(uint32_t a, uint32_t b, uint32_t d) {
  const uint32_t s = a + b;
  return s >= d ? s - d : s;
}
SDIM_DEV uint32_t negm(uint32_t a, uint32_t d) { return a ? d - a : 0u; }

// Walks the gate list of one item for the shot of this thread.  F points at the
// item's frame block [n][NT] (LDS or device memory); this thread owns column tid.
// (Inlined into each branch of the kernel, so LDS accesses compile to ds_* instructions.)
SDIM_DEV void frame_walk(const FrameParams& P, const ItemDesc& it, uint8_t* F, uint32_t tid, uint32_t NT) {
  const uint32_t d = it.d;
  const uint64_t shot = (uint64_t)it.shot0 + tid;
  const bool live = tid < it.S;
  uint8_t* rec = P.arena + it.rec_off;
  for (uint32_t q = 0; q < it.n; q++)
    F[(size_t)q * NT + tid] = (uint8_t)(frame_value(it.seed, it.stream, q, shot, 0, d) << 4);
  const DevOp* ops = P.ops + it.ops_off;
  for (uint32_t k = 0; k < it.n_ops; k++) {
    const DevOp o = ops[k];
    const uint32_t kind = op_kind(o);
    uint8_t* fq = &F[(size_t)o.q * NT + tid];
    if (kind >= FOP_M) {
      // one Philox call site for every op that draws: N1 (k, shot, 0, 0), else (k, shot, 1, 1)
      const uint64_t dom = kind == FOP_NOISE ? 0 : FRAME_DOMAIN;
      const U64x4 r = philox4x64_10(o.src, shot, dom, dom, it.seed, it.stream);
      const uint32_t f = *fq;
      if (kind == FOP_NOISE) {
        const uint64_t keep53 = ((uint64_t)(o.w0 >> 10) << 32) | o.t;
        const uint32_t ab = noise_ab(r, d, (o.w0 >> 8) & 3u, keep53);
        if (ab) *fq = (uint8_t)(addm(f & 15u, ab & 15u, d) | (addm(f >> 4, (ab >> 4) & 15u, d) << 4));
        continue;
      }
      const uint32_t znew = (uint32_t)__umul64hi(r.v[0], (uint64_t)d) << 4;
      const uint32_t ref = op_a(o);
      if (kind == FOP_RESET) {
        if (live) rec[(size_t)o.t * it.S + tid] = (uint8_t)ref;
        *fq = (uint8_t)znew;
        continue;
      }
      // M, M_X (M_X first applies the H^-1 rule); then z is drawn again
      const uint32_t x = kind == FOP_MX ? (f >> 4) : (f & 15u);
      if (live) rec[(size_t)o.t * it.S + tid] = (uint8_t)(addm(ref & 0x7Fu, x, d) | (ref & 0x80u));
      *fq = (uint8_t)(x | znew);
      continue;
    }
    switch (kind) {
      case FOP_H: {          // (x, z) <- (-z, x)
        const uint32_t f = *fq;
        *fq = (uint8_t)(negm(f >> 4, d) | ((f & 15u) << 4));
        break;
      }
      case FOP_HI: {         // (x, z) <- (z, -x)
        const uint32_t f = *fq;
        *fq = (uint8_t)((f >> 4) | (negm(f & 15u, d) << 4));
        break;
      }
      case FOP_P: {          // z += x
        const uint32_t f = *fq;
        *fq = (uint8_t)((f & 15u) | (addm(f >> 4, f & 15u, d) << 4));
        break;
      }
      case FOP_PI: {         // z -= x
        const uint32_t f = *fq;
        *fq = (uint8_t)((f & 15u) | (addm(f >> 4, negm(f & 15u, d), d) << 4));
        break;
      }
      // two-qudit updates in the reference's statement order (sdim/program.py:105-112), each
      // statement reading what the previous one wrote, so a gate whose two qudits coincide ends
      // as the reference's in-place update does
      case FOP_CNOT: case FOP_CNOTI: {   // x_t += x_c, then z_c -= z_t (inverse: signs flipped)
        uint8_t* ft = &F[(size_t)o.t * NT + tid];
        const bool inv = kind == FOP_CNOTI;
        uint32_t fc = *fq, fT = *ft;
        *ft = (uint8_t)(addm(fT & 15u, inv ? negm(fc & 15u, d) : (fc & 15u), d) | (fT & 0xF0u));
        fc = *fq; fT = *ft;
        *fq = (uint8_t)((fc & 15u) | (addm(fc >> 4, inv ? (fT >> 4) : negm(fT >> 4, d), d) << 4));
        break;
      }
      case FOP_CZ: case FOP_CZI: {       // z_t += x_c, then z_c += x_t (inverse: -=)
        uint8_t* ft = &F[(size_t)o.t * NT + tid];
        const bool inv = kind == FOP_CZI;
        uint32_t fc = *fq, fT = *ft;
        *ft = (uint8_t)((fT & 15u) | (addm(fT >> 4, inv ? negm(fc & 15u, d) : (fc & 15u), d) << 4));
        fc = *fq; fT = *ft;
        *fq = (uint8_t)((fc & 15u) | (addm(fc >> 4, inv ? negm(fT & 15u, d) : (fT & 15u), d) << 4));
        break;
      }
      case FOP_SWAP: {
        uint8_t* ft = &F[(size_t)o.t * NT + tid];
        const uint8_t a = *fq, b = *ft;
        *fq = b;
        *ft = a;
        break;
      }
      default: break;
    }
  }
}

}  // namespace

template <int NT>
__global__ void __launch_bounds__(NT) frame_kernel(FrameParams P) {
  extern __shared__ uint8_t lds[];
  const ItemDesc it = P.items[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  if (it.flags & ITEM_LDS) {
    frame_walk(P, it, lds, tid, NT);
  } else {
    frame_walk(P, it, P.arena + it.tab_off, tid, NT);
  }
}

int launch_frame(const FrameParams& P, uint32_t n_items, uint32_t lds_bytes, hipStream_t st) {
  if (!n_items) return 0;
  if (P.nt != FRAME_NT) return 1;
  hipLaunchKernelGGL(frame_kernel<FRAME_NT>, dim3(n_items), dim3(FRAME_NT), lds_bytes, st, P);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int prepare_frame_kernel() {
  return hipFuncSetAttribute((const void*)frame_kernel<FRAME_NT>, hipFuncAttributeMaxDynamicSharedMemorySize, 163840) ==
                 hipSuccess ? 0 : 1;
}

}  // namespace sdim
