// Device engine shared by the batch, gang and plain kernels: gate segments and
// measurement of one item by one workgroup (included by batch.hip, gang.hip, plain.hip).
//
// Algorithm (reference files read as text):
//   gates          sdim/tableau/tableau_optimized.py:5-117 and the X/Z words of
//                  sdim/tableau/tableau_gates.py:27-137 in closed form
//   measurement    sdim/tableau/tableau_prime.py:262-363 (pivot, exponentiate,
//                  random rowsum, deterministic ancilla)
//   RESET fix-up   sdim/program.py:334-339
// Random outcomes / noise draws: Philox4x64-10 keyed (seed, stream | op, shot).
//
// Execution model.  An item is one circuit with S shots; its X/Z blocks are
// shared by the shots and each shot owns a phase vector (sdim_types.h).  One
// workgroup processes one item at a time.  The circuit arrives as segments of
// gates (layered on the host: the gates of a layer touch disjoint qudits) and
// measurement ops.
//   * Gate segment: layer by layer, thread (word w, slot k) applies the k-th,
//     (k+NS)-th, ... gate of the layer to column word w, loading the rows of a
//     batch of independent gates before computing (latency hiding).  Gate phase
//     increments do not depend on the shot, so each thread sums them in
//     registers (delta) and the shots' phase vectors are touched once, at the
//     end of the segment.  Noise ops snapshot their row; their shot-dependent
//     phase terms are added in the same end-of-segment pass.
//   * Measurement: four dependent steps separated by barriers.
//       A  read row q: pivot candidate (atomic-min over the stabilizer half),
//          rowsum factors f = -x[q, .], list of words with f != 0
//       B  gather pivot column p: list of touched rows (x_p or z_p != 0),
//          <x_p, z_p>, pivot phases of every shot
//       C  sweep (touched row) x (word with f != 0): X += x_p f, Z += z_p f with
//          x_p, z_p scaled by e = x[q,p]^-1 on the fly (exponentiation), column
//          dots <z_i, x_p> from the old Z accumulated through LDS atomics
//       D  columns p / n+p rewritten (stabilizer p <- Z_q, destabilizer p <- old
//          stabilizer p), per-shot phases ph += f ph_p + f dot + xz f(f-1)/2
//     then per-shot outcomes.  Without a pivot the outcome is the deterministic
//     ancilla, computed per row as an exclusive scan over the columns with f != 0.
//   The work of a rowsum is (rows in the support of the pivot generator) x
//   (words where f != 0), which is what the sweep iterates over.
// The tableau lives in LDS for the whole item when it fits (ITEM_LDS) and in
// HBM otherwise; the same code runs on both (address spaces are inferred
// after inlining).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>

#include "sdim_types.h"
#include "sdim_arith.h"
#include "engine.h"

namespace sdim {

// Measurement variants (compile-time switches, for measuring their effect):
//   SDIM_EARLY_DRAWS  for items whose tableau is in HBM, draw random outcomes during step A
//                     (hiding the draws behind the loads of row q) instead of step E; items in
//                     LDS always draw in step E, and only when the outcome is random
//   SDIM_DET_BATCH    listed words (f != 0) per load batch in the deterministic scan
#ifndef SDIM_EARLY_DRAWS
#define SDIM_EARLY_DRAWS 1
#endif
// Per-launch variants (template parameters, chosen per launch class at run time):
//   RB   rows of the rowsum sweep each thread keeps in flight (16 bytes of X and Z per row)
//   OPB  gate batches of a layer whose ops are loaded in one round trip
#ifndef SDIM_DET_BATCH
#define SDIM_DET_BATCH 4
#endif

// ------------------------------------------------------------------ scratch carve
struct Scr {
  int* misc;         // two counter sets of 16 ints (double-buffered per measurement)
  MulTab* mtab;      // tables of c * v mod d, c < 16
  uint8_t* xp;       // pivot column X, n bytes (0/1 for d = 2)
  uint8_t* zp;       // pivot column Z, n bytes
  uint16_t* TR;      // touched rows of a rowsum
  uint16_t* FL;      // words with f != 0
  uint8_t* RQ;       // row q of X, rs bytes
  uint8_t* F;        // rowsum factors, rs bytes
  uint32_t* DOT;     // column dots: two u32 per word
  uint8_t* DW;       // segment phase delta: 2 * rs bytes
  uint32_t* red;     // NT * 4 words
  uint32_t* shot;    // 2 * S words
  uint8_t* noise;    // noise tables of a segment (over the measurement-only scratch)
  uint8_t* lds_end;  // end of this workgroup's dynamic LDS
  uint8_t* gap;      // LDS reserved by the caller in front of the measurement-only scratch
  bool tiling;       // stage column tiles of an HBM-resident tableau through LDS for gate segments
  uint32_t tr_cap;   // entries of TR
  uint32_t fl_cap;   // entries of FL
  unsigned long long* err;   // the launch's error word (counters[SDIM_ERR_SLOT])
  uint32_t item;     // id of the item being processed
};

// Records the first capacity violation of a launch (code << 32 | item id); the host then
// fails the ticket.  The list write that would have overflowed is dropped.
SDIM_DEV void report_error(const Scr& S, uint32_t code) {
  if (S.err) atomicCAS(S.err, 0ull, ((unsigned long long)code << 32) | S.item);
}

template <int NT>
SDIM_DEV Scr carve(uint8_t* base, uint32_t n, uint32_t rs, uint32_t S, uint32_t gap = 0) {
  Scr s;
  uint8_t* p = base;
  s.misc = (int*)p; p += 256;
  s.mtab = (MulTab*)p; p += 16 * 16;
  s.DW = p; p += 2 * rs;
  s.red = (uint32_t*)p; p += NT * 16;
  s.gap = p; p += (gap + 15u) & ~15u;
  // measurement-only scratch, last: a gate segment overlays it (and whatever LDS follows)
  // with its noise tables and column tiles
  const uint32_t nb = (n + 15u) & ~15u;
  s.xp = p; p += nb;
  s.zp = p; p += nb;
  s.TR = (uint16_t*)p; p += (2 * n + 15u) & ~15u;
  s.tr_cap = ((2 * n + 15u) & ~15u) / 2;
  s.FL = (uint16_t*)p; p += ((rs / 2) + 15u) & ~15u;
  s.fl_cap = (((rs / 2) + 15u) & ~15u) / 2;
  s.RQ = p; p += rs;
  s.F = p; p += rs;
  s.DOT = (uint32_t*)p; p += 2 * rs;
  s.shot = (uint32_t*)p; p += ((2 * S * 4) + 15u) & ~15u;
  s.noise = s.xp;
  s.lds_end = nullptr;
  s.tiling = false;
  s.err = nullptr;
  s.item = 0;
  return s;
}

// tableau pointers of an item
struct Tab {
  uint8_t* X;
  uint8_t* Z;
  uint8_t* Ph;
  uint32_t n, rs, S;   // rs: row stride of X, Z and of the per-shot phase vectors
  uint32_t wofs;      // first column word of this view (column tiles of an HBM-resident tableau)
  uint32_t rs_full;   // width in bytes of the per-column LDS arrays (segment deltas, noise snapshots)
  uint32_t nw;        // column words covered by the view (gate sweeps and phase passes)
  uint32_t il;        // X/Z row interleave (see sdim_types.h)
};

// word w (4 bytes, odd d) / word w (8 bytes, d = 2) / byte b of row r of an X/Z block
SDIM_DEV uint32_t* w32(uint8_t* base, uint32_t r, uint32_t w, uint32_t rs, uint32_t il) { return (uint32_t*)(base + cell(r, 4 * w, rs, il)); }
SDIM_DEV uint64_t* w64(uint8_t* base, uint32_t r, uint32_t w, uint32_t rs, uint32_t il) { return (uint64_t*)(base + cell(r, 8 * w, rs, il)); }
SDIM_DEV uint8_t* b8(uint8_t* base, uint32_t r, uint32_t b, uint32_t rs, uint32_t il) { return base + cell(r, b, rs, il); }
SDIM_DEV uint8_t* phase_vec(const Tab& T, uint32_t s) { return T.Ph + (size_t)s * T.rs; }
SDIM_DEV uint64_t* plane(const Tab& T, uint32_t s, int k) { return (uint64_t*)(T.Ph + (size_t)(2 * s + k) * T.rs); }

// Optional per-phase timers (build with -DSDIM_PHASE_TIMERS): thread 0 of each
// workgroup accumulates wall-clock ticks per phase into counters[2 + phase].
// Switched off: a timer build of batch_kernel<1024> stopped with a memory-aperture
// violation on the uneven batch (the cause is not known; the kernel spills far more
// registers in that build), so the timers stay out until the cause is found.
#ifdef SDIM_PHASE_TIMERS
#error "SDIM_PHASE_TIMERS is switched off (see the comment above)"
#define SDIM_T0() const uint64_t t_start_ = (threadIdx.x == 0) ? wall_clock64() : 0
#define SDIM_TMARK(var) uint64_t var = (threadIdx.x == 0) ? wall_clock64() : 0
#define SDIM_TADD(P, ph, from) do { if (threadIdx.x == 0) atomicAdd(&(P).counters[2 + (ph)], (unsigned long long)(wall_clock64() - (from))); } while (0)
#else
#define SDIM_T0()
#define SDIM_TMARK(var)
#define SDIM_TADD(P, ph, from) do {} while (0)
#endif

// fold the two 16-bit partial sums of a 32-bit word mod D
template <int D> SDIM_DEV uint32_t fold16(uint32_t v) { return ((v & 0xFFFFu) % D) | (((v >> 16) % D) << 16); }

// ------------------------------------------------------------------ block reductions
template <int NT>
SDIM_DEV uint32_t block_sum(uint32_t v, uint32_t* red) {
  #pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  uint32_t r = 0;
  #pragma unroll
  for (int i = 0; i < NT / 64; i++) r += red[i];
  __syncthreads();
  return r;
}

template <int NT>
SDIM_DEV unsigned long long block_sum64(unsigned long long v, uint32_t* red) {
  #pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) { red[2 * wid] = (uint32_t)v; red[2 * wid + 1] = (uint32_t)(v >> 32); }
  __syncthreads();
  unsigned long long r = 0;
  #pragma unroll
  for (int i = 0; i < NT / 64; i++) r += (unsigned long long)red[2 * i] | ((unsigned long long)red[2 * i + 1] << 32);
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------ tableau init / copy
template <int D, int NT>
SDIM_DEV void init_identity(const Tab& T, uint64_t xz_bytes, uint64_t ph_bytes) {
  // zero X, Z and the phase vectors, then the diagonals: stabilizer i = Z_i, destabilizer i = X_i
  uint4* p = (uint4*)T.X;
  for (uint64_t i = threadIdx.x; i < xz_bytes / 16; i += NT) p[i] = make_uint4(0, 0, 0, 0);
  uint4* q = (uint4*)T.Ph;
  for (uint64_t i = threadIdx.x; i < ph_bytes / 16; i += NT) q[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  const uint32_t n = T.n;
  for (uint32_t r = threadIdx.x; r < n; r += NT) {
    if constexpr (D == 2) {
      const uint32_t g = n + r;
      *w64(T.X, r, g >> 6, T.rs, T.il) |= 1ull << (g & 63);
      *w64(T.Z, r, r >> 6, T.rs, T.il) |= 1ull << (r & 63);
    } else {
      *b8(T.X, r, n + r, T.rs, T.il) = 1;
      *b8(T.Z, r, r, T.rs, T.il) = 1;
    }
  }
  __syncthreads();
}

template <int NT>
SDIM_DEV void copy16(uint8_t* dst, const uint8_t* src, uint64_t bytes) {
  const uint4* s = (const uint4*)src;
  uint4* d = (uint4*)dst;
  for (uint64_t i = threadIdx.x; i < bytes / 16; i += NT) d[i] = s[i];
}

// ------------------------------------------------------------------ one gate on one column word
// The caller has loaded the row words (xq, zq of qudit q; xt, zt of the target
// for SUM).  Stores the updated words and accumulates the phase increment.
template <int D>
SDIM_DEV uint32_t apply_odd(const DevOp& o, uint32_t xq, uint32_t zq, uint32_t xt, uint32_t zt,
                            const Tab& T, uint32_t w, uint32_t& dl, const Scr& S) {
  const uint32_t rs = T.rs;
  uint32_t* pxq = w32(T.X, o.q, w, rs, T.il);
  uint32_t* pzq = w32(T.Z, o.q, w, rs, T.il);
  switch (op_kind(o)) {
    case OP_H:                                     // (x, z) <- (-z, x); phase += po (-z) x
      *pxq = negm<D>(zq); *pzq = xq;
      dl = subm<D>(dl, mulv<D>(xq, zq));
      return 16;
    case OP_HI:                                    // (x, z) <- (z, -x); phase += po z (-x)
      *pxq = zq; *pzq = negm<D>(xq);
      dl = subm<D>(dl, mulv<D>(xq, zq));
      return 16;
    case OP_P: {                                   // z += x; phase += x (x - 1) / 2
      const MulTab h = make_half_tab<D>();
      *pzq = addm<D>(zq, xq);
      dl = addm<D>(dl, mult<D>(h, xq));
      return 12;
    }
    case OP_PI: {
      const MulTab h = make_half_tab<D>();
      *pzq = subm<D>(zq, xq);
      dl = subm<D>(dl, mult<D>(h, xq));
      return 12;
    }
    case OP_SUM:                                   // x[t] += x[c]; z[c] += (d - 1) z[t]
      *w32(T.X, o.t, w, rs, T.il) = addm<D>(xt, xq);
      *pzq = subm<D>(zq, zt);
      return 24;
    case OP_SUMI:
      *w32(T.X, o.t, w, rs, T.il) = subm<D>(xt, xq);
      *pzq = addm<D>(zq, zt);
      return 24;
    case OP_PAULI:                                 // X^a Z^b: phase += b x - a z
      dl = addm<D>(dl, subm<D>(mult<D>(S.mtab[op_b(o)], xq), mult<D>(S.mtab[op_a(o)], zq)));
      return 8;
    case OP_CZ:                                    // z[t] += x[c], z[c] += x[t]; phase += x[c] x[t]
      *w32(T.Z, o.t, w, rs, T.il) = addm<D>(zt, xq);
      *pzq = addm<D>(zq, xt);
      dl = addm<D>(dl, mulv<D>(xq, xt));
      return 24;
    case OP_CZI:
      *w32(T.Z, o.t, w, rs, T.il) = subm<D>(zt, xq);
      *pzq = subm<D>(zq, xt);
      dl = subm<D>(dl, mulv<D>(xq, xt));
      return 24;
    case OP_SWAP:
      *pxq = xt; *pzq = zt;
      *w32(T.X, o.t, w, rs, T.il) = xq;
      *w32(T.Z, o.t, w, rs, T.il) = zq;
      return 32;
    case OP_NOISE: {                               // snapshot of row q for the end-of-segment pass
      uint8_t* snap = S.noise + op_noise_slot(o) * 2 * T.rs_full;
      ((uint32_t*)snap)[T.wofs + w] = xq;
      ((uint32_t*)(snap + T.rs_full))[T.wofs + w] = zq;
      return 8;
    }
  }
  return 0;
}

SDIM_DEV uint32_t apply_bin(const DevOp& o, uint64_t xq, uint64_t zq, uint64_t xt, uint64_t zt,
                            const Tab& T, uint32_t w, uint64_t& d0, uint64_t& d1, const Scr& S) {
  const uint32_t rs = T.rs;
  uint64_t* pxq = w64(T.X, o.q, w, rs, T.il);
  uint64_t* pzq = w64(T.Z, o.q, w, rs, T.il);
  switch (op_kind(o)) {
    case OP_H:
    case OP_HI:                                    // swap; phase += 2 x z
      *pxq = zq; *pzq = xq;
      d1 ^= xq & zq;
      return 32;
    case OP_P:                                     // phase += x; z ^= x
      *pzq = zq ^ xq;
      d1 ^= d0 & xq; d0 ^= xq;
      return 24;
    case OP_PI:                                    // phase += 3 x; z ^= x
      *pzq = zq ^ xq;
      d1 ^= xq & ~d0; d0 ^= xq;
      return 24;
    case OP_SUM:
    case OP_SUMI:
      *w64(T.X, o.t, w, rs, T.il) = xt ^ xq;
      *pzq = zq ^ zt;
      return 48;
    case OP_PAULI:                                 // phase += 2 (b x - a z)
      d1 ^= ((op_b(o) & 1) ? xq : 0ull) ^ ((op_a(o) & 1) ? zq : 0ull);
      return 16;
    case OP_CZ:
    case OP_CZI:                                   // z[t] ^= x[c], z[c] ^= x[t]; phase += 2 x[c] x[t] (mod 4)
      *w64(T.Z, o.t, w, rs, T.il) = zt ^ xq;
      *pzq = zq ^ xt;
      d1 ^= xq & xt;
      return 48;
    case OP_SWAP:
      *pxq = xt; *pzq = zt;
      *w64(T.X, o.t, w, rs, T.il) = xq;
      *w64(T.Z, o.t, w, rs, T.il) = zq;
      return 64;
    case OP_NOISE: {
      uint8_t* snap = S.noise + op_noise_slot(o) * 2 * T.rs_full;
      ((uint64_t*)snap)[T.wofs + w] = xq;
      ((uint64_t*)(snap + T.rs_full))[T.wofs + w] = zq;
      return 16;
    }
  }
  return 0;
}

SDIM_DEV bool two_rows(const DevOp& o) {
  const uint32_t k = op_kind(o);
  return k == OP_SUM || k == OP_SUMI || k == OP_CZ || k == OP_CZI || k == OP_SWAP;
}

// ------------------------------------------------------------------ serial layers of small tableaux
// A tableau of at most TINY_ROWS rows whose row has at most 64 column words: in a serial layer
// the thread of word w keeps word w of every row in registers (a vector indexed by the op's
// qudit, which is the same in every lane), so an op costs no LDS round trip.  The updates are
// those of apply_odd / apply_bin on the register copies.  For d = 2 a row's G = 2n <= 32
// generators sit in the low half of word 0; the high half holds padding columns, which are zero
// from the start and stay zero (every update is column by column: gates and rowsums combine the
// entries of one column, and a rowsum adds f_i times the pivot column with f_i = -x[q, i] = 0
// there), so the register copy keeps 32 bits per row.
constexpr uint32_t TINY_ROWS = SDIM_TINY_ROWS;
typedef uint32_t RowsU32 __attribute__((ext_vector_type(TINY_ROWS)));

// Each op reads rows q and t, computes their new words in the switch (an op leaves a row's word
// alone unless it writes it) and writes row t, then row q, after it: the register vectors are
// written in one place, so no branch of the switch holds a copy of them.  Row t before row q is
// the order in which the LDS form stores them, so an op whose two qudits coincide ends the same.
template <int D>
SDIM_DEV uint32_t apply_odd_rows(const DevOp& o, RowsU32& X, RowsU32& Z, const Tab& T, uint32_t w, uint32_t& dl,
                                 const Scr& S) {
  const uint32_t q = o.q, t = o.t;
  const uint32_t xq = X[q], zq = Z[q], xt = X[t], zt = Z[t];
  uint32_t vxt = xt, vzt = zt, vxq = xq, vzq = zq, nb = 0;
  bool wxq = false, wzq = false;
  switch (op_kind(o)) {
    case OP_H:                                     // (x, z) <- (-z, x); phase += po (-z) x
      vxq = negm<D>(zq); vzq = xq; wxq = wzq = true;
      dl = subm<D>(dl, mulv<D>(xq, zq));
      nb = 16;
      break;
    case OP_HI:                                    // (x, z) <- (z, -x); phase += po z (-x)
      vxq = zq; vzq = negm<D>(xq); wxq = wzq = true;
      dl = subm<D>(dl, mulv<D>(xq, zq));
      nb = 16;
      break;
    case OP_P: {                                   // z += x; phase += x (x - 1) / 2
      const MulTab h = make_half_tab<D>();
      vzq = addm<D>(zq, xq); wzq = true;
      dl = addm<D>(dl, mult<D>(h, xq));
      nb = 12;
      break;
    }
    case OP_PI: {
      const MulTab h = make_half_tab<D>();
      vzq = subm<D>(zq, xq); wzq = true;
      dl = subm<D>(dl, mult<D>(h, xq));
      nb = 12;
      break;
    }
    case OP_SUM:                                   // x[t] += x[c]; z[c] += (d - 1) z[t]
      vxt = addm<D>(xt, xq);
      vzq = subm<D>(zq, zt); wzq = true;
      nb = 24;
      break;
    case OP_SUMI:
      vxt = subm<D>(xt, xq);
      vzq = addm<D>(zq, zt); wzq = true;
      nb = 24;
      break;
    case OP_PAULI:                                 // X^a Z^b: phase += b x - a z
      dl = addm<D>(dl, subm<D>(mult<D>(S.mtab[op_b(o)], xq), mult<D>(S.mtab[op_a(o)], zq)));
      nb = 8;
      break;
    case OP_CZ:                                    // z[t] += x[c], z[c] += x[t]; phase += x[c] x[t]
      vzt = addm<D>(zt, xq);
      vzq = addm<D>(zq, xt); wzq = true;
      dl = addm<D>(dl, mulv<D>(xq, xt));
      nb = 24;
      break;
    case OP_CZI:
      vzt = subm<D>(zt, xq);
      vzq = subm<D>(zq, xt); wzq = true;
      dl = subm<D>(dl, mulv<D>(xq, xt));
      nb = 24;
      break;
    case OP_SWAP:
      vxq = xt; vzq = zt; vxt = xq; vzt = zq; wxq = wzq = true;
      nb = 32;
      break;
    case OP_NOISE: {                               // snapshot of row q for the end-of-segment pass
      uint8_t* snap = S.noise + op_noise_slot(o) * 2 * T.rs_full;
      ((uint32_t*)snap)[T.wofs + w] = xq;
      ((uint32_t*)(snap + T.rs_full))[T.wofs + w] = zq;
      nb = 8;
      break;
    }
  }
  if (q == t) {                                    // row q holds what was just written to row t
    if (!wxq) vxq = vxt;
    if (!wzq) vzq = vzt;
  }
  X[t] = vxt; Z[t] = vzt;
  X[q] = vxq; Z[q] = vzq;
  return nb;
}

SDIM_DEV uint32_t apply_bin_rows(const DevOp& o, RowsU32& X, RowsU32& Z, const Tab& T, uint32_t w, uint64_t& d0,
                                 uint64_t& d1, const Scr& S) {
  const uint32_t q = o.q, t = o.t, k = op_kind(o);
  const uint32_t xq = X[q], zq = Z[q], xt = X[t], zt = Z[t];
  uint32_t vxt = xt, vzt = zt, vxq = xq, vzq = zq, nb = 0;
  bool wxq = false, wzq = false;
  if (k <= OP_SUMI) {
    // the common gates without a branch per kind: H, H^-1 swap x and z (phase += 2 x z); P, P^-1
    // add x to z (phase += x, 3x); SUM, SUM^-1 add x[c] to x[t] and z[t] to z[c]
    const bool h = k <= OP_HI, p = k == OP_P || k == OP_PI, s = k >= OP_SUM;
    const uint64_t pp = p ? (uint64_t)xq : 0ull;
    vxq = h ? zq : xq;
    vzq = h ? xq : (zq ^ (p ? xq : (s ? zt : 0u)));
    vxt = s ? (xt ^ xq) : xt;
    d1 ^= (h ? (uint64_t)(xq & zq) : 0ull) ^ (k == OP_PI ? (pp & ~d0) : (pp & d0));
    d0 ^= pp;
    wxq = h;
    wzq = true;
    nb = h ? 32 : (p ? 24 : 48);
  } else {
    switch (k) {
      case OP_PAULI:                               // phase += 2 (b x - a z)
        d1 ^= (uint64_t)(((op_b(o) & 1) ? xq : 0u) ^ ((op_a(o) & 1) ? zq : 0u));
        nb = 16;
        break;
      case OP_CZ:
      case OP_CZI:                                 // z[t] ^= x[c], z[c] ^= x[t]; phase += 2 x[c] x[t]
        vzt = zt ^ xq;
        vzq = zq ^ xt; wzq = true;
        d1 ^= (uint64_t)(xq & xt);
        nb = 48;
        break;
      case OP_SWAP:
        vxq = xt; vzq = zt; vxt = xq; vzt = zq; wxq = wzq = true;
        nb = 64;
        break;
      case OP_NOISE: {
        uint8_t* snap = S.noise + op_noise_slot(o) * 2 * T.rs_full;
        ((uint64_t*)snap)[T.wofs + w] = xq;
        ((uint64_t*)(snap + T.rs_full))[T.wofs + w] = zq;
        nb = 16;
        break;
      }
    }
  }
  if (q == t) {
    if (!wxq) vxq = vxt;
    if (!wzq) vzq = vzt;
  }
  X[t] = vxt; Z[t] = vzt;
  X[q] = vxq; Z[q] = vzq;
  return nb;
}

// One serial layer (L ops in program order) of a view with V.n <= TINY_ROWS rows and NWd <= 64
// words, run by the first wave: lane w loads word w of every row, applies the ops, stores the
// rows back.  Ops are fetched 64 per global round trip (one per lane) and broadcast by readlane.
// Phase increments accumulate in dl (odd d) or d0 / d1 (d = 2) as in the LDS form.
template <int D>
SDIM_DEV void ser_layer_rows(const DevOp* lops, uint32_t L, const Tab& V, uint32_t NWd, const Scr& S,
                             uint32_t& dl, uint64_t& d0, uint64_t& d1, unsigned long long& gb) {
  constexpr bool BIN = (D == 2);
  if (threadIdx.x >= 64) return;
  const uint32_t lane = threadIdx.x, w = lane, n = V.n;
  const bool mine = w < NWd;
  RowsU32 X = (RowsU32)0, Z = (RowsU32)0;
  if (mine) {
    #pragma unroll
    for (uint32_t r = 0; r < TINY_ROWS; r++) {
      if (r < n) {
        if constexpr (BIN) { X[r] = (uint32_t)*w64(V.X, r, w, V.rs, V.il); Z[r] = (uint32_t)*w64(V.Z, r, w, V.rs, V.il); }
        else { X[r] = *w32(V.X, r, w, V.rs, V.il); Z[r] = *w32(V.Z, r, w, V.rs, V.il); }
      }
    }
  }
  uint32_t nw0 = 0, nq = 0, ntg = 0;
  if (lane < L) { const DevOp& src = lops[lane]; nw0 = src.w0; nq = src.q; ntg = src.t; }
  for (uint32_t c0 = 0; c0 < L; c0 += 64) {
    const uint32_t cw0 = nw0, cq = nq, ctg = ntg;
    if (c0 + 64 + lane < L) { const DevOp& src = lops[c0 + 64 + lane]; nw0 = src.w0; nq = src.q; ntg = src.t; }
    const uint32_t cnt = (L - c0 < 64) ? L - c0 : 64;
    for (uint32_t b = 0; b < cnt; b++) {
      DevOp o;
      o.w0 = (uint32_t)__builtin_amdgcn_readlane((int)cw0, (int)b);
      o.q = (uint32_t)__builtin_amdgcn_readlane((int)cq, (int)b);
      o.t = (uint32_t)__builtin_amdgcn_readlane((int)ctg, (int)b);
      o.src = 0;
      if (!mine) continue;
      if constexpr (BIN) gb += apply_bin_rows(o, X, Z, V, w, d0, d1, S);
      else gb += apply_odd_rows<D>(o, X, Z, V, w, dl, S);
    }
  }
  if (mine) {
    #pragma unroll
    for (uint32_t r = 0; r < TINY_ROWS; r++) {
      if (r < n) {
        if constexpr (BIN) { *w64(V.X, r, w, V.rs, V.il) = (uint64_t)X[r]; *w64(V.Z, r, w, V.rs, V.il) = (uint64_t)Z[r]; }
        else { *w32(V.X, r, w, V.rs, V.il) = X[r]; *w32(V.Z, r, w, V.rs, V.il) = Z[r]; }
      }
    }
  }
}

// ------------------------------------------------------------------ gate segment
// The layers of one gate segment over the column words [0, NWd) of view V (the whole
// tableau, or one column tile staged in LDS); ops[kl] is the first layer header.  Returns
// the index after the last layer; leaves the segment's phase increments in DW.  Called
// once per view origin so that the accesses of each call compile to one address space.
template <int D, int NT, int OPB>
SDIM_DEV uint32_t seg_layers(const DevOp* ops, uint32_t kl, uint32_t nl, const Tab& V, uint32_t NWd,
                             const Scr& S, uint32_t lw, unsigned long long& gb) {
  constexpr bool BIN = (D == 2);
  constexpr int GB = 4;     // gates per load batch
  constexpr int OB = OPB;   // batches whose ops are loaded together
  const uint32_t tid = threadIdx.x;
  // word ownership: (word, slot) pairs when a row has at most NT words (the
  // slot partial deltas are reduced at the end); otherwise each thread owns
  // several whole words and adds each (layer, word) delta into DW directly.
  uint32_t NS, my_w, my_slot;
  const bool many = NWd > NT;
  if (NWd == 0) { NS = 1; my_w = 0; my_slot = 1; }           // a gang member without columns
  else if (!many) { NS = NT / NWd; my_w = tid % NWd; my_slot = tid / NWd; }
  else { NS = 1; my_w = tid; my_slot = 0; }
  const bool active = my_slot < NS;
  if (many) {
    for (uint32_t w = tid; w < NWd; w += NT) {
      if constexpr (BIN) { ((uint64_t*)S.DW)[V.wofs + w] = 0; ((uint64_t*)(S.DW + lw))[V.wofs + w] = 0; }
      else ((uint32_t*)S.DW)[V.wofs + w] = 0;
    }
  }
  uint32_t dl = 0;
  uint64_t b0 = 0, b1 = 0;

  for (uint32_t l = 0; l < nl; l++) {
    const DevOp lh = ops[kl];
    const uint32_t L = lh.q;
    const DevOp* lops = ops + kl + 1;
    // a serial layer holds ops in program order that may share qudits: the thread of every
    // word applies them one after another (a word's generators are independent of the other
    // words', so no barrier is needed between them).  Each wave that holds such threads
    // fetches 64 ops per global round trip, one per lane, the next 64 in flight, and
    // broadcasts them with readlane.
    const bool ser = lh.t != 0;
    // one-wave workgroups (the launch class of small tableaux) keep a small tableau's rows in
    // registers through a serial layer
    if (NT == 64 && ser && V.n <= TINY_ROWS && NWd <= 64) {
      ser_layer_rows<D>(lops, L, V, NWd, S, dl, b0, b1, gb);
    } else if (ser) {
      const uint32_t nwv = ((NWd < NT ? NWd : NT) + 63) / 64;
      if (tid / 64 < nwv) {
        const uint32_t lane = tid & 63;
        for (uint32_t wb = 0; wb < NWd; wb += NT) {
          const uint32_t w = wb + tid;
          const bool mine = w < NWd;
          if (many) { dl = 0; b0 = 0; b1 = 0; }
          uint32_t nw0 = 0, nq = 0, ntg = 0;
          if (lane < L) { const DevOp& src = lops[lane]; nw0 = src.w0; nq = src.q; ntg = src.t; }
          for (uint32_t c0 = 0; c0 < L; c0 += 64) {
            const uint32_t cw0 = nw0, cq = nq, ctg = ntg;
            if (c0 + 64 + lane < L) { const DevOp& src = lops[c0 + 64 + lane]; nw0 = src.w0; nq = src.q; ntg = src.t; }
            const uint32_t cnt = (L - c0 < 64) ? L - c0 : 64;
            for (uint32_t b = 0; b < cnt; b++) {
              DevOp o;
              o.w0 = (uint32_t)__builtin_amdgcn_readlane((int)cw0, (int)b);
              o.q = (uint32_t)__builtin_amdgcn_readlane((int)cq, (int)b);
              o.t = (uint32_t)__builtin_amdgcn_readlane((int)ctg, (int)b);
              o.src = 0;
              if (!mine) continue;
              if constexpr (BIN) {
                const uint64_t xq = *w64(V.X, o.q, w, V.rs, V.il), zq = *w64(V.Z, o.q, w, V.rs, V.il);
                uint64_t xt = 0, zt = 0;
                if (two_rows(o)) { xt = *w64(V.X, o.t, w, V.rs, V.il); zt = *w64(V.Z, o.t, w, V.rs, V.il); }
                gb += apply_bin(o, xq, zq, xt, zt, V, w, b0, b1, S);
              } else {
                const uint32_t xq = *w32(V.X, o.q, w, V.rs, V.il), zq = *w32(V.Z, o.q, w, V.rs, V.il);
                uint32_t xt = 0, zt = 0;
                if (two_rows(o)) { xt = *w32(V.X, o.t, w, V.rs, V.il); zt = *w32(V.Z, o.t, w, V.rs, V.il); }
                gb += apply_odd<D>(o, xq, zq, xt, zt, V, w, dl, S);
              }
            }
          }
          if (many && mine) {
            if constexpr (BIN) {
              uint64_t* p0 = (uint64_t*)S.DW + V.wofs + w;
              uint64_t* p1 = (uint64_t*)(S.DW + lw) + V.wofs + w;
              const uint64_t a0 = *p0, a1 = *p1;
              *p1 = a1 ^ b1 ^ (a0 & b0);
              *p0 = a0 ^ b0;
            } else {
              uint32_t* pd = (uint32_t*)S.DW + V.wofs + w;
              *pd = addm<D>(*pd, dl);
            }
          }
        }
      }
    } else if (active) {
      for (uint32_t w = my_w; w < NWd; w += NT) {
        if (many) { dl = 0; b0 = 0; b1 = 0; }
        // the ops of OB consecutive batches are loaded first (one global round trip), then the
        // rows of each batch of GB gates (LDS or global) before computing it
        for (uint32_t g0 = my_slot; g0 < L; g0 += NS * GB * OB) {
          DevOp oa[GB * OB];
          #pragma unroll
          for (int b = 0; b < GB * OB; b++) {
            const uint32_t g = g0 + (uint32_t)b * NS;
            if (g < L) { const DevOp& src = lops[g]; oa[b].w0 = src.w0; oa[b].q = src.q; oa[b].t = src.t; oa[b].src = 0; }
          }
          #pragma unroll
          for (int h = 0; h < OB; h++) {
          const uint32_t gh = g0 + (uint32_t)(h * GB) * NS;
          if (gh >= L) break;
          const DevOp* ob = oa + h * GB;
          if constexpr (BIN) {
            uint64_t xq[GB], zq[GB], xt[GB], zt[GB];
            #pragma unroll
            for (int b = 0; b < GB; b++) {
              const uint32_t g = gh + (uint32_t)b * NS;
              if (g < L) {
                xq[b] = *w64(V.X, ob[b].q, w, V.rs, V.il);
                zq[b] = *w64(V.Z, ob[b].q, w, V.rs, V.il);
                if (two_rows(ob[b])) { xt[b] = *w64(V.X, ob[b].t, w, V.rs, V.il); zt[b] = *w64(V.Z, ob[b].t, w, V.rs, V.il); }
                else { xt[b] = 0; zt[b] = 0; }
              }
            }
            #pragma unroll
            for (int b = 0; b < GB; b++)
              if (gh + (uint32_t)b * NS < L) gb += apply_bin(ob[b], xq[b], zq[b], xt[b], zt[b], V, w, b0, b1, S);
          } else {
            uint32_t xq[GB], zq[GB], xt[GB], zt[GB];
            #pragma unroll
            for (int b = 0; b < GB; b++) {
              const uint32_t g = gh + (uint32_t)b * NS;
              if (g < L) {
                xq[b] = *w32(V.X, ob[b].q, w, V.rs, V.il);
                zq[b] = *w32(V.Z, ob[b].q, w, V.rs, V.il);
                if (two_rows(ob[b])) { xt[b] = *w32(V.X, ob[b].t, w, V.rs, V.il); zt[b] = *w32(V.Z, ob[b].t, w, V.rs, V.il); }
                else { xt[b] = 0; zt[b] = 0; }
              }
            }
            #pragma unroll
            for (int b = 0; b < GB; b++)
              if (gh + (uint32_t)b * NS < L) gb += apply_odd<D>(ob[b], xq[b], zq[b], xt[b], zt[b], V, w, dl, S);
          }
          }
        }
        if (many) {
          if constexpr (BIN) {
            uint64_t* p0 = (uint64_t*)S.DW + V.wofs + w;
            uint64_t* p1 = (uint64_t*)(S.DW + lw) + V.wofs + w;
            const uint64_t a0 = *p0, a1 = *p1;
            *p1 = a1 ^ b1 ^ (a0 & b0);
            *p0 = a0 ^ b0;
          } else {
            uint32_t* pd = (uint32_t*)S.DW + V.wofs + w;
            *pd = addm<D>(*pd, dl);
          }
        }
        if (!many) break;
      }
    }
    kl += 1 + L;
    __syncthreads();
  }

  // reduce the slot deltas into DW
  if (!many) {
    if constexpr (BIN) {
      uint64_t* r = (uint64_t*)S.red;
      if (active) { r[2 * tid] = b0; r[2 * tid + 1] = b1; }
      __syncthreads();
      for (uint32_t w = tid; w < NWd; w += NT) {
        uint64_t a0 = 0, a1 = 0;
        for (uint32_t s = 0; s < NS; s++) {
          const uint32_t src = s * NWd + w;
          const uint64_t c0 = r[2 * src], c1 = r[2 * src + 1];
          a1 ^= c1 ^ (a0 & c0); a0 ^= c0;
        }
        ((uint64_t*)S.DW)[V.wofs + w] = a0;
        ((uint64_t*)(S.DW + lw))[V.wofs + w] = a1;
      }
    } else {
      if (active) S.red[tid] = dl;
      __syncthreads();
      for (uint32_t w = tid; w < NWd; w += NT) {
        uint32_t a = 0;
        for (uint32_t s = 0; s < NS; s++) a = addm<D>(a, S.red[s * NWd + w]);
        ((uint32_t*)S.DW)[V.wofs + w] = a;
      }
    }
  }
  __syncthreads();
  return kl;
}

// ops[k] is the OP_SEG header.  Returns the index after the segment.
template <int D, int NT, int OPB = 2>
SDIM_DEV uint32_t run_segment(const DevOp* ops, uint32_t k, const Tab& T, const Scr& S,
                              const ItemDesc& it, const KParams& P, unsigned long long& bytes,
                              unsigned long long& lbytes) {
  constexpr bool BIN = (D == 2);
  constexpr int GB = 4;     // gates per load batch
  constexpr int OB = OPB;   // batches whose ops are loaded together
  SDIM_TMARK(tSeg);
  const DevOp hdr = ops[k];
  const uint32_t nl = hdr.q, nn = hdr.t;
  const uint32_t n = T.n, rs = T.rs, Sh = T.S;
  const uint32_t tid = threadIdx.x;
  const uint32_t ndesc = k + 1;
  k += 1 + nn;
  unsigned long long gb = 0;                   // bytes of the gate updates

  // noise draws of the segment: mask[j][S/32] bits + ab[j][S] bytes
  const uint32_t mw = (Sh + 31) / 32;
  const uint32_t lw = T.rs_full;                // width of the LDS per-column arrays
  uint32_t* nmask = (uint32_t*)(S.noise + nn * 2 * lw);
  uint8_t* nab = (uint8_t*)(nmask + nn * mw);
  if (nn) {
    for (uint32_t i = tid; i < nn * mw; i += NT) nmask[i] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < nn * Sh; i += NT) {
      const uint32_t j = i / Sh, s = i - j * Sh;
      const DevOp nd = ops[ndesc + j];
      const U64x4 o = draw(it.seed, it.stream, nd.src, it.shot0 + s);
      const uint32_t ab = noise_ab(o, D, op_a(nd), P.keep53[it.keep_off + nd.t]);
      nab[i] = (uint8_t)(ab & 0xFF);
      if (ab) atomicOr(&nmask[j * mw + (s >> 5)], 1u << (s & 31));
    }
    __syncthreads();
  }

  // Column tiles: an HBM-resident tableau is staged through LDS, TW bytes of every
  // row at a time (gates act on each generator column independently, so the tiles
  // need no coordination); an LDS-resident tableau is a single tile.
  uint32_t TW = rs;
  uint8_t* tile = nullptr;
  // phases cached in LDS are parked at their global place while the tiles run (a segment
  // without noise ops touches them only in its final phase pass); the tiles then also
  // cover their LDS
  const bool park = S.tiling && (it.flags & ITEM_PH_LDS) && nn == 0;
  uint8_t* ph_home = P.arena + it.tab_off + 2ull * rows_padded(n) * rs;
  const uint32_t ph_bytes = (BIN ? 2u : 1u) * Sh * rs;
  if (park) {
    copy16<NT>(ph_home, S.gap, ph_bytes);
    __syncthreads();
  }
  if (S.tiling) {
    tile = park ? S.gap : S.noise + ((noise_bytes(nn, lw, Sh) + 15u) & ~15u);
    const int64_t left = (int64_t)(S.lds_end - tile);
    const uint64_t room = left > 0 ? (uint64_t)left : 0ull;
    uint64_t tw = room / (2ull * rows_padded(n));
    tw &= ~15ull;
    TW = tw >= 16 ? (uint32_t)(tw < rs ? tw : rs) : 0;
  }
  const uint32_t CPWb = BIN ? 8 : 4;            // bytes per column word
  for (uint32_t c0 = 0; c0 < rs; c0 += (TW ? TW : rs)) {
    uint32_t vw = rs;                            // bytes of each row in this view
    if (S.tiling && TW) {
      vw = (rs - c0 < TW) ? rs - c0 : TW;
      // load the tile: pieces of 2^il rows x 16 bytes (one block of the interleave) of the X
      // and Z blocks; the tile keeps the interleave with row stride vw
      const uint32_t il = T.il;
      const uint32_t cpr = vw / 16, nrb = (2 * rows_padded(n)) >> il;
      for (uint32_t i = tid; i < (nrb * cpr) << il; i += NT) {
        const uint32_t ch = i >> il, k = i & ((1u << il) - 1u);
        const uint32_t rb = ch / cpr, cc = ch - rb * cpr;
        const uint8_t* src = T.X + (((size_t)rb * rs + c0 + 16 * cc) << il) + 16 * k;
        *(uint4*)(tile + (((size_t)rb * vw + 16 * cc) << il) + 16 * k) = *(const uint4*)src;
        bytes += 16;
      }
      __syncthreads();
    }
    uint32_t kl;
    if (S.tiling && TW) {
      Tab Vt = T;                                // a view of LDS only
      Vt.X = tile; Vt.Z = tile + (size_t)rows_padded(n) * vw; Vt.rs = vw; Vt.wofs = c0 / CPWb;
      kl = seg_layers<D, NT, OPB>(ops, k, nl, Vt, vw / CPWb, S, lw, gb);
    } else {
      kl = seg_layers<D, NT, OPB>(ops, k, nl, T, T.nw, S, lw, gb);
    }

    if (S.tiling && TW) {
      // write the tile back
      const uint32_t il = T.il;
      const uint32_t cpr = vw / 16, nrb = (2 * rows_padded(n)) >> il;
      for (uint32_t i = tid; i < (nrb * cpr) << il; i += NT) {
        const uint32_t ch = i >> il, k = i & ((1u << il) - 1u);
        const uint32_t rb = ch / cpr, cc = ch - rb * cpr;
        uint8_t* dst = T.X + (((size_t)rb * rs + c0 + 16 * cc) << il) + 16 * k;
        *(uint4*)dst = *(const uint4*)(tile + (((size_t)rb * vw + 16 * cc) << il) + 16 * k);
        bytes += 16;
      }
      __syncthreads();
    }
    if (c0 + vw >= rs) { k = kl; break; }
  }
  // gate traffic lands in LDS when the tableau was staged through tiles
  if (S.tiling && TW) lbytes += gb; else bytes += gb;
  if (park) {
    copy16<NT>(S.gap, ph_home, ph_bytes);
    __syncthreads();
    if (tid == 0) bytes += 2ull * ph_bytes;
  }
  const uint32_t NWd = T.nw;
  SDIM_TADD(P, 7, tSeg);
  SDIM_TMARK(tPh);
  // per-shot phase pass: Ph_s += delta + noise terms of shot s
  for (uint32_t i = tid; i < Sh * NWd; i += NT) {
    const uint32_t s = i / NWd, w = i - s * NWd;
    if constexpr (BIN) {
      uint64_t* p0 = plane(T, s, 0) + w;
      uint64_t* p1 = plane(T, s, 1) + w;
      const uint64_t c0 = ((const uint64_t*)S.DW)[w];
      uint64_t c1 = ((const uint64_t*)(S.DW + lw))[w];
      for (uint32_t j = 0; j < nn; j++) {
        if (nmask[j * mw + (s >> 5)] & (1u << (s & 31))) {
          const uint32_t ab = nab[j * Sh + s];
          const uint8_t* snap = S.noise + j * 2 * lw;
          const uint64_t sx = ((const uint64_t*)snap)[w], sz = ((const uint64_t*)(snap + lw))[w];
          c1 ^= (((ab >> 4) & 1) ? sx : 0ull) ^ ((ab & 1) ? sz : 0ull);
        }
      }
      if (c0 | c1) {
        const uint64_t a0 = *p0, a1 = *p1;
        *p1 = a1 ^ c1 ^ (a0 & c0);
        *p0 = a0 ^ c0;
        bytes += 32;
      }
    } else {
      uint32_t* ph = (uint32_t*)phase_vec(T, s) + w;
      uint32_t c = ((const uint32_t*)S.DW)[w];
      for (uint32_t j = 0; j < nn; j++) {
        if (nmask[j * mw + (s >> 5)] & (1u << (s & 31))) {
          const uint32_t ab = nab[j * Sh + s];
          const uint8_t* snap = S.noise + j * 2 * lw;
          const uint32_t sx = ((const uint32_t*)snap)[w], sz = ((const uint32_t*)(snap + lw))[w];
          c = addm<D>(c, subm<D>(mult<D>(S.mtab[(ab >> 4) & 15], sx), mult<D>(S.mtab[ab & 15], sz)));
        }
      }
      if (c) {
        *ph = addm<D>(*ph, c);
        bytes += 8;
      }
    }
  }
  __syncthreads();
  SDIM_TADD(P, 8, tPh);
  return k;
}

// ------------------------------------------------------------------ rowsum steps on 16-byte column groups
// Shared by the one-workgroup measurement and the gang measurement.  The view V
// covers column words [w0, w0 + V.nw) of every row (V.nw a multiple of the words per
// 16-byte group); LDS arrays (F, DOT, FL) are indexed by the view's own words.

// Step A: groups of row q -> rowsum factors f = -x[q, .] in F, DOT cleared, FL = groups
// with some f != 0 (count in M[2]), pivot candidate over the stabilizer columns
// (atomic-min into M[0]).  rq != nullptr also keeps a copy of row q.
template <int D, int NT>
SDIM_DEV void rowq_groups(const Tab& V, uint32_t w0, uint32_t q, const Scr& S, int* M, uint8_t* rq,
                          unsigned long long& bytes) {
  constexpr bool BIN = (D == 2);
  constexpr uint32_t CPW = BIN ? 64 : 4;
  constexpr uint32_t WPG = BIN ? 2 : 4;              // words per 16-byte group
  const uint32_t tid = threadIdx.x;
  const uint32_t n = V.n, rs = V.rs, nw = V.nw;
  const uint32_t nsw = (n + CPW - 1) / CPW;
  const uint32_t ngr = nw / WPG;
  for (uint32_t g = tid; g < ngr; g += NT) {
    const uint4 v4 = *(const uint4*)b8(V.X, q, 16 * g, rs, V.il);
    if (rq) *(uint4*)(rq + 16 * g) = v4;
    uint4 f4;
    bool any;
    if constexpr (BIN) {
      const uint64_t v0 = (uint64_t)v4.x | ((uint64_t)v4.y << 32), v1 = (uint64_t)v4.z | ((uint64_t)v4.w << 32);
      f4 = v4;                                        // f = -x = x mod 2
      any = (v0 | v1) != 0;
      #pragma unroll
      for (int k = 0; k < 2; k++) {
        const uint32_t gw = w0 + 2 * g + k;
        uint64_t vs = k ? v1 : v0;
        if (gw < nsw && vs) {
          const uint32_t lim = n - 64 * gw;
          if (lim < 64) vs &= (1ull << lim) - 1;
          if (vs) atomicMin(&M[0], (int)(64 * gw + __builtin_ctzll(vs)));
        }
      }
    } else {
      const uint32_t vv[4] = {v4.x, v4.y, v4.z, v4.w};
      f4 = make_uint4(negm<D>(v4.x), negm<D>(v4.y), negm<D>(v4.z), negm<D>(v4.w));
      any = (f4.x | f4.y | f4.z | f4.w) != 0;
      #pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint32_t gw = w0 + 4 * g + k;
        uint32_t vs = vv[k];
        if (gw < nsw && vs) {
          const uint32_t lim = n - 4 * gw;
          if (lim < 4) vs &= (1u << (8 * lim)) - 1;
          if (vs) atomicMin(&M[0], (int)(4 * gw + (__builtin_ctz(vs) >> 3)));
        }
      }
    }
    *(uint4*)(S.F + 16 * g) = f4;
    *(uint4*)(S.DOT + 2 * WPG * g) = make_uint4(0, 0, 0, 0);                // two dot words per column word
    if constexpr (!BIN) *(uint4*)(S.DOT + 2 * WPG * g + 4) = make_uint4(0, 0, 0, 0);
    if (any) {
      const uint32_t k = (uint32_t)atomicAdd(&M[2], 1);
      if (k < S.fl_cap) S.FL[k] = (uint16_t)g;
      else report_error(S, KERR_FACTOR_GROUPS);
    }
    bytes += 16;
  }
}

// Step C: X += e x_p f, Z += e z_p f over (touched rows TR[0, nT)) x (groups FL[0, nF)),
// column dots <z_i, x_p> of the old Z accumulated into DOT through LDS atomics.
template <int D, int NT, int RB>
SDIM_DEV void sweep_groups(const Tab& V, const Scr& S, uint32_t nT, uint32_t nF, uint32_t e,
                           unsigned long long& bytes) {
  constexpr bool BIN = (D == 2);
  const uint32_t tid = threadIdx.x;
  const uint32_t rs = V.rs;
    if (nT && nF) {
      // NRG threads per group.  With row blocks (il > 0) adjacent lanes take adjacent touched
      // rows of one group, so touched rows of one block share a line within a load
      // instruction; with contiguous rows (il = 0) adjacent lanes take adjacent groups of a row.
      uint32_t NRG, j0, g0, jstep;
      if (nF >= NT) { NRG = 1; j0 = tid; g0 = 0; jstep = NT; }
      else if (V.il) { NRG = NT / nF; j0 = tid / NRG; g0 = tid % NRG; jstep = nF; }
      else { NRG = NT / nF; j0 = tid % nF; g0 = tid / nF; jstep = nF; }
      if (g0 < NRG) {
        for (uint32_t j = j0; j < nF; j += jstep) {
          const uint32_t gi = S.FL[j];
          const uint4 f4 = *(const uint4*)(S.F + 16 * gi);
          if constexpr (BIN) {
            const uint64_t f0 = (uint64_t)f4.x | ((uint64_t)f4.y << 32), f1 = (uint64_t)f4.z | ((uint64_t)f4.w << 32);
            uint64_t da0 = 0, da1 = 0;
            for (uint32_t k0 = g0; k0 < nT; k0 += NRG * RB) {
              uint32_t rr[RB], sxv[RB], szv[RB];
              uint4 xv[RB], zv[RB];
              #pragma unroll
              for (int b = 0; b < RB; b++) {
                const uint32_t kk = k0 + (uint32_t)b * NRG;
                if (kk < nT) {
                  rr[b] = S.TR[kk];
                  sxv[b] = S.xp[rr[b]]; szv[b] = S.zp[rr[b]];
                  zv[b] = *(const uint4*)b8(V.Z, rr[b], 16 * gi, rs, V.il);
                  xv[b] = sxv[b] ? *(const uint4*)b8(V.X, rr[b], 16 * gi, rs, V.il) : make_uint4(0, 0, 0, 0);
                }
              }
              #pragma unroll
              for (int b = 0; b < RB; b++) {
                const uint32_t kk = k0 + (uint32_t)b * NRG;
                if (kk < nT) {
                  const uint64_t z0 = (uint64_t)zv[b].x | ((uint64_t)zv[b].y << 32), z1 = (uint64_t)zv[b].z | ((uint64_t)zv[b].w << 32);
                  if (sxv[b]) {
                    da0 ^= z0; da1 ^= z1;
                    uint4 o = xv[b];
                    o.x ^= f4.x; o.y ^= f4.y; o.z ^= f4.z; o.w ^= f4.w;
                    *(uint4*)b8(V.X, rr[b], 16 * gi, rs, V.il) = o;
                    bytes += 32;
                  }
                  if (szv[b]) {
                    uint4 o = zv[b];
                    o.x ^= f4.x; o.y ^= f4.y; o.z ^= f4.z; o.w ^= f4.w;
                    *(uint4*)b8(V.Z, rr[b], 16 * gi, rs, V.il) = o;
                    bytes += 16;
                  }
                  bytes += 16;
                }
              }
            }
            da0 &= f0; da1 &= f1;
            if (da0) { atomicXor(&S.DOT[4 * gi], (uint32_t)da0); atomicXor(&S.DOT[4 * gi + 1], (uint32_t)(da0 >> 32)); }
            if (da1) { atomicXor(&S.DOT[4 * gi + 2], (uint32_t)da1); atomicXor(&S.DOT[4 * gi + 3], (uint32_t)(da1 >> 32)); }
          } else {
            const uint32_t fw[4] = {f4.x, f4.y, f4.z, f4.w};
            uint32_t de[4] = {0, 0, 0, 0}, dO[4] = {0, 0, 0, 0};
            uint32_t cnt = 0;
            for (uint32_t k0 = g0; k0 < nT; k0 += NRG * RB) {
              uint32_t rr[RB], sxv[RB], szv[RB];
              uint4 xv[RB], zv[RB];
              #pragma unroll
              for (int b = 0; b < RB; b++) {
                const uint32_t kk = k0 + (uint32_t)b * NRG;
                if (kk < nT) {
                  rr[b] = S.TR[kk];
                  sxv[b] = (e * S.xp[rr[b]]) % D;
                  szv[b] = (e * S.zp[rr[b]]) % D;
                  zv[b] = *(const uint4*)b8(V.Z, rr[b], 16 * gi, rs, V.il);
                  xv[b] = sxv[b] ? *(const uint4*)b8(V.X, rr[b], 16 * gi, rs, V.il) : make_uint4(0, 0, 0, 0);
                }
              }
              #pragma unroll
              for (int b = 0; b < RB; b++) {
                const uint32_t kk = k0 + (uint32_t)b * NRG;
                if (kk < nT) {
                  const uint32_t zw[4] = {zv[b].x, zv[b].y, zv[b].z, zv[b].w};
                  if (sxv[b]) {
                    const MulTab& t = S.mtab[sxv[b]];
                    const uint32_t xw[4] = {xv[b].x, xv[b].y, xv[b].z, xv[b].w};
                    uint32_t ow[4];
                    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                      const uint32_t pz = mult<D>(t, zw[k]);
                      de[k] += pz & 0x00FF00FFu;
                      dO[k] += (pz >> 8) & 0x00FF00FFu;
                      ow[k] = addm<D>(xw[k], mult<D>(t, fw[k]));
                    }
                    *(uint4*)b8(V.X, rr[b], 16 * gi, rs, V.il) = make_uint4(ow[0], ow[1], ow[2], ow[3]);
                    bytes += 32;
                  }
                  if (szv[b]) {
                    const MulTab& t = S.mtab[szv[b]];
                    *(uint4*)b8(V.Z, rr[b], 16 * gi, rs, V.il) =
                        make_uint4(addm<D>(zw[0], mult<D>(t, fw[0])), addm<D>(zw[1], mult<D>(t, fw[1])),
                                   addm<D>(zw[2], mult<D>(t, fw[2])), addm<D>(zw[3], mult<D>(t, fw[3])));
                    bytes += 16;
                  }
                  bytes += 16;
                }
              }
              if (++cnt == 512) {
                #pragma unroll
                for (int k = 0; k < 4; k++) { de[k] = fold16<D>(de[k]); dO[k] = fold16<D>(dO[k]); }
                cnt = 0;
              }
            }
            #pragma unroll
            for (int k = 0; k < 4; k++) {
              const uint32_t a = fold16<D>(de[k]), b = fold16<D>(dO[k]);
              if (a) atomicAdd(&S.DOT[2 * (4 * gi + k)], a);
              if (b) atomicAdd(&S.DOT[2 * (4 * gi + k) + 1], b);
            }
          }
        }
      }
    }
}

// Step D (phases): ph_i += f_i ph_p' + f_i <z_i, x_p> + xz f_i (f_i - 1) / 2 for every shot and
// every word of the groups in FL (ph_p' = exponentiated pivot phase of the shot).  For odd d
// the last two terms come from DOT as rowsum_shared_terms left them.
template <int D, int NT>
SDIM_DEV void rowsum_phases(const Tab& V, const Scr& S, uint32_t nF, uint32_t e, uint32_t xz, uint32_t pinc,
                            unsigned long long& bytes) {
  constexpr bool BIN = (D == 2);
  constexpr uint32_t WPG = BIN ? 2 : 4;
  const uint32_t tid = threadIdx.x;
  const uint32_t Sh = V.S;
    for (uint32_t i = tid; i < Sh * nF * WPG; i += NT) {
      const uint32_t s = i / (nF * WPG), jk = i - s * nF * WPG;
      const uint32_t w = S.FL[jk / WPG] * WPG + jk % WPG;
      if constexpr (BIN) {
        const uint32_t k2 = S.shot[s];
        const uint64_t f = ((const uint64_t*)S.F)[w];
        if (!f) continue;
        const uint64_t cdot = ((uint64_t)S.DOT[2 * w] | ((uint64_t)S.DOT[2 * w + 1] << 32)) & f;
        uint64_t* p0 = plane(V, s, 0) + w;
        uint64_t* p1 = plane(V, s, 1) + w;
        uint64_t a0 = *p0, a1 = *p1;
        const uint64_t k0m = (k2 & 1) ? f : 0ull, k1m = (k2 & 2) ? f : 0ull;
        a1 ^= k1m ^ (a0 & k0m);
        a0 ^= k0m;
        a1 ^= cdot;
        *p0 = a0; *p1 = a1;
        bytes += 32;
      } else {
        const uint32_t f = ((const uint32_t*)S.F)[w];
        if (!f) continue;
        const uint32_t php = (e * S.shot[s] + pinc) % D;
        uint32_t* ph = (uint32_t*)phase_vec(V, s) + w;
        *ph = addm<D>(addm<D>(*ph, mult<D>(S.mtab[php], f)), S.DOT[2 * w]);
        bytes += 8;
      }
    }
}

// Odd d, before step D: the shot-independent part of each listed word's increment,
// f <z_i, x_p> + xz f (f - 1) / 2 (bytes mod d), replaces the word's first dot word.
template <int D, int NT>
SDIM_DEV void rowsum_shared_terms(const Scr& S, uint32_t nF, uint32_t xz) {
  constexpr uint32_t WPG = 4;
  for (uint32_t i = threadIdx.x; i < nF * WPG; i += NT) {
    const uint32_t w = S.FL[i / WPG] * WPG + i % WPG;
    const uint32_t f = ((const uint32_t*)S.F)[w];
    if (!f) continue;
    const uint32_t de = S.DOT[2 * w], dO = S.DOT[2 * w + 1];
    uint32_t c = 0;
    #pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const uint32_t fb = byte_of(f, kb);
      const uint32_t dot = ((kb & 1) ? (dO >> (16 * (kb >> 1))) : (de >> (16 * (kb >> 1)))) & 0xFFFFu;
      c |= ((fb * (dot % D) + xz * ((fb * (fb - 1) / 2) % D)) % D) << (8 * kb);
    }
    S.DOT[2 * w] = c;
  }
}

// ------------------------------------------------------------------ snapshots
// The item's whole tableau (X/Z blocks, then every shot's phase vector) copied to snapshot
// `slot` in the item's own layout (the host converts it like a final tableau).
template <int D, int NT>
SDIM_DEV void snapshot(const Tab& T, const ItemDesc& it, const KParams& P, uint32_t slot) {
  __syncthreads();
  const uint64_t xz = 2ull * rows_padded(T.n) * T.rs, ph = (D == 2 ? 2ull : 1ull) * T.S * T.rs;
  uint8_t* dst = P.arena + it.snap_off + (uint64_t)slot * snapshot_stride(T.n, D, T.S);
  copy16<NT>(dst, T.X, xz);
  copy16<NT>(dst + xz, T.Ph, ph);
  __syncthreads();   // the copy reads the tableau before the next step writes it
}

// ------------------------------------------------------------------ measurement
// Z-basis measurement of qudit q (tableau_prime.py:262-363) followed, for
// RESET, by X^((-m) mod d) (program.py:334-339).  Writes one record per shot.
// mk counts the measurements of the item (selects the counter set).
template <int D, int NT, int RB = 1>
SDIM_DEV void measure(const Tab& T, uint32_t q, uint32_t slot, uint32_t src, bool reset, uint32_t mk,
                      const Scr& S, const ItemDesc& it, const KParams& P, unsigned long long& bytes) {
  constexpr bool BIN = (D == 2);
  constexpr uint32_t CPW = BIN ? 64 : 4;
  const uint32_t tid = threadIdx.x;
  const uint32_t n = T.n, rs = T.rs, Sh = T.S;
  const uint32_t G = 2 * n;
  const uint32_t NWd = (G + CPW - 1) / CPW;
  const uint32_t nsw = (n + CPW - 1) / CPW;
  int* M = S.misc + 16 * (mk & 1);
  int* Mn = S.misc + 16 * ((mk + 1) & 1);
  uint8_t* rec = P.arena + it.rec_off + (size_t)slot * Sh;
  SDIM_TMARK(tA);

  // random outcomes do not depend on the tableau: draw them first (step E uses them when
  // the measurement turns out to be random)
  const bool early = SDIM_EARLY_DRAWS && !(it.flags & ITEM_LDS);
  if (early)
    for (uint32_t s = tid; s < Sh; s += NT) S.shot[Sh + s] = outcome_of(draw(it.seed, it.stream, src, it.shot0 + s), D);
  // ---- A: row q -> pivot candidate, f, groups with f != 0, copy of row q
  rowq_groups<D, NT>(T, 0, q, S, M, S.RQ, bytes);
  __syncthreads();                                                        // B1
  SDIM_TADD(P, 0, tA);
  SDIM_TMARK(tB);
  const int piv = M[0];
  if (tid == 0) { Mn[0] = INT_MAX; Mn[1] = 0; Mn[2] = 0; Mn[3] = 0; }

  if (piv != INT_MAX) {
    const uint32_t p = (uint32_t)piv;
    // exponentiation of stabilizer p by e = x[q,p]^-1 (tableau_prime.py:278-282, 365-380),
    // folded into every later use of column p
    uint32_t e = 1;
    if constexpr (!BIN) {
      const uint32_t xv = S.RQ[p];
      #pragma unroll
      for (uint32_t c = 1; c < D; c++) if ((xv * c) % D == 1) e = c;
    }
    // ---- B: gather the pivot column, touched rows, <x_p, z_p>, pivot phases.  Columns p and
    // n+p get their final values here, from the values just read: stabilizer p becomes Z_q at
    // the touched rows (it is zero elsewhere), destabilizer n+p becomes e times stabilizer p at
    // every row (its old support is not known).  The sweep reads these groups after the full
    // barrier below and keeps both columns, f being 0 there.
    const uint32_t np_ = n + p;
    for (uint32_t r = tid; r < n; r += NT) {
      uint32_t xr, zr;
      if constexpr (BIN) {
        const uint32_t wp = p >> 6, wn = np_ >> 6;
        uint64_t* xpw = w64(T.X, r, wp, rs, T.il);
        uint64_t* zpw = w64(T.Z, r, wp, rs, T.il);
        uint64_t* xnw = w64(T.X, r, wn, rs, T.il);
        uint64_t* znw = w64(T.Z, r, wn, rs, T.il);
        const bool sep = wn != wp;
        const uint64_t xa = *xpw, za = *zpw;
        const uint64_t xb = sep ? *xnw : xa, zb = sep ? *znw : za;
        xr = (uint32_t)((xa >> (p & 63)) & 1);
        zr = (uint32_t)((za >> (p & 63)) & 1);
        const uint64_t mp = 1ull << (p & 63), mn = 1ull << (np_ & 63);
        uint64_t xn = (xb & ~mn) | (xr ? mn : 0ull), zn = (zb & ~mn) | (zr ? mn : 0ull);
        bytes += sep ? 32 : 16;
        if (!sep && (xr | zr)) {
          xn &= ~mp;
          zn = (zn & ~mp) | (r == q ? mp : 0ull);
        }
        if (xn != xb) { *xnw = xn; bytes += 8; }
        if (zn != zb) { *znw = zn; bytes += 8; }
        if (sep && (xr | zr)) {
          *xpw = xa & ~mp;
          *zpw = (za & ~mp) | (r == q ? mp : 0ull);
          bytes += 16;
        }
      } else {
        xr = *b8(T.X, r, p, rs, T.il);
        zr = *b8(T.Z, r, p, rs, T.il);
        const uint32_t xo = *b8(T.X, r, np_, rs, T.il), zo = *b8(T.Z, r, np_, rs, T.il);
        const uint32_t xn = (e * xr) % D, zn = (e * zr) % D;
        if (xn != xo) { *b8(T.X, r, np_, rs, T.il) = (uint8_t)xn; bytes += 1; }
        if (zn != zo) { *b8(T.Z, r, np_, rs, T.il) = (uint8_t)zn; bytes += 1; }
        bytes += 4;
        if (xr | zr) {
          *b8(T.X, r, p, rs, T.il) = 0;
          *b8(T.Z, r, p, rs, T.il) = (uint8_t)(r == q);
          bytes += 2;
        }
      }
      S.xp[r] = (uint8_t)xr;
      S.zp[r] = (uint8_t)zr;
      if (xr | zr) {
        const uint32_t k = (uint32_t)atomicAdd(&M[1], 1);
        if (k < S.tr_cap) S.TR[k] = (uint16_t)r;
        else report_error(S, KERR_TOUCHED_ROWS);
        if constexpr (!BIN) atomicAdd(&M[3], (int)(xr * zr));
      }
    }
    if (tid == 0) {
      if constexpr (BIN) {
        ((uint64_t*)S.F)[p >> 6] &= ~(1ull << (p & 63));
        ((uint64_t*)S.F)[(n + p) >> 6] &= ~(1ull << ((n + p) & 63));
      } else {
        S.F[p] = 0;
        S.F[n + p] = 0;
      }
    }
    for (uint32_t s = tid; s < Sh; s += NT) {
      if constexpr (BIN) {
        const uint32_t v0 = (uint32_t)((plane(T, s, 0)[p >> 6] >> (p & 63)) & 1);
        const uint32_t v1 = (uint32_t)((plane(T, s, 1)[p >> 6] >> (p & 63)) & 1);
        S.shot[s] = v0 | (v1 << 1);
      } else {
        S.shot[s] = phase_vec(T, s)[p];
      }
    }
    __syncthreads();                                                      // B2
    SDIM_TADD(P, 1, tB);
    SDIM_TMARK(tC);
    const uint32_t nT = min((uint32_t)M[1], S.tr_cap), nF = min((uint32_t)M[2], S.fl_cap);
    uint32_t xz = 0, pinc = 0;
    if constexpr (!BIN) {
      const uint32_t xz_old = (uint32_t)M[3] % D;
      xz = ((e * e) % D) * xz_old % D;
      pinc = xz_old * ((e * (e - 1) / 2) % D) % D;
    }

    // ---- C: sweep (touched rows) x (groups with f != 0)
    sweep_groups<D, NT, RB>(T, S, nT, nF, e, bytes);
    __syncthreads();                                                      // B3
    SDIM_TADD(P, 2, tC);
    SDIM_TMARK(tD);

    // ---- D: per-shot phases over words with f != 0
    if constexpr (!BIN) {
      rowsum_shared_terms<D, NT>(S, nF, xz);
      __syncthreads();
    }
    rowsum_phases<D, NT>(T, S, nF, e, xz, pinc, bytes);
    __syncthreads();                                                      // B4
    SDIM_TADD(P, 3, tD);
    SDIM_TMARK(tE);
    // ---- E: outcome per shot (drawn in step A); ph_{n+p} <- ph_p, ph_p <- -m po
    for (uint32_t s = tid; s < Sh; s += NT) {
      uint32_t m;
      if (early) m = S.shot[Sh + s];
      else { m = outcome_of(draw(it.seed, it.stream, src, it.shot0 + s), D); S.shot[Sh + s] = m; }
      if constexpr (BIN) {
        const uint32_t k2 = S.shot[s];
        const uint32_t g = n + p;
        uint64_t* p0 = plane(T, s, 0);
        uint64_t* p1 = plane(T, s, 1);
        p0[g >> 6] = (p0[g >> 6] & ~(1ull << (g & 63))) | ((uint64_t)(k2 & 1) << (g & 63));
        p1[g >> 6] = (p1[g >> 6] & ~(1ull << (g & 63))) | ((uint64_t)(k2 >> 1) << (g & 63));
        p0[p >> 6] &= ~(1ull << (p & 63));
        p1[p >> 6] = (p1[p >> 6] & ~(1ull << (p & 63))) | ((uint64_t)(m & 1) << (p & 63));
      } else {
        uint8_t* ph = phase_vec(T, s);
        ph[n + p] = (uint8_t)((e * S.shot[s] + pinc) % D);
        ph[p] = (uint8_t)((D - m) % D);
      }
      rec[s] = (uint8_t)m;
    }
    SDIM_TADD(P, 4, tE);
  } else {
    SDIM_TMARK(tDet);
    // ---- deterministic outcome: prod_i S_i^{f_i}, f_i = dx[q, i] = X[q][n + i]
    // F holds f aligned to the stabilizer columns (word w = columns w*CPW ...)
    for (uint32_t w = tid; w < nsw; w += NT) {
      if constexpr (BIN) {
        const uint64_t* rq = (const uint64_t*)S.RQ;
        const uint32_t g = n + 64 * w;
        const uint64_t lo = rq[g >> 6] >> (g & 63);
        const uint64_t hi = ((g & 63) && (g >> 6) + 1 < rs / 8) ? rq[(g >> 6) + 1] << (64 - (g & 63)) : 0ull;
        uint64_t f = lo | hi;
        const uint32_t lim = n - 64 * w;
        if (lim < 64) f &= (1ull << lim) - 1;
        ((uint64_t*)S.F)[w] = f;
      } else {
        uint32_t f = 0;
        #pragma unroll
        for (int b = 0; b < 4; b++) {
          const uint32_t c = 4 * w + b;
          if (c < n) f |= (uint32_t)S.RQ[n + c] << (8 * b);
        }
        ((uint32_t*)S.F)[w] = f;
      }
    }
    // ordered list of the words with f != 0 in FL (f has few nonzero words: the rows and
    // the shots below visit only those); each thread tests the words it wrote itself
    uint32_t nfw = 0;
    {
      const uint32_t lane = tid & 63, wid = tid >> 6;
      for (uint32_t w0 = 0; w0 < nsw; w0 += NT) {
        const uint32_t w = w0 + tid;
        bool nz = false;
        if (w < nsw) nz = BIN ? ((const uint64_t*)S.F)[w] != 0ull : ((const uint32_t*)S.F)[w] != 0u;
        const uint64_t bal = __ballot(nz);
        if (lane == 0) S.red[wid] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t before = 0, tot = 0;
        #pragma unroll
        for (uint32_t i = 0; i < NT / 64; i++) {
          const uint32_t c = S.red[i];
          before += i < wid ? c : 0u;
          tot += c;
        }
        if (nz) S.FL[nfw + before + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = (uint16_t)w;
        nfw += tot;
        __syncthreads();
      }
    }
    // cross term sum_r sum_i (f x)_ri * prefix_{j<i} (f z)_rj and the exponent term
    // sum_i f_i (f_i - 1) / 2 <x_i, z_i>, one thread per row, over the listed words in
    // order (the loads of DB listed words are issued before their use)
    constexpr uint32_t DB = SDIM_DET_BATCH;
    uint32_t part = 0;
    for (uint32_t r = tid; r < n; r += NT) {
      if constexpr (BIN) {
        uint64_t carry = 0, par = 0;
        for (uint32_t k0 = 0; k0 < nfw; k0 += DB) {
          uint64_t xa[DB], za[DB];
          #pragma unroll
          for (uint32_t u = 0; u < DB; u++) {
            xa[u] = 0ull; za[u] = 0ull;
            if (k0 + u < nfw) {
              const uint32_t w = S.FL[k0 + u];
              xa[u] = *w64(T.X, r, w, rs, T.il);
              za[u] = *w64(T.Z, r, w, rs, T.il);
            }
          }
          #pragma unroll
          for (uint32_t u = 0; u < DB; u++) {
            if (k0 + u >= nfw) break;
            const uint64_t f = ((const uint64_t*)S.F)[S.FL[k0 + u]];
            const uint64_t a = xa[u] & f, b = za[u] & f;
            uint64_t y = b;
            y ^= y << 1; y ^= y << 2; y ^= y << 4; y ^= y << 8; y ^= y << 16; y ^= y << 32;
            const uint64_t pre = (y << 1) ^ (carry ? ~0ull : 0ull);
            par ^= (uint64_t)(__builtin_popcountll(a & pre) & 1);
            carry ^= (uint64_t)(__builtin_popcountll(b) & 1);
            bytes += 16;
          }
        }
        part += (uint32_t)par;
      } else {
        uint32_t U = 0, acc = 0;
        for (uint32_t k0 = 0; k0 < nfw; k0 += DB) {
          uint32_t xa[DB], za[DB];
          #pragma unroll
          for (uint32_t u = 0; u < DB; u++) {
            xa[u] = 0u; za[u] = 0u;
            if (k0 + u < nfw) {
              const uint32_t w = S.FL[k0 + u];
              xa[u] = *w32(T.X, r, w, rs, T.il);
              za[u] = *w32(T.Z, r, w, rs, T.il);
            }
          }
          #pragma unroll
          for (uint32_t u = 0; u < DB; u++) {
            if (k0 + u >= nfw) break;
            const uint32_t f = ((const uint32_t*)S.F)[S.FL[k0 + u]];
            #pragma unroll
            for (int b = 0; b < 4; b++) {
              const uint32_t fb = byte_of(f, b);
              const uint32_t xb = byte_of(xa[u], b), zb = byte_of(za[u], b);
              acc += fb * xb * U + ((fb * (fb - 1) / 2) % D) * xb * zb;
              U = (U + fb * zb) % D;
            }
            bytes += 8;
          }
        }
        part += acc % D;
      }
    }
    const uint32_t K = BIN ? (block_sum<NT>(part, S.red) & 1) : (block_sum<NT>(part % D, S.red) % D);
    for (uint32_t s = tid; s < Sh; s += NT) {
      uint32_t m;
      if constexpr (BIN) {
        uint32_t c0 = 0, c1 = 0;
        for (uint32_t k = 0; k < nfw; k++) {
          const uint32_t w = S.FL[k];
          const uint64_t f = ((const uint64_t*)S.F)[w];
          c0 += __builtin_popcountll(plane(T, s, 0)[w] & f);
          c1 += __builtin_popcountll(plane(T, s, 1)[w] & f);
          bytes += 16;
        }
        const uint32_t ap = (c0 + 2 * c1 + 2 * K) & 3;
        m = ((ap + 1) >> 1) & 1;                       // (-ap // 2) mod 2
      } else {
        const uint32_t* ph = (const uint32_t*)phase_vec(T, s);
        uint32_t acc = 0;
        for (uint32_t k = 0; k < nfw; k++) {
          const uint32_t w = S.FL[k];
          const uint32_t f = ((const uint32_t*)S.F)[w];
          const uint32_t v = ph[w];
          #pragma unroll
          for (int b = 0; b < 4; b++) acc += byte_of(f, b) * byte_of(v, b);
          bytes += 4;
        }
        m = (D - (acc + K) % D) % D;
      }
      S.shot[Sh + s] = m;
      rec[s] = (uint8_t)(m | 0x80u);
    }
    SDIM_TADD(P, 5, tDet);
  }

  if (it.flags & ITEM_SNAP_MEAS) snapshot<D, NT>(T, it, P, slot);
  if (reset) {
    SDIM_TMARK(tR);
    __syncthreads();
    // X^k with k = (-m) mod d: phase -= k po z[q, .]   (program.py:334-339)
    for (uint32_t i = tid; i < Sh * NWd; i += NT) {
      const uint32_t s = i / NWd, w = i - s * NWd;
      const uint32_t m = S.shot[Sh + s];
      if constexpr (BIN) {
        if (m & 1) {
          const uint64_t z = *w64(T.Z, q, w, rs, T.il);
          if (z) { plane(T, s, 1)[w] ^= z; bytes += 24; }
        }
      } else {
        const uint32_t kk = (D - m) % D;
        if (kk) {
          const uint32_t z = *w32(T.Z, q, w, rs, T.il);
          if (z) {
            uint32_t* ph = (uint32_t*)phase_vec(T, s) + w;
            *ph = subm<D>(*ph, mult<D>(S.mtab[kk], z));
            bytes += 12;
          }
        }
      }
    }
    SDIM_TADD(P, 6, tR);
  }
  // the per-shot loops above read F, shot[] and the counters; the next op rewrites them
  __syncthreads();
}

// ------------------------------------------------------------------ one item
// RES: the whole tableau in LDS (a compile-time placement lets its accesses compile to LDS
// instructions instead of generic ones); otherwise X/Z in HBM and the phases in HBM or LDS.
template <int D, int NT, bool RES, int RB = 1, int OPB = 2>
SDIM_DEV void run_item_at(const ItemDesc& it, const KParams& P, uint8_t* lds) {
  constexpr bool res = RES;
  const bool phl = !res && (it.flags & ITEM_PH_LDS) != 0;     // phases kept in LDS
  const uint32_t n = it.n, rs = it.rs, Sh = it.S;
  const uint64_t T_bytes = tableau_bytes(n, D, Sh);
  const uint64_t XZ_bytes = 2ull * rows_padded(n) * rs, PH_bytes = T_bytes - XZ_bytes;
  uint8_t* gtab = P.arena + it.tab_off;
  uint8_t* tab = res ? lds : gtab;
  // LDS: [whole tableau (res)] counters, tables, segment delta, reductions, [phases (phl)],
  // measurement-only scratch, free
  uint8_t* sbase = res ? lds + ((T_bytes + 15) & ~15ull) : lds;
  Scr S = carve<NT>(sbase, n, rs, Sh, phl ? (uint32_t)PH_bytes : 0u);
  S.err = P.counters + SDIM_ERR_SLOT;
  S.item = it.id;
  if (P.list_cap) { S.tr_cap = min(S.tr_cap, P.list_cap); S.fl_cap = min(S.fl_cap, P.list_cap); }
  S.lds_end = lds + P.lds_bytes;
  S.tiling = !res;
  Tab T;
  T.X = tab; T.Z = tab + (size_t)rows_padded(n) * rs; T.Ph = phl ? S.gap : tab + XZ_bytes; T.n = n; T.rs = rs; T.S = Sh;
  T.wofs = 0; T.rs_full = rs; T.nw = rs / (D == 2 ? 8 : 4); T.il = it.il;
  unsigned long long bytes = 0, lbytes = 0;
  SDIM_TMARK(tItem);

  if (threadIdx.x < 16) S.mtab[threadIdx.x] = make_tab<D>(threadIdx.x < D ? threadIdx.x : 0);
  if (threadIdx.x < 32) S.misc[threadIdx.x] = ((threadIdx.x & 15) == 0) ? INT_MAX : 0;
  if (it.flags & ITEM_INIT_GIVEN) {
    if (res) copy16<NT>(lds, gtab, T_bytes);
    if (phl) copy16<NT>(S.gap, gtab + XZ_bytes, PH_bytes);
    __syncthreads();
  } else {
    init_identity<D, NT>(T, XZ_bytes, PH_bytes);
  }
  if (threadIdx.x == 0) bytes += T_bytes;
  SDIM_TADD(P, 9, tItem);

  const DevOp* ops = P.ops + it.ops_off;
  uint32_t k = 0, mk = 0;
  while (k < it.n_ops) {
    const DevOp o = ops[k];
    const uint32_t kind = op_kind(o);
    if (kind == OP_SEG) {
      k = run_segment<D, NT, OPB>(ops, k, T, S, it, P, bytes, lbytes);
    } else if (kind == OP_MEAS || kind == OP_RESET) {
      measure<D, NT, RB>(T, o.q, o.t, o.src, kind == OP_RESET, mk++, S, it, P, bytes);
      k++;
    } else if (kind == OP_SNAP) {
      snapshot<D, NT>(T, it, P, o.t);
      k++;
    } else {
      break;   // OP_END
    }
  }
  __syncthreads();
  SDIM_TMARK(tOut);
  if (res) {
    copy16<NT>(gtab, lds, T_bytes);
    if (threadIdx.x == 0) bytes += T_bytes;
  }
  if (phl) {
    copy16<NT>(gtab + XZ_bytes, S.gap, PH_bytes);
    if (threadIdx.x == 0) bytes += PH_bytes;
  }
  SDIM_TADD(P, 10, tOut);
  SDIM_TADD(P, 11, tItem);
  const unsigned long long tot = block_sum64<NT>(bytes, S.red);
  const unsigned long long ltot = block_sum64<NT>(lbytes, S.red);
  if (threadIdx.x == 0) {
    atomicAdd(&P.counters[res ? 1 : 0], tot);
    if (ltot) atomicAdd(&P.counters[1], ltot);
  }
}

}  // namespace sdim
