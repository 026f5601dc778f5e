// Device engine: gate segments, measurement, the persistent batch kernel and
// the plain one-launch-per-op kernels.
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
// gates (already layered: the gates of a layer touch disjoint qudits) and
// measurement ops:
//   * a gate segment is processed layer by layer; thread (word w, slot k)
//     applies the k-th, (k+NS)-th, ... gate of the layer to column word w.
//     The phase increments of gates do not depend on the shot, so each thread
//     accumulates them in registers (delta) and the per-shot phase vectors are
//     touched once at the end of the segment.  Noise ops snapshot their row;
//     their shot-dependent phase terms are added in the same end-of-segment pass.
//   * a measurement is a cooperative workgroup operation: ballot/atomic-min
//     pivot search over the stabilizer half of row q, gather of the pivot
//     column, a single sweep over all rows that applies the rank-1 update
//     X += x_p f^T, Z += z_p f^T and accumulates the column dots <z_i, x_p>
//     from the old Z, then per-shot phase updates; or, without a pivot, the
//     deterministic ancilla as a per-row exclusive scan.
// The tableau lives in LDS for the whole item when it fits (ITEM_LDS) and in
// HBM otherwise; the same code runs on both (address spaces are inferred
// after inlining).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>

#include "sdim_types.h"
#include "sdim_arith.h"
#include "engine.h"

namespace sdim {

// ------------------------------------------------------------------ scratch carve
struct Scr {
  int* misc;         // 64 ints
  MulTab* mtab;      // 16 tables
  uint8_t* xp;       // pivot column X, n bytes
  uint8_t* zp;       // pivot column Z, n bytes
  uint8_t* F;        // rs bytes: rowsum factors f (odd: bytes; d=2: bits)
  uint8_t* C;        // rs bytes: shot-invariant phase term per column
  uint8_t* DW;       // 2*rs bytes: segment phase delta (d = 2 uses two planes)
  uint32_t* red;     // NT * 4 words
  uint32_t* shot;    // 2 * S words
  uint8_t* noise;    // noise tables of a segment
};

SDIM_DEV uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }

template <int NT>
SDIM_DEV Scr carve(uint8_t* base, uint32_t n, uint32_t rs, uint32_t S) {
  Scr s;
  uint8_t* p = base;
  s.misc = (int*)p; p += 256;
  s.mtab = (MulTab*)p; p += 16 * 16;
  uint32_t nb = (n + 15u) & ~15u;
  s.xp = p; p += nb;
  s.zp = p; p += nb;
  s.F = p; p += rs;
  s.C = p; p += rs;
  s.DW = p; p += 2 * rs;
  s.red = (uint32_t*)p; p += NT * 16;
  s.shot = (uint32_t*)p; p += ((2 * S * 4) + 15u) & ~15u;
  s.noise = p;
  return s;
}

template <int D> struct Cfg {
  static constexpr int PO = (D == 2) ? 2 : 1;
  static constexpr int ORD = (D == 2) ? 4 : D;
};

// tableau pointers of an item
struct Tab {
  uint8_t* X;
  uint8_t* Z;
  uint8_t* Ph;
  uint32_t n, rs, S;
};

SDIM_DEV uint32_t* row32(uint8_t* base, uint32_t r, uint32_t rs) { return (uint32_t*)(base + (size_t)r * rs); }
SDIM_DEV uint64_t* row64(uint8_t* base, uint32_t r, uint32_t rs) { return (uint64_t*)(base + (size_t)r * rs); }

// per-shot phase vector (odd d) / bit-plane (d = 2)
SDIM_DEV uint8_t* phase_vec(const Tab& T, uint32_t s) { return T.Ph + (size_t)s * T.rs; }
SDIM_DEV uint64_t* plane(const Tab& T, uint32_t s, int k) { return (uint64_t*)(T.Ph + (size_t)(2 * s + k) * T.rs); }

// ------------------------------------------------------------------ block reductions
template <int NT>
SDIM_DEV uint32_t block_sum(uint32_t v, uint32_t* red) {
  // wave reduction then across waves
  #pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
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
SDIM_DEV void init_identity(const Tab& T, uint64_t total) {
  // zero everything, then the diagonals: stabilizer i = Z_i, destabilizer i = X_i
  uint4* p = (uint4*)T.X;
  const uint64_t nv = total / 16;
  for (uint64_t i = threadIdx.x; i < nv; i += NT) p[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  const uint32_t n = T.n;
  for (uint32_t r = threadIdx.x; r < n; r += NT) {
    if constexpr (D == 2) {
      uint32_t g = n + r;
      row64(T.X, r, T.rs)[g >> 6] |= 1ull << (g & 63);
      row64(T.Z, r, T.rs)[r >> 6] |= 1ull << (r & 63);
    } else {
      T.X[(size_t)r * T.rs + n + r] = 1;
      T.Z[(size_t)r * T.rs + r] = 1;
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

// ------------------------------------------------------------------ gate on one column word
// odd d: one 32-bit word = 4 generators.  Returns bytes touched.
template <int D>
SDIM_DEV uint32_t gate_odd(uint32_t kind, uint32_t q, uint32_t t, uint32_t a, uint32_t b,
                           const Tab& T, uint32_t w, uint32_t& dl, const Scr& S, uint32_t nslot) {
  const uint32_t rs = T.rs;
  uint32_t* xq = row32(T.X, q, rs) + w;
  uint32_t* zq = row32(T.Z, q, rs) + w;
  switch (kind) {
    case OP_H: {
      uint32_t x = *xq, z = *zq;
      *xq = negm<D>(z); *zq = x;
      dl = subm<D>(dl, mulv<D>(x, z));   // phase += po * (-z) * x
      return 16;
    }
    case OP_HI: {
      uint32_t x = *xq, z = *zq;
      *xq = z; *zq = negm<D>(x);
      dl = subm<D>(dl, mulv<D>(x, z));   // phase += po * z * (-x)
      return 16;
    }
    case OP_P: {
      const MulTab h = make_half_tab<D>();
      uint32_t x = *xq, z = *zq;
      *zq = addm<D>(z, x);
      dl = addm<D>(dl, mult<D>(h, x));   // phase += x(x-1)/2
      return 12;
    }
    case OP_PI: {
      const MulTab h = make_half_tab<D>();
      uint32_t x = *xq, z = *zq;
      *zq = subm<D>(z, x);
      dl = subm<D>(dl, mult<D>(h, x));
      return 12;
    }
    case OP_SUM: {                       // x[t] += x[c]; z[c] += (d-1) z[t]
      uint32_t* xt = row32(T.X, t, rs) + w;
      uint32_t* zt = row32(T.Z, t, rs) + w;
      *xt = addm<D>(*xt, *xq);
      *zq = subm<D>(*zq, *zt);
      return 24;
    }
    case OP_SUMI: {
      uint32_t* xt = row32(T.X, t, rs) + w;
      uint32_t* zt = row32(T.Z, t, rs) + w;
      *xt = subm<D>(*xt, *xq);
      *zq = addm<D>(*zq, *zt);
      return 24;
    }
    case OP_PAULI: {                     // X^a Z^b: phase += b x - a z
      uint32_t x = *xq, z = *zq;
      dl = addm<D>(dl, subm<D>(mult<D>(S.mtab[b], x), mult<D>(S.mtab[a], z)));
      return 8;
    }
    case OP_NOISE: {                     // snapshot of row q for the end-of-segment pass
      uint8_t* snap = S.noise + nslot * 2 * rs;
      ((uint32_t*)snap)[w] = *xq;
      ((uint32_t*)(snap + rs))[w] = *zq;
      return 8;
    }
  }
  return 0;
}

// d = 2: one 64-bit word = 64 generators; delta is the plane pair (d0, d1).
SDIM_DEV uint32_t gate_bin(uint32_t kind, uint32_t q, uint32_t t, uint32_t a, uint32_t b,
                           const Tab& T, uint32_t w, uint64_t& d0, uint64_t& d1, const Scr& S, uint32_t nslot) {
  const uint32_t rs = T.rs;
  uint64_t* xq = row64(T.X, q, rs) + w;
  uint64_t* zq = row64(T.Z, q, rs) + w;
  switch (kind) {
    case OP_H:
    case OP_HI: {                        // (x, z) <- (z, x); phase += 2 x z
      uint64_t x = *xq, z = *zq;
      *xq = z; *zq = x;
      d1 ^= x & z;
      return 32;
    }
    case OP_P: {                         // phase += x; z ^= x
      uint64_t x = *xq, z = *zq;
      *zq = z ^ x;
      d1 ^= d0 & x; d0 ^= x;
      return 24;
    }
    case OP_PI: {                        // phase += 3 x; z ^= x
      uint64_t x = *xq, z = *zq;
      *zq = z ^ x;
      d1 ^= x & ~d0; d0 ^= x;
      return 24;
    }
    case OP_SUM:
    case OP_SUMI: {
      uint64_t* xt = row64(T.X, t, rs) + w;
      uint64_t* zt = row64(T.Z, t, rs) + w;
      *xt ^= *xq;
      *zq ^= *zt;
      return 48;
    }
    case OP_PAULI: {                     // phase += 2 (b x - a z)
      uint64_t x = *xq, z = *zq;
      d1 ^= ((b & 1) ? x : 0ull) ^ ((a & 1) ? z : 0ull);
      return 16;
    }
    case OP_NOISE: {
      uint8_t* snap = S.noise + nslot * 2 * rs;
      ((uint64_t*)snap)[w] = *xq;
      ((uint64_t*)(snap + rs))[w] = *zq;
      return 16;
    }
  }
  return 0;
}

// ------------------------------------------------------------------ gate segment
// ops[k] is the OP_SEG header.  Returns the index after the segment.
template <int D, int NT>
SDIM_DEV uint32_t run_segment(const DevOp* ops, uint32_t k, const Tab& T, const Scr& S,
                              const ItemDesc& it, const KParams& P, unsigned long long& bytes) {
  const DevOp hdr = ops[k];
  const uint32_t nl = hdr.q, nn = hdr.t;
  const uint32_t n = T.n, rs = T.rs, Sh = T.S;
  const uint32_t tid = threadIdx.x;
  const uint32_t G = 2 * n;
  const uint32_t NWd = (D == 2) ? (G + 63) / 64 : (G + 3) / 4;
  const uint32_t ndesc = k + 1;
  k += 1 + nn;

  // noise draws of the segment: mask[j][S/32] bits + ab[j][S] bytes
  const uint32_t mw = (Sh + 31) / 32;
  uint32_t* nmask = (uint32_t*)(S.noise + nn * 2 * rs);
  uint8_t* nab = (uint8_t*)(nmask + nn * mw);
  if (nn) {
    for (uint32_t i = tid; i < nn * mw; i += NT) nmask[i] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < nn * Sh; i += NT) {
      const uint32_t j = i / Sh, s = i - j * Sh;
      const DevOp nd = ops[ndesc + j];
      U64x4 o = draw(it.seed, it.stream, nd.src, it.shot0 + s);
      uint32_t ab = noise_ab(o, D, op_a(nd), P.keep53[it.keep_off + nd.t]);
      nab[i] = (uint8_t)(ab & 0xFF);
      if (ab) atomicOr(&nmask[j * mw + (s >> 5)], 1u << (s & 31));
    }
    __syncthreads();
  }

  // word ownership
  uint32_t NS, my_w, my_slot;
  if (NWd <= NT) { NS = NT / NWd; my_w = tid % NWd; my_slot = tid / NWd; }
  else { NS = 1; my_w = tid; my_slot = 0; }
  const bool active = my_slot < NS;
  constexpr int KW = 4;  // words per thread supported when NWd > NT
  uint32_t dl[KW];
  uint64_t b0[KW], b1[KW];
  #pragma unroll
  for (int i = 0; i < KW; i++) { dl[i] = 0; b0[i] = 0; b1[i] = 0; }

  for (uint32_t l = 0; l < nl; l++) {
    const uint32_t L = ops[k].q;
    const DevOp* lops = ops + k + 1;
    if (active) {
      #pragma unroll
      for (int kw = 0; kw < KW; kw++) {
        const uint32_t w = my_w + (uint32_t)kw * NT;
        if (w < NWd && (kw == 0 || NWd > NT)) {
          for (uint32_t g = my_slot; g < L; g += NS) {
            const DevOp o = lops[g];
            if constexpr (D == 2)
              bytes += gate_bin(op_kind(o), o.q, o.t, op_a(o), op_b(o), T, w, b0[kw], b1[kw], S, op_noise_slot(o));
            else
              bytes += gate_odd<D>(op_kind(o), o.q, o.t, op_a(o), op_b(o), T, w, dl[kw], S, op_noise_slot(o));
          }
        }
      }
    }
    k += 1 + L;
    __syncthreads();
  }

  // reduce the deltas over slots into DW
  if (NWd <= NT) {
    if constexpr (D == 2) {
      uint64_t* r = (uint64_t*)S.red;   // NT * 16 bytes: two words per thread
      if (active) { r[2 * tid] = b0[0]; r[2 * tid + 1] = b1[0]; }
      __syncthreads();
      for (uint32_t w = tid; w < NWd; w += NT) {
        uint64_t a0 = 0, a1 = 0;
        for (uint32_t s = 0; s < NS; s++) {
          const uint32_t src = s * NWd + w;
          uint64_t c0 = r[2 * src], c1 = r[2 * src + 1];
          a1 ^= c1 ^ (a0 & c0); a0 ^= c0;
        }
        ((uint64_t*)S.DW)[w] = a0;
        ((uint64_t*)(S.DW + rs))[w] = a1;
      }
    } else {
      if (active) S.red[tid] = dl[0];
      __syncthreads();
      for (uint32_t w = tid; w < NWd; w += NT) {
        uint32_t a = 0;
        for (uint32_t s = 0; s < NS; s++) a = addm<D>(a, S.red[s * NWd + w]);
        ((uint32_t*)S.DW)[w] = a;
      }
    }
  } else {
    #pragma unroll
    for (int kw = 0; kw < KW; kw++) {
      const uint32_t w = my_w + (uint32_t)kw * NT;
      if (w < NWd) {
        if constexpr (D == 2) { ((uint64_t*)S.DW)[w] = b0[kw]; ((uint64_t*)(S.DW + rs))[w] = b1[kw]; }
        else ((uint32_t*)S.DW)[w] = dl[kw];
      }
    }
  }
  __syncthreads();

  // per-shot phase pass: Ph_s += delta + noise terms of shot s
  for (uint32_t i = tid; i < Sh * NWd; i += NT) {
    const uint32_t s = i / NWd, w = i - s * NWd;
    if constexpr (D == 2) {
      uint64_t* p0 = plane(T, s, 0) + w;
      uint64_t* p1 = plane(T, s, 1) + w;
      uint64_t c0 = ((uint64_t*)S.DW)[w], c1 = ((uint64_t*)(S.DW + rs))[w];
      for (uint32_t j = 0; j < nn; j++) {
        if (nmask[j * mw + (s >> 5)] & (1u << (s & 31))) {
          const uint32_t ab = nab[j * Sh + s];
          const uint8_t* snap = S.noise + j * 2 * rs;
          uint64_t sx = ((const uint64_t*)snap)[w], sz = ((const uint64_t*)(snap + rs))[w];
          c1 ^= (((ab >> 4) & 1) ? sx : 0ull) ^ ((ab & 1) ? sz : 0ull);
        }
      }
      uint64_t a0 = *p0, a1 = *p1;
      *p1 = a1 ^ c1 ^ (a0 & c0);
      *p0 = a0 ^ c0;
      bytes += 32;
    } else {
      uint32_t* ph = (uint32_t*)phase_vec(T, s) + w;
      uint32_t c = ((uint32_t*)S.DW)[w];
      for (uint32_t j = 0; j < nn; j++) {
        if (nmask[j * mw + (s >> 5)] & (1u << (s & 31))) {
          const uint32_t ab = nab[j * Sh + s];
          const uint8_t* snap = S.noise + j * 2 * rs;
          uint32_t sx = ((const uint32_t*)snap)[w], sz = ((const uint32_t*)(snap + rs))[w];
          c = addm<D>(c, subm<D>(mult<D>(S.mtab[(ab >> 4) & 15], sx), mult<D>(S.mtab[ab & 15], sz)));
        }
      }
      *ph = addm<D>(*ph, c);
      bytes += 8;
    }
  }
  __syncthreads();
  return k;
}

// ------------------------------------------------------------------ measurement
// Z-basis measurement of qudit q (tableau_prime.py:262-363) followed, for
// RESET, by X^((-m) mod d) (program.py:334-339).  Writes one record per shot.
template <int D, int NT>
SDIM_DEV void measure(const Tab& T, uint32_t q, uint32_t slot, uint32_t src, bool reset,
                      const Scr& S, const ItemDesc& it, const KParams& P, unsigned long long& bytes) {
  constexpr int PO = Cfg<D>::PO;
  const uint32_t tid = threadIdx.x;
  const uint32_t n = T.n, rs = T.rs, Sh = T.S;
  const uint32_t G = 2 * n;
  uint8_t* rec = P.arena + it.rec_off + (size_t)slot * Sh;

  // ---- pivot: first stabilizer column with x[q, i] != 0
  if (tid == 0) S.misc[0] = INT_MAX;
  __syncthreads();
  if constexpr (D == 2) {
    const uint64_t* xr = row64(T.X, q, rs);
    const uint32_t nw = (n + 63) / 64;
    for (uint32_t w = tid; w < nw; w += NT) {
      uint64_t v = xr[w];
      const uint32_t lim = n - 64 * w;
      if (lim < 64) v &= (1ull << lim) - 1;
      if (v) atomicMin(&S.misc[0], (int)(64 * w + __builtin_ctzll(v)));
      bytes += 8;
    }
  } else {
    const uint32_t* xr = row32(T.X, q, rs);
    const uint32_t nw = (n + 3) / 4;
    for (uint32_t w = tid; w < nw; w += NT) {
      uint32_t v = xr[w];
      const uint32_t lim = n - 4 * w;
      if (lim < 4) v &= (1u << (8 * lim)) - 1;
      if (v) atomicMin(&S.misc[0], (int)(4 * w + (__builtin_ctz(v) >> 3)));
      bytes += 4;
    }
  }
  __syncthreads();
  const int piv = S.misc[0];

  if (piv != INT_MAX) {
    const uint32_t p = (uint32_t)piv;
    // ---- gather the pivot column
    for (uint32_t r = tid; r < n; r += NT) {
      if constexpr (D == 2) {
        S.xp[r] = (uint8_t)((row64(T.X, r, rs)[p >> 6] >> (p & 63)) & 1);
        S.zp[r] = (uint8_t)((row64(T.Z, r, rs)[p >> 6] >> (p & 63)) & 1);
      } else {
        S.xp[r] = T.X[(size_t)r * rs + p];
        S.zp[r] = T.Z[(size_t)r * rs + p];
      }
      bytes += (D == 2 ? 16 : 2);
    }
    __syncthreads();
    uint32_t xz = 0;
    if constexpr (D != 2) {
      // ---- exponentiate stabilizer p by e = x[q,p]^-1 when x[q,p] != 1 (tableau_prime.py:278-282, 365-380)
      const uint32_t xv = S.xp[q];
      uint32_t e = 1;
      for (uint32_t c = 1; c < D; c++) if ((xv * c) % D == 1) e = c;
      if (e != 1) {
        uint32_t part = 0;
        for (uint32_t r = tid; r < n; r += NT) part += (uint32_t)S.xp[r] * S.zp[r];
        const uint32_t xz_old = block_sum<NT>(part, S.red) % D;
        const uint32_t inc = (xz_old * ((e * (e - 1) / 2) % D)) % D;
        for (uint32_t r = tid; r < n; r += NT) {
          S.xp[r] = (uint8_t)((S.xp[r] * e) % D);
          S.zp[r] = (uint8_t)((S.zp[r] * e) % D);
        }
        for (uint32_t s = tid; s < Sh; s += NT) {
          uint8_t* ph = phase_vec(T, s) + p;
          *ph = (uint8_t)((*ph * e + inc) % D);
        }
        __syncthreads();
      }
      uint32_t part = 0;
      for (uint32_t r = tid; r < n; r += NT) part += (uint32_t)S.xp[r] * S.zp[r];
      xz = block_sum<NT>(part, S.red) % D;
    }

    // ---- rowsum factors f = -x[q, .] (f_p = f_{n+p} = 0) and per-shot pivot phases
    const uint32_t NWd = (D == 2) ? (G + 63) / 64 : (G + 3) / 4;
    for (uint32_t w = tid; w < NWd; w += NT) {
      if constexpr (D == 2) {
        uint64_t f = row64(T.X, q, rs)[w];
        if ((p >> 6) == w) f &= ~(1ull << (p & 63));
        if (((n + p) >> 6) == w) f &= ~(1ull << ((n + p) & 63));
        ((uint64_t*)S.F)[w] = f;
      } else {
        uint32_t f = negm<D>(row32(T.X, q, rs)[w]);
        if ((p >> 2) == w) f &= ~(0xFFu << (8 * (p & 3)));
        if (((n + p) >> 2) == w) f &= ~(0xFFu << (8 * ((n + p) & 3)));
        ((uint32_t*)S.F)[w] = f;
      }
      bytes += (D == 2 ? 8 : 4);
    }
    for (uint32_t s = tid; s < Sh; s += NT) {
      if constexpr (D == 2) {
        const uint32_t b0 = (plane(T, s, 0)[p >> 6] >> (p & 63)) & 1;
        const uint32_t b1 = (plane(T, s, 1)[p >> 6] >> (p & 63)) & 1;
        S.shot[s] = b0 | (b1 << 1);
      } else {
        S.shot[s] = phase_vec(T, s)[p];
      }
    }
    __syncthreads();

    // ---- sweep: rank-1 update of X and Z, column dots <z_i, x_p> from the old Z
    uint32_t NRG, my_w, my_rg;
    if (NWd <= NT) { NRG = NT / NWd; my_w = tid % NWd; my_rg = tid / NWd; }
    else { NRG = 1; my_w = tid; my_rg = 0; }
    const bool act = my_rg < NRG;
    constexpr int KW = 4;
    if constexpr (D == 2) {
      uint64_t acc[KW];
      #pragma unroll
      for (int i = 0; i < KW; i++) acc[i] = 0;
      if (act) {
        #pragma unroll
        for (int kw = 0; kw < KW; kw++) {
          const uint32_t w = my_w + (uint32_t)kw * NT;
          if (w < NWd && (kw == 0 || NWd > NT)) {
            const uint64_t f = ((const uint64_t*)S.F)[w];
            uint64_t a = 0;
            for (uint32_t r = my_rg; r < n; r += NRG) {
              const uint32_t sx = S.xp[r], sz = S.zp[r];
              if (sx | sz) {
                uint64_t* xr = row64(T.X, r, rs) + w;
                uint64_t* zr = row64(T.Z, r, rs) + w;
                const uint64_t z = *zr;
                if (sx) { a ^= z; *xr ^= f; bytes += 16; }
                if (sz) { *zr = z ^ f; bytes += 8; }
                bytes += 8;
              }
            }
            acc[kw] = a;
          }
        }
      }
      uint64_t* r64 = (uint64_t*)S.red;
      if (NWd <= NT) {
        if (act) r64[tid] = acc[0];
        __syncthreads();
        for (uint32_t w = tid; w < NWd; w += NT) {
          uint64_t a = 0;
          for (uint32_t g = 0; g < NRG; g++) a ^= r64[g * NWd + w];
          ((uint64_t*)S.C)[w] = a & ((const uint64_t*)S.F)[w];   // po * f * <z_i, x_p> mod 4 -> bit of p1
        }
      } else {
        #pragma unroll
        for (int kw = 0; kw < KW; kw++) {
          const uint32_t w = my_w + (uint32_t)kw * NT;
          if (w < NWd) ((uint64_t*)S.C)[w] = acc[kw] & ((const uint64_t*)S.F)[w];
        }
      }
    } else {
      uint32_t ae[KW], ao[KW];
      #pragma unroll
      for (int i = 0; i < KW; i++) { ae[i] = 0; ao[i] = 0; }
      if (act) {
        #pragma unroll
        for (int kw = 0; kw < KW; kw++) {
          const uint32_t w = my_w + (uint32_t)kw * NT;
          if (w < NWd && (kw == 0 || NWd > NT)) {
            const uint32_t f = ((const uint32_t*)S.F)[w];
            uint32_t e_ = 0, o_ = 0;
            uint32_t cnt = 0;
            for (uint32_t r = my_rg; r < n; r += NRG) {
              const uint32_t sx = S.xp[r], sz = S.zp[r];
              if (sx | sz) {
                uint32_t* xr = row32(T.X, r, rs) + w;
                uint32_t* zr = row32(T.Z, r, rs) + w;
                const uint32_t z = *zr;
                if (sx) {
                  const uint32_t pz = mult<D>(S.mtab[sx], z);
                  e_ += pz & 0x00FF00FFu; o_ += (pz >> 8) & 0x00FF00FFu;
                  *xr = addm<D>(*xr, mult<D>(S.mtab[sx], f));
                  bytes += 8;
                  if (++cnt == 4096) {   // keep the 16-bit partial sums from overflowing
                    e_ = ((e_ & 0xFFFFu) % D) | (((e_ >> 16) % D) << 16);
                    o_ = ((o_ & 0xFFFFu) % D) | (((o_ >> 16) % D) << 16);
                    cnt = 0;
                  }
                }
                if (sz) { *zr = addm<D>(z, mult<D>(S.mtab[sz], f)); bytes += 4; }
                bytes += 4;
              }
            }
            ae[kw] = ((e_ & 0xFFFFu) % D) | (((e_ >> 16) % D) << 16);
            ao[kw] = ((o_ & 0xFFFFu) % D) | (((o_ >> 16) % D) << 16);
          }
        }
      }
      // reduce partial dots over row groups; C = f * dot + xz * f (f - 1) / 2 per column
      if (NWd <= NT) {
        if (act) { S.red[2 * tid] = ae[0]; S.red[2 * tid + 1] = ao[0]; }
        __syncthreads();
        for (uint32_t w = tid; w < NWd; w += NT) {
          uint32_t e_ = 0, o_ = 0;
          for (uint32_t g = 0; g < NRG; g++) { e_ += S.red[2 * (g * NWd + w)]; o_ += S.red[2 * (g * NWd + w) + 1]; }
          const uint32_t f = ((const uint32_t*)S.F)[w];
          uint32_t c = 0;
          #pragma unroll
          for (int kb = 0; kb < 4; kb++) {
            const uint32_t fb = byte_of(f, kb);
            const uint32_t dot = ((kb & 1) ? (o_ >> (16 * (kb >> 1))) : (e_ >> (16 * (kb >> 1)))) & 0xFFFFu;
            const uint32_t cb = (fb * (dot % D) + xz * ((fb * (fb - 1) / 2) % D)) % D;
            c |= cb << (8 * kb);
          }
          ((uint32_t*)S.C)[w] = c;
        }
      } else {
        #pragma unroll
        for (int kw = 0; kw < KW; kw++) {
          const uint32_t w = my_w + (uint32_t)kw * NT;
          if (w < NWd) {
            const uint32_t f = ((const uint32_t*)S.F)[w];
            uint32_t c = 0;
            #pragma unroll
            for (int kb = 0; kb < 4; kb++) {
              const uint32_t fb = byte_of(f, kb);
              const uint32_t dot = ((kb & 1) ? (ao[kw] >> (16 * (kb >> 1))) : (ae[kw] >> (16 * (kb >> 1)))) & 0xFFFFu;
              const uint32_t cb = (fb * (dot % D) + xz * ((fb * (fb - 1) / 2) % D)) % D;
              c |= cb << (8 * kb);
            }
            ((uint32_t*)S.C)[w] = c;
          }
        }
      }
    }
    __syncthreads();

    // ---- columns p and n+p: destabilizer p <- stabilizer p, stabilizer p <- Z_q
    for (uint32_t r = tid; r < n; r += NT) {
      const uint32_t xv = S.xp[r], zv = S.zp[r];
      if constexpr (D == 2) {
        uint64_t* xr = row64(T.X, r, rs);
        uint64_t* zr = row64(T.Z, r, rs);
        const uint32_t g = n + p;
        if ((p >> 6) == (g >> 6)) {
          uint64_t x = xr[p >> 6], z = zr[p >> 6];
          x &= ~(1ull << (p & 63)); x = (x & ~(1ull << (g & 63))) | ((uint64_t)xv << (g & 63));
          z = (z & ~(1ull << (p & 63))) | ((uint64_t)(r == q) << (p & 63));
          z = (z & ~(1ull << (g & 63))) | ((uint64_t)zv << (g & 63));
          xr[p >> 6] = x; zr[p >> 6] = z;
        } else {
          xr[p >> 6] &= ~(1ull << (p & 63));
          zr[p >> 6] = (zr[p >> 6] & ~(1ull << (p & 63))) | ((uint64_t)(r == q) << (p & 63));
          xr[g >> 6] = (xr[g >> 6] & ~(1ull << (g & 63))) | ((uint64_t)xv << (g & 63));
          zr[g >> 6] = (zr[g >> 6] & ~(1ull << (g & 63))) | ((uint64_t)zv << (g & 63));
        }
        bytes += 32;
      } else {
        T.X[(size_t)r * rs + p] = 0;
        T.Z[(size_t)r * rs + p] = (uint8_t)(r == q);
        T.X[(size_t)r * rs + n + p] = (uint8_t)xv;
        T.Z[(size_t)r * rs + n + p] = (uint8_t)zv;
        bytes += 4;
      }
    }
    // ---- per-shot phases: ph_i += f_i ph_p + C_i
    for (uint32_t i = tid; i < Sh * NWd; i += NT) {
      const uint32_t s = i / NWd, w = i - s * NWd;
      if constexpr (D == 2) {
        const uint32_t k2 = S.shot[s];
        const uint64_t f = ((const uint64_t*)S.F)[w];
        uint64_t* p0 = plane(T, s, 0) + w;
        uint64_t* p1 = plane(T, s, 1) + w;
        uint64_t a0 = *p0, a1 = *p1;
        const uint64_t k0m = (k2 & 1) ? f : 0ull, k1m = (k2 & 2) ? f : 0ull;
        a1 ^= k1m ^ (a0 & k0m); a0 ^= k0m;
        a1 ^= ((const uint64_t*)S.C)[w];
        *p0 = a0; *p1 = a1;
        bytes += 32;
      } else {
        uint32_t* ph = (uint32_t*)phase_vec(T, s) + w;
        const uint32_t f = ((const uint32_t*)S.F)[w];
        *ph = addm<D>(addm<D>(*ph, mult<D>(S.mtab[S.shot[s]], f)), ((const uint32_t*)S.C)[w]);
        bytes += 8;
      }
    }
    __syncthreads();
    // ---- outcome per shot; ph_{n+p} <- ph_p, ph_p <- -m po
    for (uint32_t s = tid; s < Sh; s += NT) {
      const U64x4 o = draw(it.seed, it.stream, src, it.shot0 + s);
      const uint32_t m = outcome_of(o, D);
      if constexpr (D == 2) {
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
        ph[n + p] = (uint8_t)S.shot[s];
        ph[p] = (uint8_t)((D - m) % D);
      }
      S.shot[Sh + s] = m;
      rec[s] = (uint8_t)m;
    }
  } else {
    // ---- deterministic: f_i = dx[q, i] = X[q][n + i]
    uint32_t part = 0;
    if constexpr (D == 2) {
      // f bits aligned to stabilizer columns: F[w] = bits n + 64 w .. of row q
      const uint32_t nw = (n + 63) / 64;
      const uint64_t* xr = row64(T.X, q, rs);
      for (uint32_t w = tid; w < nw; w += NT) {
        const uint32_t g = n + 64 * w;
        uint64_t lo = xr[g >> 6] >> (g & 63);
        uint64_t hi = ((g & 63) && (g >> 6) + 1 < rs / 8) ? xr[(g >> 6) + 1] << (64 - (g & 63)) : 0ull;
        uint64_t f = lo | hi;
        const uint32_t lim = n - 64 * w;
        if (lim < 64) f &= (1ull << lim) - 1;
        ((uint64_t*)S.F)[w] = f;
      }
      __syncthreads();
      // cross term per row: parity of sum_i (f x)_i * prefix_{j<i} (f z)_j
      for (uint32_t r = tid; r < n; r += NT) {
        const uint64_t* xr2 = row64(T.X, r, rs);
        const uint64_t* zr2 = row64(T.Z, r, rs);
        uint64_t carry = 0, par = 0;
        for (uint32_t w = 0; w < nw; w++) {
          const uint64_t f = ((const uint64_t*)S.F)[w];
          const uint64_t a = xr2[w] & f, b = zr2[w] & f;
          uint64_t y = b;
          y ^= y << 1; y ^= y << 2; y ^= y << 4; y ^= y << 8; y ^= y << 16; y ^= y << 32;
          const uint64_t pre = (y << 1) ^ (carry ? ~0ull : 0ull);
          par ^= (uint64_t)(__builtin_popcountll(a & pre) & 1);
          carry ^= (uint64_t)(__builtin_popcountll(b) & 1);
        }
        part += (uint32_t)par;
        bytes += 16ull * nw;
      }
      const uint32_t K = block_sum<NT>(part, S.red) & 1;
      for (uint32_t s = tid; s < Sh; s += NT) {
        uint32_t c0 = 0, c1 = 0;
        for (uint32_t w = 0; w < nw; w++) {
          const uint64_t f = ((const uint64_t*)S.F)[w];
          c0 += __builtin_popcountll(plane(T, s, 0)[w] & f);
          c1 += __builtin_popcountll(plane(T, s, 1)[w] & f);
        }
        const uint32_t ap = (c0 + 2 * c1 + 2 * K) & 3;
        const uint32_t m = ((ap + 1) >> 1) & 1;     // (-ap // 2) mod 2
        S.shot[Sh + s] = m;
        rec[s] = (uint8_t)(m | 0x80u);
        bytes += 16ull * nw;
      }
    } else {
      // f and g = f (f - 1) / 2 per stabilizer column, in F / C
      for (uint32_t i = tid; i < n; i += NT) {
        const uint32_t f = T.X[(size_t)q * rs + n + i];
        S.F[i] = (uint8_t)f;
        S.C[i] = (uint8_t)((f * (f - 1) / 2) % D);
        bytes += 1;
      }
      __syncthreads();
      // per row: exclusive scan along the stabilizer columns, one wave per row
      const uint32_t lane = tid & 63, wid = tid >> 6;
      const uint32_t cs = (n + 63) / 64;
      const uint32_t c0 = lane * cs, c1 = (c0 + cs < n) ? c0 + cs : n;
      for (uint32_t r = wid; r < n; r += NT / 64) {
        const uint8_t* xr = T.X + (size_t)r * rs;
        const uint8_t* zr = T.Z + (size_t)r * rs;
        uint32_t U = 0, V = 0, Lc = 0, SF = 0;
        for (uint32_t j = c0; j < c1; j++) {
          const uint32_t f = S.F[j], x = xr[j], z = zr[j];
          const uint32_t fx = (f * x) % D;
          Lc = (Lc + fx * U) % D;
          U = (U + f * z) % D;
          V = (V + fx) % D;
          SF = (SF + S.C[j] * ((x * z) % D)) % D;
        }
        bytes += 2ull * (c1 > c0 ? c1 - c0 : 0);
        // exclusive scan of U across lanes (mod D)
        uint32_t incl = U;
        #pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const uint32_t v = __shfl_up(incl, off, 64);
          if (lane >= (uint32_t)off) incl = (incl + v) % D;
        }
        const uint32_t excl = (incl + D - U) % D;
        part += (Lc + V * excl + SF) % D;
      }
      const uint32_t K = block_sum<NT>(part % D, S.red) % D;
      for (uint32_t s = tid; s < Sh; s += NT) {
        const uint8_t* ph = phase_vec(T, s);
        uint32_t acc = 0;
        for (uint32_t i = 0; i < n; i++) acc += (uint32_t)S.F[i] * ph[i];
        const uint32_t ap = (acc + K) % D;
        const uint32_t m = (D - ap) % D;
        S.shot[Sh + s] = m;
        rec[s] = (uint8_t)(m | 0x80u);
        bytes += n;
      }
    }
  }
  __syncthreads();

  if (reset) {
    // X^k with k = (-m) mod d: phase -= k po z[q, .]   (program.py:334-339)
    const uint32_t NWd = (D == 2) ? (G + 63) / 64 : (G + 3) / 4;
    for (uint32_t i = tid; i < Sh * NWd; i += NT) {
      const uint32_t s = i / NWd, w = i - s * NWd;
      const uint32_t m = S.shot[Sh + s];
      if constexpr (D == 2) {
        if (m & 1) { plane(T, s, 1)[w] ^= row64(T.Z, q, rs)[w]; bytes += 24; }
      } else {
        const uint32_t k = (D - m) % D;
        if (k) {
          uint32_t* ph = (uint32_t*)phase_vec(T, s) + w;
          *ph = subm<D>(*ph, mult<D>(S.mtab[k], row32(T.Z, q, rs)[w]));
          bytes += 12;
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ one item
template <int D, int NT>
SDIM_DEV void run_item(const ItemDesc& it, const KParams& P, uint8_t* lds) {
  const bool res = (it.flags & ITEM_LDS) != 0;
  const uint32_t n = it.n, rs = it.rs, Sh = it.S;
  const uint64_t T_bytes = tableau_bytes(n, D, Sh);
  uint8_t* gtab = P.arena + it.tab_off;
  uint8_t* tab = res ? lds : gtab;
  uint8_t* sbase = res ? lds + ((T_bytes + 15) & ~15ull) : lds;
  const Scr S = carve<NT>(sbase, n, rs, Sh);
  Tab T;
  T.X = tab; T.Z = tab + (size_t)n * rs; T.Ph = tab + 2 * (size_t)n * rs; T.n = n; T.rs = rs; T.S = Sh;
  unsigned long long bytes = 0;

  if (threadIdx.x < 16) S.mtab[threadIdx.x] = make_tab<D>(threadIdx.x < D ? threadIdx.x : 0);
  if (it.flags & ITEM_INIT_GIVEN) {
    if (res) copy16<NT>(lds, gtab, T_bytes);
    __syncthreads();
  } else {
    init_identity<D, NT>(T, T_bytes);
  }
  if (threadIdx.x == 0) bytes += T_bytes;

  const DevOp* ops = P.ops + it.ops_off;
  uint32_t k = 0;
  while (k < it.n_ops) {
    const DevOp o = ops[k];
    const uint32_t kind = op_kind(o);
    if (kind == OP_SEG) {
      k = run_segment<D, NT>(ops, k, T, S, it, P, bytes);
    } else if (kind == OP_MEAS || kind == OP_RESET) {
      measure<D, NT>(T, o.q, o.t, o.src, kind == OP_RESET, S, it, P, bytes);
      k++;
    } else {
      break;   // OP_END
    }
  }
  __syncthreads();
  if (res) {
    copy16<NT>(gtab, lds, T_bytes);
    if (threadIdx.x == 0) bytes += T_bytes;
    __syncthreads();
  }
  const unsigned long long tot = block_sum64<NT>(bytes, S.red);
  if (threadIdx.x == 0) atomicAdd(&P.counters[res ? 1 : 0], tot);
}

// ------------------------------------------------------------------ persistent batch kernel
template <int NT>
__global__ __launch_bounds__(NT) void batch_kernel(KParams P) {
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

// ------------------------------------------------------------------ plain path: one launch per op
// Every item of one circuit (its shot groups) is processed by the same launch;
// the tableau stays in HBM.  Gates update the phases immediately.
template <int D>
__global__ __launch_bounds__(256) void plain_init_kernel(KParams P) {
  const ItemDesc it = P.items[blockIdx.x];
  uint8_t* tab = P.arena + it.tab_off;
  Tab T;
  T.X = tab; T.Z = tab + (size_t)it.n * it.rs; T.Ph = tab + 2 * (size_t)it.n * it.rs;
  T.n = it.n; T.rs = it.rs; T.S = it.S;
  init_identity<D, 256>(T, tableau_bytes(it.n, D, it.S));
}

template <int D>
__global__ __launch_bounds__(256) void plain_gate_kernel(KParams P, DevOp o) {
  const ItemDesc it = P.items[blockIdx.y];
  const uint32_t n = it.n, rs = it.rs, Sh = it.S;
  const uint32_t G = 2 * n;
  const uint32_t NWd = (D == 2) ? (G + 63) / 64 : (G + 3) / 4;
  const uint32_t w = blockIdx.x * 256 + threadIdx.x;
  __shared__ MulTab mt[16];
  if (threadIdx.x < 16) mt[threadIdx.x] = make_tab<D>(threadIdx.x < D ? threadIdx.x : 0);
  __syncthreads();
  if (w >= NWd) return;
  uint8_t* tab = P.arena + it.tab_off;
  Tab T;
  T.X = tab; T.Z = tab + (size_t)n * rs; T.Ph = tab + 2 * (size_t)n * rs; T.n = n; T.rs = rs; T.S = Sh;
  Scr S;
  S.mtab = mt;
  const uint32_t kind = op_kind(o);
  if (kind == OP_NOISE) {
    // per shot draw and apply: phase += po (b x - a z)
    const uint32_t q = o.q;
    for (uint32_t s = 0; s < Sh; s++) {
      const U64x4 dr = draw(it.seed, it.stream, o.src, it.shot0 + s);
      const uint32_t ab = noise_ab(dr, D, op_a(o), P.keep53[it.keep_off + o.t]);
      if (!ab) continue;
      const uint32_t a = ab & 15, b = (ab >> 4) & 15;
      if constexpr (D == 2) {
        const uint64_t x = row64(T.X, q, rs)[w], z = row64(T.Z, q, rs)[w];
        plane(T, s, 1)[w] ^= ((b & 1) ? x : 0ull) ^ ((a & 1) ? z : 0ull);
      } else {
        const uint32_t x = row32(T.X, q, rs)[w], z = row32(T.Z, q, rs)[w];
        uint32_t* ph = (uint32_t*)phase_vec(T, s) + w;
        *ph = addm<D>(*ph, subm<D>(mult<D>(mt[b], x), mult<D>(mt[a], z)));
      }
    }
    return;
  }
  if constexpr (D == 2) {
    uint64_t d0 = 0, d1 = 0;
    gate_bin(kind, o.q, o.t, op_a(o), op_b(o), T, w, d0, d1, S, 0);
    if (d0 | d1) {
      for (uint32_t s = 0; s < Sh; s++) {
        uint64_t* p0 = plane(T, s, 0) + w;
        uint64_t* p1 = plane(T, s, 1) + w;
        uint64_t a0 = *p0, a1 = *p1;
        *p1 = a1 ^ d1 ^ (a0 & d0);
        *p0 = a0 ^ d0;
      }
    }
  } else {
    uint32_t dl = 0;
    gate_odd<D>(kind, o.q, o.t, op_a(o), op_b(o), T, w, dl, S, 0);
    if (dl) {
      for (uint32_t s = 0; s < Sh; s++) {
        uint32_t* ph = (uint32_t*)phase_vec(T, s) + w;
        *ph = addm<D>(*ph, dl);
      }
    }
  }
}

template <int D, int NT>
__global__ __launch_bounds__(NT) void plain_meas_kernel(KParams P, DevOp o) {
  extern __shared__ __align__(16) uint8_t lds[];
  const ItemDesc it = P.items[blockIdx.x];
  const uint32_t n = it.n, rs = it.rs, Sh = it.S;
  uint8_t* tab = P.arena + it.tab_off;
  Tab T;
  T.X = tab; T.Z = tab + (size_t)n * rs; T.Ph = tab + 2 * (size_t)n * rs; T.n = n; T.rs = rs; T.S = Sh;
  const Scr S = carve<NT>(lds, n, rs, Sh);
  if (threadIdx.x < 16) S.mtab[threadIdx.x] = make_tab<D>(threadIdx.x < D ? threadIdx.x : 0);
  __syncthreads();
  unsigned long long bytes = 0;
  measure<D, NT>(T, o.q, o.t, o.src, op_kind(o) == OP_RESET, S, it, P, bytes);
}

}  // namespace sdim

// ======================================================================== host launchers
namespace sdim {

__global__ void reset_kernel(uint32_t* queue, unsigned long long* counters) {
  if (threadIdx.x < 4) queue[threadIdx.x] = 0;
  if (threadIdx.x < 2) counters[threadIdx.x] = 0;
}

int prepare_kernels() {
  const int lds = 163584;
  if (hipFuncSetAttribute((const void*)batch_kernel<256>, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void*)batch_kernel<512>, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void*)batch_kernel<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return -1;
#define SDIM_SET_PLAIN(D) \
  if (hipFuncSetAttribute((const void*)plain_meas_kernel<D, 256>, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return -1;
  SDIM_SET_PLAIN(2) SDIM_SET_PLAIN(3) SDIM_SET_PLAIN(5) SDIM_SET_PLAIN(7) SDIM_SET_PLAIN(11)
#undef SDIM_SET_PLAIN
  return 0;
}

int launch_reset(uint32_t* queue, unsigned long long* counters, hipStream_t st) {
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, st, queue, counters);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_batch(const KParams& P, int nt, int grid, hipStream_t st) {
  const size_t lds = P.lds_bytes;
  switch (nt) {
    case 256: hipLaunchKernelGGL(batch_kernel<256>, dim3(grid), dim3(256), lds, st, P); break;
    case 512: hipLaunchKernelGGL(batch_kernel<512>, dim3(grid), dim3(512), lds, st, P); break;
    case 1024: hipLaunchKernelGGL(batch_kernel<1024>, dim3(grid), dim3(1024), lds, st, P); break;
    default: return -2;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <int D>
static int plain_op_d(const KParams& P, const DevOp& o, uint32_t n, uint32_t n_items, uint32_t S, hipStream_t st) {
  const uint32_t kind = op_kind(o);
  if (kind == OP_MEAS || kind == OP_RESET) {
    const uint32_t lds = scratch_bytes(n, row_stride(n, D), S, 256);
    hipLaunchKernelGGL((plain_meas_kernel<D, 256>), dim3(n_items), dim3(256), lds, st, P, o);
  } else if (kind == OP_END) {
    hipLaunchKernelGGL(plain_init_kernel<D>, dim3(n_items), dim3(256), 0, st, P);
  } else {
    const uint32_t G = 2 * n;
    const uint32_t NWd = (D == 2) ? (G + 63) / 64 : (G + 3) / 4;
    hipLaunchKernelGGL(plain_gate_kernel<D>, dim3((NWd + 255) / 256, n_items), dim3(256), 0, st, P, o);
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// one plain launch; OP_END means "initialise the items' tableaux"
int launch_plain_op(const KParams& P, const DevOp& o, uint32_t n, uint32_t d, uint32_t n_items, uint32_t S, hipStream_t st) {
  switch (d) {
    case 2: return plain_op_d<2>(P, o, n, n_items, S, st);
    case 3: return plain_op_d<3>(P, o, n, n_items, S, st);
    case 5: return plain_op_d<5>(P, o, n, n_items, S, st);
    case 7: return plain_op_d<7>(P, o, n, n_items, S, st);
    case 11: return plain_op_d<11>(P, o, n, n_items, S, st);
  }
  return -2;
}

}  // namespace sdim
