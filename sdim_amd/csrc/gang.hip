// Gang kernel of one dimension: one large tableau processed by a gang of
// workgroups.  This file is compiled once per dimension (-DSDIM_GANG_D=2, 3, 5, 7,
// 11); each object defines launch_gang_d / prepare_gang for its dimension.
#include "tableau_dev.h"

#ifndef SDIM_GANG_D
#error "compile gang.hip with -DSDIM_GANG_D=<dimension>"
#endif

namespace sdim {

// ------------------------------------------------------------------ gangs: one item on many workgroups
// A tableau with thousands of qudits is processed by a gang of K workgroups.
// Member m owns a contiguous range of column words (in every row of X and Z
// and in every shot's phase vector), so gate segments need no coordination
// (gates act on each generator column independently) and run with the
// one-workgroup segment code on a view of the member's columns.  Measurements
// synchronise the gang:
//   random  A  own part of row q -> pivot candidate -> gang atomic-min  [barrier]
//           B  each member gathers the pivot column and pivot phases    [barrier]
//           C, D, E on own columns (the owners of p and n+p rewrite them)
//   det     per row, the local scan over own stabilizer columns; per-row sums
//           SX = sum f x, SZ = sum f z and per-shot partial phase sums are
//           published                                                   [barrier]
//           cross term sum_r sum_w SX_w[r] * sum_{v<w} SZ_v[r] over a row range,
//           per-shot totals over a shot range                           [barrier]
// The barrier is an arrival counter in global memory, published and observed
// with agent-scope fences.  All members of a launch are resident together (the
// grid is sized from the kernel's occupancy) and every wait is bounded: a
// member that waits past a deadline raises the gang's error word, and every
// member leaves the kernel when it sees it.
struct GangSync {
  uint32_t ctr;        // barrier arrivals (one per XCC group after the first barrier)
  int err;             // nonzero: a member timed out
  uint32_t item[2];    // item index broadcast (double-buffered)
  int piv[2];          // pivot candidates (double-buffered by measurement count)
  uint32_t xcnt[16];   // members on each XCC (counted once, before the first barrier)
  uint32_t pad[42];
  struct { uint32_t v; uint32_t pad[31]; } xarr[16];   // per-XCC arrivals, one 128-byte line each
  uint32_t tail[448];
};
static_assert(sizeof(GangSync) == GANG_SYNC_BYTES, "GangSync size");


struct GangCtx {
  GangSync* sy;
  uint8_t* sxz;        // [n][Kp] bytes: SX | SZ << 4 of member v for row r
  uint8_t* ps;         // [S][Kp] bytes: per-shot partial phase sums of member v
  uint32_t* acc;       // [K] local scan sums
  uint32_t* crs;       // [K] cross-term sums
  uint32_t* tot;       // [S] per-shot totals
  uint8_t* stage[2];   // pivot column staging (by measurement parity): x | z << 4 per row, then pivot phases
  uint32_t K, Kp, me;
  int* gerr;
  uint32_t target;     // arrivals expected at the next barrier
  uint32_t xcc;        // this member's XCC (shared L2)
  uint32_t xc, ngrp;   // members on this XCC, XCCs holding members (0 until counted)
  uint32_t xtarget;    // arrivals expected at this XCC's next barrier
  uint64_t timeout;    // barrier wait bound (constant-clock ticks)
};

// XCC (accelerator die, one L2 each) of the executing wave
SDIM_DEV uint32_t xcc_id() {
  uint32_t v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return v & 15u;
}

// Barrier of a gang.  The first one (g.ngrp == 0) counts every member and releases each
// member's writes on its own.  Later ones are two-level: members on one XCC share an L2,
// so each member only waits until its own stores are acknowledged by that L2 and counts
// itself in its XCC's record; the XCC's last arriver releases the L2 (one write-back for
// the whole group) and counts the group once; everyone polls the group count.
SDIM_DEV bool gang_barrier(GangCtx& g, int* s_ok) {
  const bool grouped = g.ngrp != 0;
  if (grouped) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // this wave's stores are in L2
  __syncthreads();
  g.target += grouped ? g.ngrp : g.K;
  if (grouped) g.xtarget += g.xc;
  if (threadIdx.x == 0) {
    int ok = 1;
    if (grouped) {
      const uint32_t old = __hip_atomic_fetch_add(&g.sy->xarr[g.xcc].v, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (old == g.xtarget - 1)   // last of the XCC: write back its L2, then count the group
        __hip_atomic_fetch_add(&g.sy->ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      // release: this workgroup's writes (ordered before thread 0 by the barrier above) become
      // visible device-wide before the arrival is counted
      __hip_atomic_fetch_add(&g.sy->ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(&g.sy->ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < g.target) {
      if (__hip_atomic_load(&g.sy->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { ok = 0; break; }
      if (wall_clock64() - t0 > g.timeout) {
        __hip_atomic_store(&g.sy->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(g.gerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // later reads see the other members' writes
    *s_ok = ok;
  }
  __syncthreads();
  return *s_ok != 0;
}

// identity on the member's columns: V covers bytes [c_lo, c_lo + cb) of every row
template <int D, int NT>
SDIM_DEV void gang_init_identity(const Tab& V, uint32_t c_lo, uint32_t cb) {
  const uint32_t n = V.n, rs = V.rs;
  const uint32_t q16 = cb / 16;
  const uint32_t xzr = 2 * rows_padded(n);                  // rows of X then Z
  for (uint32_t i = threadIdx.x; i < xzr * q16; i += NT) {
    const uint32_t r = i / q16, c = i - r * q16;
    *(uint4*)b8(V.X, r, 16 * c, rs, V.il) = make_uint4(0, 0, 0, 0);
  }
  const uint32_t prow = (D == 2 ? 2 : 1) * V.S;             // phase rows (row-major)
  for (uint32_t i = threadIdx.x; i < prow * q16; i += NT) {
    const uint32_t r = i / q16, c = i - r * q16;
    *(uint4*)(V.Ph + (size_t)r * rs + 16 * c) = make_uint4(0, 0, 0, 0);
  }
  __syncthreads();
  for (uint32_t r = threadIdx.x; r < n; r += NT) {
    if constexpr (D == 2) {
      const uint32_t g = n + r;
      if (g >= c_lo * 8 && g < (c_lo + cb) * 8) *w64(V.X, r, (g >> 6) - c_lo / 8, rs, V.il) |= 1ull << (g & 63);
      if (r >= c_lo * 8 && r < (c_lo + cb) * 8) *w64(V.Z, r, (r >> 6) - c_lo / 8, rs, V.il) |= 1ull << (r & 63);
    } else {
      if (n + r >= c_lo && n + r < c_lo + cb) *b8(V.X, r, n + r - c_lo, rs, V.il) = 1;
      if (r >= c_lo && r < c_lo + cb) *b8(V.Z, r, r - c_lo, rs, V.il) = 1;
    }
  }
  __syncthreads();
}

// Cross term of one row over the K members' published bytes b_v = SX | SZ << 4:
// sum_v SX_v * (sum_{u<v} SZ_u), mod d (parity for d = 2).
template <int D>
SDIM_DEV uint32_t gang_row_cross(const uint8_t* row, uint32_t K) {
  uint32_t pre = 0, c = 0;
  for (uint32_t v0 = 0; v0 < K; v0 += 16) {
    const uint4 q = *(const uint4*)(row + v0);
    const uint32_t wd[4] = {q.x, q.y, q.z, q.w};
    #pragma unroll
    for (int j = 0; j < 16; j++) {
      if (v0 + (uint32_t)j < K) {
        const uint32_t b = (wd[j >> 2] >> (8 * (j & 3))) & 0xFFu;
        if constexpr (D == 2) { c ^= (b & 1u) & pre; pre ^= (b >> 4) & 1u; }
        else { c += (b & 15u) * pre; pre = (pre + (b >> 4)) % D; }
      }
    }
    if constexpr (D != 2) c %= D;
  }
  return c;
}

// plain sum of the K bytes of one row
template <int D>
SDIM_DEV uint32_t gang_row_sum(const uint8_t* row, uint32_t K) {
  uint32_t t = 0;
  for (uint32_t v0 = 0; v0 < K; v0 += 16) {
    const uint4 q = *(const uint4*)(row + v0);
    const uint32_t wd[4] = {q.x, q.y, q.z, q.w};
    #pragma unroll
    for (int j = 0; j < 16; j++)
      if (v0 + (uint32_t)j < K) t += (wd[j >> 2] >> (8 * (j & 3))) & 0xFFu;
  }
  return t;
}

// Measurement of qudit q by one member.  Tg: the whole tableau; V: the member's
// columns (words w0 .. w0 + V.nw).  Returns false when the gang was aborted.
template <int D, int NT>
SDIM_DEV bool gang_measure(const Tab& Tg, const Tab& V, uint32_t w0, uint32_t q, uint32_t slot, uint32_t src,
                           bool reset, uint32_t mk, const Scr& S, const ItemDesc& it, const KParams& P,
                           GangCtx& Gc, int* s_ok, unsigned long long& bytes) {
  constexpr bool BIN = (D == 2);
  constexpr uint32_t CPW = BIN ? 64 : 4;
  constexpr int RB = 8;     // rows per load batch in the sweep (16 bytes of X and Z each)
  const uint32_t tid = threadIdx.x;
  const uint32_t n = Tg.n, rs = Tg.rs, Sh = Tg.S;
  const uint32_t nw = V.nw;
  const uint32_t nsw = (n + CPW - 1) / CPW;
  const uint32_t c_lo = w0 * CPW, c_hi = (w0 + nw) * CPW;   // own columns
  int* M = S.misc + 16 * (mk & 1);
  int* Mn = S.misc + 16 * ((mk + 1) & 1);
  uint8_t* rec = P.arena + it.rec_off + (size_t)slot * Sh;
  const bool lead = Gc.me == 0;

  // phase timers (gang): 0 A, 1 barrier 1, 2 column pack, 3 barrier 2, 4 unpack, 5 sweep, 6 D, 10 E, 12 det
  SDIM_TMARK(tA);
  // ---- A: own 16-byte groups of row q -> pivot candidate, f, groups with f != 0
  rowq_groups<D, NT>(V, w0, q, S, M, nullptr, bytes);
  __syncthreads();
  // the gang's pivot is the least candidate column; its code carries x[q, p] in the low four
  // bits so that every member knows the exponent e right after the barrier
  if (tid == 0 && M[0] != INT_MAX) {
    const uint32_t pc = (uint32_t)M[0];
    uint32_t xqp = 1;
    if constexpr (!BIN) xqp = (D - S.F[pc - c_lo]) % D;      // F = -x[q, .] on own columns
    __hip_atomic_fetch_min(&Gc.sy->piv[mk & 1], (int)(pc * 16u + xqp), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  SDIM_TADD(P, 0, tA);
  SDIM_TMARK(tG1);
  if (!gang_barrier(Gc, s_ok)) return false;                                // G1
  SDIM_TADD(P, 1, tG1);
  const int pcode = __hip_atomic_load(&Gc.sy->piv[mk & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int piv = pcode == INT_MAX ? INT_MAX : pcode >> 4;
  if (tid == 0) {
    Mn[0] = INT_MAX; Mn[1] = 0; Mn[2] = 0; Mn[3] = 0;
    // the other slot is next written in measurement mk + 1, after this one's last barrier
    if (lead) __hip_atomic_store(&Gc.sy->piv[(mk + 1) & 1], INT_MAX, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  if (piv != INT_MAX) {
    const uint32_t p = (uint32_t)piv;
    const bool own_p = p >= c_lo && p < c_hi;
    const bool own_g = n + p >= c_lo && n + p < c_hi;
    uint8_t* stg = Gc.stage[mk & 1];
    // exponentiation of stabilizer p by e = x[q,p]^-1 (tableau_prime.py:278-282, 365-380)
    uint32_t e = 1;
    if constexpr (!BIN) {
      const uint32_t xv = (uint32_t)pcode & 15u;
      #pragma unroll
      for (uint32_t c = 1; c < D; c++) if ((xv * c) % D == 1) e = c;
    }
    SDIM_TMARK(tB);
    // ---- B: every member packs its slice of rows of column p (x | z << 4 per row) and of the
    // pivot phases into the staging buffer: one round of strided loads per member.  At the
    // same rows it gives columns p and n+p their final values (destabilizer n+p becomes e
    // times stabilizer p at every row, stabilizer p becomes Z_q at the touched rows), so that
    // no member rewrites n scattered rows alone; the owners' sweep keeps both columns, f
    // being 0 there.
    {
      const uint32_t r0 = (uint32_t)(((uint64_t)n * Gc.me) / Gc.K), r1 = (uint32_t)(((uint64_t)n * (Gc.me + 1)) / Gc.K);
      const uint32_t np_ = n + p;
      for (uint32_t r = r0 + tid; r < r1; r += NT) {
        uint32_t xr, zr;
        if constexpr (BIN) {
          const uint32_t wp = p >> 6, wn = np_ >> 6;
          uint64_t* xpw = w64(Tg.X, r, wp, rs, Tg.il);
          uint64_t* zpw = w64(Tg.Z, r, wp, rs, Tg.il);
          uint64_t* xnw = w64(Tg.X, r, wn, rs, Tg.il);
          uint64_t* znw = w64(Tg.Z, r, wn, rs, Tg.il);
          const bool sep = wn != wp;
          const uint64_t xa = *xpw, za = *zpw;
          const uint64_t xb = sep ? *xnw : xa, zb = sep ? *znw : za;
          xr = (uint32_t)((xa >> (p & 63)) & 1);
          zr = (uint32_t)((za >> (p & 63)) & 1);
          const uint64_t mp = 1ull << (p & 63), mn = 1ull << (np_ & 63);
          uint64_t xn = (xb & ~mn) | (xr ? mn : 0ull), zn = (zb & ~mn) | (zr ? mn : 0ull);
          if (!sep && (xr | zr)) {
            xn &= ~mp;
            zn = (zn & ~mp) | (r == q ? mp : 0ull);
          }
          if (xn != xb) *xnw = xn;
          if (zn != zb) *znw = zn;
          if (sep && (xr | zr)) {
            *xpw = xa & ~mp;
            *zpw = (za & ~mp) | (r == q ? mp : 0ull);
          }
          bytes += 32;
        } else {
          xr = *b8(Tg.X, r, p, rs, Tg.il);
          zr = *b8(Tg.Z, r, p, rs, Tg.il);
          const uint32_t xo = *b8(Tg.X, r, np_, rs, Tg.il), zo = *b8(Tg.Z, r, np_, rs, Tg.il);
          const uint32_t xn = (e * xr) % D, zn = (e * zr) % D;
          if (xn != xo) *b8(Tg.X, r, np_, rs, Tg.il) = (uint8_t)xn;
          if (zn != zo) *b8(Tg.Z, r, np_, rs, Tg.il) = (uint8_t)zn;
          if (xr | zr) {
            *b8(Tg.X, r, p, rs, Tg.il) = 0;
            *b8(Tg.Z, r, p, rs, Tg.il) = (uint8_t)(r == q);
          }
          bytes += 6;
        }
        stg[r] = (uint8_t)(xr | (zr << 4));
      }
      const uint32_t s0 = (uint32_t)(((uint64_t)Sh * Gc.me) / Gc.K), s1 = (uint32_t)(((uint64_t)Sh * (Gc.me + 1)) / Gc.K);
      for (uint32_t s = s0 + tid; s < s1; s += NT) {
        if constexpr (BIN) {
          const uint32_t v0 = (uint32_t)((plane(Tg, s, 0)[p >> 6] >> (p & 63)) & 1);
          const uint32_t v1 = (uint32_t)((plane(Tg, s, 1)[p >> 6] >> (p & 63)) & 1);
          stg[n + s] = (uint8_t)(v0 | (v1 << 1));
        } else {
          stg[n + s] = phase_vec(Tg, s)[p];
        }
      }
    }
    if (tid == 0) {
      if constexpr (BIN) {
        if (own_p) ((uint64_t*)S.F)[(p >> 6) - w0] &= ~(1ull << (p & 63));
        if (own_g) ((uint64_t*)S.F)[((n + p) >> 6) - w0] &= ~(1ull << ((n + p) & 63));
      } else {
        if (own_p) S.F[p - c_lo] = 0;
        if (own_g) S.F[n + p - c_lo] = 0;
      }
    }
    SDIM_TADD(P, 2, tB);
    SDIM_TMARK(tG2);
    if (!gang_barrier(Gc, s_ok)) return false;                              // G2: column p staged
    SDIM_TADD(P, 3, tG2);
    SDIM_TMARK(tU);
    // every member: unpack the pivot column, touched rows (compacted per wave), <x_p, z_p>,
    // pivot phases.  Four rows per 32-bit word, UW words per thread in flight.
    {
      constexpr uint32_t UW = 4;
      const uint32_t nwr = (n + 3) / 4;                 // staging words holding rows
      const uint32_t lane = tid & 63;
      uint32_t xzs = 0;
      for (uint32_t j0 = 0; j0 < nwr; j0 += NT * UW) {
        uint32_t wv[UW];
        #pragma unroll
        for (uint32_t u = 0; u < UW; u++) {
          const uint32_t j = j0 + u * NT + tid;
          wv[u] = j < nwr ? ((const uint32_t*)stg)[j] : 0u;
        }
        #pragma unroll
        for (uint32_t u = 0; u < UW; u++) {
          const uint32_t j = j0 + u * NT + tid;
          uint32_t v = wv[u];
          if (4 * j + 4 > n && j < nwr) v &= (1u << (8 * (n - 4 * j))) - 1;   // bytes past row n-1
          uint32_t cnt = 0;
          #pragma unroll
          for (int b = 0; b < 4; b++) cnt += ((v >> (8 * b)) & 0xFFu) ? 1u : 0u;
          // wave-exclusive prefix of the counts, one LDS atomic per wave
          uint32_t x = cnt;
          #pragma unroll
          for (int o = 1; o < 64; o <<= 1) { const uint32_t y = __shfl_up(x, o, 64); if (lane >= (uint32_t)o) x += y; }
          const uint32_t tot = __shfl(x, 63, 64);
          uint32_t base = 0;
          if (lane == 0 && tot) base = (uint32_t)atomicAdd(&M[1], (int)tot);
          base = __shfl(base, 0, 64) + x - cnt;
          if (j < nwr) {
            #pragma unroll
            for (int b = 0; b < 4; b++) {
              const uint32_t r = 4 * j + b;
              if (r < n) {
                const uint32_t e8 = (v >> (8 * b)) & 0xFFu;
                const uint32_t xr = e8 & 15u, zr = e8 >> 4;
                S.xp[r] = (uint8_t)xr;
                S.zp[r] = (uint8_t)zr;
                if (e8) {
                  if (base < S.tr_cap) S.TR[base] = (uint16_t)r;
                  else report_error(S, KERR_TOUCHED_ROWS);
                  base++;
                }
                xzs += xr * zr;
              }
            }
          }
        }
      }
      if constexpr (!BIN) if (xzs) atomicAdd(&M[3], (int)(xzs % D));
    }
    for (uint32_t s = tid; s < Sh; s += NT) S.shot[s] = stg[n + s];
    __syncthreads();
    SDIM_TADD(P, 4, tU);
    SDIM_TMARK(tC);
    const uint32_t nT = min((uint32_t)M[1], S.tr_cap), nF = min((uint32_t)M[2], S.fl_cap);
    uint32_t xz = 0, pinc = 0;
    if constexpr (!BIN) {
      const uint32_t xz_old = (uint32_t)M[3] % D;
      xz = ((e * e) % D) * xz_old % D;
      pinc = xz_old * ((e * (e - 1) / 2) % D) % D;
    }

    // ---- C: sweep (touched rows) x (own 16-byte groups with f != 0)
    sweep_groups<D, NT, RB>(V, S, nT, nF, e, bytes);
    __syncthreads();
    SDIM_TADD(P, 5, tC);
    SDIM_TMARK(tD);

    // ---- D: per-shot phases over own words with f != 0.
    if constexpr (D != 2) {
      rowsum_shared_terms<D, NT>(S, nF, xz);
      __syncthreads();
    }
    rowsum_phases<D, NT>(V, S, nF, e, xz, pinc, bytes);
    __syncthreads();
    SDIM_TADD(P, 6, tD);
    SDIM_TMARK(tE);
    // ---- E: outcomes; ph_{n+p} <- ph_p, ph_p <- -m po on the owners' columns
    for (uint32_t s = tid; s < Sh; s += NT) {
      const U64x4 o = draw(it.seed, it.stream, src, it.shot0 + s);
      const uint32_t m = outcome_of(o, D);
      if constexpr (BIN) {
        const uint32_t k2 = S.shot[s];
        const uint32_t g = n + p;
        uint64_t* p0 = plane(V, s, 0);
        uint64_t* p1 = plane(V, s, 1);
        if (own_g) {
          const uint32_t wl = (g >> 6) - w0;
          p0[wl] = (p0[wl] & ~(1ull << (g & 63))) | ((uint64_t)(k2 & 1) << (g & 63));
          p1[wl] = (p1[wl] & ~(1ull << (g & 63))) | ((uint64_t)(k2 >> 1) << (g & 63));
        }
        if (own_p) {
          const uint32_t wl = (p >> 6) - w0;
          p0[wl] &= ~(1ull << (p & 63));
          p1[wl] = (p1[wl] & ~(1ull << (p & 63))) | ((uint64_t)(m & 1) << (p & 63));
        }
      } else {
        uint8_t* ph = phase_vec(V, s);
        if (own_g) ph[n + p - c_lo] = (uint8_t)((e * S.shot[s] + pinc) % D);
        if (own_p) ph[p - c_lo] = (uint8_t)((D - m) % D);
      }
      S.shot[Sh + s] = m;
      if (lead) rec[s] = (uint8_t)m;
    }
    SDIM_TADD(P, 10, tE);
  } else {
    SDIM_TMARK(tDet);
    // ---- deterministic outcome: prod_i S_i^{f_i}, f_i = dx[q, i] = X[q][n + i]
    const uint32_t ns_own = (w0 < nsw) ? ((nsw - w0) < nw ? (nsw - w0) : nw) : 0u;
    for (uint32_t w = tid; w < ns_own; w += NT) {
      const uint32_t gw = w0 + w;
      if constexpr (BIN) {
        const uint32_t g = n + 64 * gw;
        const uint64_t lo = *w64(Tg.X, q, g >> 6, rs, Tg.il) >> (g & 63);
        const uint64_t hi = ((g & 63) && (g >> 6) + 1 < rs / 8) ? *w64(Tg.X, q, (g >> 6) + 1, rs, Tg.il) << (64 - (g & 63)) : 0ull;
        uint64_t f = lo | hi;
        const uint32_t lim = n - 64 * gw;
        if (lim < 64) f &= (1ull << lim) - 1;
        ((uint64_t*)S.F)[w] = f;
        bytes += 16;
      } else {
        uint32_t f = 0;
        #pragma unroll
        for (int b = 0; b < 4; b++) {
          const uint32_t c = 4 * gw + b;
          if (c < n) f |= (uint32_t)*b8(Tg.X, q, n + c, rs, Tg.il) << (8 * b);
        }
        ((uint32_t*)S.F)[w] = f;
        bytes += 4;
      }
    }
    __syncthreads();
    // local scan of every row over own stabilizer columns; publish SX, SZ of the row
    uint32_t part = 0;
    for (uint32_t r = tid; r < n; r += NT) {
      uint32_t sx = 0, sz = 0;
      if constexpr (BIN) {
        uint64_t carry = 0, par = 0, pa = 0;
        for (uint32_t w = 0; w < ns_own; w++) {
          const uint64_t f = ((const uint64_t*)S.F)[w];
          if (!f) continue;
          const uint64_t a = *w64(V.X, r, w, rs, V.il) & f, b = *w64(V.Z, r, w, rs, V.il) & f;
          uint64_t y = b;
          y ^= y << 1; y ^= y << 2; y ^= y << 4; y ^= y << 8; y ^= y << 16; y ^= y << 32;
          const uint64_t pre = (y << 1) ^ (carry ? ~0ull : 0ull);
          par ^= (uint64_t)(__builtin_popcountll(a & pre) & 1);
          carry ^= (uint64_t)(__builtin_popcountll(b) & 1);
          pa ^= (uint64_t)(__builtin_popcountll(a) & 1);
          bytes += 16;
        }
        part += (uint32_t)par;
        sx = (uint32_t)pa; sz = (uint32_t)carry;
      } else {
        uint32_t U = 0, acc = 0, ax = 0;
        for (uint32_t w = 0; w < ns_own; w++) {
          const uint32_t f = ((const uint32_t*)S.F)[w];
          if (!f) continue;
          const uint32_t x = *w32(V.X, r, w, rs, V.il), z = *w32(V.Z, r, w, rs, V.il);
          #pragma unroll
          for (int b = 0; b < 4; b++) {
            const uint32_t fb = byte_of(f, b);
            const uint32_t xb = byte_of(x, b), zb = byte_of(z, b);
            acc += fb * xb * U + ((fb * (fb - 1) / 2) % D) * xb * zb;
            U = (U + fb * zb) % D;
            ax += fb * xb;
          }
          bytes += 8;
        }
        part += acc % D;
        sx = ax % D; sz = U;
      }
      Gc.sxz[(size_t)r * Gc.Kp + Gc.me] = (uint8_t)(sx | (sz << 4));
    }
    const uint32_t loc = BIN ? (block_sum<NT>(part, S.red) & 1) : (block_sum<NT>(part % D, S.red) % D);
    if (tid == 0) Gc.acc[Gc.me] = loc;
    // per-shot partial sums of f_i ph_i over own stabilizer columns
    for (uint32_t s = tid; s < Sh; s += NT) {
      uint32_t v;
      if constexpr (BIN) {
        uint32_t c0 = 0, c1 = 0;
        for (uint32_t w = 0; w < ns_own; w++) {
          const uint64_t f = ((const uint64_t*)S.F)[w];
          if (!f) continue;
          c0 += __builtin_popcountll(plane(V, s, 0)[w] & f);
          c1 += __builtin_popcountll(plane(V, s, 1)[w] & f);
          bytes += 16;
        }
        v = (c0 + 2 * c1) & 3;
      } else {
        const uint32_t* ph = (const uint32_t*)phase_vec(V, s);
        uint32_t acc = 0;
        for (uint32_t w = 0; w < ns_own; w++) {
          const uint32_t f = ((const uint32_t*)S.F)[w];
          if (!f) continue;
          const uint32_t pv = ph[w];
          #pragma unroll
          for (int b = 0; b < 4; b++) acc += byte_of(f, b) * byte_of(pv, b);
          bytes += 4;
        }
        v = acc % D;
      }
      Gc.ps[(size_t)s * Gc.Kp + Gc.me] = (uint8_t)v;
    }
    if (!gang_barrier(Gc, s_ok)) return false;                              // G2
    // cross term over this member's row range, per-shot totals over its shot range
    const uint32_t r0 = (uint32_t)(((uint64_t)n * Gc.me) / Gc.K), r1 = (uint32_t)(((uint64_t)n * (Gc.me + 1)) / Gc.K);
    uint32_t cr = 0;
    for (uint32_t r = r0 + tid; r < r1; r += NT) cr += gang_row_cross<D>(Gc.sxz + (size_t)r * Gc.Kp, Gc.K);
    const uint32_t crs = BIN ? (block_sum<NT>(cr, S.red) & 1) : (block_sum<NT>(cr, S.red) % D);
    if (tid == 0) Gc.crs[Gc.me] = crs;
    const uint32_t s0 = (uint32_t)(((uint64_t)Sh * Gc.me) / Gc.K), s1 = (uint32_t)(((uint64_t)Sh * (Gc.me + 1)) / Gc.K);
    for (uint32_t s = s0 + tid; s < s1; s += NT) {
      const uint32_t t = gang_row_sum<D>(Gc.ps + (size_t)s * Gc.Kp, Gc.K);
      Gc.tot[s] = BIN ? (t & 3) : (t % D);
    }
    if (!gang_barrier(Gc, s_ok)) return false;                              // G3
    uint32_t kk = 0;
    for (uint32_t v = tid; v < Gc.K; v += NT) kk += Gc.acc[v] + Gc.crs[v];
    const uint32_t Kc = BIN ? (block_sum<NT>(kk, S.red) & 1) : (block_sum<NT>(kk, S.red) % D);
    for (uint32_t s = tid; s < Sh; s += NT) {
      const uint32_t t = Gc.tot[s];
      uint32_t m;
      if constexpr (BIN) m = ((((t + 2 * Kc) & 3) + 1) >> 1) & 1;
      else m = (D - (t + Kc) % D) % D;
      S.shot[Sh + s] = m;
      if (lead) rec[s] = (uint8_t)(m | 0x80u);
    }
    SDIM_TADD(P, 12, tDet);
  }

  if (reset) {
    __syncthreads();
    // X^k with k = (-m) mod d on own columns: phase -= k po z[q, .]
    for (uint32_t i = tid; i < Sh * nw; i += NT) {
      const uint32_t s = i / nw, w = i - s * nw;
      const uint32_t m = S.shot[Sh + s];
      if constexpr (BIN) {
        if (m & 1) {
          const uint64_t z = *w64(V.Z, q, w, rs, V.il);
          if (z) { plane(V, s, 1)[w] ^= z; bytes += 24; }
        }
      } else {
        const uint32_t kk = (D - m) % D;
        if (kk) {
          const uint32_t z = *w32(V.Z, q, w, rs, V.il);
          if (z) {
            uint32_t* ph = (uint32_t*)phase_vec(V, s) + w;
            *ph = subm<D>(*ph, mult<D>(S.mtab[kk], z));
            bytes += 12;
          }
        }
      }
    }
  }
  __syncthreads();
  return true;
}

// One item on one member of a gang.  Returns false when the gang was aborted.
template <int D, int NT>
SDIM_DEV bool gang_item(const ItemDesc& it, const KParams& P, GangCtx& Gc, uint8_t* lds, int* s_ok, uint32_t& mk) {
  constexpr uint32_t CPWb = (D == 2) ? 8 : 4;
  const uint32_t n = it.n, rs = it.rs, Sh = it.S;
  const uint32_t NWall = rs / CPWb;
  const uint32_t cwn = gang_words(NWall, Gc.K, CPWb);
  const uint32_t w0 = (Gc.me * cwn < NWall) ? Gc.me * cwn : NWall;
  const uint32_t w1 = (w0 + cwn < NWall) ? w0 + cwn : NWall;
  uint8_t* tab = P.arena + it.tab_off;
  Tab Tg;
  Tg.X = tab; Tg.Z = tab + (size_t)rows_padded(n) * rs; Tg.Ph = tab + 2 * (size_t)rows_padded(n) * rs;
  Tg.n = n; Tg.rs = rs; Tg.S = Sh; Tg.wofs = 0; Tg.rs_full = rs; Tg.nw = NWall; Tg.il = it.il;
  Tab V = Tg;
  // column bytes [c, c + k) of every row start at byte c << il of an X/Z block (c a
  // multiple of 16); the phase vectors are row-major
  V.X += ((size_t)w0 * CPWb) << Tg.il; V.Z += ((size_t)w0 * CPWb) << Tg.il; V.Ph += (size_t)w0 * CPWb;
  V.rs_full = cwn * CPWb; V.nw = w1 - w0;
  Scr S = carve<NT>(lds, n, cwn * CPWb, Sh);
  S.err = P.counters + SDIM_ERR_SLOT;
  S.item = it.id;
  if (P.list_cap) { S.tr_cap = min(S.tr_cap, P.list_cap); S.fl_cap = min(S.fl_cap, P.list_cap); }
  S.lds_end = lds + P.lds_bytes;
  S.tiling = false;
  unsigned long long bytes = 0, lbytes = 0;
  if (threadIdx.x < 16) S.mtab[threadIdx.x] = make_tab<D>(threadIdx.x < D ? threadIdx.x : 0);
  if (threadIdx.x < 32) S.misc[threadIdx.x] = ((threadIdx.x & 15) == 0) ? INT_MAX : 0;
  if (!(it.flags & ITEM_INIT_GIVEN)) {
    gang_init_identity<D, NT>(V, w0 * CPWb, (w1 - w0) * CPWb);
    if (threadIdx.x == 0) bytes += (uint64_t)(2 * n + (D == 2 ? 2 : 1) * Sh) * (w1 - w0) * CPWb;
  } else {
    __syncthreads();
  }
  const DevOp* ops = P.ops + it.ops_off;
  uint32_t k = 0;
  bool ok = true;
  SDIM_TMARK(tItem);
  while (k < it.n_ops) {
    const DevOp o = ops[k];
    const uint32_t kind = op_kind(o);
    if (kind == OP_SEG) {
      SDIM_TMARK(tSg);
      k = run_segment<D, NT>(ops, k, V, S, it, P, bytes, lbytes);
      SDIM_TADD(P, 9, tSg);
    } else if (kind == OP_MEAS || kind == OP_RESET) {
      ok = gang_measure<D, NT>(Tg, V, w0, o.q, o.t, o.src, kind == OP_RESET, mk++, S, it, P, Gc, s_ok, bytes);
      if (!ok) break;
      k++;
    } else {
      break;
    }
  }
  SDIM_TADD(P, 11, tItem);
  const unsigned long long tot = block_sum64<NT>(bytes + lbytes, S.red);
  if (threadIdx.x == 0) atomicAdd(&P.counters[0], tot);
  return ok;
}

// Persistent gang kernel for dimension D: gang g = blockIdx.x / K pulls items from P.queue.
template <int D, int NT>
__global__ __launch_bounds__(NT) void gang_kernel(KParams P, GangParams G) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ int s_ok;
  __shared__ uint32_t s_item;
  const uint32_t g = blockIdx.x / G.K;
  GangCtx Gc;
  Gc.sy = G.sync + g;
  uint8_t* base = G.scr + (size_t)g * G.scr_stride;
  Gc.Kp = gang_kp(G.K);
  Gc.sxz = base;
  Gc.ps = base + gang_align(G.n_max * (uint64_t)Gc.Kp);
  Gc.acc = (uint32_t*)(Gc.ps + gang_align(G.S_max * (uint64_t)Gc.Kp));
  Gc.crs = Gc.acc + G.K;
  Gc.tot = Gc.crs + G.K;
  Gc.stage[0] = (uint8_t*)Gc.acc + gang_align(4ull * (2 * G.K + G.S_max));
  Gc.stage[1] = Gc.stage[0] + gang_stage_bytes(G.n_max, G.S_max);
  Gc.K = G.K;
  Gc.me = blockIdx.x % G.K;
  Gc.gerr = G.err;
  Gc.timeout = G.timeout;
  Gc.target = 0;
  Gc.xcc = xcc_id();
  Gc.xc = 0; Gc.ngrp = 0; Gc.xtarget = 0;
  // members per XCC: counted before the first (flat) barrier, read after it
  if (threadIdx.x == 0) __hip_atomic_fetch_add(&Gc.sy->xcnt[Gc.xcc], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!gang_barrier(Gc, &s_ok)) return;
  {
    uint32_t xc = 0, ng = 0;
    for (uint32_t i = 0; i < 16; i++) {
      const uint32_t c = __hip_atomic_load(&Gc.sy->xcnt[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ng += c ? 1u : 0u;
      if (i == Gc.xcc) xc = c;
    }
    Gc.xc = xc;
    Gc.ngrp = ng;
  }
  uint32_t mk = 0, round = 0;
  for (;;) {
    if (Gc.me == 0 && threadIdx.x == 0)
      __hip_atomic_store(&Gc.sy->item[round & 1], atomicAdd(P.queue, 1u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!gang_barrier(Gc, &s_ok)) return;
    if (threadIdx.x == 0) s_item = __hip_atomic_load(&Gc.sy->item[round & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t idx = s_item;
    round++;
    if (idx >= P.n_items) return;
    const ItemDesc it = P.items[P.order[idx]];
    if (!gang_item<D, NT>(it, P, Gc, lds, &s_ok, mk)) return;
  }
}

static __global__ void gang_reset_kernel(GangSync* sync, uint32_t n_gangs, int* err) {
  for (uint32_t i = threadIdx.x; i < n_gangs; i += blockDim.x) {
    sync[i].ctr = 0;
    sync[i].err = 0;
    sync[i].item[0] = 0; sync[i].item[1] = 0;
    sync[i].piv[0] = INT_MAX; sync[i].piv[1] = INT_MAX;
    for (int x = 0; x < 16; x++) { sync[i].xcnt[x] = 0; sync[i].xarr[x].v = 0; }
  }
  if (threadIdx.x == 0) *err = 0;
}

template <int D>
int prepare_gang() {
  const int lds = 163584;
  if (hipFuncSetAttribute((const void*)gang_kernel<D, 256>, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void*)gang_kernel<D, 512>, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void*)gang_kernel<D, 1024>, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return -1;
  return 0;
}

// resets the gangs' barrier records, then launches the gang kernel of dimension D
template <int D>
int launch_gang_d(const KParams& P, const GangParams& G, int nt, int grid, hipStream_t st) {
  const size_t lds = P.lds_bytes;
  const uint32_t n_gangs = (uint32_t)grid / G.K;
  hipLaunchKernelGGL(gang_reset_kernel, dim3(1), dim3(256), 0, st, G.sync, n_gangs, G.err);
  switch (nt) {
    case 256: hipLaunchKernelGGL((gang_kernel<D, 256>), dim3(grid), dim3(256), lds, st, P, G); break;
    case 512: hipLaunchKernelGGL((gang_kernel<D, 512>), dim3(grid), dim3(512), lds, st, P, G); break;
    case 1024: hipLaunchKernelGGL((gang_kernel<D, 1024>), dim3(grid), dim3(1024), lds, st, P, G); break;
    default: return -2;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template int prepare_gang<SDIM_GANG_D>();
template int launch_gang_d<SDIM_GANG_D>(const KParams& P, const GangParams& G, int nt, int grid, hipStream_t st);

}  // namespace sdim
