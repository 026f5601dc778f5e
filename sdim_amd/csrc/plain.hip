// The plain path (one kernel launch per primitive op), the reset kernel of the
// batch launches, and the host dispatchers over the per-size batch objects and
// per-dimension gang objects.
#include "tableau_dev.h"

namespace sdim {

// ------------------------------------------------------------------ plain path: one launch per op
// Every item of one circuit (its shot groups) is processed by the same launch;
// the tableau stays in HBM.  Gates update the phases immediately.
template <int D>
__global__ __launch_bounds__(256) void plain_init_kernel(KParams P) {
  const ItemDesc it = P.items[blockIdx.x];
  uint8_t* tab = P.arena + it.tab_off;
  Tab T;
  T.X = tab; T.Z = tab + (size_t)rows_padded(it.n) * it.rs; T.Ph = tab + 2 * (size_t)rows_padded(it.n) * it.rs;
  T.n = it.n; T.rs = it.rs; T.S = it.S; T.wofs = 0; T.rs_full = it.rs; T.nw = it.rs / (D == 2 ? 8 : 4); T.il = it.il;
  const uint64_t xz = 2ull * rows_padded(it.n) * it.rs;
  init_identity<D, 256>(T, xz, tableau_bytes(it.n, D, it.S) - xz);
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
  T.X = tab; T.Z = tab + (size_t)rows_padded(n) * rs; T.Ph = tab + 2 * (size_t)rows_padded(n) * rs; T.n = n; T.rs = rs; T.S = Sh;
  T.wofs = 0; T.rs_full = rs; T.nw = rs / (D == 2 ? 8 : 4); T.il = it.il;
  Scr S;
  S.mtab = mt;
  S.tiling = false;
  S.lds_end = nullptr;
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
        const uint64_t x = *w64(T.X, q, w, rs, T.il), z = *w64(T.Z, q, w, rs, T.il);
        plane(T, s, 1)[w] ^= ((b & 1) ? x : 0ull) ^ ((a & 1) ? z : 0ull);
      } else {
        const uint32_t x = *w32(T.X, q, w, rs, T.il), z = *w32(T.Z, q, w, rs, T.il);
        uint32_t* ph = (uint32_t*)phase_vec(T, s) + w;
        *ph = addm<D>(*ph, subm<D>(mult<D>(mt[b], x), mult<D>(mt[a], z)));
      }
    }
    return;
  }
  const bool two = two_rows(o);
  if constexpr (D == 2) {
    uint64_t d0 = 0, d1 = 0;
    const uint64_t xq = *w64(T.X, o.q, w, rs, T.il), zq = *w64(T.Z, o.q, w, rs, T.il);
    const uint64_t xt = two ? *w64(T.X, o.t, w, rs, T.il) : 0ull, zt = two ? *w64(T.Z, o.t, w, rs, T.il) : 0ull;
    apply_bin(o, xq, zq, xt, zt, T, w, d0, d1, S);
    if (d0 | d1) {
      for (uint32_t s = 0; s < Sh; s++) {
        uint64_t* p0 = plane(T, s, 0) + w;
        uint64_t* p1 = plane(T, s, 1) + w;
        const uint64_t a0 = *p0, a1 = *p1;
        *p1 = a1 ^ d1 ^ (a0 & d0);
        *p0 = a0 ^ d0;
      }
    }
  } else {
    uint32_t dl = 0;
    const uint32_t xq = *w32(T.X, o.q, w, rs, T.il), zq = *w32(T.Z, o.q, w, rs, T.il);
    const uint32_t xt = two ? *w32(T.X, o.t, w, rs, T.il) : 0u, zt = two ? *w32(T.Z, o.t, w, rs, T.il) : 0u;
    apply_odd<D>(o, xq, zq, xt, zt, T, w, dl, S);
    if (dl) {
      for (uint32_t s = 0; s < Sh; s++) {
        uint32_t* ph = (uint32_t*)phase_vec(T, s) + w;
        *ph = addm<D>(*ph, dl);
      }
    }
  }
}

// snapshot `o.t` of every item of the launch (one workgroup per item)
template <int D>
__global__ __launch_bounds__(256) void plain_snap_kernel(KParams P, DevOp o) {
  const ItemDesc it = P.items[blockIdx.x];
  uint8_t* tab = P.arena + it.tab_off;
  Tab T;
  T.X = tab; T.Z = tab + (size_t)rows_padded(it.n) * it.rs; T.Ph = tab + 2 * (size_t)rows_padded(it.n) * it.rs;
  T.n = it.n; T.rs = it.rs; T.S = it.S; T.wofs = 0; T.rs_full = it.rs; T.nw = it.rs / (D == 2 ? 8 : 4); T.il = it.il;
  snapshot<D, 256>(T, it, P, o.t);
}

template <int D, int NT>
__global__ __launch_bounds__(NT) void plain_meas_kernel(KParams P, DevOp o) {
  extern __shared__ __align__(16) uint8_t lds[];
  const ItemDesc it = P.items[blockIdx.x];
  const uint32_t n = it.n, rs = it.rs, Sh = it.S;
  uint8_t* tab = P.arena + it.tab_off;
  Tab T;
  T.X = tab; T.Z = tab + (size_t)rows_padded(n) * rs; T.Ph = tab + 2 * (size_t)rows_padded(n) * rs; T.n = n; T.rs = rs; T.S = Sh;
  T.wofs = 0; T.rs_full = rs; T.nw = rs / (D == 2 ? 8 : 4); T.il = it.il;
  Scr S = carve<NT>(lds, n, rs, Sh);
  S.err = P.counters + SDIM_ERR_SLOT;
  S.item = it.id;
  if (P.list_cap) { S.tr_cap = min(S.tr_cap, P.list_cap); S.fl_cap = min(S.fl_cap, P.list_cap); }
  if (threadIdx.x < 16) S.mtab[threadIdx.x] = make_tab<D>(threadIdx.x < D ? threadIdx.x : 0);
  if (threadIdx.x < 32) S.misc[threadIdx.x] = ((threadIdx.x & 15) == 0) ? INT_MAX : 0;
  __syncthreads();
  unsigned long long bytes = 0;
  measure<D, NT>(T, o.q, o.t, o.src, op_kind(o) == OP_RESET, 0, S, it, P, bytes);
}

__global__ void reset_kernel(uint32_t* queue, unsigned long long* counters) {
  if (threadIdx.x < SDIM_MAX_QUEUES) queue[threadIdx.x] = 0;
  if (threadIdx.x < SDIM_NUM_COUNTERS) counters[threadIdx.x] = 0;
}

int prepare_kernels() {
  const int lds = 163584;
  if (prepare_batch<64>() || prepare_batch<128>() || prepare_batch<256>() || prepare_batch<512>() ||
      prepare_batch<1024>())
    return -1;
  if (prepare_gang<2>() || prepare_gang<3>() || prepare_gang<5>() || prepare_gang<7>() || prepare_gang<11>()) return -1;
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

int launch_batch(const KParams& P, int nt, bool resident, int rb, int opb, int grid, hipStream_t st) {
  switch (nt) {
    case 64: return launch_batch_nt<64>(P, resident, rb, opb, grid, st);
    case 128: return launch_batch_nt<128>(P, resident, rb, opb, grid, st);
    case 256: return launch_batch_nt<256>(P, resident, rb, opb, grid, st);
    case 512: return launch_batch_nt<512>(P, resident, rb, opb, grid, st);
    case 1024: return launch_batch_nt<1024>(P, resident, rb, opb, grid, st);
  }
  return -2;
}

// every item of a gang launch has dimension d
int launch_gang(const KParams& P, const GangParams& G, int d, int nt, int grid, hipStream_t st) {
  if (nt != 256 && nt != 512 && nt != 1024) return -2;
  switch (d) {
    case 2: return launch_gang_d<2>(P, G, nt, grid, st);
    case 3: return launch_gang_d<3>(P, G, nt, grid, st);
    case 5: return launch_gang_d<5>(P, G, nt, grid, st);
    case 7: return launch_gang_d<7>(P, G, nt, grid, st);
    case 11: return launch_gang_d<11>(P, G, nt, grid, st);
  }
  return -2;
}

template <int D>
static int plain_op_d(const KParams& P, const DevOp& o, uint32_t n, uint32_t n_items, uint32_t S, hipStream_t st) {
  const uint32_t kind = op_kind(o);
  if (kind == OP_MEAS || kind == OP_RESET) {
    const uint32_t lds = scratch_bytes(n, row_stride(n, D), S, 256);
    hipLaunchKernelGGL((plain_meas_kernel<D, 256>), dim3(n_items), dim3(256), lds, st, P, o);
  } else if (kind == OP_END) {
    hipLaunchKernelGGL(plain_init_kernel<D>, dim3(n_items), dim3(256), 0, st, P);
  } else if (kind == OP_SNAP) {
    hipLaunchKernelGGL(plain_snap_kernel<D>, dim3(n_items), dim3(256), 0, st, P, o);
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
