// Host runtime of the engine: C ABI (include/sdim_hip.h), circuit lowering,
// item planning, device memory and result conversion.
//
// Lowering restates the reference's gate decompositions for prime d
// (sdim/tableau/tableau_gates.py) in closed form: X/X_INV/Z/Z_INV become one
// Pauli conjugation (their H/P words of :27-137 net to a phase-only update),
// CZ/CZ_INV = H^-1(t) SUM(+-1)(c,t) H(t) (:229-261) become one two-row op
// (z[t] +-= x[c], z[c] +-= x[t], phase +-= po x[c] x[t]), SWAP's nine-primitive
// prime word (:319-328) becomes a row exchange (it changes no phase), M_X =
// H^-1 then M (:280-296), RESET = M then
// X^((-m) mod d) (sdim/program.py:334-339), N1 = X^a Z^b drawn with the
// channel semantics of Program._build_ir (sdim/program.py:484-510).
//
// Gate runs between measurements become segments; each segment is layered
// ASAP so that the gates of a layer act on disjoint qudits (gates on disjoint
// qudits commute exactly, phases included, because their phase updates add).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "../../include/sdim_hip.h"
#include "engine.h"
#include "sdim_types.h"

using namespace sdim;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// C++ exceptions do not cross the C ABI: the entry points that allocate host memory end in this
// handler (a function-try-block), so running out of host memory returns SDIM_ENOMEM and any other
// exception SDIM_EINTERNAL, with the message in sdim_last_error().
#define SDIM_CATCH                                                                                 \
  catch (const std::bad_alloc&) {                                                                  \
    return fail(SDIM_ENOMEM, "host memory exhausted");                                             \
  }                                                                                                \
  catch (const std::exception& ex_) {                                                              \
    return fail(SDIM_EINTERNAL, std::string("internal error: ") + ex_.what());                     \
  }

// f() with an exception turned into an error code, for threads of the library's own (the planner's
// workers, the serving loop), from which an exception must not escape either.
template <class F>
int no_throw(F&& f) {
  try {
    return f();
  }
  SDIM_CATCH
}

#define HIP_OK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return fail(SDIM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

bool supported_dim(int d) { return d == 2 || d == 3 || d == 5 || d == 7 || d == 11; }

uint64_t keep_threshold53(double prob) {
  // identity iff (u >> 11) < keep53, u the Philox word; matches `u < 1 - prob` (program.py:500-502)
  const double x = (1.0 - prob) * 9007199254740992.0;
  if (!(x > 0.0)) return 0;
  if (x >= 9007199254740992.0) return 1ull << 53;
  return (uint64_t)std::ceil(x);
}

struct Prim {
  uint32_t kind, q, t, a, b, src, aux;  // aux: keep index (noise) / measurement slot
};

struct JobPlan {
  int32_t n = 0, d = 0;
  int64_t shots = 0;
  int64_t shot_base = 0;     // shot index of the job's first shot (frame sampling starts after the reference shot)
  int64_t n_meas = 0;
  int64_t n_snap = 0;        // tableau snapshots kept per item (events or instructions)
  std::vector<int32_t> meas_q, meas_round;
  std::vector<uint8_t> meas_kind;   // frame tickets: gate id of each event (14 M, 15 M_X, 16 RESET)
  uint32_t S_item = 1;
  bool resident = false;
  uint32_t first_item = 0, n_items = 0;
  uint32_t ops_off = 0, n_ops = 0;
  uint64_t keep_off = 0;
  std::vector<Prim> prims;   // kept for the plain path
};

struct Alloc {
  void* p = nullptr;
  size_t bytes = 0;
};

// Launch classes of the persistent kernel.  Resident classes keep the whole
// tableau in LDS; the others keep it in HBM and use LDS for scratch only.
// An item goes to the first class whose LDS budget holds what it needs.
struct LaunchClass {
  bool resident;
  int nt;          // threads per workgroup
  uint32_t lds;    // dynamic LDS bytes per workgroup
  int per_cu;      // workgroups resident per CU
  uint32_t gang = 0;  // > 0: items processed by gangs of this many workgroups (gang_kernel)
  int d = 0;          // gang classes: the dimension of their items
  int rb = 1;         // HBM-resident classes: sweep rows in flight per thread
  int opb = 2;        //                       gate batches per op load
  uint32_t max_n = 0;  // > 0: only items of at most max_n qudits and max_s shots
  uint32_t max_s = 0;
};
// default class table; the small HBM class is configurable (options "hbm_threads", "hbm_per_cu"),
// and the HBM classes' sweep depth and op batching (options "sweep_rows", "seg_batches")
std::vector<LaunchClass> make_classes(int hbm_nt, int hbm_per_cu, int rb, int opb) {
  auto budget = [](int per_cu) { return (uint32_t)((163840 / per_cu - 256) & ~15); };
  return {
      // small tableaux: one-wave workgroups whose serial gate layers keep the rows in registers
      // (the 64-thread kernel is built for two waves per SIMD, hence eight per CU)
      {true, 64, budget(8), 8, 0, 0, 1, 2, SDIM_TINY_ROWS, SDIM_TINY_SHOTS},
      {true, 128, budget(8), 8},
      {true, 256, budget(4), 4},
      {true, 512, budget(2), 2},
      {true, 1024, 163584, 1},
      {false, hbm_nt, budget(hbm_per_cu), hbm_per_cu, 0, 0, rb, opb},
      {false, 256, budget(4), 4, 0, 0, rb, opb},
      {false, 1024, 163584, 1, 0, 0, rb, opb},
  };
}

struct Group {
  int branch = 0;      // which parallel branch launches it (batch groups)
  uint64_t cost = 0;   // estimated work (LPT cost sum)
  int cls = 0;
  int d = 0;           // gang groups: the dimension of their items
  uint32_t first = 0, count = 0;
  GangParams gp{};     // gang classes: barrier records, scratch, error word
  int gang_grid = 0;   // gang classes: workgroups launched (gangs x members)
};

struct Ticket {
  uint64_t id = 0;
  int mode = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipStream_t> branches;                // parallel branches of the batch launches
  std::vector<hipEvent_t> fork_join;                // [0] fork, [1..] joins
  Alloc arena, meta;
  uint64_t rec_bytes = 0;
  std::vector<JobPlan> jobs;
  std::vector<ItemDesc> items;
  std::vector<uint8_t> h_rec;
  std::vector<uint8_t> h_meta;                      // staging kept alive until the ticket finishes
  Alloc gang;                                       // gang barrier records and scratch
  std::vector<int> h_gang_err;                      // one error word per gang group
  std::vector<std::vector<uint8_t>> h_inits;
  unsigned long long* d_counters = nullptr;
  unsigned long long h_counters[SDIM_NUM_COUNTERS] = {0};
  bool finished = false;
  int err = 0;                                      // sticky: nonzero once the ticket's results are invalid
  std::string err_msg;
  int64_t n_lds_items = 0;
  // launch plan (re-enqueued by sdim_replay)
  KParams P{};
  uint32_t* d_queue = nullptr;
  std::vector<Group> groups;
  std::vector<LaunchClass> classes;
  int n_cu = 256, grid_override = 0, threads = 0;
  hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
  // frame-sampler tickets
  FrameParams FP{};
  uint32_t frame_lds = 0;
  // frame tickets that simulate their own reference shots: the tableau ticket of those shots
  // (shot 0 of every job, run concurrently on another stream); frame records then hold the
  // frames' x components only and are combined with the reference records when fetched
  std::unique_ptr<Ticket> child;
};

// ticket mode of frame-sampler submissions (sdim_frame_submit)
constexpr int MODE_FRAME = 2;

}  // namespace

struct sdim_ctx {
  int device = 0;
  int n_cu = 0;
  hipStream_t stream = nullptr;               // streams[0]
  std::vector<hipStream_t> streams;           // all default priority
  std::mutex mu;
  uint64_t next_ticket = 1;
  std::map<uint64_t, std::unique_ptr<Ticket>> tickets;
  std::multimap<size_t, void*> free_blocks;   // caching allocator
  // options
  int64_t threads = 0;       // 0: per-class workgroup size
  int64_t lds_bytes = 163584;
  int64_t shots_per_item = 0;
  int64_t grid = 0;
  int64_t allow_lds = 1;
  int64_t hbm_threads = 256;
  int64_t hbm_per_cu = 4;     // LDS budget of the HBM class = 160 KB / 4, 256 threads (chosen by the autotuner:
                              // four items per CU in flight, room left for LDS-class workgroups)
  int64_t serve_hbm_threads = 512;  // the same two for serving submissions: fewer, larger workgroups
  int64_t serve_hbm_per_cu = 2;     // halve p50/p99 request latency there
  int64_t gang_min_n = 1536;  // circuits with at least this many qudits run on gangs (0: never)
  int64_t gang_chunk = 128;   // minimum bytes of each member's share of a row (one cache line)
  int64_t gang_threads = 512;
  int64_t gang_max = 0;       // most members per gang (0: every CU)
  int64_t frame_lds_bytes = 65536;  // frame sampler: LDS for the frames of one workgroup (else device memory)
  int64_t gang_timeout = (int64_t)GANG_TIMEOUT_TICKS;  // barrier wait bound of a gang member (100 MHz ticks)
  int64_t list_cap = 0;       // nonzero: caps the kernels' LDS row / group lists (tests the overflow check)
  int64_t snapshots = 0;      // tableau snapshots: 1 at every measurement event, 2 after every instruction
  int64_t sweep_rows = 1;     // HBM-resident items: rows in flight per thread in the rowsum sweep (1, 2, 4)
  int64_t seg_batches = 2;    // HBM-resident items: gate batches whose ops are loaded together (1, 2)
  int64_t interleave = 3;     // one-workgroup batched items: X/Z rows in blocks of 2^interleave (0..3)
  int64_t serial_max = 4;     // runs of >= 2 layers of at most this many gates become one serial layer (0: off)
  hipEvent_t gang_done = nullptr; // completion of the last gang launch on any stream
  std::vector<std::vector<hipStream_t>> branch_streams;  // per submission stream: extra streams for parallel launches
};

namespace {

uint32_t nt_of(const sdim_ctx* c, const LaunchClass& L) { return c->threads ? (uint32_t)c->threads : (uint32_t)L.nt; }

size_t round_block(size_t b) {
  size_t r = 1 << 20;
  while (r < b) r <<= 1;
  return r;
}

int ctx_alloc(sdim_ctx* c, size_t bytes, Alloc& a) {
  const size_t rb = round_block(bytes ? bytes : 1);
  auto it = c->free_blocks.find(rb);
  if (it != c->free_blocks.end()) {
    a.p = it->second;
    a.bytes = rb;
    c->free_blocks.erase(it);
    return 0;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, rb);
  if (e != hipSuccess) {
    // drop the cache and retry once
    for (auto& kv : c->free_blocks) (void)hipFree(kv.second);
    c->free_blocks.clear();
    e = hipMalloc(&p, rb);
    if (e != hipSuccess) return fail(SDIM_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  a.p = p;
  a.bytes = rb;
  return 0;
}

void ctx_free(sdim_ctx* c, Alloc& a) {
  if (a.p) c->free_blocks.emplace(a.bytes, a.p);
  a.p = nullptr;
  a.bytes = 0;
}

// ---------------------------------------------------------------- lowering
int expand(const sdim_job& j, JobPlan& jp, std::vector<uint64_t>& keep, bool snap_every_op = false) {
  const int64_t n = j.n;
  const uint32_t d = (uint32_t)j.d;
  std::vector<int32_t> counts((size_t)n, 0);
  jp.prims.clear();
  jp.meas_q.clear();
  jp.meas_round.clear();
  auto push = [&](uint32_t kind, uint32_t q, uint32_t t, uint32_t a, uint32_t b, uint32_t src, uint32_t aux) {
    jp.prims.push_back(Prim{kind, q, t, a, b, src, aux});
  };
  for (int64_t k = 0; k < j.n_ops; k++) {
    const int64_t gid = j.ir[3 * k], q = j.ir[3 * k + 1], t = j.ir[3 * k + 2];
    if (gid < 0 || gid > 17) return fail(SDIM_EINVAL, "Invalid gate value");
    if (q < 0 || q >= n) return fail(SDIM_EINVAL, "qudit index out of range at op " + std::to_string(k));
    const bool two = (gid >= 9 && gid <= 13);
    if (two && (t < 0 || t >= n)) return fail(SDIM_EINVAL, "target index out of range at op " + std::to_string(k));
    const uint32_t Q = (uint32_t)q, Tq = (uint32_t)(two ? t : 0), src = (uint32_t)k;
    // CZ, CZ^-1 and SWAP on one qudit: the closed forms assume two rows, so the reference's own
    // words (sdim/tableau/tableau_gates.py:241-243, 258-260, 320-328) are lowered instead; SUM
    // and SUM^-1 on one qudit already equal the reference's in-place update
    if (two && Q == Tq && gid >= 11) {
      if (gid == 13) {
        push(OP_SUM, Q, Q, 0, 0, src, 0); push(OP_H, Q, 0, 0, 0, src, 0); push(OP_H, Q, 0, 0, 0, src, 0);
        push(OP_SUM, Q, Q, 0, 0, src, 0); push(OP_H, Q, 0, 0, 0, src, 0); push(OP_H, Q, 0, 0, 0, src, 0);
        push(OP_SUM, Q, Q, 0, 0, src, 0); push(OP_H, Q, 0, 0, 0, src, 0); push(OP_H, Q, 0, 0, 0, src, 0);
      } else {
        push(OP_HI, Q, 0, 0, 0, src, 0);
        push(gid == 11 ? OP_SUM : OP_SUMI, Q, Q, 0, 0, src, 0);
        push(OP_H, Q, 0, 0, 0, src, 0);
      }
      continue;
    }
    switch (gid) {
      case 0: break;
      case 1: push(OP_PAULI, Q, 0, 1 % d, 0, src, 0); break;
      case 2: push(OP_PAULI, Q, 0, (d - 1) % d, 0, src, 0); break;
      case 3: push(OP_PAULI, Q, 0, 0, 1 % d, src, 0); break;
      case 4: push(OP_PAULI, Q, 0, 0, (d - 1) % d, src, 0); break;
      case 5: push(OP_H, Q, 0, 0, 0, src, 0); break;
      case 6: push(OP_HI, Q, 0, 0, 0, src, 0); break;
      case 7: push(OP_P, Q, 0, 0, 0, src, 0); break;
      case 8: push(OP_PI, Q, 0, 0, 0, src, 0); break;
      case 9: push(OP_SUM, Q, Tq, 0, 0, src, 0); break;
      case 10: push(OP_SUMI, Q, Tq, 0, 0, src, 0); break;
      case 11: push(OP_CZ, Q, Tq, 0, 0, src, 0); break;
      case 12: push(OP_CZI, Q, Tq, 0, 0, src, 0); break;
      case 13: push(OP_SWAP, Q, Tq, 0, 0, src, 0); break;
      case 14: case 15: case 16: {
        if (gid == 15) push(OP_HI, Q, 0, 0, 0, src, 0);
        const uint32_t slot = (uint32_t)jp.meas_q.size();
        push(gid == 16 ? OP_RESET : OP_MEAS, Q, 0, 0, 0, src, slot);
        jp.meas_q.push_back((int32_t)Q);
        jp.meas_round.push_back(counts[Q]++);
        break;
      }
      case 17: {
        if (!j.noise_channel || !j.noise_prob) return fail(SDIM_EINVAL, "N1 op without noise parameters");
        const int32_t ch = j.noise_channel[k];
        if (ch < 0 || ch > 2) return fail(SDIM_EINVAL, "Must specify a valid noise channel.");
        const double p = j.noise_prob[k];
        const uint32_t kidx = (uint32_t)keep.size();
        keep.push_back(keep_threshold53(p));
        push(OP_NOISE, Q, 0, (uint32_t)ch, 0, src, kidx);
        break;
      }
    }
    if (snap_every_op) push(OP_SNAP, 0, 0, 0, 0, src, src);   // the tableau after instruction k
  }
  jp.n_meas = (int64_t)jp.meas_q.size();
  return 0;
}

DevOp mkop(uint32_t kind, uint32_t a, uint32_t b, uint32_t q, uint32_t t, uint32_t src) {
  DevOp o;
  o.w0 = kind | (a << 8) | (b << 16);
  o.q = q;
  o.t = t;
  o.src = src;
  return o;
}

// Frame-sampler lowering (sdim/program.py:75-165): gates that change a frame, the
// reference record of every measurement event (checked against the event's qudit
// and round), and each N1's keep53 threshold inside its op.
int frame_expand(const sdim_job& j, const uint8_t* ref, JobPlan& jp, std::vector<DevOp>& ops) {
  const int64_t n = j.n;
  const uint32_t d = (uint32_t)j.d;
  std::vector<int32_t> counts((size_t)n, 0);
  jp.meas_q.clear();
  jp.meas_round.clear();
  jp.meas_kind.clear();
  for (int64_t k = 0; k < j.n_ops; k++) {
    const int64_t gid = j.ir[3 * k], q = j.ir[3 * k + 1], t = j.ir[3 * k + 2];
    if (gid < 0 || gid > 17) return fail(SDIM_EINVAL, "Invalid gate value");
    if (q < 0 || q >= n) return fail(SDIM_EINVAL, "qudit index out of range at op " + std::to_string(k));
    const bool two = (gid >= 9 && gid <= 13);
    if (two && (t < 0 || t >= n)) return fail(SDIM_EINVAL, "target index out of range at op " + std::to_string(k));
    DevOp o{(uint32_t)gid, (uint32_t)q, two ? (uint32_t)t : 0u, (uint32_t)k};
    if (gid <= 4) continue;                       // I, X, Z and inverses leave a frame unchanged
    if (gid >= 14 && gid <= 16) {
      const int64_t e = (int64_t)jp.meas_q.size();
      if (ref) {   // without reference records the kernel writes the frames' x components only
        const uint8_t* r = ref + 33 * (size_t)e;
        int64_t rq, rround, rval;
        memcpy(&rq, r, 8);
        memcpy(&rround, r + 8, 8);
        memcpy(&rval, r + 25, 8);
        if (rq != q || rround != counts[(size_t)q] || rval < 0 || rval >= (int64_t)d || r[24] > 1)
          return fail(SDIM_EINVAL, "reference record " + std::to_string(e) + " does not match measurement op " +
                                       std::to_string(k) + " (qudit " + std::to_string(q) + ", round " +
                                       std::to_string(counts[(size_t)q]) + ")");
        o.w0 |= (uint32_t)(rval | (r[24] << 7)) << 8;
      }
      o.t = (uint32_t)e;
      jp.meas_q.push_back((int32_t)q);
      jp.meas_round.push_back(counts[(size_t)q]++);
      jp.meas_kind.push_back((uint8_t)gid);
    } else if (gid == 17) {
      if (!j.noise_channel || !j.noise_prob) return fail(SDIM_EINVAL, "N1 op without noise parameters");
      const int32_t ch = j.noise_channel[k];
      if (ch < 0 || ch > 2) return fail(SDIM_EINVAL, "Must specify a valid noise channel.");
      const uint64_t keep = keep_threshold53(j.noise_prob[k]);
      o.w0 |= ((uint32_t)ch << 8) | ((uint32_t)(keep >> 32) << 10);
      o.t = (uint32_t)(keep & 0xFFFFFFFFull);
    }
    ops.push_back(o);
  }
  jp.n_meas = (int64_t)jp.meas_q.size();
  return 0;
}

// segments + ASAP layers for the batch kernel
// A run of at least two consecutive layers of at most serial_max ops each becomes one serial
// layer (their ops in layer order, which respects every dependency): the kernel applies it
// without a barrier per layer.  serial_max = 0 keeps every layer parallel.
void build_segments(const std::vector<Prim>& prims, uint32_t n, uint32_t nn_max, uint64_t keep_base,
                    uint32_t serial_max, std::vector<DevOp>& out) {
  (void)keep_base;
  std::vector<int32_t> last(n, -1);
  std::vector<uint32_t> touched;
  size_t i = 0;
  while (i < prims.size()) {
    const Prim& p0 = prims[i];
    if (p0.kind == OP_MEAS || p0.kind == OP_RESET || p0.kind == OP_SNAP) {
      out.push_back(mkop(p0.kind, 0, 0, p0.q, p0.aux, p0.src));
      i++;
      continue;
    }
    size_t j = i;
    uint32_t nn = 0;
    while (j < prims.size() && prims[j].kind != OP_MEAS && prims[j].kind != OP_RESET && prims[j].kind != OP_SNAP) {
      if (prims[j].kind == OP_NOISE) {
        if (nn == nn_max) break;
        nn++;
      }
      j++;
    }
    // layer assignment
    std::vector<std::vector<uint32_t>> layers;
    for (size_t k = i; k < j; k++) {
      const Prim& p = prims[k];
      const bool two = (p.kind == OP_SUM || p.kind == OP_SUMI || p.kind == OP_CZ || p.kind == OP_CZI ||
                         p.kind == OP_SWAP);
      int32_t l = last[p.q];
      if (two) l = std::max(l, last[p.t]);
      l += 1;
      if ((size_t)l >= layers.size()) layers.resize((size_t)l + 1);
      layers[(size_t)l].push_back((uint32_t)k);
      if (last[p.q] < 0) touched.push_back(p.q);
      last[p.q] = l;
      if (two) {
        if (last[p.t] < 0) touched.push_back(p.t);
        last[p.t] = l;
      }
    }
    for (uint32_t q : touched) last[q] = -1;
    touched.clear();
    std::vector<std::vector<uint32_t>> merged;
    std::vector<uint8_t> serial;
    for (size_t a = 0; a < layers.size();) {
      size_t b = a;
      while (serial_max && b < layers.size() && layers[b].size() <= serial_max) b++;
      if (b - a >= 2) {
        std::vector<uint32_t> run;
        for (size_t l = a; l < b; l++) run.insert(run.end(), layers[l].begin(), layers[l].end());
        merged.push_back(std::move(run));
        serial.push_back(1);
        a = b;
      } else {
        merged.push_back(std::move(layers[a]));
        serial.push_back(0);
        a++;
      }
    }
    layers.swap(merged);
    out.push_back(mkop(OP_SEG, 0, 0, (uint32_t)layers.size(), nn, 0));
    std::vector<uint32_t> slot_of(j - i, 0);
    uint32_t s = 0;
    for (size_t k = i; k < j; k++) {
      if (prims[k].kind == OP_NOISE) {
        slot_of[k - i] = s++;
        out.push_back(mkop(OP_NOISE, prims[k].a, 0, prims[k].q, prims[k].aux, prims[k].src));
      }
    }
    for (size_t li = 0; li < layers.size(); li++) {
      const auto& L = layers[li];
      out.push_back(mkop(OP_LAYER, 0, 0, (uint32_t)L.size(), serial[li], 0));
      for (uint32_t k : L) {
        const Prim& p = prims[k];
        if (p.kind == OP_NOISE) {
          DevOp o;
          o.w0 = OP_NOISE | ((slot_of[k - i] & 0xFFFFu) << 16);
          o.q = p.q;
          o.t = p.aux;
          o.src = p.src;
          out.push_back(o);
        } else {
          out.push_back(mkop(p.kind, p.a, p.b, p.q, p.t, p.src));
        }
      }
    }
    i = j;
  }
  out.push_back(mkop(OP_END, 0, 0, 0, 0, 0));
}

uint64_t align256(uint64_t v) { return (v + 255) & ~255ull; }

// ---------------------------------------------------------------- tableau upload / download
void pack_tableau(const JobPlan& jp, uint32_t S, uint32_t il, const sdim_job& j, std::vector<uint8_t>& buf) {
  const uint32_t n = (uint32_t)jp.n, d = (uint32_t)jp.d, rs = row_stride(n, d);
  buf.assign(tableau_bytes(n, d, S), 0);
  uint8_t* X = buf.data();                                     // X then Z: one interleaved block
  uint8_t* Ph = X + 2 * (size_t)rows_padded(n) * rs;            // phase vectors, row-major
  const uint32_t np = rows_padded(n);
  const int64_t *x = j.init[0], *z = j.init[1], *ph = j.init[2], *dx = j.init[3], *dz = j.init[4], *dph = j.init[5];
  const int64_t ord = (d == 2) ? 4 : d;
  auto md = [&](int64_t v, int64_t m) { int64_t r = v % m; return r < 0 ? r + m : r; };
  for (uint32_t r = 0; r < n; r++) {
    for (uint32_t g = 0; g < 2 * n; g++) {
      const int64_t xv = md(g < n ? x[(size_t)r * n + g] : dx[(size_t)r * n + g - n], d);
      const int64_t zv = md(g < n ? z[(size_t)r * n + g] : dz[(size_t)r * n + g - n], d);
      if (d == 2) {
        if (xv) *(uint64_t*)(X + cell(r, 8 * (g >> 6), rs, il)) |= 1ull << (g & 63);
        if (zv) *(uint64_t*)(X + cell(np + r, 8 * (g >> 6), rs, il)) |= 1ull << (g & 63);
      } else {
        X[cell(r, g, rs, il)] = (uint8_t)xv;
        X[cell(np + r, g, rs, il)] = (uint8_t)zv;
      }
    }
  }
  for (uint32_t s = 0; s < S; s++) {
    for (uint32_t g = 0; g < 2 * n; g++) {
      const int64_t pv = md(g < n ? ph[g] : dph[g - n], ord);
      if (d == 2) {
        if (pv & 1) ((uint64_t*)(Ph + (size_t)(2 * s) * rs))[g >> 6] |= 1ull << (g & 63);
        if (pv & 2) ((uint64_t*)(Ph + (size_t)(2 * s + 1) * rs))[g >> 6] |= 1ull << (g & 63);
      } else {
        Ph[(size_t)s * rs + g] = (uint8_t)pv;
      }
    }
  }
}

void unpack_tableau(const uint8_t* tab, uint32_t n, uint32_t d, uint32_t il, uint32_t s, int64_t* x, int64_t* z,
                    int64_t* ph, int64_t* dx, int64_t* dz, int64_t* dph) {
  const uint32_t rs = row_stride(n, d), np = rows_padded(n);
  const uint8_t* X = tab;                                      // X then Z: one interleaved block
  const uint8_t* Ph = X + 2 * (size_t)np * rs;
  for (uint32_t r = 0; r < n; r++) {
    for (uint32_t g = 0; g < 2 * n; g++) {
      int64_t xv, zv;
      if (d == 2) {
        xv = (*(const uint64_t*)(X + cell(r, 8 * (g >> 6), rs, il)) >> (g & 63)) & 1;
        zv = (*(const uint64_t*)(X + cell(np + r, 8 * (g >> 6), rs, il)) >> (g & 63)) & 1;
      } else {
        xv = X[cell(r, g, rs, il)];
        zv = X[cell(np + r, g, rs, il)];
      }
      if (g < n) { x[(size_t)r * n + g] = xv; z[(size_t)r * n + g] = zv; }
      else { dx[(size_t)r * n + g - n] = xv; dz[(size_t)r * n + g - n] = zv; }
    }
  }
  for (uint32_t g = 0; g < 2 * n; g++) {
    int64_t pv;
    if (d == 2) {
      const int64_t b0 = (((const uint64_t*)(Ph + (size_t)(2 * s) * rs))[g >> 6] >> (g & 63)) & 1;
      const int64_t b1 = (((const uint64_t*)(Ph + (size_t)(2 * s + 1) * rs))[g >> 6] >> (g & 63)) & 1;
      pv = b0 + 2 * b1;
    } else {
      pv = Ph[(size_t)s * rs + g];
    }
    if (g < n) ph[g] = pv; else dph[g - n] = pv;
  }
}

Ticket* find_ticket(sdim_ctx* c, uint64_t id) {
  auto it = c->tickets.find(id);
  return it == c->tickets.end() ? nullptr : it->second.get();
}

// Marks a ticket's results invalid; every later finish() of it reports the same error.
int poison(Ticket* t, int code, const std::string& msg) {
  t->finished = true;
  t->err = code;
  t->err_msg = msg;
  return fail(code, msg);
}

// The error word of the ticket's last run (a list in LDS would have overflowed: the kernel
// dropped the write, so the results are invalid).
int kernel_error(Ticket* t) {
  const unsigned long long e = t->d_counters ? t->h_counters[SDIM_ERR_SLOT] : 0ull;
  if (!e) return 0;
  const uint32_t code = (uint32_t)(e >> 32), item = (uint32_t)e;
  const char* what = code == KERR_TOUCHED_ROWS ? "touched-row list" : code == KERR_FACTOR_GROUPS ? "factor-group list"
                                                                                                  : "unknown list";
  return poison(t, SDIM_EHIP, std::string("kernel capacity check failed: ") + what + " overflow in item " +
                                  std::to_string(item) + "; results are invalid");
}

// Waits for a ticket's device work.  The first failure (a HIP error on its stream, a gang
// barrier that timed out, a failed child ticket) is kept in the ticket and returned by this
// and every later call, so no fetch hands out the records of a failed run.
int finish(sdim_ctx* c, Ticket* t) {
  if (t->finished) return t->err ? fail(t->err, t->err_msg) : 0;
  if (t->child) {
    const int rc = finish(c, t->child.get());
    if (rc) {
      (void)hipStreamSynchronize(t->stream);
      return poison(t, rc, g_err);
    }
  }
  const hipError_t e = hipStreamSynchronize(t->stream);
  if (e != hipSuccess) return poison(t, SDIM_EHIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  t->finished = true;
  for (int ge : t->h_gang_err)
    if (ge) return poison(t, SDIM_EHIP, "a gang barrier timed out (a member was not resident); results are invalid");
  return kernel_error(t);
}

// frees a ticket's (and its child's) device memory, graphs and events
void release_ticket(sdim_ctx* c, Ticket* t) {
  if (t->child) release_ticket(c, t->child.get());
  ctx_free(c, t->arena);
  ctx_free(c, t->meta);
  ctx_free(c, t->gang);
  for (auto& g : t->graph) if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  if (t->ev0) (void)hipEventDestroy(t->ev0);
  if (t->ev1) (void)hipEventDestroy(t->ev1);
  t->ev0 = t->ev1 = nullptr;
  for (auto e : t->fork_join) (void)hipEventDestroy(e);
  t->fork_join.clear();
}

// A submission's ticket until it is handed to the context: an early return (or an exception)
// hands its device memory, events and graphs back, once the work already enqueued on its
// streams (and its child's) has ended.
struct TicketUndo {
  sdim_ctx* c;
  Ticket* t;
  bool armed;
  ~TicketUndo() {
    if (!armed) return;
    for (Ticket* u = t; u; u = u->child.get()) {
      if (u->stream) (void)hipStreamSynchronize(u->stream);
      for (hipStream_t b : u->branches) (void)hipStreamSynchronize(b);
    }
    release_ticket(c, t);
  }
};

// releases a ticket and forgets it
void drop_ticket(sdim_ctx* c, std::map<uint64_t, std::unique_ptr<Ticket>>::iterator it) {
  release_ticket(c, it->second.get());
  c->tickets.erase(it);
}


// Enqueue the device work of a ticket on its stream.  which: 0 all items,
// 1 HBM-resident items only, 2 LDS-resident items only (batched mode).
// gang_gate: when set, every gang launch waits for the previous one (on any stream)
// and records its own completion there, so two gang kernels never share the device:
// a gang waits for all of its members, and two partially resident gangs could
// otherwise hold the CUs the other one needs.
int enqueue(Ticket* T, int which, hipEvent_t gang_gate = nullptr) {
  KParams& P = T->P;
  if (T->mode == MODE_FRAME) {
    if (which != 0) return fail(SDIM_EINVAL, "a frame-sampler ticket has no item classes");
    if (launch_frame(T->FP, (uint32_t)T->items.size(), T->frame_lds, T->stream))
      return fail(SDIM_EHIP, "frame kernel launch failed");
    return 0;
  }
  if (T->mode == SDIM_MODE_BATCHED) {
    if (launch_reset(T->d_queue, P.counters, T->stream)) return fail(SDIM_EHIP, "reset launch failed");
    // batch launches run as parallel branches (fork/join through events: parallel nodes
    // when captured into a graph) so that the tail of one launch overlaps the others;
    // gang launches follow the join, one at a time
    const int nb = (int)T->branches.size();
    std::vector<bool> used(nb + 1, false);
    if (nb && hipEventRecord(T->fork_join[0], T->stream) != hipSuccess) return fail(SDIM_EHIP, "hipEventRecord failed");
    for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (int b = 1; b <= nb; b++) {
        if (!used[b]) continue;
        if (hipEventRecord(T->fork_join[b], T->branches[b - 1]) != hipSuccess ||
            hipStreamWaitEvent(T->stream, T->fork_join[b], 0) != hipSuccess)
          return fail(SDIM_EHIP, "branch join failed");
      }
    }
    for (const Group& g : T->groups) {
      const LaunchClass& L = T->classes[g.cls];
      if (which == 1 && L.resident) continue;
      if (which == 2 && !L.resident) continue;
      if ((pass == 0) == (L.gang != 0)) continue;
      KParams Q = P;
      Q.order = P.order + g.first;
      Q.n_items = g.count;
      Q.queue = T->d_queue + (&g - T->groups.data());
      // lds_bytes caps the one-workgroup classes; a gang member takes a CU's whole LDS
      Q.lds_bytes = L.gang ? L.lds : std::min<uint32_t>(L.lds, P.lds_bytes);
      if (L.gang) {
        if (gang_gate && hipStreamWaitEvent(T->stream, gang_gate, 0) != hipSuccess)
          return fail(SDIM_EHIP, "hipStreamWaitEvent failed");
        if (launch_gang(Q, g.gp, L.d, L.nt, g.gang_grid, T->stream)) return fail(SDIM_EHIP, "gang kernel launch failed");
        if (gang_gate && hipEventRecord(gang_gate, T->stream) != hipSuccess) return fail(SDIM_EHIP, "hipEventRecord failed");
        continue;
      }
      const int nt = T->threads ? T->threads : L.nt;
      const int per_cu = std::max(1, std::min(L.per_cu, 1024 / nt));
      const int grid = T->grid_override ? T->grid_override
                                        : (int)std::min<uint64_t>((uint64_t)T->n_cu * per_cu, g.count);
      const int b = std::min(g.branch, nb);
      hipStream_t st = b ? T->branches[b - 1] : T->stream;
      if (b && !used[b]) {
        if (hipStreamWaitEvent(st, T->fork_join[0], 0) != hipSuccess) return fail(SDIM_EHIP, "branch fork failed");
        used[b] = true;
      }
      if (launch_batch(Q, nt, L.resident, L.rb, L.opb, grid, st)) return fail(SDIM_EHIP, "batch kernel launch failed");
    }
    }
    return 0;
  }
  if (which != 0) return fail(SDIM_EINVAL, "the plain path has no item classes");
  for (auto& jp : T->jobs) {
    KParams Q = P;
    Q.items = P.items + jp.first_item;
    Q.n_items = jp.n_items;
    uint32_t Smax = 0;
    for (uint32_t i = 0; i < jp.n_items; i++) Smax = std::max(Smax, T->items[jp.first_item + i].S);
    const bool given = T->items[jp.first_item].flags & ITEM_INIT_GIVEN;
    if (!given) {
      if (launch_plain_op(Q, mkop(OP_END, 0, 0, 0, 0, 0), (uint32_t)jp.n, (uint32_t)jp.d, jp.n_items, Smax, T->stream))
        return fail(SDIM_EHIP, "plain init launch failed");
    }
    for (const Prim& p : jp.prims) {
      DevOp o;
      if (p.kind == OP_NOISE) o = mkop(OP_NOISE, p.a, 0, p.q, p.aux, p.src);
      else if (p.kind == OP_MEAS || p.kind == OP_RESET || p.kind == OP_SNAP) o = mkop(p.kind, 0, 0, p.q, p.aux, p.src);
      else o = mkop(p.kind, p.a, p.b, p.q, p.t, p.src);
      if (launch_plain_op(Q, o, (uint32_t)jp.n, (uint32_t)jp.d, jp.n_items, Smax, T->stream))
        return fail(SDIM_EHIP, "plain op launch failed");
    }
  }
  return 0;
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

const char* sdim_last_error(void) { return g_err.c_str(); }

int sdim_ctx_create(int32_t device, sdim_ctx** out) try {
  if (!out) return fail(SDIM_EINVAL, "null output pointer");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0)
    return fail(SDIM_EHIP, "no HIP device " + std::to_string(device));
  HIP_OK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, device));
  std::unique_ptr<sdim_ctx> c(new sdim_ctx());
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  for (int i = 0; i < 4; i++) {
    hipStream_t st = nullptr;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    c->streams.push_back(st);
  }
  c->stream = c->streams[0];
  c->branch_streams.resize(c->streams.size());
  for (auto& v : c->branch_streams) {
    for (int i = 0; i < 3; i++) {
      hipStream_t st = nullptr;
      HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      v.push_back(st);
    }
  }
  HIP_OK(hipEventCreateWithFlags(&c->gang_done, hipEventDisableTiming));
  if (prepare_kernels() || prepare_frame_kernel()) return fail(SDIM_EHIP, "could not set the LDS limit of the kernels");
  if (noalloc_init_device(device)) return fail(SDIM_EHIP, "could not find the allocation-free path's device scratch");
  *out = c.release();
  return 0;
} SDIM_CATCH

int sdim_ctx_destroy(sdim_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->tickets) {
    for (Ticket* t = kv.second.get(); t; t = t->child.get()) {
      (void)hipStreamSynchronize(t->stream);
      if (t->arena.p) (void)hipFree(t->arena.p);
      if (t->meta.p) (void)hipFree(t->meta.p);
      if (t->gang.p) (void)hipFree(t->gang.p);
      for (auto& g : t->graph) if (g) (void)hipGraphExecDestroy(g);
      if (t->ev0) (void)hipEventDestroy(t->ev0);
      if (t->ev1) (void)hipEventDestroy(t->ev1);
      for (auto e : t->fork_join) (void)hipEventDestroy(e);
    }
  }
  for (auto& kv : c->free_blocks) (void)hipFree(kv.second);
  for (auto st : c->streams) (void)hipStreamDestroy(st);
  for (auto& v : c->branch_streams) for (auto st : v) (void)hipStreamDestroy(st);
  if (c->gang_done) (void)hipEventDestroy(c->gang_done);
  delete c;
  return 0;
}

int sdim_set_option(sdim_ctx* c, const char* key, int64_t v) {
  if (!c || !key) return fail(SDIM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  const std::string k(key);
  if (k == "threads") {
    if (v != 0 && v != 256 && v != 512 && v != 1024) return fail(SDIM_EINVAL, "threads must be 0, 256, 512 or 1024");
    c->threads = v;
  } else if (k == "lds_bytes") {
    if (v < 16384 || v > 163584) return fail(SDIM_EINVAL, "lds_bytes out of range");
    c->lds_bytes = v & ~15ll;
  } else if (k == "shots_per_item") {
    if (v < 0) return fail(SDIM_EINVAL, "shots_per_item must be >= 0");
    c->shots_per_item = v;
  } else if (k == "grid") {
    if (v < 0) return fail(SDIM_EINVAL, "grid must be >= 0");
    c->grid = v;
  } else if (k == "allow_lds") {
    c->allow_lds = v ? 1 : 0;
  } else if (k == "hbm_threads") {
    if (v != 64 && v != 128 && v != 256 && v != 512 && v != 1024) return fail(SDIM_EINVAL, "hbm_threads must be 64..1024");
    c->hbm_threads = v;
  } else if (k == "hbm_per_cu") {
    if (v < 1 || v > 16) return fail(SDIM_EINVAL, "hbm_per_cu must be 1..16");
    c->hbm_per_cu = v;
  } else if (k == "serve_hbm_threads") {
    if (v != 64 && v != 128 && v != 256 && v != 512 && v != 1024) return fail(SDIM_EINVAL, "serve_hbm_threads must be 64..1024");
    c->serve_hbm_threads = v;
  } else if (k == "serve_hbm_per_cu") {
    if (v < 1 || v > 16) return fail(SDIM_EINVAL, "serve_hbm_per_cu must be 1..16");
    c->serve_hbm_per_cu = v;
  } else if (k == "gang_min_n") {
    if (v < 0) return fail(SDIM_EINVAL, "gang_min_n must be >= 0");
    c->gang_min_n = v;
  } else if (k == "gang_chunk") {
    if (v < 16 || v > 65536 || (v & 15)) return fail(SDIM_EINVAL, "gang_chunk must be a multiple of 16 in [16, 65536]");
    c->gang_chunk = v;
  } else if (k == "gang_threads") {
    if (v != 256 && v != 512 && v != 1024) return fail(SDIM_EINVAL, "gang_threads must be 256, 512 or 1024");
    c->gang_threads = v;
  } else if (k == "gang_max") {
    if (v < 0) return fail(SDIM_EINVAL, "gang_max must be >= 0");
    c->gang_max = v;
  } else if (k == "frame_lds_bytes") {
    if (v < 0 || v > 163840) return fail(SDIM_EINVAL, "frame_lds_bytes must be in [0, 163840]");
    c->frame_lds_bytes = v;
  } else if (k == "gang_timeout") {
    if (v < 1) return fail(SDIM_EINVAL, "gang_timeout must be >= 1 tick");
    c->gang_timeout = v;
  } else if (k == "snapshots") {
    if (v < 0 || v > 2) return fail(SDIM_EINVAL, "snapshots must be 0, 1 or 2");
    c->snapshots = v;
  } else if (k == "sweep_rows") {
    if (v != 1 && v != 2 && v != 4) return fail(SDIM_EINVAL, "sweep_rows must be 1, 2 or 4");
    c->sweep_rows = v;
  } else if (k == "seg_batches") {
    if (v != 1 && v != 2) return fail(SDIM_EINVAL, "seg_batches must be 1 or 2");
    c->seg_batches = v;
  } else if (k == "interleave") {
    if (v < 0 || v > 3) return fail(SDIM_EINVAL, "interleave must be in [0, 3]");
    c->interleave = v;
  } else if (k == "serial_max") {
    if (v < 0 || v > 65535) return fail(SDIM_EINVAL, "serial_max must be in [0, 65535]");
    c->serial_max = v;
  } else if (k == "list_cap") {
    if (v < 0 || v > 65535) return fail(SDIM_EINVAL, "list_cap must be in [0, 65535]");
    c->list_cap = v;
  } else {
    return fail(SDIM_EINVAL, "unknown option " + k);
  }
  return 0;
}

int sdim_get_option(sdim_ctx* c, const char* key, int64_t* v) {
  if (!c || !key || !v) return fail(SDIM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  const std::string k(key);
  if (k == "threads") *v = c->threads;
  else if (k == "lds_bytes") *v = c->lds_bytes;
  else if (k == "shots_per_item") *v = c->shots_per_item;
  else if (k == "grid") *v = c->grid ? c->grid : c->n_cu;
  else if (k == "allow_lds") *v = c->allow_lds;
  else if (k == "hbm_threads") *v = c->hbm_threads;
  else if (k == "hbm_per_cu") *v = c->hbm_per_cu;
  else if (k == "serve_hbm_per_cu") *v = c->serve_hbm_per_cu;
  else if (k == "serve_hbm_threads") *v = c->serve_hbm_threads;
  else if (k == "num_cu") *v = c->n_cu;
  else if (k == "gang_min_n") *v = c->gang_min_n;
  else if (k == "gang_chunk") *v = c->gang_chunk;
  else if (k == "gang_threads") *v = c->gang_threads;
  else if (k == "gang_max") *v = c->gang_max;
  else if (k == "frame_lds_bytes") *v = c->frame_lds_bytes;
  else if (k == "gang_timeout") *v = c->gang_timeout;
  else if (k == "list_cap") *v = c->list_cap;
  else if (k == "snapshots") *v = c->snapshots;
  else if (k == "sweep_rows") *v = c->sweep_rows;
  else if (k == "seg_batches") *v = c->seg_batches;
  else if (k == "serial_max") *v = c->serial_max;
  else if (k == "interleave") *v = c->interleave;
  else return fail(SDIM_EINVAL, "unknown option " + k);
  return 0;
}

}  // extern "C"

namespace {

// What the planner decides beyond the ticket's items and groups: the host staging of the
// device metadata (item descriptors, launch order, lowered ops, noise thresholds, then
// room for the launch queues and counters) and the device layout of a submission.
struct Staging {
  std::vector<uint64_t> keep;
  std::vector<DevOp> ops;
  std::vector<uint32_t> order;
  std::vector<std::pair<uint32_t, std::vector<uint8_t>>> inits;   // item -> initial tableau bytes
  uint64_t arena_bytes = 0;   // records, then tableaux
  uint64_t gang_bytes = 0;    // gang barrier records and scratch (offsets stored in the groups)
  uint64_t o_items = 0, o_order = 0, o_ops = 0, o_keep = 0, o_queue = 0, o_cnt = 0, meta_bytes = 0;
  int used_branches = 0;      // parallel branch streams the batch groups were spread over
};

struct PlanOpts {
  bool one_item_per_job = false;   // every job is one item holding all its shots (else split by shots_per_item)
  bool no_gangs = false;           // never place a circuit on a gang
  bool branches = true;            // spread batch groups over the stream's parallel branches
};

// Lowers and places a submission: fills T's jobs, items, classes and groups and the staging
// (no device memory is touched).  Caller holds c->mu.
int plan_locked(sdim_ctx* c, const sdim_job* jobs, int32_t n_jobs, int32_t mode, int stream_idx,
                int64_t hbm_threads, int64_t hbm_per_cu, const PlanOpts& opts, Ticket* T, Staging& st) {
  if (!c || (!jobs && n_jobs) || n_jobs < 0) return fail(SDIM_EINVAL, "bad submit arguments");
  if (mode != SDIM_MODE_BATCHED && mode != SDIM_MODE_PLAIN) return fail(SDIM_EINVAL, "unknown mode");
  T->mode = mode;
  T->stream = c->streams[(size_t)stream_idx % c->streams.size()];
  T->jobs.resize((size_t)n_jobs);
  const uint32_t cap = (uint32_t)c->lds_bytes;
  T->classes = make_classes((int)(hbm_threads ? hbm_threads : c->hbm_threads),
                            (int)(hbm_per_cu ? hbm_per_cu : c->hbm_per_cu), (int)c->sweep_rows, (int)c->seg_batches);
  const std::vector<LaunchClass>& kClasses = T->classes;
  const int kNumClasses = (int)kClasses.size();
  uint64_t total_shots = 0;
  for (int32_t ji = 0; ji < n_jobs; ji++) total_shots += (uint64_t)std::max<int64_t>(jobs[ji].shots, 0);
  // automatic shots per item: enough items to keep every workgroup slot busy twice over
  const uint64_t slots = (uint64_t)c->n_cu * 4;
  const uint32_t auto_S = (uint32_t)std::min<uint64_t>(4096, std::max<uint64_t>(1, total_shots / (2 * slots)));

  std::vector<uint64_t>& keep = st.keep;
  std::vector<DevOp>& ops = st.ops;
  std::vector<ItemDesc>& items = T->items;
  std::vector<uint64_t> cost;
  std::vector<int> cls_of;
  auto& inits = st.inits;

  // Lowering, class choice and layering of each job depend on that job alone: large
  // submissions run them on worker threads; the results are merged in job order.
  struct JobWork {
    int rc = 0;
    std::string err;
    std::vector<uint64_t> keep;
    std::vector<DevOp> ops;
    std::vector<uint8_t> init_buf;
    LaunchClass gL{false, 0, 0, 1, 0, 0};
    int cls = -1;
    uint32_t S = 1, rs = 0, il = 0, gang_cb = 0;
    bool resident = false, ph_lds = false, given = false;
  };
  std::vector<JobWork> work((size_t)std::max(n_jobs, 0));
  auto plan_job = [&](int32_t ji) -> int {
    const sdim_job& j = jobs[ji];
    JobPlan& jp = T->jobs[(size_t)ji];
    JobWork& w = work[(size_t)ji];

    if (j.n < 1) return fail(SDIM_EINVAL, "Number of qudits must be greater than 0");
    if (j.d < 2) return fail(SDIM_EINVAL, "Dimension must be greater than 1");
    if (!supported_dim(j.d)) return fail(SDIM_EINVAL, "dimension " + std::to_string(j.d) + " is not supported (prime d in {2,3,5,7,11})");
    if (j.shots < 1) return fail(SDIM_EINVAL, "shots must be >= 1");
    if (j.n > 32768) return fail(SDIM_EINVAL, "at most 32768 qudits are supported");
    if (j.n_ops < 0 || (j.n_ops && !j.ir)) return fail(SDIM_EINVAL, "bad op list");
    jp.n = j.n;
    jp.d = j.d;
    jp.shots = j.shots;
    int rc = expand(j, jp, w.keep, c->snapshots == 2);
    if (rc) return rc;
    jp.n_snap = c->snapshots == 1 ? jp.n_meas : c->snapshots == 2 ? j.n_ops : 0;

    const uint32_t n = (uint32_t)j.n, d = (uint32_t)j.d, rs = row_stride(n, d);
    const bool has_noise = !w.keep.empty();
    // every item repeats the circuit's X/Z evolution, so the automatic split keeps at least
    // min(64, n / 16) shots per item: below that a large tableau's X/Z work swamps the phase
    // work it parallelises (n = 1012, d = 11, 2436 shots alone: 256 ms at one shot per item,
    // 18 ms at 64)
    const uint32_t floor_S = std::min<uint32_t>(64u, std::max<uint32_t>(1u, (uint32_t)j.n / 16u));
    uint32_t S = (uint32_t)std::min<int64_t>(j.shots, c->shots_per_item ? (int64_t)c->shots_per_item
                                                                        : (int64_t)std::max(auto_S, floor_S));
    if (opts.one_item_per_job) {
      if (j.shots > 4096) return fail(SDIM_EINVAL, "at most 4096 shots per job on this path (one item per job)");
      S = (uint32_t)j.shots;
    }
    int cls = -1;
    uint32_t gang_cb = 0;     // gang items: LDS width of a member's columns
    // snapshots are taken by one workgroup per item: no gangs then
    if (mode == SDIM_MODE_BATCHED && !opts.no_gangs && !c->snapshots && c->gang_min_n > 0 && j.n >= c->gang_min_n) {
      // a gang of K workgroups, each owning >= gang_chunk bytes of every row; gang
      // workgroups take the whole LDS of a CU, so one is resident per CU
      const int gnt = (int)c->gang_threads;
      uint64_t kmax = (uint64_t)c->n_cu;
      if (c->gang_max) kmax = std::min<uint64_t>(kmax, (uint64_t)c->gang_max);
      uint32_t K = 1;
      while ((uint64_t)K * 2 <= kmax && (uint64_t)K * 2 * (uint64_t)c->gang_chunk <= rs) K *= 2;
      if (K >= 2) {
        const uint32_t bpw = d == 2 ? 8 : 4;
        const uint32_t cb = gang_words(rs / bpw, K, bpw) * bpw;
        const uint32_t lds = 163584;      // not capped by lds_bytes (the one-workgroup classes' cut-over)
        uint32_t Sg = (uint32_t)std::min<int64_t>(j.shots, c->shots_per_item ? c->shots_per_item : 4096);
        auto gneed = [&](uint32_t s) -> uint64_t {
          return (uint64_t)scratch_bytes(n, cb, s, (uint32_t)gnt) + (has_noise ? noise_bytes(1, cb, s) : 0);
        };
        while (Sg > 1 && gneed(Sg) > lds) Sg /= 2;
        if (gneed(Sg) <= lds) {
          S = Sg;
          gang_cb = cb;
          w.gL = LaunchClass{false, gnt, lds, 1, K, (int)d};   // found or added when the jobs are merged
        }
      }
    }
    if (mode == SDIM_MODE_BATCHED && cls < 0 && !gang_cb) {
      auto need = [&](int k, uint32_t s) -> uint64_t {
        const LaunchClass& L = kClasses[k];
        return (L.resident ? ((tableau_bytes(n, d, s) + 15) & ~15ull) : 0) + scratch_bytes(n, rs, s, nt_of(c, L)) +
               (has_noise ? noise_bytes(1, rs, s) : 0);
      };
      for (int k = 0; k < kNumClasses && cls < 0; k++) {
        const LaunchClass& L = kClasses[k];
        if (L.resident && !c->allow_lds) continue;
        if (L.max_n && (n > L.max_n || S > L.max_s)) continue;
        if (need(k, S) <= std::min<uint32_t>(L.lds, cap)) cls = k;
      }
      if (cls < 0) {
        // the last class keeps the tableau in HBM with the largest scratch: shrink S until it fits
        const int k = kNumClasses - 1;
        if (opts.one_item_per_job && need(k, S) > std::min<uint32_t>(kClasses[k].lds, cap))
          return fail(SDIM_ENOMEM, "job " + std::to_string(ji) + ": its " + std::to_string(S) +
                                       " shots do not fit one item's LDS scratch on this path");
        while (S > 1 && need(k, S) > std::min<uint32_t>(kClasses[k].lds, cap)) S /= 2;
        if (need(k, S) > std::min<uint32_t>(kClasses[k].lds, cap))
          return fail(SDIM_ENOMEM, "circuit too large for the LDS scratch budget");
        cls = k;
      }
    } else {
      if (scratch_bytes(n, rs, S, 256) > 163584) return fail(SDIM_ENOMEM, "circuit too large for the plain path");
    }
    const bool resident = cls >= 0 && kClasses[cls].resident;
    bool ph_lds = false;
    uint32_t nn_max = 0;
    for (;;) {
      // one-workgroup HBM items keep the shots' phase vectors in LDS when they take at most
      // a quarter of the class budget (the rest stays for scratch, noise tables and tiles)
      ph_lds = false;
      if (mode == SDIM_MODE_BATCHED && !resident && !gang_cb) {
        const uint64_t ph = tableau_bytes(n, d, S) - 2ull * rows_padded(n) * rs;
        const uint64_t budget = std::min<uint32_t>(kClasses[cls].lds, cap);
        ph_lds = ph <= budget / 4 &&
                 ((ph + 15) & ~15ull) + scratch_bytes(n, rs, S, nt_of(c, kClasses[cls])) + (has_noise ? noise_bytes(1, rs, S) : 0) <= budget;
      }
      if (mode != SDIM_MODE_BATCHED) break;
      // noise ops per segment that fit next to the tableau / scratch
      const LaunchClass& L = gang_cb ? w.gL : kClasses[cls];
      const uint32_t cw = gang_cb ? gang_cb : rs;    // width of the per-column LDS arrays
      const uint32_t nt = gang_cb ? (uint32_t)L.nt : nt_of(c, L);
      const uint64_t base = (resident ? ((tableau_bytes(n, d, S) + 15) & ~15ull) : 0) + scratch_bytes(n, cw, S, nt) +
                            (ph_lds ? ((tableau_bytes(n, d, S) - 2ull * rows_padded(n) * rs + 15) & ~15ull) : 0);
      const uint64_t budget = std::min<uint32_t>(L.lds, cap);
      // HBM-resident one-workgroup items keep half of the remaining LDS for the column tiles of gate segments
      const uint64_t room = budget > base ? ((resident || gang_cb) ? budget - base : (budget - base) / 2) : 0;
      nn_max = 0;
      while (nn_max < 65535 && noise_bytes(nn_max + 1, cw, S) <= room) nn_max++;
      if (!has_noise || nn_max > 0) break;
      // an automatic shot count halves until a noise table fits (the noise tables grow with it)
      if (S == 1 || opts.one_item_per_job || c->shots_per_item) return fail(SDIM_ENOMEM, "no LDS room for noise tables");
      S /= 2;
    }
    jp.S_item = S;
    jp.resident = resident;
    if (mode == SDIM_MODE_BATCHED) {
      build_segments(jp.prims, n, std::max<uint32_t>(nn_max, 1), 0, (uint32_t)c->serial_max, w.ops);
      jp.n_ops = (uint32_t)w.ops.size();
    }
    const bool given = j.init[0] && j.init[1] && j.init[2] && j.init[3] && j.init[4] && j.init[5];
    // 8-row interleave for tableaux swept by one workgroup; gang members and the plain path
    // apply gates straight from HBM and keep rows contiguous
    const uint32_t il = (mode == SDIM_MODE_BATCHED && !gang_cb) ? (uint32_t)c->interleave : 0u;
    if (given) pack_tableau(jp, S, il, j, w.init_buf);
    w.cls = cls; w.S = S; w.rs = rs; w.il = il; w.gang_cb = gang_cb;
    w.resident = resident; w.ph_lds = ph_lds; w.given = given;
    return 0;
  };
  {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int32_t nth = n_jobs >= 64 ? (int32_t)std::min<unsigned>(16u, hw) : 1;
    if (nth <= 1) {
      for (int32_t ji = 0; ji < n_jobs; ji++) {
        work[(size_t)ji].rc = plan_job(ji);
        if (work[(size_t)ji].rc) { work[(size_t)ji].err = g_err; break; }
      }
    } else {
      std::atomic<int32_t> next{0};
      auto worker = [&]() {
        for (int32_t ji; (ji = next.fetch_add(1)) < n_jobs;) {
          work[(size_t)ji].rc = no_throw([&] { return plan_job(ji); });
          if (work[(size_t)ji].rc) work[(size_t)ji].err = g_err;
        }
      };
      std::vector<std::thread> pool;
      pool.reserve((size_t)nth);
      try {
        for (int32_t t = 1; t < nth; t++) pool.emplace_back(worker);
      } catch (const std::system_error&) {
        // fewer threads than asked for: the ones started and this thread share the jobs
      }
      worker();
      for (auto& th : pool) th.join();
    }
  }
  uint64_t rec_total = 0;
  for (int32_t ji = 0; ji < n_jobs; ji++) {
    const sdim_job& j = jobs[ji];
    JobPlan& jp = T->jobs[(size_t)ji];
    JobWork& w = work[(size_t)ji];
    if (w.rc) return fail(w.rc, w.err);
    int cls = w.cls;
    if (w.gang_cb) {
      for (int k = 0; k < (int)T->classes.size() && cls < 0; k++)
        if (T->classes[k].gang == w.gL.gang && T->classes[k].nt == w.gL.nt && T->classes[k].d == w.gL.d) cls = k;
      if (cls < 0) {
        if (T->classes.size() >= SDIM_MAX_CLASSES) return fail(SDIM_EINVAL, "too many launch classes");
        T->classes.push_back(w.gL);
        cls = (int)T->classes.size() - 1;
      }
    }
    jp.keep_off = keep.size();
    keep.insert(keep.end(), w.keep.begin(), w.keep.end());
    if (mode == SDIM_MODE_BATCHED) {
      jp.ops_off = (uint32_t)ops.size();
      ops.insert(ops.end(), w.ops.begin(), w.ops.end());
    }
    const uint32_t n = (uint32_t)j.n, d = (uint32_t)j.d, rs = w.rs, S = w.S, il = w.il;
    const bool resident = w.resident, ph_lds = w.ph_lds, given = w.given;
    // items
    jp.first_item = (uint32_t)items.size();
    for (int64_t s0 = 0; s0 < j.shots; s0 += S) {
      ItemDesc it;
      memset(&it, 0, sizeof(it));
      it.seed = j.seed;
      it.stream = j.stream;
      it.keep_off = jp.keep_off;
      it.ops_off = jp.ops_off;
      it.n_ops = jp.n_ops;
      it.n = n;
      it.d = d;
      it.S = (uint32_t)std::min<int64_t>(S, j.shots - s0);
      it.shot0 = (uint32_t)s0;
      it.rs = rs;
      it.flags = (resident ? ITEM_LDS : 0) | (given ? ITEM_INIT_GIVEN : 0) | (ph_lds ? ITEM_PH_LDS : 0) |
                 (c->snapshots == 1 ? ITEM_SNAP_MEAS : 0);
      it.n_meas = (uint32_t)jp.n_meas;
      it.il = il;
      it.id = (uint32_t)items.size();
      it.rec_off = rec_total;
      rec_total = align256(rec_total + (uint64_t)jp.n_meas * it.S);
      if (given) inits.emplace_back((uint32_t)items.size(), w.init_buf);
      items.push_back(it);
      cls_of.push_back(cls);
      // LPT cost estimate: measurement sweeps (rows x row bytes) + per-shot phase passes + gates
      cost.push_back((uint64_t)jp.n_meas * n * rs + (uint64_t)it.S * rs * (jp.n_meas + 4) + jp.prims.size() * 4ull * rs);
    }
    jp.n_items = (uint32_t)items.size() - jp.first_item;
  }
  // arena: records first, then tableaux
  T->rec_bytes = rec_total;
  uint64_t off = align256(rec_total);
  for (auto& it : items) {
    it.tab_off = off;
    off = align256(off + tableau_bytes(it.n, it.d, it.S));
  }
  // tableau snapshots after the final tableaux
  for (const JobPlan& jp : T->jobs) {
    for (uint32_t i = 0; i < jp.n_items; i++) {
      ItemDesc& it = items[jp.first_item + i];
      if (!jp.n_snap) continue;
      it.snap_off = off;
      off += (uint64_t)jp.n_snap * snapshot_stride(it.n, it.d, it.S);
    }
  }
  if (off > (64ull << 30)) return fail(SDIM_ENOMEM, "tableau snapshots would need more than 64 GiB");
  st.arena_bytes = std::max<uint64_t>(off, 256);
  // metadata: items, order, ops, keep, queue, counters.  Order: by launch class, longest first inside a class.
  std::vector<uint32_t>& order = st.order;
  order.resize(items.size());
  for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
  // one group (kernel launch) per launch class, longest items first
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (cls_of[a] != cls_of[b]) return cls_of[a] < cls_of[b];
    return cost[a] > cost[b];
  });
  T->groups.clear();
  if (mode == SDIM_MODE_BATCHED) {
    for (uint32_t i = 0; i < order.size();) {
      const int k = cls_of[order[i]];
      uint32_t e = i;
      while (e < order.size() && cls_of[order[e]] == k) e++;
      Group g;
      g.cls = k;
      g.d = T->classes[k].d;
      g.first = i;
      g.count = e - i;
      T->groups.push_back(g);
      i = e;
    }
  }
  // batch groups over parallel branches: longest first onto the least loaded branch
  {
    const int n_br = opts.branches ? 1 + (int)c->branch_streams[(size_t)stream_idx % c->branch_streams.size()].size() : 1;
    std::vector<size_t> gi;
    for (size_t i = 0; i < T->groups.size(); i++) {
      Group& g = T->groups[i];
      for (uint32_t k = g.first; k < g.first + g.count; k++) g.cost += cost[order[k]];
      if (!T->classes[g.cls].gang) gi.push_back(i);
    }
    std::stable_sort(gi.begin(), gi.end(), [&](size_t a, size_t b) { return T->groups[a].cost > T->groups[b].cost; });
    std::vector<uint64_t> load((size_t)n_br, 0);
    int used_br = 0;
    for (size_t i : gi) {
      const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[(size_t)b] += T->groups[i].cost;
      T->groups[i].branch = b;
      used_br = std::max(used_br, b);
    }
    const auto& bs = c->branch_streams[(size_t)stream_idx % c->branch_streams.size()];
    T->branches.assign(bs.begin(), bs.begin() + used_br);
    st.used_branches = used_br;
  }
  // gang groups: as many gangs as fit on the device at once (and no more than items)
  uint64_t& gang_bytes = st.gang_bytes;
  for (Group& g : T->groups) {
    const LaunchClass& L = T->classes[g.cls];
    if (!L.gang) continue;
    uint32_t n_max = 1, S_max = 1;
    for (uint32_t i = g.first; i < g.first + g.count; i++) {
      n_max = std::max(n_max, items[order[i]].n);
      S_max = std::max(S_max, items[order[i]].S);
    }
    const uint32_t n_gangs = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->n_cu / L.gang, g.count));
    g.gang_grid = (int)(n_gangs * L.gang);
    g.gp.K = L.gang;
    g.gp.n_max = n_max;
    g.gp.S_max = S_max;
    g.gp.scr_stride = gang_scratch_bytes(L.gang, n_max, S_max);
    g.gp.timeout = (uint64_t)c->gang_timeout;
    // layout inside the ticket's gang buffer: [error word][sync records][scratch]
    g.gp.err = (int*)(uintptr_t)gang_bytes;
    gang_bytes += 256;
    g.gp.sync = (GangSync*)(uintptr_t)gang_bytes;
    gang_bytes += GANG_SYNC_BYTES * n_gangs;
    g.gp.scr = (uint8_t*)(uintptr_t)gang_bytes;
    gang_bytes += g.gp.scr_stride * n_gangs;
  }
  st.o_items = 0;
  st.o_order = align256(st.o_items + items.size() * sizeof(ItemDesc));
  st.o_ops = align256(st.o_order + order.size() * 4);
  st.o_keep = align256(st.o_ops + ops.size() * sizeof(DevOp));
  st.o_queue = align256(st.o_keep + keep.size() * 8);
  if (T->groups.size() > SDIM_MAX_QUEUES) return fail(SDIM_EINVAL, "too many launch groups");
  st.o_cnt = align256(st.o_queue + 4 * SDIM_MAX_QUEUES);
  st.meta_bytes = align256(st.o_cnt + 8 * SDIM_NUM_COUNTERS);
  for (auto& it : items) if (it.flags & ITEM_LDS) T->n_lds_items++;
  T->n_cu = c->n_cu;
  T->grid_override = (int)c->grid;
  T->threads = (int)c->threads;
  return 0;
}

// host staging of the device metadata (queue and counters are cleared by the reset kernel)
void stage_meta(const Ticket* T, const Staging& st, std::vector<uint8_t>& h_meta) {
  h_meta.assign(st.meta_bytes, 0);
  if (!T->items.empty()) memcpy(h_meta.data() + st.o_items, T->items.data(), T->items.size() * sizeof(ItemDesc));
  if (!st.order.empty()) memcpy(h_meta.data() + st.o_order, st.order.data(), st.order.size() * 4);
  if (!st.ops.empty()) memcpy(h_meta.data() + st.o_ops, st.ops.data(), st.ops.size() * sizeof(DevOp));
  if (!st.keep.empty()) memcpy(h_meta.data() + st.o_keep, st.keep.data(), st.keep.size() * 8);
}

// kernel parameters of a planned submission whose metadata sits at `meta` and arena at `arena`
void bind_params(sdim_ctx* c, Ticket* T, const Staging& st, uint8_t* meta, uint8_t* arena) {
  KParams& P = T->P;
  P.items = (const ItemDesc*)(meta + st.o_items);
  P.order = (const uint32_t*)(meta + st.o_order);
  P.n_items = (uint32_t)T->items.size();
  P.queue = (uint32_t*)(meta + st.o_queue);
  P.ops = (const DevOp*)(meta + st.o_ops);
  P.keep53 = (const uint64_t*)(meta + st.o_keep);
  P.arena = arena;
  P.counters = (unsigned long long*)(meta + st.o_cnt);
  P.lds_bytes = (uint32_t)c->lds_bytes;
  P.list_cap = (uint32_t)c->list_cap;
  T->d_counters = P.counters;
  T->d_queue = P.queue;
}

}  // namespace

// Lowers, plans, uploads and enqueues a submission on stream `stream_idx` (caller holds c->mu).
static int submit_locked(sdim_ctx* c, const sdim_job* jobs, int32_t n_jobs, int32_t mode, int stream_idx,
                         uint64_t* ticket, uint64_t* device_bytes, int64_t hbm_threads = 0, int64_t hbm_per_cu = 0) {
  if (!c || !ticket) return fail(SDIM_EINVAL, "bad submit arguments");
  HIP_OK(hipSetDevice(c->device));
  std::unique_ptr<Ticket> T(new Ticket());
  TicketUndo undo{c, T.get(), true};
  Staging st;
  int rc = plan_locked(c, jobs, n_jobs, mode, stream_idx, hbm_threads, hbm_per_cu, PlanOpts(), T.get(), st);
  if (rc) return rc;
  for (int k = 0; k <= st.used_branches && st.used_branches > 0; k++) {
    hipEvent_t e = nullptr;
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    T->fork_join.push_back(e);
  }
  const std::vector<ItemDesc>& items = T->items;
  const uint64_t arena_bytes = st.arena_bytes, gang_bytes = st.gang_bytes, meta_bytes = st.meta_bytes;
  rc = ctx_alloc(c, arena_bytes, T->arena);
  if (rc) return rc;
  rc = ctx_alloc(c, meta_bytes, T->meta);
  if (rc) { ctx_free(c, T->arena); return rc; }
  if (gang_bytes) {
    rc = ctx_alloc(c, gang_bytes, T->gang);
    if (rc) { ctx_free(c, T->arena); ctx_free(c, T->meta); return rc; }
    uint8_t* gb = (uint8_t*)T->gang.p;
    for (Group& g : T->groups) {
      if (!T->classes[g.cls].gang) continue;
      g.gp.err = (int*)(gb + (uintptr_t)g.gp.err);
      g.gp.sync = (GangSync*)(gb + (uintptr_t)g.gp.sync);
      g.gp.scr = gb + (uintptr_t)g.gp.scr;
    }
  }
  uint8_t* meta = (uint8_t*)T->meta.p;
  uint8_t* arena = (uint8_t*)T->arena.p;
  stage_meta(T.get(), st, T->h_meta);
  HIP_OK(hipMemcpyAsync(meta, T->h_meta.data(), meta_bytes, hipMemcpyHostToDevice, T->stream));
  for (auto& kv : st.inits) {
    const ItemDesc& it = items[kv.first];
    T->h_inits.push_back(std::move(kv.second));
    HIP_OK(hipMemcpyAsync(arena + it.tab_off, T->h_inits.back().data(), T->h_inits.back().size(),
                          hipMemcpyHostToDevice, T->stream));
  }
  bind_params(c, T.get(), st, meta, arena);
  HIP_OK(hipEventCreate(&T->ev0));
  HIP_OK(hipEventCreate(&T->ev1));

  HIP_OK(hipEventRecord(T->ev0, T->stream));
  rc = enqueue(T.get(), 0, c->gang_done);
  if (rc) return rc;
  HIP_OK(hipEventRecord(T->ev1, T->stream));
  // records back to the host (async; valid after wait)
  T->h_rec.resize(T->rec_bytes);
  if (T->rec_bytes) HIP_OK(hipMemcpyAsync(T->h_rec.data(), arena, T->rec_bytes, hipMemcpyDeviceToHost, T->stream));
  HIP_OK(hipMemcpyAsync(T->h_counters, T->d_counters, sizeof(T->h_counters), hipMemcpyDeviceToHost, T->stream));
  T->h_gang_err.assign(T->groups.size(), 0);
  for (size_t gi = 0; gi < T->groups.size(); gi++)
    if (T->classes[T->groups[gi].cls].gang)
      HIP_OK(hipMemcpyAsync(&T->h_gang_err[gi], T->groups[gi].gp.err, sizeof(int), hipMemcpyDeviceToHost, T->stream));

  T->id = c->next_ticket++;
  *ticket = T->id;
  if (device_bytes) *device_bytes = T->arena.bytes + T->meta.bytes + T->gang.bytes;
  c->tickets[T->id] = std::move(T);
  undo.armed = false;
  return 0;
}

// The allocation-free variant of a batched submission (sdim_simulate_noalloc, noalloc.hip):
// one item per job, no gangs, launches in sequence on the context's first stream; the
// metadata goes to `scratch` (a device buffer of `scratch_cap` bytes owned by the
// library) and the arena is the caller's `out`.  Waits for the work before returning.
int simulate_noalloc_locked(sdim_ctx* c, const sdim_job* jobs, int32_t n_jobs, void* out, uint64_t out_capacity,
                            uint64_t* out_bytes, uint8_t* scratch, uint64_t scratch_cap) {
  if (!c || !scratch) return fail(SDIM_EINVAL, "no scratch buffer for this device");
  HIP_OK(hipSetDevice(c->device));
  Ticket T;
  Staging st;
  PlanOpts opts;
  opts.one_item_per_job = true;
  opts.no_gangs = true;
  opts.branches = false;
  int rc = plan_locked(c, jobs, n_jobs, SDIM_MODE_BATCHED, 0, 0, 0, opts, &T, st);
  if (rc) return rc;
  if (out_bytes) *out_bytes = st.arena_bytes;
  if (!out) return 0;   // size query
  if (out_capacity < st.arena_bytes)
    return fail(SDIM_EINVAL, "output buffer holds " + std::to_string(out_capacity) + " bytes; the jobs need " +
                                 std::to_string(st.arena_bytes));
  if (st.meta_bytes > scratch_cap)
    return fail(SDIM_ENOMEM, "the jobs' device metadata needs " + std::to_string(st.meta_bytes) +
                                 " bytes; the allocation-free path has " + std::to_string(scratch_cap));
  std::vector<uint8_t> h_meta;
  stage_meta(&T, st, h_meta);
  HIP_OK(hipMemcpyAsync(scratch, h_meta.data(), st.meta_bytes, hipMemcpyHostToDevice, T.stream));
  for (auto& kv : st.inits)
    HIP_OK(hipMemcpyAsync((uint8_t*)out + T.items[kv.first].tab_off, kv.second.data(), kv.second.size(),
                          hipMemcpyHostToDevice, T.stream));
  bind_params(c, &T, st, scratch, (uint8_t*)out);
  rc = enqueue(&T, 0, nullptr);   // the reset kernel clears the queues and counters first
  if (rc) {
    (void)hipStreamSynchronize(T.stream);
    return rc;
  }
  HIP_OK(hipStreamSynchronize(T.stream));
  unsigned long long e = 0;
  HIP_OK(hipMemcpy(&e, T.P.counters + SDIM_ERR_SLOT, sizeof(e), hipMemcpyDeviceToHost));
  T.d_counters = T.P.counters;
  T.h_counters[SDIM_ERR_SLOT] = e;
  return kernel_error(&T);
}

extern "C" {

int sdim_submit(sdim_ctx* c, const sdim_job* jobs, int32_t n_jobs, int32_t mode, uint64_t* ticket) try {
  if (!c) return fail(SDIM_EINVAL, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  return submit_locked(c, jobs, n_jobs, mode, 0, ticket, nullptr);
} SDIM_CATCH

int sdim_simulate_noalloc(sdim_ctx* c, const sdim_job* jobs, int32_t n_jobs, void* out, uint64_t out_capacity,
                          uint64_t* out_bytes) try {
  if (!c) return fail(SDIM_EINVAL, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  return simulate_noalloc_locked(c, jobs, n_jobs, out, out_capacity, out_bytes, noalloc_scratch(c->device),
                                 SDIM_NOALLOC_SCRATCH_BYTES);
} SDIM_CATCH

int sdim_frame_submit(sdim_ctx* c, const sdim_job* jobs, int32_t n_jobs, const void* const* refs, int64_t shot0,
                      uint64_t* ticket) try {
  if (!c) return fail(SDIM_EINVAL, "null context");
  if ((!jobs && n_jobs) || n_jobs < 0 || !ticket) return fail(SDIM_EINVAL, "bad submit arguments");
  if (shot0 < 0) return fail(SDIM_EINVAL, "shot0 must be >= 0");
  const bool own_ref = refs == nullptr;
  if (own_ref && shot0 != 0) return fail(SDIM_EINVAL, "without reference records shot0 must be 0 (shot 0 is the reference shot)");
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_OK(hipSetDevice(c->device));
  std::unique_ptr<Ticket> T(new Ticket());
  TicketUndo undo{c, T.get(), true};
  T->mode = MODE_FRAME;
  T->stream = c->stream;
  T->jobs.resize((size_t)n_jobs);
  if (own_ref) {
    // the reference shots: each job once, N1 as the identity (sdim/program.py:31, 244-247),
    // as a tableau submission on another stream, overlapping the frame launch
    for (int32_t ji = 0; ji < n_jobs; ji++) {
      const sdim_job& j = jobs[ji];
      if (j.n_ops < 0 || (j.n_ops && !j.ir)) return fail(SDIM_EINVAL, "bad op list");
      if (j.init[0] || j.init[1] || j.init[2] || j.init[3] || j.init[4] || j.init[5])
        return fail(SDIM_EINVAL, "the frame sampler starts from |0...0> (its frames start as random Z)");
    }
    std::vector<std::vector<int64_t>> irs((size_t)n_jobs);
    std::vector<sdim_job> rj((size_t)n_jobs);
    for (int32_t ji = 0; ji < n_jobs; ji++) {
      const sdim_job& j = jobs[ji];
      irs[(size_t)ji].assign(j.ir, j.ir + 3 * j.n_ops);
      for (int64_t k = 0; k < j.n_ops; k++) if (irs[(size_t)ji][3 * k] == 17) irs[(size_t)ji][3 * k] = 0;
      sdim_job r = j;
      r.ir = irs[(size_t)ji].data();
      r.noise_channel = nullptr;
      r.noise_prob = nullptr;
      r.shots = 1;
      rj[(size_t)ji] = r;
    }
    uint64_t rid = 0;
    int rc = submit_locked(c, rj.data(), n_jobs, SDIM_MODE_BATCHED, 1, &rid, nullptr);
    if (rc) return rc;
    auto it = c->tickets.find(rid);
    T->child = std::move(it->second);
    c->tickets.erase(it);
  }
  // errors from here on return through bail(): the guard above waits for the reference shots in
  // flight and frees them with the rest of the ticket
  auto bail = [](int code) { return code; };
#define FRAME_HIP_OK(call)                                                                      \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) return bail(fail(SDIM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
  } while (0)
  std::vector<DevOp> ops;
  std::vector<ItemDesc>& items = T->items;
  const uint32_t NT = FRAME_NT;
  uint64_t rec_total = 0;
  uint32_t lds = 0;
  for (int32_t ji = 0; ji < n_jobs; ji++) {
    const sdim_job& j = jobs[ji];
    JobPlan& jp = T->jobs[(size_t)ji];
    if (j.n < 1) return bail(fail(SDIM_EINVAL, "Number of qudits must be greater than 0"));
    if (j.d < 2) return bail(fail(SDIM_EINVAL, "Dimension must be greater than 1"));
    if (!supported_dim(j.d)) return bail(fail(SDIM_EINVAL, "dimension " + std::to_string(j.d) + " is not supported (prime d in {2,3,5,7,11})"));
    if (j.shots < 1) return bail(fail(SDIM_EINVAL, "shots must be >= 1"));
    if (shot0 + j.shots > 0xFFFFFFFFll) return bail(fail(SDIM_EINVAL, "shot indices must stay below 2^32"));
    if (j.n > 32768) return bail(fail(SDIM_EINVAL, "at most 32768 qudits are supported"));
    if (j.n_ops < 0 || (j.n_ops && !j.ir)) return bail(fail(SDIM_EINVAL, "bad op list"));
    if (j.init[0] || j.init[1] || j.init[2] || j.init[3] || j.init[4] || j.init[5])
      return bail(fail(SDIM_EINVAL, "the frame sampler starts from |0...0> (its frames start as random Z)"));
    jp.n = j.n;
    jp.d = j.d;
    jp.shots = j.shots;
    jp.shot_base = shot0;
    jp.ops_off = (uint32_t)ops.size();
    const uint8_t* ref = own_ref ? nullptr : (const uint8_t*)refs[ji];
    // events are counted before the records are read
    int64_t n_ev = 0;
    for (int64_t k = 0; k < j.n_ops; k++) n_ev += (j.ir[3 * k] >= 14 && j.ir[3 * k] <= 16);
    if (n_ev && !own_ref && !ref) return bail(fail(SDIM_EINVAL, "missing reference records"));
    int rc = frame_expand(j, ref, jp, ops);
    if (rc) return bail(rc);
    jp.n_ops = (uint32_t)ops.size() - jp.ops_off;
    const uint32_t n = (uint32_t)j.n;
    const bool in_lds = (uint64_t)n * NT <= (uint64_t)c->frame_lds_bytes;
    if (in_lds) lds = std::max<uint32_t>(lds, (n * NT + 15u) & ~15u);
    jp.first_item = (uint32_t)items.size();
    // frame shots: all of the job's shots, or all but the reference shot 0
    const int64_t f0 = own_ref ? 1 : 0, nfs = j.shots - f0;
    for (int64_t s0 = 0; s0 < nfs; s0 += NT) {
      ItemDesc it;
      memset(&it, 0, sizeof(it));
      it.seed = j.seed;
      it.stream = j.stream;
      it.ops_off = jp.ops_off;
      it.n_ops = jp.n_ops;
      it.n = n;
      it.d = (uint32_t)j.d;
      it.S = (uint32_t)std::min<int64_t>(NT, nfs - s0);
      it.shot0 = (uint32_t)(shot0 + f0 + s0);
      it.flags = in_lds ? ITEM_LDS : 0;
      it.n_meas = (uint32_t)jp.n_meas;
      it.id = (uint32_t)items.size();
      it.rec_off = rec_total;
      rec_total = align256(rec_total + (uint64_t)jp.n_meas * it.S);
      items.push_back(it);
    }
    jp.n_items = (uint32_t)items.size() - jp.first_item;
  }
  // arena: records, then the device-memory frames of items that do not use LDS
  T->rec_bytes = rec_total;
  uint64_t off = align256(rec_total);
  for (auto& it : items) {
    if (it.flags & ITEM_LDS) continue;
    it.tab_off = off;
    off = align256(off + (uint64_t)it.n * NT);
  }
  const uint64_t o_items = 0;
  const uint64_t o_ops = align256(o_items + items.size() * sizeof(ItemDesc));
  const uint64_t meta_bytes = std::max<uint64_t>(align256(o_ops + ops.size() * sizeof(DevOp)), 256);
  int rc = ctx_alloc(c, std::max<uint64_t>(off, 256), T->arena);
  if (rc) return bail(rc);
  rc = ctx_alloc(c, meta_bytes, T->meta);
  if (rc) { ctx_free(c, T->arena); return bail(rc); }
  uint8_t* meta = (uint8_t*)T->meta.p;
  T->h_meta.assign(meta_bytes, 0);
  if (!items.empty()) memcpy(T->h_meta.data() + o_items, items.data(), items.size() * sizeof(ItemDesc));
  if (!ops.empty()) memcpy(T->h_meta.data() + o_ops, ops.data(), ops.size() * sizeof(DevOp));
  FRAME_HIP_OK(hipMemcpyAsync(meta, T->h_meta.data(), meta_bytes, hipMemcpyHostToDevice, T->stream));
  T->FP.items = (const ItemDesc*)(meta + o_items);
  T->FP.ops = (const DevOp*)(meta + o_ops);
  T->FP.arena = (uint8_t*)T->arena.p;
  T->FP.nt = NT;
  T->frame_lds = lds;
  for (auto& it : items) if (it.flags & ITEM_LDS) T->n_lds_items++;
  FRAME_HIP_OK(hipEventCreate(&T->ev0));
  FRAME_HIP_OK(hipEventCreate(&T->ev1));
  FRAME_HIP_OK(hipEventRecord(T->ev0, T->stream));
  rc = enqueue(T.get(), 0);
  if (rc) return bail(rc);
  FRAME_HIP_OK(hipEventRecord(T->ev1, T->stream));
  T->h_rec.resize(T->rec_bytes);
  if (T->rec_bytes) FRAME_HIP_OK(hipMemcpyAsync(T->h_rec.data(), T->arena.p, T->rec_bytes, hipMemcpyDeviceToHost, T->stream));
  T->id = c->next_ticket++;
  *ticket = T->id;
  c->tickets[T->id] = std::move(T);
  undo.armed = false;
  return 0;
} SDIM_CATCH
#undef FRAME_HIP_OK

int sdim_wait(sdim_ctx* c, uint64_t ticket) try {
  if (!c) return fail(SDIM_EINVAL, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  return finish(c, t);
} SDIM_CATCH

int sdim_poll(sdim_ctx* c, uint64_t ticket, int32_t* done) try {
  if (!c || !done) return fail(SDIM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  if (t->finished) { *done = 1; return t->err ? fail(t->err, t->err_msg) : 0; }
  for (Ticket* u = t; u; u = u->child.get()) {
    const hipError_t e = hipEventQuery(u->ev1);
    if (e == hipErrorNotReady) { *done = 0; return 0; }
    if (e != hipSuccess) {
      *done = 1;
      return poison(t, SDIM_EHIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
    }
  }
  *done = 1;
  return finish(c, t);
} SDIM_CATCH

int sdim_release(sdim_ctx* c, uint64_t ticket) {
  if (!c) return fail(SDIM_EINVAL, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->tickets.find(ticket);
  if (it == c->tickets.end()) return fail(SDIM_ENOTFOUND, "unknown ticket");
  int rc = finish(c, it->second.get());
  drop_ticket(c, it);
  return rc;
}

int sdim_num_measurements(sdim_ctx* c, uint64_t ticket, int32_t job, int64_t* count) {
  if (!c || !count) return fail(SDIM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  if (job < 0 || (size_t)job >= t->jobs.size()) return fail(SDIM_ENOTFOUND, "unknown job");
  *count = t->jobs[(size_t)job].n_meas;
  return 0;
}

int sdim_fetch_records(sdim_ctx* c, uint64_t ticket, int32_t job, void* out, int64_t capacity) try {
  if (!c || !out) return fail(SDIM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  if (job < 0 || (size_t)job >= t->jobs.size()) return fail(SDIM_ENOTFOUND, "unknown job");
  int rc = finish(c, t);
  if (rc) return rc;
  const JobPlan& jp = t->jobs[(size_t)job];
  if (capacity < jp.n_meas * jp.shots) return fail(SDIM_EINVAL, "record buffer too small");
  uint8_t* o = (uint8_t*)out;
  if (t->child) {
    // shot 0: the reference shot (its tableau ticket); shots 1..: frames, whose records hold
    // the x component only: M / M_X report (reference value + x) mod d, RESET the reference
    // value, every shot the reference's deterministic flag (sdim/program.py:124-161)
    const Ticket* ch = t->child.get();
    const ItemDesc& ci = ch->items[ch->jobs[(size_t)job].first_item];
    const int64_t d = jp.d;
    for (int64_t ev = 0; ev < jp.n_meas; ev++) {
      const uint8_t rb = ch->h_rec[ci.rec_off + (size_t)ev * ci.S];
      const int64_t q = jp.meas_q[(size_t)ev], round = jp.meas_round[(size_t)ev], rv = rb & 0x7F;
      const uint8_t det = (rb >> 7) & 1;
      const bool reset = jp.meas_kind[(size_t)ev] == 16;
      auto put = [&](int64_t shot, int64_t val) {
        uint8_t* rec = o + 33 * (size_t)(ev * jp.shots + shot);
        memcpy(rec, &q, 8);
        memcpy(rec + 8, &round, 8);
        memcpy(rec + 16, &shot, 8);
        rec[24] = det;
        memcpy(rec + 25, &val, 8);
      };
      put(0, rv);
      for (uint32_t i = 0; i < jp.n_items; i++) {
        const ItemDesc& it = t->items[jp.first_item + i];
        const uint8_t* r = t->h_rec.data() + it.rec_off + (size_t)ev * it.S;
        for (uint32_t s = 0; s < it.S; s++) put((int64_t)it.shot0 + s, reset ? rv : (rv + (r[s] & 0x7F)) % d);
      }
    }
    return 0;
  }
  for (int64_t ev = 0; ev < jp.n_meas; ev++) {
    for (uint32_t i = 0; i < jp.n_items; i++) {
      const ItemDesc& it = t->items[jp.first_item + i];
      const uint8_t* r = t->h_rec.data() + it.rec_off + (size_t)ev * it.S;
      for (uint32_t s = 0; s < it.S; s++) {
        const int64_t shot = (int64_t)it.shot0 + s;
        uint8_t* rec = o + 33 * (size_t)(ev * jp.shots + (shot - jp.shot_base));
        const int64_t q = jp.meas_q[(size_t)ev], round = jp.meas_round[(size_t)ev], val = r[s] & 0x7F;
        const uint8_t det = (r[s] >> 7) & 1;
        memcpy(rec, &q, 8);
        memcpy(rec + 8, &round, 8);
        memcpy(rec + 16, &shot, 8);
        rec[24] = det;
        memcpy(rec + 25, &val, 8);
      }
    }
  }
  return 0;
} SDIM_CATCH

int sdim_fetch_tableau(sdim_ctx* c, uint64_t ticket, int32_t job, int64_t shot, int64_t* x, int64_t* z,
                       int64_t* ph, int64_t* dx, int64_t* dz, int64_t* dph) try {
  if (!c || !x || !z || !ph || !dx || !dz || !dph) return fail(SDIM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  if (job < 0 || (size_t)job >= t->jobs.size()) return fail(SDIM_ENOTFOUND, "unknown job");
  int rc = finish(c, t);
  if (rc) return rc;
  if (t->mode == MODE_FRAME) {
    // only the reference shot of a frame ticket that simulated it has a tableau
    if (!t->child || shot != 0) return fail(SDIM_EINVAL, "a frame-sampled shot keeps no tableau");
    t = t->child.get();
  }
  const JobPlan& jp = t->jobs[(size_t)job];
  if (shot < 0 || shot >= jp.shots) return fail(SDIM_EINVAL, "shot out of range");
  for (uint32_t i = 0; i < jp.n_items; i++) {
    const ItemDesc& it = t->items[jp.first_item + i];
    if (shot >= it.shot0 && shot < (int64_t)it.shot0 + it.S) {
      std::vector<uint8_t> buf(tableau_bytes(it.n, it.d, it.S));
      HIP_OK(hipMemcpy(buf.data(), (uint8_t*)t->arena.p + it.tab_off, buf.size(), hipMemcpyDeviceToHost));
      unpack_tableau(buf.data(), it.n, it.d, it.il, (uint32_t)(shot - it.shot0), x, z, ph, dx, dz, dph);
      return 0;
    }
  }
  return fail(SDIM_ENOTFOUND, "shot not found");
} SDIM_CATCH

int sdim_fetch_snapshot(sdim_ctx* c, uint64_t ticket, int32_t job, int64_t index, int64_t shot, int64_t* x,
                        int64_t* z, int64_t* ph, int64_t* dx, int64_t* dz, int64_t* dph) try {
  if (!c || !x || !z || !ph || !dx || !dz || !dph) return fail(SDIM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  if (job < 0 || (size_t)job >= t->jobs.size()) return fail(SDIM_ENOTFOUND, "unknown job");
  int rc = finish(c, t);
  if (rc) return rc;
  const JobPlan& jp = t->jobs[(size_t)job];
  if (!jp.n_snap) return fail(SDIM_EINVAL, "the ticket was submitted without tableau snapshots");
  if (index < 0 || index >= jp.n_snap) return fail(SDIM_EINVAL, "snapshot index out of range");
  if (shot < 0 || shot >= jp.shots) return fail(SDIM_EINVAL, "shot out of range");
  for (uint32_t i = 0; i < jp.n_items; i++) {
    const ItemDesc& it = t->items[jp.first_item + i];
    if (shot >= it.shot0 && shot < (int64_t)it.shot0 + it.S) {
      std::vector<uint8_t> buf(tableau_bytes(it.n, it.d, it.S));
      const uint8_t* src = (const uint8_t*)t->arena.p + it.snap_off + (uint64_t)index * snapshot_stride(it.n, it.d, it.S);
      HIP_OK(hipMemcpy(buf.data(), src, buf.size(), hipMemcpyDeviceToHost));
      unpack_tableau(buf.data(), it.n, it.d, it.il, (uint32_t)(shot - it.shot0), x, z, ph, dx, dz, dph);
      return 0;
    }
  }
  return fail(SDIM_ENOTFOUND, "shot not found");
} SDIM_CATCH

int sdim_ticket_stats(sdim_ctx* c, uint64_t ticket, double* kernel_ms, uint64_t* bytes_hbm, uint64_t* bytes_lds,
                      int64_t* n_items, int64_t* n_lds_items) {
  if (!c) return fail(SDIM_EINVAL, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  int rc = finish(c, t);
  if (rc) return rc;
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
  if (kernel_ms) *kernel_ms = ms;
  if (bytes_hbm) *bytes_hbm = t->h_counters[0];
  if (bytes_lds) *bytes_lds = t->h_counters[1];
  if (n_items) *n_items = (int64_t)t->items.size();
  if (n_lds_items) *n_lds_items = t->n_lds_items;
  return 0;
}

int sdim_ticket_info(sdim_ctx* c, uint64_t ticket, const char* key, int64_t* value) try {
  if (!c || !key || !value) return fail(SDIM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  const std::string k(key);
  if (k == "items") { *value = (int64_t)t->items.size(); return 0; }
  if (k == "lds_items") { *value = t->n_lds_items; return 0; }
  if (k == "gang_items" || k == "gang_members" || k == "gangs" || k == "launches") {
    int64_t items = 0, members = 0, gangs = 0;
    for (const Group& g : t->groups) {
      const LaunchClass& L = t->classes[g.cls];
      if (!L.gang) continue;
      items += g.count;
      members = std::max<int64_t>(members, L.gang);
      gangs += g.gang_grid / (int)L.gang;
    }
    *value = k == "gang_items" ? items : k == "gang_members" ? members : k == "gangs" ? gangs : (int64_t)t->groups.size();
    return 0;
  }
  return fail(SDIM_EINVAL, "unknown ticket key " + k);
} SDIM_CATCH

int sdim_ticket_counters(sdim_ctx* c, uint64_t ticket, uint64_t* out, int32_t count) {
  if (!c || !out || count < 0) return fail(SDIM_EINVAL, "bad arguments");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  int rc = finish(c, t);
  if (rc) return rc;
  const int32_t m = std::min<int32_t>(count, 2 + SDIM_NUM_TIMERS);
  for (int32_t i = 0; i < m; i++) out[i] = t->h_counters[i];
  return m;
}

int sdim_ticket_poison(sdim_ctx* c, uint64_t ticket, int32_t value) {
  if (!c || value < 0 || value > 255) return fail(SDIM_EINVAL, "bad arguments");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  int rc = finish(c, t);
  if (rc) return rc;
  HIP_OK(hipSetDevice(c->device));
  for (Ticket* u = t; u; u = u->child.get()) {
    if (!u->rec_bytes) continue;
    std::fill(u->h_rec.begin(), u->h_rec.end(), (uint8_t)value);
    HIP_OK(hipMemcpy(u->arena.p, u->h_rec.data(), u->rec_bytes, hipMemcpyHostToDevice));
    std::fill(std::begin(u->h_counters), std::end(u->h_counters), 0ull);
  }
  return 0;
}

int sdim_replay(sdim_ctx* c, uint64_t ticket, int32_t iters, int32_t which, double* elapsed_ms) try {
  if (!c || iters < 1 || which < 0 || which > 2) return fail(SDIM_EINVAL, "bad replay arguments");
  std::lock_guard<std::mutex> lk(c->mu);
  Ticket* t = find_ticket(c, ticket);
  if (!t) return fail(SDIM_ENOTFOUND, "unknown ticket");
  int rc = finish(c, t);
  if (rc) return rc;
  for (auto& it : t->items)
    if (it.flags & ITEM_INIT_GIVEN) return fail(SDIM_EINVAL, "replay of a job with an initial tableau is not supported");
  HIP_OK(hipSetDevice(c->device));
  if (t->mode == SDIM_MODE_BATCHED && !t->graph[which]) {
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue(t, which);
    hipError_t e = hipStreamEndCapture(t->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(SDIM_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&t->graph[which], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(SDIM_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  HIP_OK(hipEventRecord(t->ev0, t->stream));
  Ticket* ch = t->child.get();
  if (ch) HIP_OK(hipStreamWaitEvent(ch->stream, t->ev0, 0));
  for (int32_t i = 0; i < iters; i++) {
    if (ch) {                       // the reference shots, concurrently on their own stream
      rc = enqueue(ch, 0);
      if (rc) return rc;
    }
    if (t->mode == SDIM_MODE_BATCHED) {
      HIP_OK(hipGraphLaunch(t->graph[which], t->stream));
    } else {
      rc = enqueue(t, which);
      if (rc) return rc;
    }
  }
  if (ch) {
    HIP_OK(hipEventRecord(ch->ev1, ch->stream));
    HIP_OK(hipStreamWaitEvent(t->stream, ch->ev1, 0));
    if (ch->rec_bytes)
      HIP_OK(hipMemcpyAsync(ch->h_rec.data(), ch->arena.p, ch->rec_bytes, hipMemcpyDeviceToHost, ch->stream));
  }
  HIP_OK(hipEventRecord(t->ev1, t->stream));
  if (t->d_counters)
    HIP_OK(hipMemcpyAsync(t->h_counters, t->d_counters, sizeof(t->h_counters), hipMemcpyDeviceToHost, t->stream));
  if (t->rec_bytes)
    HIP_OK(hipMemcpyAsync(t->h_rec.data(), t->arena.p, t->rec_bytes, hipMemcpyDeviceToHost, t->stream));
  HIP_OK(hipStreamSynchronize(t->stream));
  if (ch) HIP_OK(hipStreamSynchronize(ch->stream));   // its record copy follows its kernels
  for (Ticket* u = t; u; u = u->child.get()) {
    for (size_t gi = 0; gi < u->groups.size(); gi++) {
      if (!u->classes[u->groups[gi].cls].gang) continue;
      int e = 0;
      HIP_OK(hipMemcpy(&e, u->groups[gi].gp.err, sizeof(int), hipMemcpyDeviceToHost));
      if (e) return poison(t, SDIM_EHIP, "a gang barrier timed out during replay; results are invalid");
    }
  }
  for (Ticket* u = t; u; u = u->child.get()) {
    rc = kernel_error(u);
    if (rc) return u == t ? rc : poison(t, rc, g_err);
  }
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
  if (elapsed_ms) *elapsed_ms = ms;
  return 0;
} SDIM_CATCH

}  // extern "C"

// ==================================================================== serving loop
// Continuous batching.  Requests (one circuit each) arrive through
// sdim_server_submit at any time.  A dispatcher thread admits pending requests
// in arrival order while their estimated device footprint fits the HBM budget,
// launches the admitted ones as one submission on a free stream (up to
// max_inflight submissions run concurrently on different streams, so new work
// overlaps work already on the device), and on completion copies each
// request's records to host memory and frees the device memory.  Every request
// keeps its own Philox key, so its results do not depend on how it was batched.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t id = 0;
  std::vector<int64_t> ir;
  std::vector<int32_t> ch;
  std::vector<double> pr;
  sdim_job job{};
  uint64_t footprint = 0;
  double t_arrive = 0, t_launch = 0, t_done = 0;
  int state = 0;                 // 0 pending, 1 in flight, 2 done, 3 failed
  std::vector<uint8_t> records;  // MEASUREMENT_DTYPE bytes
  int64_t n_records = 0;
  uint64_t gate_shots = 0;
  uint64_t cost = 0;             // work estimate: ops x n x (64-shot blocks), see serve_cost
  std::string error;
};

struct Batch {
  uint64_t ticket = 0;
  std::vector<uint64_t> reqs;
  uint64_t bytes = 0;
};

// Serving work estimate of a request: its items repeat the circuit's X/Z evolution (ops x row
// bytes ~ ops x n) once per block of at most 64 shots.
uint64_t serve_cost(uint64_t ops, int64_t n, int64_t shots) {
  return std::max<uint64_t>(1, ops) * (uint64_t)std::max<int64_t>(1, n) * (uint64_t)((shots + 63) / 64);
}
// requests of at most this estimate (about n = 300 with depth 8n and at most 64 shots) are small:
// the last free in-flight slot of the server only takes a batch of them
constexpr uint64_t kServeSmallCost = 1ull << 20;
// requests batched together differ in estimate by at most this factor
constexpr uint64_t kServeCostWindow = 16;

uint64_t estimate_footprint(const sdim_job& j) {
  const uint32_t n = (uint32_t)j.n, d = (uint32_t)j.d;
  const uint64_t S = (uint64_t)std::min<int64_t>(j.shots, 4096);
  const uint64_t items = ((uint64_t)j.shots + S - 1) / S;
  int64_t n_meas = 0;
  for (int64_t k = 0; k < j.n_ops; k++) {
    const int64_t g = j.ir[3 * k];
    if (g == 14 || g == 15 || g == 16) n_meas++;
  }
  const uint64_t tab = items * tableau_bytes(n, d, (uint32_t)S);
  const uint64_t rec = (uint64_t)n_meas * (uint64_t)j.shots;
  const uint64_t ops = (uint64_t)j.n_ops * 16 * 9 + 65536;   // worst-case primitive expansion
  return round_block(tab + rec) + round_block(ops + 4096);
}

}  // namespace

struct sdim_server {
  sdim_ctx* c = nullptr;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<uint64_t> pending;
  std::map<uint64_t, Request> reqs;
  std::vector<Batch> inflight;            // dispatcher thread only
  std::atomic<int64_t> n_inflight{0};
  uint64_t budget = 0, used = 0;
  int max_inflight = 4, max_batch = 256;
  bool stop = false;
  uint64_t next_id = 1;
  uint64_t completed = 0;
  double t_first_arrival = 0, t_last_done = 0;
  uint64_t gate_shots_done = 0;
  std::vector<double> latencies;
};

namespace {

void serve_loop(sdim_server* s) {
  sdim_ctx* c = s->c;
  for (;;) {
    // ---- completions
    std::vector<Batch> still;
    for (auto& b : s->inflight) {
      bool done = false;
      int rc = 0;
      std::string err;
      {
        std::lock_guard<std::mutex> lk(c->mu);
        Ticket* t = find_ticket(c, b.ticket);
        const hipError_t q = t ? hipEventQuery(t->ev1) : hipErrorInvalidValue;
        if (q != hipErrorNotReady) {
          // finished, or failed (an event error other than not-ready ends the batch too)
          rc = !t ? fail(SDIM_ENOTFOUND, "serving ticket vanished")
                  : q != hipSuccess ? poison(t, SDIM_EHIP, std::string("hipEventQuery: ") + hipGetErrorString(q))
                                    : finish(c, t);
          if (rc) err = g_err;
          done = true;
        }
      }
      if (!done) { still.push_back(b); continue; }
      const double t_done = now_s();
      std::vector<std::pair<uint64_t, std::vector<uint8_t>>> out;
      if (!rc) {
        // records of every request, in the layout of sdim_fetch_records
        rc = no_throw([&]() -> int {
          std::lock_guard<std::mutex> lk(c->mu);
          Ticket* t = find_ticket(c, b.ticket);
          for (size_t k = 0; k < b.reqs.size(); k++) {
            const JobPlan& jp = t->jobs[k];
            std::vector<uint8_t> buf((size_t)(jp.n_meas * jp.shots) * 33);
            for (int64_t ev = 0; ev < jp.n_meas; ev++) {
              for (uint32_t i = 0; i < jp.n_items; i++) {
                const ItemDesc& it = t->items[jp.first_item + i];
                const uint8_t* r = t->h_rec.data() + it.rec_off + (size_t)ev * it.S;
                for (uint32_t sh = 0; sh < it.S; sh++) {
                  const int64_t shot = (int64_t)it.shot0 + sh;
                  uint8_t* rec = buf.data() + 33 * (size_t)(ev * jp.shots + shot);
                  const int64_t q = jp.meas_q[(size_t)ev], round = jp.meas_round[(size_t)ev], val = r[sh] & 0x7F;
                  memcpy(rec, &q, 8); memcpy(rec + 8, &round, 8); memcpy(rec + 16, &shot, 8);
                  rec[24] = (r[sh] >> 7) & 1; memcpy(rec + 25, &val, 8);
                }
              }
            }
            out.emplace_back(b.reqs[k], std::move(buf));
          }
          return 0;
        });
        if (rc) { err = g_err; out.clear(); }
      }
      if (rc) {
        // the whole batch failed: its requests report the error instead of records
        {
          std::lock_guard<std::mutex> lk(c->mu);
          auto it = c->tickets.find(b.ticket);
          if (it != c->tickets.end()) drop_ticket(c, it);
        }
        {
          std::lock_guard<std::mutex> lk(s->mu);
          s->used -= std::min(s->used, b.bytes);
          for (uint64_t id : b.reqs) {
            Request& r = s->reqs[id];
            r.state = 3;
            r.error = err;
            r.t_done = t_done;
          }
        }
        s->cv.notify_all();
        continue;
      }
      {
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->tickets.find(b.ticket);
        if (it != c->tickets.end()) drop_ticket(c, it);
      }
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->used -= std::min(s->used, b.bytes);
        for (auto& kv : out) {
          Request& r = s->reqs[kv.first];
          r.records = std::move(kv.second);
          r.n_records = (int64_t)(r.records.size() / 33);
          r.t_done = t_done;
          r.state = 2;
          s->completed++;
          s->gate_shots_done += r.gate_shots;
          s->latencies.push_back(t_done - r.t_arrive);
          s->t_last_done = t_done;
        }
      }
      s->cv.notify_all();
    }
    s->inflight.swap(still);
    s->n_inflight = (int64_t)s->inflight.size();

    // ---- admission: arrival order while the footprint fits the budget.  A batch's records come
    // back when its last item finishes, so a batch holds requests of similar cost (within
    // kServeCostWindow of its first request; the others stay pending, in order, for a later
    // batch), and the last free in-flight slot only starts a batch of small requests, so that a
    // small request does not wait until large work releases a slot.
    std::vector<uint64_t> take;
    uint64_t take_bytes = 0;
    bool quit = false, rejected = false;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->stop && s->pending.empty() && s->inflight.empty()) quit = true;
      const int free_slots = s->max_inflight - (int)s->inflight.size();
      if (!quit && free_slots > 0) {
        const bool small_only = free_slots == 1 && s->max_inflight > 1;
        std::deque<uint64_t> rest;
        uint64_t c0 = 0;
        bool blocked = false;          // a request did not fit the budget: later ones wait behind it
        while (!s->pending.empty()) {
          const uint64_t id = s->pending.front();
          s->pending.pop_front();
          Request& r = s->reqs[id];
          if (r.footprint > s->budget) {       // can never fit: fail it
            r.state = 3;
            r.error = "request larger than the HBM budget";
            rejected = true;
            continue;
          }
          const bool same_class = take.empty()
                                      ? (!small_only || r.cost <= kServeSmallCost)
                                      : (r.cost <= c0 * kServeCostWindow && c0 <= r.cost * kServeCostWindow);
          if (blocked || (int)take.size() >= s->max_batch || !same_class) {
            rest.push_back(id);
            continue;
          }
          if (s->used + take_bytes + r.footprint > s->budget) {
            blocked = true;
            rest.push_back(id);
            continue;
          }
          if (take.empty()) c0 = r.cost;
          take_bytes += r.footprint;
          take.push_back(id);
        }
        s->pending.swap(rest);
        s->used += take_bytes;
      }
      if (rejected) s->cv.notify_all();
      if (!quit && take.empty()) {
        // nothing to launch: wait for an arrival or poll completions again shortly
        s->cv.wait_for(lk, std::chrono::microseconds(s->inflight.empty() ? 2000 : 50));
      }
    }
    if (quit) break;
    if (take.empty()) continue;
    std::vector<sdim_job> jobs;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (uint64_t id : take) jobs.push_back(s->reqs[id].job);
    }
    uint64_t ticket = 0, dev_bytes = 0;
    int rc;
    {
      std::lock_guard<std::mutex> lk(c->mu);
      rc = no_throw([&] {
        return submit_locked(c, jobs.data(), (int32_t)jobs.size(), SDIM_MODE_BATCHED,
                             (int)(s->inflight.size() % c->streams.size()), &ticket, &dev_bytes,
                             c->serve_hbm_threads, c->serve_hbm_per_cu);
      });
    }
    const double t_launch = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc) {
      for (uint64_t id : take) { s->reqs[id].state = 3; s->reqs[id].error = g_err; }
      s->used -= std::min(s->used, take_bytes);
      s->cv.notify_all();
      continue;
    }
    for (uint64_t id : take) { s->reqs[id].state = 1; s->reqs[id].t_launch = t_launch; }
    Batch b;
    b.ticket = ticket;
    b.reqs = take;
    b.bytes = take_bytes;
    s->inflight.push_back(b);
    s->n_inflight = (int64_t)s->inflight.size();
  }
}

}  // namespace

extern "C" {

int sdim_server_create(sdim_ctx* c, uint64_t hbm_budget_bytes, int32_t max_inflight, int32_t max_batch,
                       sdim_server** out) try {
  if (!c || !out || max_inflight < 1 || max_batch < 1) return fail(SDIM_EINVAL, "bad server arguments");
  std::unique_ptr<sdim_server> s(new sdim_server());
  s->c = c;
  s->budget = hbm_budget_bytes;
  s->max_inflight = std::min<int>(max_inflight, (int)c->streams.size());
  s->max_batch = max_batch;
  s->th = std::thread(serve_loop, s.get());
  *out = s.release();
  return 0;
} SDIM_CATCH

int sdim_server_submit(sdim_server* s, const sdim_job* j, uint64_t* request_id) try {
  if (!s || !j || !request_id) return fail(SDIM_EINVAL, "null argument");
  if (j->n < 1 || !supported_dim(j->d) || j->shots < 1 || j->n_ops < 0 || (j->n_ops && !j->ir))
    return fail(SDIM_EINVAL, "invalid request");
  for (int k = 0; k < 6; k++)
    if (j->init[k]) return fail(SDIM_EINVAL, "serving requests start from |0...0>");
  Request r;
  r.ir.assign(j->ir, j->ir + 3 * j->n_ops);
  if (j->noise_channel) r.ch.assign(j->noise_channel, j->noise_channel + j->n_ops);
  if (j->noise_prob) r.pr.assign(j->noise_prob, j->noise_prob + j->n_ops);
  r.job = *j;
  r.job.ir = r.ir.data();
  r.job.noise_channel = j->noise_channel ? r.ch.data() : nullptr;
  r.job.noise_prob = j->noise_prob ? r.pr.data() : nullptr;
  r.footprint = estimate_footprint(r.job);
  uint64_t ops = 0;
  for (int64_t k = 0; k < j->n_ops; k++) ops += (j->ir[3 * k] != 0);
  r.gate_shots = ops * (uint64_t)j->shots;
  r.cost = serve_cost(ops, j->n, j->shots);
  r.t_arrive = now_s();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r.id = s->next_id++;
    if (s->t_first_arrival == 0) s->t_first_arrival = r.t_arrive;
    *request_id = r.id;
    const uint64_t id = r.id;
    s->reqs.emplace(id, std::move(r));
    // vectors moved: re-point the job at the stored copies
    Request& st = s->reqs[id];
    st.job.ir = st.ir.data();
    st.job.noise_channel = st.ch.empty() ? nullptr : st.ch.data();
    st.job.noise_prob = st.pr.empty() ? nullptr : st.pr.data();
    s->pending.push_back(id);
  }
  s->cv.notify_all();
  return 0;
} SDIM_CATCH

int sdim_server_poll(sdim_server* s, uint64_t id, int32_t* state) {
  if (!s || !state) return fail(SDIM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(SDIM_ENOTFOUND, "unknown request");
  *state = it->second.state;
  return 0;
}

int sdim_server_wait(sdim_server* s, uint64_t id, double timeout_s) {
  if (!s) return fail(SDIM_EINVAL, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
  for (;;) {
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(SDIM_ENOTFOUND, "unknown request");
    if (it->second.state == 2) return 0;
    if (it->second.state == 3) return fail(SDIM_EINVAL, it->second.error);
    if (std::chrono::steady_clock::now() >= deadline) return fail(SDIM_EBUSY, "timeout");
    s->cv.wait_until(lk, deadline);   // the state is checked again after every wake-up
  }
}

int sdim_server_result(sdim_server* s, uint64_t id, void* out, int64_t capacity, int64_t* n_records,
                       double* latency_s) {
  if (!s) return fail(SDIM_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(SDIM_ENOTFOUND, "unknown request");
  Request& r = it->second;
  if (r.state == 3) return fail(SDIM_EINVAL, r.error);
  if (r.state != 2) return fail(SDIM_EBUSY, "request not finished");
  if (n_records) *n_records = r.n_records;
  if (latency_s) *latency_s = r.t_done - r.t_arrive;
  if (!out) return 0;                                  // size query
  if (capacity < r.n_records) return fail(SDIM_EINVAL, "record buffer too small");
  memcpy(out, r.records.data(), r.records.size());
  s->reqs.erase(it);
  return 0;
}

int sdim_server_stats(sdim_server* s, int64_t* completed, int64_t* pending, int64_t* inflight,
                      uint64_t* gate_shots, double* span_s, double* p50_s, double* p99_s) {
  if (!s) return fail(SDIM_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (completed) *completed = (int64_t)s->completed;
  if (pending) *pending = (int64_t)s->pending.size();
  if (inflight) *inflight = s->n_inflight.load();
  if (gate_shots) *gate_shots = s->gate_shots_done;
  if (span_s) *span_s = s->t_last_done > s->t_first_arrival ? s->t_last_done - s->t_first_arrival : 0.0;
  std::vector<double> l = s->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double q) {
    if (l.empty()) return 0.0;
    size_t k = (size_t)std::ceil(q * (double)l.size()) - 1;
    return l[std::min(k, l.size() - 1)];
  };
  if (p50_s) *p50_s = pct(0.50);
  if (p99_s) *p99_s = pct(0.99);
  return 0;
}

int sdim_server_destroy(sdim_server* s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->th.joinable()) s->th.join();
  delete s;
  return 0;
}

}  // extern "C"
