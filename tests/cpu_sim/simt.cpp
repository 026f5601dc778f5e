// Fiber scheduler and HIP host API of the CPU build of the engine (see
// hip/hip_runtime.h in this directory for the execution model).  Test
// infrastructure: it runs the engine's own kernels on the host so that
// AddressSanitizer and ThreadSanitizer can check them.
#include <hip/hip_runtime.h>

#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define SIMT_ASAN 1
#endif
#if __has_feature(thread_sanitizer)
#define SIMT_TSAN 1
#endif
#endif
#ifdef SIMT_ASAN
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>
#endif
#ifdef SIMT_TSAN
#include <sanitizer/tsan_interface.h>
#endif

// the kernels' dynamic LDS (extern __shared__ uint8_t lds[] in namespace sdim)
#define SIMT_LDS_MAX (160 * 1024)
namespace sdim {
thread_local __attribute__((aligned(64))) uint8_t lds[SIMT_LDS_MAX];
}

// x86-64 stack switch: saves the callee-saved registers on the current stack,
// stores the stack pointer in *from and continues on stack `to`.
extern "C" void simt_switch(void** from, void* to);
asm(R"(
  .text
  .globl simt_switch
  .type simt_switch,@function
simt_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
  .size simt_switch, .-simt_switch
)");

namespace simt {
namespace {

size_t stack_bytes() {
  static size_t b = [] {
    const char* e = getenv("SDIM_SIM_STACK_KB");
    size_t kb = e ? (size_t)atoi(e) : 0;
    if (!kb) kb = 128;
    return kb * 1024;
  }();
  return b;
}

enum FState { RUNNABLE = 0, AT_BARRIER = 1, AT_WAVE = 2, DONE = 3 };

struct Group;

struct Fiber {
  void* sp = nullptr;
  uint8_t* stack = nullptr;
  dim3 tid;
  uint32_t linear = 0;
  int state = RUNNABLE;
  int wave_gen = 0;
  Group* g = nullptr;
#ifdef SIMT_ASAN
  void* fake = nullptr;
#endif
#ifdef SIMT_TSAN
  void* tsan = nullptr;
  char death_obj = 0;
#endif
};

struct Wave {
  int arrived = 0;
  int gen = 0;
  uint64_t vals[2][64];
  char sync_obj = 0;        // ThreadSanitizer release/acquire object of the wave's exchanges
};

struct Group {
  dim3 bid, bdim, gdim;
  uint32_t nt = 0, live = 0, at_barrier = 0;
  std::vector<Fiber> f;
  std::vector<Wave> waves;
  void* sched_sp = nullptr;
  const std::function<void()>* body = nullptr;
  char barrier_obj = 0;     // ThreadSanitizer release/acquire object of __syncthreads
  uint64_t barriers = 0;
#ifdef SIMT_ASAN
  void* sched_fake = nullptr;
  const void* sched_bottom = nullptr;
  size_t sched_size = 0;
#endif
#ifdef SIMT_TSAN
  void* sched_tsan = nullptr;
#endif
};

thread_local Group* t_group = nullptr;
thread_local Fiber* t_fiber_ = nullptr;
std::atomic<uint64_t> g_clock{0};

// The scheduler's bookkeeping is shared by the fibers of a workgroup and the scheduler,
// which ThreadSanitizer sees as separate threads that do not synchronise when a fiber
// yields (only barriers and wave exchanges order the fibers' own accesses): it is
// accessed with relaxed atomics.
template <class T> T ld(const T& x) { return __atomic_load_n(&x, __ATOMIC_RELAXED); }
template <class T> void st(T& x, T v) { __atomic_store_n(&x, v, __ATOMIC_RELAXED); }
template <class T> void add(T& x, T v) { __atomic_fetch_add(&x, v, __ATOMIC_RELAXED); }
inline Fiber* cur_fiber() { return ld(t_fiber_); }

// fiber -> scheduler
void to_scheduler(Fiber* me, bool dying) {
  Group* g = me->g;
  void* const sched_sp = g->sched_sp;   // read before a dying fiber's final release
#ifdef SIMT_TSAN
  void* const sched_tsan = g->sched_tsan;
  // no happens-before from a yielding fiber: fibers synchronise only where the device's
  // threads do (barriers, wave exchanges, atomics).  A finished fiber releases its own
  // object, which the scheduler acquires once the whole workgroup has finished (before it
  // reuses or frees the workgroup's bookkeeping), so no fiber is ordered after another.
  if (dying) __tsan_release(&me->death_obj);
  __tsan_switch_to_fiber(sched_tsan, __tsan_switch_to_fiber_no_sync);
#endif
#ifdef SIMT_ASAN
  __sanitizer_start_switch_fiber(dying ? nullptr : &me->fake, g->sched_bottom, g->sched_size);
#else
  (void)dying;
#endif
  simt_switch(&me->sp, sched_sp);
#ifdef SIMT_ASAN
  __sanitizer_finish_switch_fiber(me->fake, &g->sched_bottom, &g->sched_size);
#endif
}

void fiber_entry() {
  Fiber* me = cur_fiber();
  Group* g = me->g;
#ifdef SIMT_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &g->sched_bottom, &g->sched_size);
#endif
  (*g->body)();
  st(me->state, (int)DONE);
  add(g->live, (uint32_t)-1);
  to_scheduler(me, true);
  abort();   // never resumed
}

// scheduler -> fiber
void run_fiber(Group* g, Fiber* f) {
  st(t_fiber_, f);
  void* const to = f->sp;
#ifdef SIMT_TSAN
  __tsan_switch_to_fiber(f->tsan, 0);
#endif
#ifdef SIMT_ASAN
  __sanitizer_start_switch_fiber(&g->sched_fake, f->stack, stack_bytes());
#endif
  simt_switch(&g->sched_sp, to);
#ifdef SIMT_ASAN
  __sanitizer_finish_switch_fiber(g->sched_fake, nullptr, nullptr);
#endif
  st(t_fiber_, (Fiber*)nullptr);
}

[[noreturn]] void deadlock(Group* g) {
  fprintf(stderr, "simt: workgroup (%u,%u,%u) cannot make progress after %llu barriers:\n", g->bid.x, g->bid.y,
          g->bid.z, (unsigned long long)g->barriers);
  const char* names[] = {"runnable", "__syncthreads", "wave operation", "exited"};
  for (uint32_t i = 0; i < g->nt; i++)
    if (i < 4 || (i % 64) == 0 || ld(g->f[i].state) != ld(g->f[0].state))
      fprintf(stderr, "  thread %u: %s\n", i, names[ld(g->f[i].state)]);
  fflush(stderr);
  abort();
}

void run_group(Group* g) {
  g->live = g->nt;
  g->at_barrier = 0;
#ifdef SIMT_TSAN
  g->sched_tsan = __tsan_get_current_fiber();
#endif
  for (uint32_t i = 0; i < g->nt; i++) {
    Fiber& f = g->f[i];
    f.state = RUNNABLE;
    f.wave_gen = 0;
    f.g = g;
    // initial frame: six callee-saved registers, then fiber_entry as the return address
    uintptr_t top = ((uintptr_t)f.stack + stack_bytes()) & ~(uintptr_t)15;
    void** sp = (void**)top;
    *--sp = nullptr;                      // fake return address of fiber_entry
    *--sp = (void*)&fiber_entry;
    for (int r = 0; r < 6; r++) *--sp = nullptr;
    f.sp = sp;
#ifdef SIMT_TSAN
    f.tsan = __tsan_create_fiber(0);
#endif
  }
  for (auto& w : g->waves) { w.arrived = 0; w.gen = 0; }
  for (;;) {
    bool progressed = false;
    for (uint32_t i = 0; i < g->nt; i++) {
      if (ld(g->f[i].state) == RUNNABLE) {
        run_fiber(g, &g->f[i]);
        progressed = true;
      }
    }
    if (ld(g->live) == 0) break;
    // a barrier completes when every live fiber is waiting at it
    if (ld(g->at_barrier) == ld(g->live)) {
      for (auto& f : g->f)
        if (ld(f.state) == AT_BARRIER) st(f.state, (int)RUNNABLE);
      st(g->at_barrier, 0u);
      g->barriers++;
      continue;
    }
    // a wave operation completes when every live lane of the wave has arrived
    for (size_t w = 0; w < g->waves.size(); w++) {
      Wave& wv = g->waves[w];
      if (!ld(wv.arrived)) continue;
      uint32_t live = 0;
      for (uint32_t l = 0; l < 64 && w * 64 + l < g->nt; l++) live += ld(g->f[w * 64 + l].state) != DONE;
      if ((uint32_t)ld(wv.arrived) == live) {
        for (uint32_t l = 0; l < 64 && w * 64 + l < g->nt; l++)
          if (ld(g->f[w * 64 + l].state) == AT_WAVE) st(g->f[w * 64 + l].state, (int)RUNNABLE);
        st(wv.arrived, 0);
        st(wv.gen, ld(wv.gen) + 1);
        progressed = true;
      }
    }
    if (!progressed) deadlock(g);
  }
#ifdef SIMT_TSAN
  for (auto& f : g->f) {
    __tsan_acquire(&f.death_obj);
    __tsan_destroy_fiber(f.tsan);
  }
#endif
}

// one OS thread's fibers (stacks kept across workgroups of the same size)
struct Worker {
  std::vector<Fiber> fibers;
  uint8_t* stacks = nullptr;
  size_t n_stacks = 0;
  void ensure(uint32_t nt) {
    if (n_stacks >= nt) return;
    if (stacks) munmap(stacks, n_stacks * stack_bytes());
    stacks = (uint8_t*)mmap(nullptr, nt * stack_bytes(), PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (stacks == MAP_FAILED) { fprintf(stderr, "simt: out of memory for fiber stacks\n"); abort(); }
    // a guard page below every stack: an overflow faults instead of writing the next stack
    for (uint32_t i = 0; i < nt; i++) mprotect(stacks + (size_t)i * stack_bytes(), 4096, PROT_NONE);
    n_stacks = nt;
  }
  ~Worker() { if (stacks) munmap(stacks, n_stacks * stack_bytes()); }
};
thread_local Worker t_worker;

void run_block(const std::function<void()>& body, dim3 bid, dim3 grid, dim3 block, size_t shmem) {
  const uint32_t nt = block.x * block.y * block.z;
  if (shmem > SIMT_LDS_MAX) { fprintf(stderr, "simt: %zu bytes of dynamic LDS requested\n", shmem); abort(); }
  Worker& wk = t_worker;
  wk.ensure(nt);
  Group g;
  g.bid = bid; g.bdim = block; g.gdim = grid; g.nt = nt; g.body = &body;
  g.f.resize(nt);
  g.waves.resize((nt + 63) / 64);
  for (uint32_t i = 0; i < nt; i++) {
    g.f[i].stack = wk.stacks + (size_t)i * stack_bytes();
    g.f[i].linear = i;
    g.f[i].tid = dim3(i % block.x, (i / block.x) % block.y, i / (block.x * block.y));
  }
#ifdef SIMT_ASAN
  // LDS beyond the launch's dynamic size is out of bounds
  ASAN_POISON_MEMORY_REGION(sdim::lds + shmem, SIMT_LDS_MAX - shmem);
#endif
  // fresh LDS holds garbage on the device; make reads of unwritten LDS visible as such
  memset(sdim::lds, 0xA5, shmem);
  t_group = &g;
  run_group(&g);
  t_group = nullptr;
#ifdef SIMT_ASAN
  ASAN_UNPOISON_MEMORY_REGION(sdim::lds, SIMT_LDS_MAX);
#endif
}

int n_workers() {
  static int n = [] {
    const char* e = getenv("SDIM_SIM_THREADS");
    int v = e ? atoi(e) : 0;
    if (v <= 0) v = (int)std::max(1u, std::thread::hardware_concurrency());
    return v;
  }();
  return n;
}

// Persistent pool of OS threads that run the workgroups of one grid at a time (launches
// are serialised by the caller).  Each worker keeps its fiber stacks and its LDS.
class Pool {
 public:
  void run(const std::function<void()>& body, dim3 grid, dim3 block, size_t shmem) {
    const uint64_t nb = (uint64_t)grid.x * grid.y * grid.z;
    std::unique_lock<std::mutex> lk(mu_);
    if (threads_.empty()) start(lk);
    job_ = Job{&body, grid, block, shmem, nb};
    next_.store(0);
    pending_ = (int)threads_.size();
    gen_++;
    cv_.notify_all();
    done_cv_.wait(lk, [&] { return pending_ == 0; });
    job_.body = nullptr;
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }

 private:
  struct Job {
    const std::function<void()>* body = nullptr;
    dim3 grid, block;
    size_t shmem = 0;
    uint64_t nb = 0;
  };
  void start(std::unique_lock<std::mutex>&) {
    for (int i = 0; i < n_workers(); i++) threads_.emplace_back([this] { loop(); });
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        j = job_;
      }
      for (;;) {
        const uint64_t b = next_.fetch_add(1);
        if (b >= j.nb) break;
        const dim3 bid((uint32_t)(b % j.grid.x), (uint32_t)((b / j.grid.x) % j.grid.y),
                       (uint32_t)(b / ((uint64_t)j.grid.x * j.grid.y)));
        run_block(*j.body, bid, j.grid, j.block, j.shmem);
      }
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_cv_.notify_all();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::vector<std::thread> threads_;
  Job job_;
  std::atomic<uint64_t> next_{0};
  uint64_t gen_ = 0;
  int pending_ = 0;
  bool stop_ = false;
};

void run_grid(const std::function<void()>& body, dim3 grid, dim3 block, size_t shmem) {
  static Pool* pool = new Pool();   // lives until the process ends
  pool->run(body, grid, block, shmem);
}

}  // namespace

const dim3& thread_idx() { return cur_fiber()->tid; }
const dim3& block_idx() { return t_group->bid; }
const dim3& block_dim() { return t_group->bdim; }
const dim3& grid_dim() { return t_group->gdim; }
uint32_t lane_id() { return cur_fiber()->linear & 63u; }
uint64_t clock_ticks() { return g_clock.fetch_add(1, std::memory_order_relaxed); }

void syncthreads() {
  Fiber* me = cur_fiber();
  Group* g = me->g;
#ifdef SIMT_TSAN
  __tsan_release(&g->barrier_obj);
#endif
  st(me->state, (int)AT_BARRIER);
  add(g->at_barrier, 1u);
  to_scheduler(me, false);
#ifdef SIMT_TSAN
  __tsan_acquire(&g->barrier_obj);
#endif
}

const uint64_t* wave_exchange(uint64_t v) {
  Fiber* me = cur_fiber();
  Group* g = me->g;
  Wave& w = g->waves[me->linear / 64];
  const int buf = ld(w.gen) & 1;
  w.vals[buf][me->linear & 63] = v;
#ifdef SIMT_TSAN
  __tsan_release(&w.sync_obj);
#endif
  st(me->state, (int)AT_WAVE);
  add(w.arrived, 1);
  to_scheduler(me, false);
#ifdef SIMT_TSAN
  __tsan_acquire(&w.sync_obj);
#endif
  return w.vals[buf];
}

void yield() {
  // only meaningful between workgroups (spin waits); a fiber yields its OS thread
  sched_yield();
}

}  // namespace simt

// ==================================================================== host API
struct ihipEvent_t {
  double t = 0;
  ihipGraph* capture = nullptr;   // recorded into a capturing stream
};
struct ihipGraph {
  std::vector<std::function<void()>> nodes;
  std::vector<hipStream_t> streams;
};
struct ihipGraphExec {
  std::vector<std::function<void()>> nodes;
};
struct ihipStream_t {
  ihipGraph* capture = nullptr;
};

namespace {
thread_local hipError_t t_last = hipSuccess;
std::mutex g_launch_mu;   // launches of different host threads run one at a time (one device)
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
hipError_t ret(hipError_t e) {
  if (e != hipSuccess) t_last = e;
  return e;
}
}  // namespace

namespace simt {
void submit_launch(hipStream_t s, dim3 grid, dim3 block, size_t shmem, std::function<void()> body) {
  if (s && s->capture) {
    s->capture->nodes.push_back([grid, block, shmem, body] {
      std::lock_guard<std::mutex> lk(g_launch_mu);
      run_grid(body, grid, block, shmem);
    });
    return;
  }
  std::lock_guard<std::mutex> lk(g_launch_mu);
  run_grid(body, grid, block, shmem);
}
}  // namespace simt

hipError_t hipGetDeviceCount(int* count) { *count = 1; return hipSuccess; }
hipError_t hipSetDevice(int device) { return ret(device == 0 ? hipSuccess : hipErrorInvalidDevice); }
hipError_t hipGetDevice(int* device) { *device = 0; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* prop, int device) {
  if (device != 0) return ret(hipErrorInvalidDevice);
  memset(prop, 0, sizeof(*prop));
  strcpy(prop->name, "simt CPU emulation");
  strcpy(prop->gcnArchName, "simt");
  const char* e = getenv("SDIM_SIM_CUS");
  prop->multiProcessorCount = e ? std::max(1, atoi(e)) : 4;
  prop->totalGlobalMem = (size_t)16 << 30;
  prop->maxThreadsPerBlock = 1024;
  prop->sharedMemPerBlock = 65536;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) {
  switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorInvalidValue: return "hipErrorInvalidValue";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorNotReady: return "hipErrorNotReady";
    case hipErrorNotSupported: return "hipErrorNotSupported";
    default: return "hipError (simt)";
  }
}
hipError_t hipGetLastError() { hipError_t e = t_last; t_last = hipSuccess; return e; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new ihipStream_t(); return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t* s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  if (e && e->capture && s && !s->capture) {   // a stream that waits on a captured event joins the capture
    s->capture = e->capture;
    e->capture->streams.push_back(s);
  }
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) { *e = new ihipEvent_t(); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  if (s && s->capture) {
    e->capture = s->capture;
    s->capture->nodes.push_back([e] { e->t = now_ms(); });
    return hipSuccess;
  }
  e->capture = nullptr;
  e->t = now_ms();
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(b->t - a->t); return hipSuccess; }
hipError_t hipMalloc(void** p, size_t bytes) {
  *p = aligned_alloc(256, (bytes + 255) & ~(size_t)255);
  return ret(*p ? hipSuccess : hipErrorOutOfMemory);
}
hipError_t hipFree(void* p) { free(p); return hipSuccess; }
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) { memcpy(dst, src, bytes); return hipSuccess; }
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k, hipStream_t s) {
  if (s && s->capture) {
    s->capture->nodes.push_back([dst, src, bytes] { memcpy(dst, src, bytes); });
    return hipSuccess;
  }
  return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemset(void* dst, int v, size_t bytes) { memset(dst, v, bytes); return hipSuccess; }
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t s) {
  if (s && s->capture) {
    s->capture->nodes.push_back([dst, v, bytes] { memset(dst, v, bytes); });
    return hipSuccess;
  }
  return hipMemset(dst, v, bytes);
}
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (!s || s->capture) return ret(hipErrorInvalidValue);
  s->capture = new ihipGraph();
  s->capture->streams.push_back(s);
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  if (!s || !s->capture) return ret(hipErrorInvalidValue);
  ihipGraph* G = s->capture;
  for (hipStream_t t : G->streams) t->capture = nullptr;
  *g = G;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *x = new ihipGraphExec();
  (*x)->nodes = g->nodes;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
  if (s && s->capture) {
    s->capture->nodes.push_back([x] { for (auto& n : x->nodes) n(); });
    return hipSuccess;
  }
  for (auto& n : x->nodes) n();
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t x) { delete x; return hipSuccess; }
hipError_t hipGetSymbolAddress(void** p, const void* symbol) { *p = const_cast<void*>(symbol); return hipSuccess; }
