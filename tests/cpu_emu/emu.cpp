// CPU emulation runtime for the HIP subset in include/hip/hip_runtime.h
// (test infrastructure).  Workgroups run one after another; the threads of
// a workgroup are fibers on private stacks, switched with _setjmp/_longjmp
// and scheduled round-robin until each reaches a barrier, a wavefront
// operation or its end.  Launches from several host threads are serialised.
#include <hip/hip_runtime.h>
#include <setjmp.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <new>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/asan_interface.h>
#define FLEMU_ASAN 1
#else
#define FLEMU_ASAN 0
#endif

namespace flemu {

dim3 g_block, g_bdim, g_gdim;
dim3* g_tid = nullptr;

namespace {

enum State { RUNNABLE, AT_BARRIER, AT_WAVE, DONE };
enum WaveOp { OP_ANY, OP_BALLOT, OP_READLANE };

struct Fiber {
  jmp_buf env;
  dim3 tid;
  State state = RUNNABLE;
  bool started = false;
  WaveOp op = OP_ANY;
  int value = 0;   // deposited operand
  int arg = 0;     // readlane source lane
  unsigned long long result = 0;  // result of the wavefront operation
};

constexpr size_t kStack = 256 * 1024;
constexpr uint32_t kMaxThreads = 1024;
constexpr uint32_t kWave = 64;

std::mutex g_launch_mu;
jmp_buf g_sched;
Fiber* g_cur = nullptr;
const std::function<void()>* g_body = nullptr;
std::vector<Fiber> g_fibers(kMaxThreads);
char* g_stacks = nullptr;
#if FLEMU_ASAN
void* g_fake_stack = nullptr;
#endif

[[noreturn]] void die(const char* msg) {
  std::fprintf(stderr, "flemu: %s (block %u, thread %u)\n", msg, g_block.x, g_cur ? g_cur->tid.x : 0u);
  std::abort();
}

char* stack_base(const Fiber* f) { return g_stacks + (size_t)(f - g_fibers.data()) * kStack; }

[[noreturn]] void fiber_main() {
#if FLEMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
#endif
  (*g_body)();
  g_cur->state = DONE;
#if FLEMU_ASAN
  __sanitizer_start_switch_fiber(nullptr, nullptr, 0);
#endif
  _longjmp(g_sched, 1);
  __builtin_unreachable();
}

__attribute__((noinline)) void start_on(char* top) {
  asm volatile(
      "movq %0, %%rsp\n\t"
      "xorq %%rbp, %%rbp\n\t"
      "callq *%1\n\t"
      "ud2\n\t" ::"r"(top),
      "r"(&fiber_main)
      : "memory");
  __builtin_unreachable();
}

void resume(Fiber* f) {
  g_cur = f;
  g_tid = &f->tid;
  if (!_setjmp(g_sched)) {
#if FLEMU_ASAN
    __sanitizer_start_switch_fiber(&g_fake_stack, stack_base(f), kStack);
#endif
    if (!f->started) {
      f->started = true;
      char* top = stack_base(f) + kStack;
      top = (char*)((uintptr_t)top & ~(uintptr_t)15);
      start_on(top);
    }
    _longjmp(f->env, 1);
  }
#if FLEMU_ASAN
  __sanitizer_finish_switch_fiber(g_fake_stack, nullptr, nullptr);
#endif
}

void yield_to_scheduler() {
  if (!_setjmp(g_cur->env)) {
#if FLEMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, nullptr, 0);
#endif
    _longjmp(g_sched, 1);
  }
#if FLEMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
#endif
}

void run_block(uint32_t nthreads) {
  for (uint32_t t = 0; t < nthreads; t++) {
    Fiber& f = g_fibers[t];
    f.tid = dim3(t, 0, 0);
    f.state = RUNNABLE;
    f.started = false;
  }
  uint32_t nwaves = (nthreads + kWave - 1) / kWave;
  for (;;) {
    bool any_live = false;
    for (uint32_t t = 0; t < nthreads; t++)
      if (g_fibers[t].state == RUNNABLE) resume(&g_fibers[t]);
    // wavefront operations whose live lanes have all arrived
    bool released = false;
    for (uint32_t w = 0; w < nwaves; w++) {
      uint32_t l0 = w * kWave, l1 = std::min(nthreads, l0 + kWave);
      uint32_t live = 0, at = 0;
      for (uint32_t t = l0; t < l1; t++) {
        if (g_fibers[t].state != DONE) live++;
        if (g_fibers[t].state == AT_WAVE) at++;
      }
      if (live == 0 || at == 0) continue;
      if (at != live) continue;  // some lanes are at the workgroup barrier: resolved below
      const Fiber& first = g_fibers[l0 + 0];
      (void)first;
      WaveOp op = OP_ANY;
      int arg = -1;
      bool set = false;
      for (uint32_t t = l0; t < l1; t++) {
        Fiber& f = g_fibers[t];
        if (f.state != AT_WAVE) continue;
        if (!set) {
          op = f.op;
          arg = f.arg;
          set = true;
        } else if (f.op != op || (op == OP_READLANE && f.arg != arg)) {
          g_cur = &f;
          die("divergent wavefront operation");
        }
      }
      unsigned long long result = 0;
      if (op == OP_ANY) {
        for (uint32_t t = l0; t < l1; t++)
          if (g_fibers[t].state == AT_WAVE && g_fibers[t].value) result = 1;
      } else if (op == OP_BALLOT) {
        for (uint32_t t = l0; t < l1; t++)
          if (g_fibers[t].state == AT_WAVE && g_fibers[t].value) result |= 1ull << (t - l0);
      } else {
        if (arg < 0 || (uint32_t)arg >= kWave) die("readlane lane out of range");
        uint32_t src = l0 + (uint32_t)arg;
        if (src >= l1 || g_fibers[src].state != AT_WAVE) die("readlane of an inactive lane");
        result = (unsigned long long)(unsigned)g_fibers[src].value;
      }
      for (uint32_t t = l0; t < l1; t++) {
        Fiber& f = g_fibers[t];
        if (f.state == AT_WAVE) {
          f.result = result;
          f.state = RUNNABLE;
        }
      }
      released = true;
    }
    if (released) continue;
    uint32_t live = 0, at_barrier = 0, at_wave = 0;
    for (uint32_t t = 0; t < nthreads; t++) {
      State s = g_fibers[t].state;
      if (s != DONE) live++;
      if (s == AT_BARRIER) at_barrier++;
      if (s == AT_WAVE) at_wave++;
    }
    any_live = live > 0;
    if (!any_live) break;
    if (at_wave) die("wavefront operation reached by only part of a wavefront");
    if (at_barrier == live) {
      for (uint32_t t = 0; t < nthreads; t++)
        if (g_fibers[t].state == AT_BARRIER) g_fibers[t].state = RUNNABLE;
      continue;
    }
    die("scheduler stalled");
  }
}

}  // namespace

// Work recorded by a stream capture on this host thread (null: none).
thread_local std::vector<std::function<void()>>* g_capture = nullptr;

void run_launch(dim3 grid, dim3 block, const std::function<void()>& body);

void launch(dim3 grid, dim3 block, const std::function<void()>& body) {
  if (g_capture) {
    std::function<void()> b = body;
    g_capture->push_back([grid, block, b]() { run_launch(grid, block, b); });
    return;
  }
  run_launch(grid, block, body);
}

void run_launch(dim3 grid, dim3 block, const std::function<void()>& body) {
  std::lock_guard<std::mutex> lk(g_launch_mu);
  if (block.y != 1 || block.z != 1 || grid.y != 1 || grid.z != 1) die("only 1-D launches are emulated");
  if (block.x == 0 || block.x > kMaxThreads) die("block size out of range");
  if (!g_stacks) {
    g_stacks = (char*)std::aligned_alloc(4096, kStack * kMaxThreads);
    if (!g_stacks) die("out of memory for fiber stacks");
  }
  g_body = &body;
  g_bdim = block;
  g_gdim = grid;
  for (uint32_t b = 0; b < grid.x; b++) {
    g_block = dim3(b, 0, 0);
    run_block(block.x);
  }
  g_cur = nullptr;
  g_body = nullptr;
}

void block_barrier() {
  g_cur->state = AT_BARRIER;
  yield_to_scheduler();
}

int wave_any(int pred) {
  g_cur->state = AT_WAVE;
  g_cur->op = OP_ANY;
  g_cur->value = pred;
  yield_to_scheduler();
  return (int)g_cur->result;
}

unsigned long long wave_ballot(int pred) {
  g_cur->state = AT_WAVE;
  g_cur->op = OP_BALLOT;
  g_cur->value = pred;
  yield_to_scheduler();
  return g_cur->result;
}

int readlane(int v, int lane) {
  g_cur->state = AT_WAVE;
  g_cur->op = OP_READLANE;
  g_cur->value = v;
  g_cur->arg = lane;
  yield_to_scheduler();
  return (int)(unsigned)g_cur->result;
}

}  // namespace flemu

// ------------------------------------------------------------- runtime ---

struct ihipStream_t {
  int dummy;
};
struct ihipEvent_t {
  std::chrono::steady_clock::time_point t;
};

namespace {
// Fresh device allocations hold a NaN-like pattern, so a kernel that reads
// memory nobody wrote gives visibly wrong results.
void poison(void* p, size_t bytes) { std::memset(p, 0xff, bytes); }
}  // namespace

hipError_t hipGetDeviceCount(int* count) {
  *count = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int device) { return device == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipMalloc(void** p, size_t bytes) {
  *p = std::malloc(bytes ? bytes : 1);
  if (!*p) return hipErrorOutOfMemory;
  poison(*p, bytes);
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
struct ihipGraph_t {
  std::vector<std::function<void()>> work;
};
struct ihipGraphExec_t {
  std::vector<std::function<void()>> work;
};

hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t) {
  if (flemu::g_capture) {
    flemu::g_capture->push_back([=]() {
      if (bytes) std::memmove(dst, src, bytes);
    });
    return hipSuccess;
  }
  if (bytes) std::memmove(dst, src, bytes);
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  if (flemu::g_capture) return hipErrorInvalidValue;
  flemu::g_capture = new std::vector<std::function<void()>>();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* graph) {
  if (!flemu::g_capture) return hipErrorInvalidValue;
  *graph = new ihipGraph_t{std::move(*flemu::g_capture)};
  delete flemu::g_capture;
  flemu::g_capture = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t graph, hipGraphNode_t*, char*, size_t) {
  *exec = new ihipGraphExec_t{graph->work};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t) {
  for (auto& f : exec->work) f();
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t graph) {
  delete graph;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
  delete exec;
  return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  return hipMemcpyAsync(dst, src, bytes, kind, nullptr);
}
hipError_t hipMemsetAsync(void* dst, int value, size_t bytes, hipStream_t) {
  if (flemu::g_capture) return hipErrorInvalidValue;  // no memset nodes in captured work
  if (bytes) std::memset(dst, value, bytes);
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = new ihipStream_t();
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new ihipEvent_t();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) {
  switch (e) {
    case hipSuccess:
      return "hipSuccess";
    case hipErrorOutOfMemory:
      return "hipErrorOutOfMemory";
    case hipErrorInvalidValue:
      return "hipErrorInvalidValue";
    default:
      return "hipError (emulated)";
  }
}
