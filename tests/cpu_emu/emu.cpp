// Thread-index state, launch order and block-level barriers of the CPU stand-in
// (TEST INFRASTRUCTURE).  The threads of one workgroup run as fibers (ucontext) scheduled
// round-robin; __syncthreads() and the cross-lane __shfl_down() switch to the next fiber,
// so a workgroup's threads meet at every barrier as they do on the GPU.
#include "hip/hip_runtime.h"

#include <sanitizer/asan_interface.h>
#include <ucontext.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

// hipMalloc'd blocks (start -> size) for hipMemGetAddressRange
static std::map<uintptr_t, size_t> g_allocs;
static std::mutex g_alloc_mu;
void emu_track(void *p, size_t n) {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_allocs[(uintptr_t)p] = n;
}
void emu_untrack(void *p) {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_allocs.erase((uintptr_t)p);
}
hipError_t hipMemGetAddressRange(hipDeviceptr_t *base, size_t *size, hipDeviceptr_t p) {
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  auto it = g_allocs.upper_bound((uintptr_t)p);
  if (it == g_allocs.begin()) return 1;
  --it;
  if ((uintptr_t)p >= it->first + it->second) return 1;
  *base = (void *)it->first;
  *size = it->second;
  return hipSuccess;
}

thread_local emu_uint3 blockIdx, threadIdx;
thread_local dim3 blockDim, gridDim;

static int g_order = 0;
int emu_launch_order() { return g_order; }
extern "C" void emu_set_launch_order(int order) { g_order = order; }

namespace {
struct Fiber {
  ucontext_t ctx;
  char *stack = nullptr;
  emu_uint3 tid;
  bool done = false;
};
const size_t kStack = 128 * 1024;
std::vector<char *> g_stack_pool;  // one stack per thread slot, allocated once
ucontext_t g_sched;
std::vector<Fiber> g_fibers;
int g_cur = -1;
const std::function<void()> *g_body = nullptr;
void *g_fake_stack = nullptr;
float g_shfl[1024];

void fiber_entry() {
  Fiber &f = g_fibers[g_cur];
  __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
  threadIdx = f.tid;
  (*g_body)();
  f.done = true;
  __sanitizer_start_switch_fiber(nullptr, nullptr, 0);
  swapcontext(&f.ctx, &g_sched);
}

void yield_to_scheduler() {
  Fiber &f = g_fibers[g_cur];
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, nullptr, 0);
  swapcontext(&f.ctx, &g_sched);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
  threadIdx = f.tid;
}
}  // namespace

void emu_run_block(long nt, bool rev, dim3 b, const std::function<void()> &body) {
  while ((long)g_stack_pool.size() < nt) g_stack_pool.push_back((char *)malloc(kStack));
  g_fibers.assign(nt, Fiber());
  for (long ot = 0; ot < nt; ot++) {
    const long tt = rev ? nt - 1 - ot : ot;
    Fiber &f = g_fibers[ot];
    f.tid = emu_uint3{(unsigned)(tt % b.x), (unsigned)((tt / b.x) % b.y), (unsigned)(tt / ((long)b.x * b.y))};
    f.stack = g_stack_pool[ot];
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack;
    f.ctx.uc_stack.ss_size = kStack;
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, fiber_entry, 0);
  }
  g_body = &body;
  bool any = true;
  while (any) {
    any = false;
    for (long n = 0; n < nt; n++) {
      if (g_fibers[n].done) continue;
      g_cur = (int)n;
      void *fake = nullptr;
      __sanitizer_start_switch_fiber(&fake, g_fibers[n].stack, kStack);
      swapcontext(&g_sched, &g_fibers[n].ctx);
      __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
      if (!g_fibers[n].done) any = true;
    }
  }
  g_fibers.clear();
  g_cur = -1;
}

void __syncthreads() { yield_to_scheduler(); }

float __shfl_down(float v, unsigned delta, int width) {
  const unsigned tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  g_shfl[tid] = v;
  __syncthreads();
  const unsigned lane = tid % 64u;
  const unsigned src = (lane % (unsigned)width) + delta < (unsigned)width ? tid + delta : tid;
  const float r = g_shfl[src];
  __syncthreads();
  return r;
}
