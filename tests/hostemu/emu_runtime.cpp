// Fiber scheduler of the host emulation build (test infrastructure; see
// hip/hip_runtime.h for the model).  x86-64 System V only.
#include "hip/hip_runtime.h"

#include <algorithm>
#include <cstdio>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#else
#define EMU_ASAN 0
#endif

// Save the callee-saved registers on the current stack, store its stack
// pointer in *save_sp, load new_sp and restore the registers saved there.
asm(R"(
  .text
  .globl hipemu_switch
  .type hipemu_switch,@function
hipemu_switch:
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
  .size hipemu_switch, .-hipemu_switch
)");

namespace hipemu {

static thread_local Group* t_group = nullptr;

Group& group() { return *t_group; }

static void to_scheduler()
{
  Group& g = *t_group;
  Fiber& f = g.fib[g.cur];
#if EMU_ASAN
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, g.sched_stack_bottom, g.sched_stack_size);
  hipemu_switch(&f.sp, g.sched_sp);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#else
  hipemu_switch(&f.sp, g.sched_sp);
#endif
}

void yield() { to_scheduler(); }

static void fiber_main()
{
#if EMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &t_group->sched_stack_bottom, &t_group->sched_stack_size);
#endif
  Group& g = *t_group;
  (*g.body)();
  // the work-item has left the kernel
  const int id = g.cur;
  Wave& w = g.waves[id / kLanes];
  w.alive[id % kLanes] = false;
  drop(w.bar);
  drop(g.sync);
  g.fib[id].done = true;
  for (;;) to_scheduler();
}

static void prepare(Fiber& f)
{
  if (!f.stack) f.stack = static_cast<char*>(std::malloc(kFiberStack));
  uintptr_t top = (reinterpret_cast<uintptr_t>(f.stack) + kFiberStack) & ~uintptr_t(15);
  void** sp = reinterpret_cast<void**>(top);
  *--sp = nullptr;                                  // fake return address of fiber_main
  *--sp = reinterpret_cast<void*>(&fiber_main);     // popped by ret
  for (int r = 0; r < 6; ++r) *--sp = nullptr;      // rbp rbx r12 r13 r14 r15
  f.sp = sp;
  f.done = false;
}

static void run_group(Group& g, unsigned block)
{
  g.sync = Bar();
  g.sync.expected = g.nthreads;
  const int nw = (g.nthreads + kLanes - 1) / kLanes;
  for (int w = 0; w < nw; ++w) {
    const int lanes = std::min(kLanes, g.nthreads - w * kLanes);
    g.waves[w].bar = Bar();
    g.waves[w].bar.expected = lanes;
    for (int l = 0; l < kLanes; ++l) g.waves[w].alive[l] = l < lanes;
  }
  for (int i = 0; i < g.nthreads; ++i) prepare(g.fib[i]);
  blockIdx = dim3(block, 0, 0);
  int live = g.nthreads;
  while (live > 0) {
    for (int i = 0; i < g.nthreads; ++i) {
      Fiber& f = g.fib[i];
      if (f.done) continue;
      g.cur = i;
      threadIdx = dim3((unsigned)i, 0, 0);
#if EMU_ASAN
      void* fake = nullptr;
      __sanitizer_start_switch_fiber(&fake, f.stack, kFiberStack);
      hipemu_switch(&g.sched_sp, f.sp);
      __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#else
      hipemu_switch(&g.sched_sp, f.sp);
#endif
      if (f.done) --live;
    }
  }
}

// OS worker pool: each worker runs whole workgroups.
class Pool {
 public:
  static Pool& get() { static Pool* p = new Pool; return *p; }   // never destroyed
  void run(std::function<void()>& body, unsigned nblocks, int nthreads)
  {
    std::unique_lock<std::mutex> lk(mu_);
    body_ = &body;
    nblocks_ = nblocks;
    nthreads_ = nthreads;
    next_.store(0);
    dim_ = blockDim;
    remaining_ = (int)workers_.size();
    ++gen_;
    cv_.notify_all();
    done_cv_.wait(lk, [this] { return remaining_ == 0; });
    body_ = nullptr;
  }

 private:
  Pool()
  {
    int nw = (int)std::thread::hardware_concurrency();
    if (const char* e = std::getenv("HIPEMU_THREADS")) nw = std::atoi(e);
    nw = std::max(1, std::min(nw, 64));
    for (int i = 0; i < nw; ++i) workers_.emplace_back([this] { worker(); });
  }
  void worker()
  {
    Group* g = new Group();
    t_group = g;
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
      }
      blockDim = dim_;
      g->body = body_;
      g->nthreads = nthreads_;
      for (;;) {
        const unsigned b = next_.fetch_add(1);
        if (b >= nblocks_) break;
        run_group(*g, b);
      }
      std::lock_guard<std::mutex> lk(mu_);
      if (--remaining_ == 0) done_cv_.notify_all();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::vector<std::thread> workers_;
  std::function<void()>* body_ = nullptr;
  unsigned nblocks_ = 0;
  int nthreads_ = 0;
  dim3 dim_;
  std::atomic<unsigned> next_{0};
  uint64_t gen_ = 0;
  int remaining_ = 0;
};

void run_launch(std::function<void()>& body, unsigned nblocks, int nthreads)
{
  Pool::get().run(body, nblocks, nthreads);
}

}  // namespace hipemu
