// CPU runtime behind tests/cpuemu/hip/hip_runtime.h (test infrastructure only).
//
// Kernels: every work-item of a block is a fiber with its own stack; the
// fibers of a block run until each waits at a barrier, at a wave operation or
// has returned.  A wave operation runs once every live lane of the wave waits
// at it (lanes that returned do not take part, like lanes outside EXEC); a
// barrier releases once every live work-item of the block waits at it.  A
// block in which neither can happen is reported as a deadlock.
//
// Device memory: every allocation gets its own mapping, with the buffer placed
// against a PROT_NONE guard page (its end by default, its start with
// W2K_EMU_ALIGN=start) and filled with the byte 0x41 (doubles of about 2e6,
// ints of about 1.1e9, so a never-written index runs far out of bounds), or
// with 0xFF (NaN doubles, ints of -1) when W2K_EMU_POISON=nan.  A
// freed buffer is unmapped to PROT_NONE and its address range is never
// reused.  A fault prints the kernel, block, work-item and the buffer that
// the address belongs to (or lies past), then ends the process with status 86.
#include <cerrno>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <map>
#include <mutex>
#include <vector>
#include <sys/mman.h>
#include <unistd.h>

#include "hip/hip_runtime.h"

// ------------------------------------------------------------------ fiber switch (x86-64 SysV)
extern "C" void emu_switch(void **save_sp, void *new_sp);
asm(R"(
  .text
  .globl emu_switch
  .type emu_switch,@function
emu_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  subq $8, %rsp
  stmxcsr (%rsp)
  fnstcw 4(%rsp)
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  ldmxcsr (%rsp)
  fldcw 4(%rsp)
  addq $8, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
  .size emu_switch, .-emu_switch
)");

namespace emu {

thread_local Item *cur = nullptr;
thread_local dim3 cur_bdim, cur_gdim;
thread_local emu_uint3 cur_bidx;

namespace {

constexpr size_t STACK = 256 * 1024;
constexpr int MAXT = 1024;

enum State { READY = 0, BARRIER = 1, WAVEOP = 2, DONE = 3 };

enum Op {
  OP_SHFL = 1, OP_SHFL_DOWN, OP_SHFL_XOR, OP_BALLOT, OP_READLANE, OP_DPP, OP_SWIZZLE, OP_MFMA_F64, OP_MFMA_I8
};

struct Fiber {
  Item item;
  void *sp = nullptr;
  char *stack = nullptr;
  int state = READY;
  int op = 0;
  uint64_t in[12];
  uint64_t out[4];
};

struct BlockRun {
  Fiber *f = nullptr;
  int nthreads = 0;
  void *sched_sp = nullptr;
  LaunchBody *body = nullptr;
  Fiber *running = nullptr;
};

thread_local BlockRun *g_block = nullptr;
thread_local std::vector<char *> *g_stacks = nullptr;
thread_local const char *g_kernel = "(host)";

// kernel and work-item being run, for the fault report (last writer wins
// across host threads; the tests that use the emulation run one at a time)
const char *volatile g_fault_kernel = "(host)";
volatile unsigned g_fault_blk[3], g_fault_tid[3];

void fiber_entry()
{
  BlockRun *b = g_block;
  b->body->run();
  b->running->state = DONE;
  emu_switch(&b->running->sp, b->sched_sp);
  __builtin_trap();
}

char *new_stack()
{
  void *p = mmap(nullptr, STACK, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  if (p == MAP_FAILED) {
    perror("emu stack");
    abort();
  }
  mprotect(p, 4096, PROT_NONE);  // stack overflow guard
  return (char *)p;
}

void fiber_init(Fiber &f)
{
  // initial frame for emu_switch: mxcsr/fcw slot, six callee-saved registers,
  // the return address (fiber_entry) and one pad slot, so that fiber_entry
  // starts with rsp = 8 (mod 16) as after a call
  uint64_t *top = (uint64_t *)(f.stack + STACK);
  uint64_t *sp = top;
  *--sp = 0;                              // pad
  *--sp = (uint64_t)(void *)&fiber_entry; // return address
  for (int r = 0; r < 6; r++) *--sp = 0;  // rbp rbx r12 r13 r14 r15
  --sp;
  uint32_t mx;
  uint16_t cw;
  asm volatile("stmxcsr %0" : "=m"(mx));
  asm volatile("fnstcw %0" : "=m"(cw));
  std::memcpy((char *)sp, &mx, 4);
  std::memcpy((char *)sp + 4, &cw, 2);
  f.sp = sp;
  f.state = READY;
}

void yield(int state)
{
  BlockRun *b = g_block;
  Fiber *f = b->running;
  f->state = state;
  emu_switch(&f->sp, b->sched_sp);
}

[[noreturn]] void fail(const char *what)
{
  fprintf(stderr, "EMU ERROR: %s in kernel %s block (%u,%u,%u)\n", what, g_kernel, cur_bidx.x, cur_bidx.y,
          cur_bidx.z);
  fflush(stderr);
  _exit(87);
}

uint64_t run_op(Fiber *f)
{
  // called in the fiber: deposit, wait, return the result
  yield(WAVEOP);
  return f->out[0];
}

// perform the wave operation of the live lanes [w0, w0 + 64) that wait at it
void do_waveop(BlockRun &b, int w0, int wn)
{
  Fiber *L[64] = {nullptr};
  int op = 0;
  for (int q = 0; q < wn; q++) {
    Fiber &f = b.f[w0 + q];
    if (f.state == DONE) continue;
    L[q] = &f;
    if (op == 0) op = f.op;
    else if (f.op != op) fail("lanes of one wave wait at different wave operations");
  }
  auto val = [&](int lane, int k) -> uint64_t { return (lane >= 0 && lane < 64 && L[lane]) ? L[lane]->in[k] : 0; };
  auto live = [&](int lane) { return lane >= 0 && lane < 64 && L[lane] != nullptr; };
  switch (op) {
    case OP_SHFL:
      for (int q = 0; q < 64; q++)
        if (L[q]) {
          const int src = (int)L[q]->in[1] & 63;
          if (!live(src)) fail("__shfl from a lane that has returned");
          L[q]->out[0] = val(src, 0);
        }
      break;
    case OP_SHFL_DOWN:
      for (int q = 0; q < 64; q++)
        if (L[q]) {
          const int src = q + (int)L[q]->in[1];
          L[q]->out[0] = (src < 64) ? val(src, 0) : L[q]->in[0];
        }
      break;
    case OP_SHFL_XOR:
      for (int q = 0; q < 64; q++)
        if (L[q]) {
          const int src = q ^ ((int)L[q]->in[1] & 63);
          if (!live(src)) fail("__shfl_xor from a lane that has returned");
          L[q]->out[0] = val(src, 0);
        }
      break;
    case OP_BALLOT: {
      uint64_t m = 0;
      for (int q = 0; q < 64; q++)
        if (L[q] && L[q]->in[0]) m |= 1ull << q;
      for (int q = 0; q < 64; q++)
        if (L[q]) L[q]->out[0] = m;
      break;
    }
    case OP_READLANE: {
      int src = -1;
      for (int q = 0; q < 64; q++)
        if (L[q]) {
          if (src < 0) src = (int)L[q]->in[1];
          else if ((int)L[q]->in[1] != src) fail("readlane with a lane index that differs across the wave");
        }
      if (src < 0 || src > 63) fail("readlane lane index out of range");
      // a lane that has returned keeps its last value on the device; here it reads 0
      for (int q = 0; q < 64; q++)
        if (L[q]) L[q]->out[0] = val(src, 0);
      break;
    }
    case OP_DPP:
      for (int q = 0; q < 64; q++)
        if (L[q]) {
          const int ctrl = (int)L[q]->in[2];
          const int r = q & 15, row = q & ~15;
          int src = -1;
          if (ctrl >= 0 && ctrl <= 0xFF) {  // quad_perm
            const int sel = (ctrl >> (2 * (r & 3))) & 3;
            src = (q & ~3) + sel;
          } else if (ctrl == 0x140) {  // row_mirror
            src = row + (15 - r);
          } else if (ctrl == 0x141) {  // row_half_mirror
            src = row + (r & 8) + (7 - (r & 7));
          } else {
            fail("unsupported DPP control");
          }
          if (!live(src)) fail("DPP move from a lane that has returned");
          L[q]->out[0] = val(src, 1);
        }
      break;
    case OP_SWIZZLE:
      for (int q = 0; q < 64; q++)
        if (L[q]) {
          const int p = (int)L[q]->in[1];
          if (p & 0x8000) fail("unsupported swizzle mode");
          const int andm = p & 31, orm = (p >> 5) & 31, xorm = (p >> 10) & 31;
          const int src = (q & ~31) + (((q & 31) & andm | orm) ^ xorm);
          if (!live(src)) fail("swizzle from a lane that has returned");
          L[q]->out[0] = val(src, 0);
        }
      break;
    case OP_MFMA_F64: {
      // A(i,k): lane i + 16 k (in[0]); B(k,j): lane j + 16 k (in[1]);
      // D(i,j): lane j + 16 (i % 4), item i / 4 (in[2..5] -> out[0..3])
      for (int q = 0; q < 64; q++)
        if (!L[q]) fail("MFMA with lanes that have returned");
      double D[16][16];
      for (int i = 0; i < 16; i++)
        for (int j = 0; j < 16; j++) D[i][j] = of_bits<double>(L[j + 16 * (i % 4)]->in[2 + i / 4]);
      for (int i = 0; i < 16; i++)
        for (int j = 0; j < 16; j++) {
          double s = D[i][j];
          for (int k = 0; k < 4; k++)
            s += of_bits<double>(L[i + 16 * k]->in[0]) * of_bits<double>(L[j + 16 * k]->in[1]);
          D[i][j] = s;
        }
      for (int q = 0; q < 64; q++)
        for (int r = 0; r < 4; r++) L[q]->out[r] = bits_of(D[4 * r + q / 16][q % 16]);
      break;
    }
    case OP_MFMA_I8: {
      // A(i,k): lane i + 16 (k / 16), byte k % 16 (in[0..1]); B likewise
      // (in[2..3]); D(i,j): lane j + 16 (i / 4), item i % 4 (in[4..7])
      for (int q = 0; q < 64; q++)
        if (!L[q]) fail("MFMA with lanes that have returned");
      auto byte = [&](int lane, int base, int k) -> int {
        const uint64_t w = L[lane]->in[base + k / 8];
        return (int)(int8_t)((w >> (8 * (k % 8))) & 0xff);
      };
      for (int i = 0; i < 16; i++)
        for (int j = 0; j < 16; j++) {
          const int dl = j + 16 * (i / 4), di = i % 4;
          int64_t s = (int32_t)(uint32_t)L[dl]->in[4 + di];
          for (int k = 0; k < 64; k++) s += (int64_t)byte(i + 16 * (k / 16), 0, k % 16) * byte(j + 16 * (k / 16), 2, k % 16);
          L[dl]->out[di] = (uint64_t)(uint32_t)(int32_t)s;
        }
      break;
    }
    default:
      fail("unknown wave operation");
  }
  for (int q = 0; q < 64; q++)
    if (L[q]) L[q]->state = READY;
}

void run_block(BlockRun &b)
{
  for (;;) {
    // every ready work-item runs until it waits or returns
    for (int t = 0; t < b.nthreads; t++) {
      Fiber &f = b.f[t];
      if (f.state != READY) continue;
      b.running = &f;
      cur = &f.item;
      g_fault_tid[0] = f.item.tid.x;
      g_fault_tid[1] = f.item.tid.y;
      g_fault_tid[2] = f.item.tid.z;
      emu_switch(&b.sched_sp, f.sp);
    }
    b.running = nullptr;
    cur = nullptr;
    // wave operations whose live lanes all wait at them
    bool op_done = false;
    for (int w0 = 0; w0 < b.nthreads; w0 += 64) {
      const int wn = std::min(64, b.nthreads - w0);
      int live = 0, wait = 0;
      for (int q = 0; q < wn; q++) {
        const int s = b.f[w0 + q].state;
        if (s != DONE) live++;
        if (s == WAVEOP) wait++;
      }
      if (live > 0 && wait == live) {
        do_waveop(b, w0, wn);
        op_done = true;
      }
    }
    if (op_done) continue;
    int alive = 0, at_bar = 0;
    for (int t = 0; t < b.nthreads; t++) {
      if (b.f[t].state != DONE) alive++;
      if (b.f[t].state == BARRIER) at_bar++;
    }
    if (alive == 0) return;
    if (at_bar == alive) {
      for (int t = 0; t < b.nthreads; t++)
        if (b.f[t].state == BARRIER) b.f[t].state = READY;
      continue;
    }
    fail("deadlock: work-items wait at a barrier while others of the block wait at a wave operation");
  }
}

}  // namespace

void syncthreads()
{
  if (!g_block || !g_block->running) fail("__syncthreads outside a kernel");
  yield(BARRIER);
}

static uint64_t wave_op(int op, const uint64_t *in, int nin, uint64_t *out4 = nullptr)
{
  BlockRun *b = g_block;
  if (!b || !b->running) fail("wave operation outside a kernel");
  Fiber *f = b->running;
  f->op = op;
  for (int k = 0; k < nin; k++) f->in[k] = in[k];
  yield(WAVEOP);
  if (out4)
    for (int r = 0; r < 4; r++) out4[r] = f->out[r];
  return f->out[0];
}

u2 permlane_swap(unsigned x, unsigned y, int s)
{
  const int lane = cur->lane;
  const unsigned px = (unsigned)wave_shfl((uint64_t)x, lane ^ s);
  const unsigned py = (unsigned)wave_shfl((uint64_t)y, lane ^ s);
  u2 r;
  if (lane & s) {
    r[0] = py;  // the first register's upper group receives the second's lower group
    r[1] = y;
  } else {
    r[0] = x;
    r[1] = px;  // the second register's lower group receives the first's upper group
  }
  return r;
}

uint64_t wave_shfl(uint64_t v, int src)
{
  const uint64_t in[2] = {v, (uint64_t)(uint32_t)src};
  return wave_op(OP_SHFL, in, 2);
}
uint64_t wave_shfl_down(uint64_t v, int delta)
{
  const uint64_t in[2] = {v, (uint64_t)(uint32_t)delta};
  return wave_op(OP_SHFL_DOWN, in, 2);
}
uint64_t wave_shfl_xor(uint64_t v, int mask)
{
  const uint64_t in[2] = {v, (uint64_t)(uint32_t)mask};
  return wave_op(OP_SHFL_XOR, in, 2);
}
uint64_t wave_ballot(int pred)
{
  const uint64_t in[1] = {(uint64_t)(pred != 0)};
  return wave_op(OP_BALLOT, in, 1);
}
int wave_readlane(int v, int src)
{
  const uint64_t in[2] = {(uint64_t)(uint32_t)v, (uint64_t)(uint32_t)src};
  return (int)(uint32_t)wave_op(OP_READLANE, in, 2);
}
int wave_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl)
{
  if (row_mask != 0xF || bank_mask != 0xF) fail("DPP with partial row or bank mask");
  const uint64_t in[3] = {(uint64_t)(uint32_t)old, (uint64_t)(uint32_t)src, (uint64_t)(uint32_t)ctrl};
  return (int)(uint32_t)wave_op(OP_DPP, in, 3);
}
int wave_swizzle(int v, int pattern)
{
  const uint64_t in[2] = {(uint64_t)(uint32_t)v, (uint64_t)(uint32_t)pattern};
  return (int)(uint32_t)wave_op(OP_SWIZZLE, in, 2);
}
d4 wave_mfma_f64(double a, double b, d4 c)
{
  const uint64_t in[6] = {bits_of(a), bits_of(b), bits_of((double)c[0]), bits_of((double)c[1]),
                          bits_of((double)c[2]), bits_of((double)c[3])};
  uint64_t o[4];
  wave_op(OP_MFMA_F64, in, 6, o);
  return d4{of_bits<double>(o[0]), of_bits<double>(o[1]), of_bits<double>(o[2]), of_bits<double>(o[3])};
}
v4i wave_mfma_i8(v4i a, v4i b, v4i c)
{
  auto pack = [](int lo, int hi) { return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32); };
  const uint64_t in[8] = {pack(a[0], a[1]), pack(a[2], a[3]), pack(b[0], b[1]), pack(b[2], b[3]),
                          (uint64_t)(uint32_t)c[0], (uint64_t)(uint32_t)c[1], (uint64_t)(uint32_t)c[2],
                          (uint64_t)(uint32_t)c[3]};
  uint64_t o[4];
  wave_op(OP_MFMA_I8, in, 8, o);
  return v4i{(int)(uint32_t)o[0], (int)(uint32_t)o[1], (int)(uint32_t)o[2], (int)(uint32_t)o[3]};
}

static thread_local hipError_t g_last = hipSuccess;

void launch(const char *name, dim3 grid, dim3 block, hipStream_t, LaunchBody &body)
{
  const long nthreads = (long)block.x * block.y * block.z;
  if (grid.x == 0 || grid.y == 0 || grid.z == 0 || nthreads <= 0 || nthreads > MAXT || grid.y > 65535 ||
      grid.z > 65535) {
    fprintf(stderr, "EMU: invalid launch configuration of %s: grid (%u,%u,%u) block (%u,%u,%u)\n", name, grid.x,
            grid.y, grid.z, block.x, block.y, block.z);
    g_last = hipErrorInvalidConfiguration;
    return;
  }
  if (g_block) fail("kernel launch inside a kernel");
  if (!g_stacks) g_stacks = new std::vector<char *>();
  while ((long)g_stacks->size() < nthreads) g_stacks->push_back(new_stack());
  std::vector<Fiber> fibers((size_t)nthreads);
  BlockRun b;
  b.f = fibers.data();
  b.nthreads = (int)nthreads;
  b.body = &body;
  g_block = &b;
  g_kernel = name;
  g_fault_kernel = name;
  cur_bdim = block;
  cur_gdim = grid;
  for (unsigned bz = 0; bz < grid.z; bz++)
    for (unsigned by = 0; by < grid.y; by++)
      for (unsigned bx = 0; bx < grid.x; bx++) {
        cur_bidx = {bx, by, bz};
        g_fault_blk[0] = bx;
        g_fault_blk[1] = by;
        g_fault_blk[2] = bz;
        int t = 0;
        for (unsigned z = 0; z < block.z; z++)
          for (unsigned y = 0; y < block.y; y++)
            for (unsigned x = 0; x < block.x; x++, t++) {
              Fiber &f = fibers[t];
              f.item.tid = {x, y, z};
              f.item.lane = t & 63;
              f.stack = (*g_stacks)[t];
              fiber_init(f);
            }
        run_block(b);
      }
  g_block = nullptr;
  g_kernel = "(host)";
  g_fault_kernel = "(host)";
}

hipError_t take_last()
{
  const hipError_t e = g_last;
  g_last = hipSuccess;
  return e;
}

}  // namespace emu

// ------------------------------------------------------------------ device memory
namespace {

struct Region {
  char *base;
  size_t total;
  char *user;
  size_t bytes;
  bool freed;
};
std::mutex g_mem_mu;
std::map<char *, Region> g_regions;  // by base
const size_t PAGE = 4096;

bool align_start()
{
  static const bool s = getenv("W2K_EMU_ALIGN") && strcmp(getenv("W2K_EMU_ALIGN"), "start") == 0;
  return s;
}

void segv_handler(int sig, siginfo_t *si, void *)
{
  char *a = (char *)si->si_addr;
  char msg[768];
  int len = snprintf(msg, sizeof msg, "EMU FAULT: signal %d at address %p in kernel %s block (%u,%u,%u) item (%u,%u,%u)\n",
                     sig, (void *)a, emu::g_fault_kernel, emu::g_fault_blk[0], emu::g_fault_blk[1],
                     emu::g_fault_blk[2], emu::g_fault_tid[0], emu::g_fault_tid[1], emu::g_fault_tid[2]);
  if (write(2, msg, len) < 0) {}
  // the region the address belongs to (no locking: the process ends here)
  for (auto &kv : g_regions) {
    const Region &r = kv.second;
    if (a >= r.base && a < r.base + r.total) {
      const long off = (long)(a - r.user);
      len = snprintf(msg, sizeof msg, "EMU FAULT: buffer of %zu bytes at %p%s: byte offset %ld (%s)\n", r.bytes,
                     (void *)r.user, r.freed ? " (FREED)" : "", off,
                     r.freed ? "use after free" : (off < 0 ? "before the start" : "past the end"));
      if (write(2, msg, len) < 0) {}
    }
  }
  _exit(86);
}

struct Init {
  Init()
  {
    static char altstack[1 << 16];
    stack_t ss;
    ss.ss_sp = altstack;
    ss.ss_size = sizeof altstack;
    ss.ss_flags = 0;
    sigaltstack(&ss, nullptr);
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = segv_handler;
    sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
    sigaction(SIGSEGV, &sa, nullptr);
    sigaction(SIGBUS, &sa, nullptr);
  }
} g_init;

}  // namespace

// fill byte of new allocations: 0x41 by default; W2K_EMU_POISON=nan fills
// with 0xFF (every double a NaN, every int -1), so that a value read before
// it was written and then "masked" by a multiplication with zero still shows
static int poison_byte()
{
  static const int b = [] {
    const char *e = getenv("W2K_EMU_POISON");
    return (e && strcmp(e, "nan") == 0) ? 0xFF : 0x41;
  }();
  return b;
}

hipError_t hipMalloc(void **p, size_t bytes)
{
  if (!p) return hipErrorInvalidValue;
  const size_t data = ((bytes ? bytes : 1) + 255) & ~(size_t)255;
  const size_t pages = (data + PAGE - 1) / PAGE * PAGE;
  const size_t total = pages + 2 * PAGE;
  char *base = (char *)mmap(nullptr, total, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  if (base == MAP_FAILED) return hipErrorOutOfMemory;
  if (mprotect(base + PAGE, pages, PROT_READ | PROT_WRITE) != 0) {
    munmap(base, total);
    return hipErrorOutOfMemory;
  }
  char *user = align_start() ? base + PAGE : base + PAGE + pages - data;
  memset(base + PAGE, poison_byte(), pages);
  {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_regions[base] = Region{base, total, user, bytes, false};
  }
  *p = user;
  return hipSuccess;
}

hipError_t hipFree(void *p)
{
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  for (auto &kv : g_regions) {
    Region &r = kv.second;
    if (r.user == p && !r.freed) {
      madvise(r.base, r.total, MADV_DONTNEED);
      mprotect(r.base, r.total, PROT_NONE);
      r.freed = true;
      return hipSuccess;
    }
  }
  fprintf(stderr, "EMU: hipFree of %p, which is not a live allocation\n", p);
  return hipErrorInvalidValue;
}

hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind)
{
  if (bytes) memmove(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t)
{
  return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemcpy2DAsync(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                            hipMemcpyKind, hipStream_t)
{
  if (width > dpitch || width > spitch) return hipErrorInvalidValue;
  for (size_t r = 0; r < height; r++) memmove((char *)dst + r * dpitch, (const char *)src + r * spitch, width);
  return hipSuccess;
}
hipError_t hipMemset(void *dst, int v, size_t bytes)
{
  if (bytes) memset(dst, v, bytes);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t) { return hipMemset(dst, v, bytes); }
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetDevice(int *d)
{
  *d = 0;
  return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n)
{
  *n = 1;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return emu::take_last(); }
hipError_t hipGetLastError() { return emu::take_last(); }
const char *hipGetErrorString(hipError_t e)
{
  switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorInvalidConfiguration: return "hipErrorInvalidConfiguration";
    case hipErrorNotSupported: return "hipErrorNotSupported (emulation)";
    default: return "hip error (emulation)";
  }
}

struct emu_stream {
  int dummy;
};
struct emu_event {
  double t = 0.0;
};
static double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
  *s = new emu_stream();
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return emu::take_last(); }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e)
{
  *e = new emu_event();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
  if (!e) return hipErrorInvalidValue;
  e->t = now_ms();
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
  if (!a || !b) return hipErrorInvalidValue;
  *ms = (float)(b->t - a->t);
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g)
{
  *g = nullptr;
  return hipErrorNotSupported;
}
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus *st)
{
  *st = hipStreamCaptureStatusNone;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t)
{
  return hipErrorNotSupported;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t emu_get_symbol_address(void **p, const void *sym)
{
  *p = const_cast<void *>(sym);
  return hipSuccess;
}
