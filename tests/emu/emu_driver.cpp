// CPU emulation driver: runs the engine's column/row kernels (compiled from
// gpulucas_amd/csrc/lucas_kernels.hip and lucas_line.hip against
// tests/emu/hip/hip_runtime.h) on host memory for one job.  Tests only.
#include <sys/mman.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "lucas_desc.h"
#include "lucas_host.h"

thread_local emu_idx threadIdx, blockIdx;
emu_idx blockDim;
// the running workgroup's LDS (`extern __shared__` is thread_local here)
thread_local double2 lucas_lds[163840 / sizeof(double2)];

// x86-64 System V context switch: pushes the callee-saved registers and the
// FP control words of the running context, stores its stack pointer in
// *from, and resumes the context whose stack pointer is `to`.
extern "C" void emu_ctx_swap(void** from, void* to);
// first frame of a fiber: the fiber's Fiber* arrives in rbx
extern "C" void emu_fiber_entry();
asm(R"(
    .pushsection .text
    .globl emu_ctx_swap
    .type emu_ctx_swap, @function
emu_ctx_swap:
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
    .size emu_ctx_swap, .-emu_ctx_swap

    .globl emu_fiber_entry
    .type emu_fiber_entry, @function
emu_fiber_entry:
    movq %rbx, %rdi
    andq $-16, %rsp
    call emu_fiber_main@PLT
    ud2
    .size emu_fiber_entry, .-emu_fiber_entry
    .popsection
)");

namespace emu {
std::mutex atom_mu;

namespace {
constexpr size_t kStack = 128 * 1024;  // per fiber (reserved, committed as touched)

struct Group;
struct Fiber {
    Group* g;
    unsigned t;
};
// One host thread's workgroup: its fibers, their saved stack pointers, the
// scheduler's context and the barrier state.
struct Group {
    unsigned n = 0;
    const std::function<void()>* body = nullptr;
    std::vector<void*> sp;
    std::vector<Fiber> fib;
    std::vector<char> done;
    void* sched = nullptr;
    unsigned cur = 0;
    unsigned bar_count = 0, bar_gen = 0;
    std::vector<unsigned> w_count, w_gen;
    std::vector<int> lanes;
    std::vector<float> shfl;
    char* stacks = nullptr;
    size_t stacks_n = 0;  // fibers the stack block holds
};
thread_local Group* t_group = nullptr;

void yield_to_scheduler() {
    Group& G = *t_group;
    emu_ctx_swap(&G.sp[G.cur], G.sched);
}

// a fresh context that starts emu_fiber_entry with rbx = &fib
void* fresh_context(char* stack_top, Fiber* fib) {
    uint64_t* sp = reinterpret_cast<uint64_t*>(reinterpret_cast<uintptr_t>(stack_top) & ~uintptr_t(15));
    *--sp = 0;                                               // alignment slot
    *--sp = reinterpret_cast<uint64_t>(&emu_fiber_entry);    // return address
    *--sp = 0;                                               // rbp
    *--sp = reinterpret_cast<uint64_t>(fib);                 // rbx
    *--sp = 0;                                               // r12
    *--sp = 0;                                               // r13
    *--sp = 0;                                               // r14
    *--sp = 0;                                               // r15
    *--sp = (uint64_t)0x1F80 | ((uint64_t)0x037F << 32);     // mxcsr, x87 control word
    return sp;
}

void run_blocks(Group& G, unsigned first, unsigned step, unsigned nblocks) {
    t_group = &G;
    for (unsigned blk = first; blk < nblocks; blk += step) {
        for (unsigned t = 0; t < G.n; ++t) {
            G.fib[t] = {&G, t};
            G.sp[t] = fresh_context(G.stacks + (size_t)(t + 1) * kStack, &G.fib[t]);
            G.done[t] = 0;
        }
        G.bar_count = 0;
        std::fill(G.w_count.begin(), G.w_count.end(), 0u);
        unsigned left = G.n;
        while (left) {
            for (unsigned t = 0; t < G.n; ++t) {
                if (G.done[t]) continue;
                G.cur = t;
                threadIdx = {t, 0, 0};
                blockIdx = {blk, 0, 0};
                emu_ctx_swap(&G.sched, G.sp[t]);
                if (G.done[t]) --left;
            }
        }
    }
    t_group = nullptr;
}

std::vector<Group>& groups() {
    static std::vector<Group> g;
    return g;
}
}  // namespace

void group_barrier() {
    Group& G = *t_group;
    const unsigned gen = G.bar_gen;
    if (++G.bar_count == G.n) {
        G.bar_count = 0;
        ++G.bar_gen;
        return;
    }
    while (G.bar_gen == gen) yield_to_scheduler();
}

void wave_barrier() {
    Group& G = *t_group;
    const unsigned w = G.cur >> 6, size = std::min(64u, G.n - 64 * w);
    const unsigned gen = G.w_gen[w];
    if (++G.w_count[w] == size) {
        G.w_count[w] = 0;
        ++G.w_gen[w];
        return;
    }
    while (G.w_gen[w] == gen) yield_to_scheduler();
}

int* lane_buf() { return t_group->lanes.data(); }
float* shfl_buf() { return t_group->shfl.data(); }

// Runs `body` for every work-item of nblocks workgroups of nthreads: up to
// 8 host threads, each taking every 8th workgroup, its work-items as fibers.
void run_grid(unsigned nblocks, unsigned nthreads, const std::function<void()>& body) {
    blockDim = {nthreads, 1, 1};
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nw = std::max(1u, std::min({8u, hw, nblocks}));
    auto& gs = groups();
    if (gs.size() < nw) gs.resize(nw);
    for (unsigned g = 0; g < nw; ++g) {
        Group& G = gs[g];
        G.n = nthreads;
        G.body = &body;
        G.sp.assign(nthreads, nullptr);
        G.fib.assign(nthreads, Fiber{});
        G.done.assign(nthreads, 0);
        G.w_count.assign((nthreads + 63) / 64, 0u);
        G.w_gen.assign((nthreads + 63) / 64, 0u);
        G.lanes.assign(nthreads, 0);
        G.shfl.assign(nthreads, 0.f);
        if (G.stacks_n < nthreads) {
            if (G.stacks) munmap(G.stacks, G.stacks_n * kStack);
            void* m = mmap(nullptr, (size_t)nthreads * kStack, PROT_READ | PROT_WRITE,
                           MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
            if (m == MAP_FAILED) std::abort();
            G.stacks = static_cast<char*>(m);
            G.stacks_n = nthreads;
        }
    }
    std::vector<std::thread> ts;
    for (unsigned g = 1; g < nw; ++g) ts.emplace_back([&, g] { run_blocks(gs[g], g, nw, nblocks); });
    run_blocks(gs[0], 0, nw, nblocks);
    for (auto& th : ts) th.join();
}
}  // namespace emu

// body of a fiber: the kernel for its work-item, then back to the scheduler
extern "C" void emu_fiber_main(emu::Fiber* f) {
    (*f->g->body)();
    f->g->done[f->t] = 1;
    void* dead;
    emu_ctx_swap(&dead, f->g->sched);
    std::abort();
}

namespace lucas {
hipError_t launch_col_kernel(int, int, int, const LucasJobDesc*, const int2*, int, size_t, int, hipStream_t);
hipError_t launch_row_kernel(int, int, const LucasJobDesc*, const int2*, int, size_t, hipStream_t);
hipError_t launch_col1_kernel(int, int, int, const LucasJobDesc&, int, size_t, int, hipStream_t);
hipError_t launch_row1_kernel(int, int, const LucasJobDesc&, int, size_t, hipStream_t);
hipError_t launch_line_col(int, const LucasJobDesc*, const int2*, const LucasJobDesc*, int, int, hipStream_t);
hipError_t launch_line_row(int, const LucasJobDesc*, const int2*, const LucasJobDesc*, int, hipStream_t);
}

// 1: launch the single-job kernels (descriptor by value, tile from the
// workgroup id) instead of the batched ones (map + descriptor array).
static int g_single = 0;
extern "C" void emu_set_single(int on) { g_single = on; }
// Wave factors the plan must use (-1: the planner's first choice).
static int g_wA = -1, g_wB = -1;
extern "C" void emu_set_waves(int wA, int wB) { g_wA = wA; g_wB = wB; }

// Line kernels (A rows, K own columns; A < 0: the tiled kernels).
static int g_line_A = -1, g_line_K = 0, g_line_split = -1, g_line_col_split = -1;
extern "C" void emu_set_line(int A, int K, int split, int col_split) {
    g_line_A = A;
    g_line_K = K;
    g_line_split = split;
    g_line_col_split = col_split;
}

// Product-size diagnostic of the last emu_run: max bit length, max 64-bit
// leading-zero count of the rounded products.
static unsigned g_pbits[2] = {0, 0};
extern "C" void emu_product_bits(int* out) {
    out[0] = (int)g_pbits[0];
    out[1] = (int)g_pbits[1];
}

extern "C" int emu_run(int p, int N, int A, int C, int T, const int32_t* digits_in, long long iters,
                       int32_t* digits_out, double* maxerr) {
    using namespace lucas;
    Plan pl;
    if (g_line_A >= 0) {
        if (!make_line_plan(p, N, g_line_A, g_line_K, &pl, 0, g_line_split, g_line_col_split).empty()) return -1;
    } else if (!(T || g_wA >= 0 || g_wB >= 0 ? make_plan_exact(p, N, A, C, T, 0, &pl, g_wA, g_wB)
                                              : make_plan(p, N, A, C, &pl)).empty() &&
               !make_plan_exact(p, N, A, C, T, 0, &pl, g_wA, g_wB, true).empty()) {
        return -1;
    }
    Tables t;
    make_tables(pl, &t);
    std::vector<double2> spec(spec_bytes(pl) / sizeof(double2)), spec2(spec_bytes(pl) / sizeof(double2));
    std::vector<int> dig(digits_in, digits_in + N);
    unsigned err = 0;
    LucasJobDesc d{};
    fill_desc_shape(pl, &d);
    d.twA = (const double2*)t.twA.data(); d.twB = (const double2*)t.twB.data();
    d.twMA = (const double2*)t.twMA.data(); d.twMB = (const double2*)t.twMB.data();
    d.ct1 = (const double2*)t.ct1.data(); d.ct2 = (const double2*)t.ct2.data();
    d.rt1 = (const double2*)t.rt1.data(); d.rt2 = (const double2*)t.rt2.data();
    d.twsA = (const double2*)t.twsA.data(); d.twsB = (const double2*)t.twsB.data();
    d.lnA = (const double2*)t.lnA.data(); d.lnB = (const double2*)t.lnB.data();
    d.rlo = t.rlo.data(); d.rhi = t.rhi.data();
    d.wlo = t.wlo.data(); d.whi = t.whi.data(); d.iwlo = t.iwlo.data(); d.iwhi = t.iwhi.data();
    d.spec = spec.data(); d.spec2 = spec2.data(); d.digits = dig.data(); d.maxerr = &err;
    g_pbits[0] = g_pbits[1] = 0;
    d.pbits = g_pbits;
    std::vector<int2> cmap, rmap;
    for (int w = 0; w < pl.col_wgs; ++w) cmap.push_back({0, w});
    for (int w = 0; w < pl.row_wgs; ++w) rmap.push_back({0, w});
    auto col = [&](int mode) {
        if (pl.line)
            launch_line_col(pl.line_col, &d, cmap.data(), g_single ? &d : nullptr, pl.col_wgs, mode, nullptr);
        else if (g_single)
            launch_col1_kernel(pl.threads, pl.wA / 64, pl.col_dense, d, pl.col_wgs, 0, mode, nullptr);
        else
            launch_col_kernel(pl.threads, pl.wA / 64, pl.col_dense, &d, cmap.data(), pl.col_wgs, 0, mode, nullptr);
    };
    auto row = [&]() {
        if (pl.line)
            launch_line_row(pl.line_row, &d, rmap.data(), g_single ? &d : nullptr, pl.row_wgs, nullptr);
        else if (g_single)
            launch_row1_kernel(pl.threads, pl.wB / 64, d, pl.row_wgs, 0, nullptr);
        else
            launch_row_kernel(pl.threads, pl.wB / 64, &d, rmap.data(), pl.row_wgs, 0, nullptr);
    };
    if (iters > 0) {
        col(LUCAS_IN_DIGITS);
        for (long long i = 0; i < iters; ++i) {
            row();
            col(i + 1 == iters ? LUCAS_OUT_DIGITS : 0);
        }
    }
    std::memcpy(digits_out, dig.data(), sizeof(int32_t) * N);
    float f;
    std::memcpy(&f, &err, 4);
    *maxerr = f;
    return 0;
}
