// CPU emulation driver: runs the engine's column/row kernels (compiled from
// gpulucas_amd/csrc/lucas_kernels.hip against tests/emu/hip/hip_runtime.h)
// on host memory for one job.  Tests only.
#include <cstdint>
#include <cstring>
#include <vector>

#include "lucas_desc.h"
#include "lucas_host.h"

thread_local emu_idx threadIdx, blockIdx;
emu_idx blockDim;
double2 lucas_lds[163840 / sizeof(double2)];

namespace emu {
std::barrier<>* bar = nullptr;
std::vector<float> shfl;
std::mutex atom_mu;
std::vector<std::unique_ptr<std::barrier<>>> wave_bar;
std::vector<int> lane_buf;

void run_grid(unsigned nblocks, unsigned nthreads, const std::function<void()>& body) {
    std::barrier<> b(nthreads);
    bar = &b;
    shfl.assign(nthreads, 0.f);
    lane_buf.assign(nthreads, 0);
    wave_bar.clear();
    for (unsigned w = 0; w < (nthreads + 63) / 64; ++w)
        wave_bar.push_back(std::make_unique<std::barrier<>>(std::min(64u, nthreads - 64 * w)));
    blockDim = {nthreads, 1, 1};
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < nthreads; ++t)
        ts.emplace_back([&, t] {
            threadIdx = {t, 0, 0};
            for (unsigned blk = 0; blk < nblocks; ++blk) {
                blockIdx = {blk, 0, 0};
                body();
                bar->arrive_and_wait();  // the workgroup's LDS is reused by the next one
            }
        });
    for (auto& th : ts) th.join();
    bar = nullptr;
}
}  // namespace emu

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
    std::vector<double2> spec(pl.M), spec2(pl.M);
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
