// lucas_noalloc.hip -- an entry point that allocates no device memory.
//
// lucas_step_digits_noalloc runs the same kernels as lucas_ctx_step on a
// digit vector the caller keeps in its own device memory.  Every
// intermediate of the iteration -- the two spectrum buffers, the weight,
// width and twiddle tables, and the roundoff / product-size accumulators --
// lives in `scratch`, one file-scope device array per GPU, so the call
// needs no hipMalloc: for code under HIP graph capture, tight serving loops
// and harnesses that forbid allocation.  The array's device address comes
// from hipGetSymbolAddress, once per device at the first call on it; host
// code uses only that address (the array's name is a host-side stand-in).
//
// Every call on one GPU shares that GPU's array and overwrites it: calls that
// overlap on one GPU race.  The accumulators are cleared by a small kernel
// at the start of each call (no memset node, so the launch sequence is
// capture-safe); the tables are re-uploaded whenever (p, N) differs from the
// previous call's on that device.
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "gpulucas.h"
#include "lucas_desc.h"
#include "lucas_host.h"

namespace lucas {
hipError_t launch_col1_kernel(int threads, int p, int dense, const LucasJobDesc& job, int nwg, size_t lds, int mode,
                              hipStream_t stream);
hipError_t launch_row1_kernel(int threads, int p, const LucasJobDesc& job, int nwg, size_t lds, hipStream_t stream);
hipError_t configure_kernels(size_t max_lds);
}  // namespace lucas

using namespace lucas;

// Sized for N = LUCAS_NOALLOC_MAX_N: two spectra of 8N bytes each, the
// tables (at most 4096-entry twiddle and weight tables per dimension) and
// the accumulators.
static __device__ __align__(256) unsigned char scratch[LUCAS_NOALLOC_SCRATCH_BYTES];

// Zeroes the roundoff and product-size accumulators of a call.
extern "C" __global__ void lucas_noalloc_clear(unsigned int* acc, int n) {
    if ((int)threadIdx.x < n) acc[threadIdx.x] = 0u;
}

namespace {

const int kMaxDevices = 64;

struct DeviceScratch {
    char* ptr = nullptr;  // device address of `scratch` on this device
    int p = 0, N = 0;     // job whose tables the array holds
    LucasJobDesc desc{};
    Plan plan;
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Set-up of `device` (current): the device address of its `scratch`.
int device_setup(int device, DeviceScratch* ds) {
    if (ds->ptr) return LUCAS_OK;
    if (configure_kernels(kMaxLdsBytes) != hipSuccess) return LUCAS_ERR_DEVICE;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return LUCAS_ERR_DEVICE;
    }
    (void)device;
    ds->ptr = (char*)ptr;
    return LUCAS_OK;
}

// Lays out (p, N)'s tables in the device's scratch and uploads them.
int load_tables(DeviceScratch* ds, int p, int N, hipStream_t stream) {
    Plan pl;
    if (!make_plan(p, N, 0, 0, &pl).empty()) return LUCAS_ERR_INVALID_ARG;
    Tables t;
    make_tables(pl, &t);
    size_t off = 0;
    auto take = [&](size_t n) {
        size_t o = off;
        off += align256(n);
        return o;
    };
    const size_t o_spec = take(sizeof(double2) * pl.M);
    const size_t o_spec2 = take(sizeof(double2) * pl.M);
    const size_t o_acc = take(4 * sizeof(unsigned int));
    const size_t o_tab = off;
    const size_t o_twA = take(sizeof(double) * t.twA.size());
    const size_t o_twMA = take(sizeof(double) * t.twMA.size());
    const size_t o_twB = take(sizeof(double) * t.twB.size());
    const size_t o_twMB = take(sizeof(double) * t.twMB.size());
    const size_t o_rlo = take(sizeof(int) * t.rlo.size());
    const size_t o_rhi = take(sizeof(int) * t.rhi.size());
    const size_t o_wlo = take(sizeof(double) * t.wlo.size());
    const size_t o_iwlo = take(sizeof(double) * t.iwlo.size());
    const size_t o_whi = take(sizeof(double) * t.whi.size());
    const size_t o_iwhi = take(sizeof(double) * t.iwhi.size());
    const size_t o_twsA = take(sizeof(double) * t.twsA.size());
    const size_t o_twsB = take(sizeof(double) * t.twsB.size());
    if (off > (size_t)LUCAS_NOALLOC_SCRATCH_BYTES) return LUCAS_ERR_INVALID_ARG;
    std::vector<char> host(off - o_tab, 0);
    auto put = [&](size_t o, const void* src, size_t n) { memcpy(host.data() + (o - o_tab), src, n); };
    put(o_twA, t.twA.data(), sizeof(double) * t.twA.size());
    put(o_twMA, t.twMA.data(), sizeof(double) * t.twMA.size());
    put(o_twB, t.twB.data(), sizeof(double) * t.twB.size());
    put(o_twMB, t.twMB.data(), sizeof(double) * t.twMB.size());
    put(o_rlo, t.rlo.data(), sizeof(int) * t.rlo.size());
    put(o_rhi, t.rhi.data(), sizeof(int) * t.rhi.size());
    put(o_wlo, t.wlo.data(), sizeof(double) * t.wlo.size());
    put(o_iwlo, t.iwlo.data(), sizeof(double) * t.iwlo.size());
    put(o_whi, t.whi.data(), sizeof(double) * t.whi.size());
    put(o_iwhi, t.iwhi.data(), sizeof(double) * t.iwhi.size());
    put(o_twsA, t.twsA.data(), sizeof(double) * t.twsA.size());
    put(o_twsB, t.twsB.data(), sizeof(double) * t.twsB.size());
    char* m = ds->ptr;
    hipError_t e = hipMemcpyAsync(m + o_tab, host.data(), host.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        ds->p = ds->N = 0;
        return LUCAS_ERR_DEVICE;
    }
    LucasJobDesc d{};
    fill_desc_shape(pl, &d);
    d.spec = (double2*)(m + o_spec);
    d.spec2 = (double2*)(m + o_spec2);
    d.digits = nullptr;  // the caller's, per call
    d.twA = (const double2*)(m + o_twA);
    d.twMA = (const double2*)(m + o_twMA);
    d.twB = (const double2*)(m + o_twB);
    d.twMB = (const double2*)(m + o_twMB);
    d.rlo = (const int*)(m + o_rlo);
    d.rhi = (const int*)(m + o_rhi);
    d.wlo = (const double*)(m + o_wlo);
    d.iwlo = (const double*)(m + o_iwlo);
    d.whi = (const double*)(m + o_whi);
    d.iwhi = (const double*)(m + o_iwhi);
    d.twsA = (const double2*)(m + o_twsA);
    d.twsB = (const double2*)(m + o_twsB);
    d.maxerr = (unsigned int*)(m + o_acc);
    d.pbits = d.maxerr + 1;
    ds->desc = d;
    ds->plan = pl;
    ds->p = p;
    ds->N = N;
    return LUCAS_OK;
}

}  // namespace

extern "C" int lucas_step_digits_noalloc(int p, int N, int32_t* digits, long long iters, int device, void* stream,
                                         double* max_roundoff) {
    if (p < 7 || N <= 0 || N > LUCAS_NOALLOC_MAX_N || !digits || iters < 0 || device < 0 || device >= kMaxDevices)
        return LUCAS_ERR_INVALID_ARG;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return LUCAS_ERR_NO_DEVICE;
    }
    if (device >= count) return LUCAS_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return LUCAS_ERR_NO_DEVICE;
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch* ds = &g_dev[device];
    int rc = device_setup(device, ds);
    if (rc != LUCAS_OK) return rc;
    if (ds->p != p || ds->N != N) {
        rc = load_tables(ds, p, N, s);
        if (rc != LUCAS_OK) return rc;
    }
    LucasJobDesc d = ds->desc;
    d.digits = digits;
    const Plan& pl = ds->plan;
    hipLaunchKernelGGL(lucas_noalloc_clear, dim3(1), dim3(64), 0, s, d.maxerr, 4);
    if (hipGetLastError() != hipSuccess) return LUCAS_ERR_DEVICE;
    if (iters > 0) {
        // the launch order of Engine::step: col(digits in), [row; col] x
        // (iters - 1), row, col(digits out)
        auto col = [&](int mode) {
            return launch_col1_kernel(pl.threads, pl.wA / 64, pl.col_dense, d, pl.col_wgs, pl.lds_col, mode, s);
        };
        auto row = [&]() { return launch_row1_kernel(pl.threads, pl.wB / 64, d, pl.row_wgs, pl.lds_row, s); };
        hipError_t e = col(LUCAS_IN_DIGITS);
        for (long long i = 1; i < iters && e == hipSuccess; ++i) {
            e = row();
            if (e == hipSuccess) e = col(0);
        }
        if (e == hipSuccess) e = row();
        if (e == hipSuccess) e = col(LUCAS_OUT_DIGITS);
        if (e != hipSuccess) return LUCAS_ERR_DEVICE;
    }
    if (max_roundoff) {
        unsigned int bits = 0;
        hipError_t e = hipMemcpyAsync(&bits, d.maxerr, sizeof(bits), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return LUCAS_ERR_DEVICE;
        float f;
        memcpy(&f, &bits, sizeof(f));
        *max_roundoff = (double)f;
    }
    return LUCAS_OK;
}
