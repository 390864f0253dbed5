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
hipError_t configure_line_kernels(size_t max_lds);
hipError_t launch_line_col(int idx, const LucasJobDesc* jobs, const int2* map, const LucasJobDesc* one, int nwg,
                           int mode, hipStream_t stream);
hipError_t launch_line_row(int idx, const LucasJobDesc* jobs, const int2* map, const LucasJobDesc* one, int nwg,
                           hipStream_t stream);
}  // namespace lucas

using namespace lucas;

// Sized for N = LUCAS_NOALLOC_MAX_N: two spectra of 8N bytes each, then
// the tables and the accumulators of the length's plan.
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
    if (configure_kernels(kMaxLdsBytes) != hipSuccess || configure_line_kernels(kMaxLdsBytes) != hipSuccess)
        return LUCAS_ERR_DEVICE;
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
    const size_t o_spec = 0;
    const size_t o_spec2 = align256(spec_bytes(pl));
    const size_t o_tab = 2 * o_spec2;
    const TableLayout lay = table_layout(pl);
    if (o_tab + lay.bytes > (size_t)LUCAS_NOALLOC_SCRATCH_BYTES) return LUCAS_ERR_INVALID_ARG;
    std::vector<char> host(lay.bytes);
    table_image(pl, t, lay, host.data());
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
    desc_tables(lay, m + o_tab, &d);
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
            if (pl.line) return launch_line_col(pl.line_col, nullptr, nullptr, &d, pl.col_wgs, mode, s);
            return launch_col1_kernel(pl.threads, pl.wA / 64, pl.col_dense, d, pl.col_wgs, pl.lds_col, mode, s);
        };
        auto row = [&]() {
            if (pl.line) return launch_line_row(pl.line_row, nullptr, nullptr, &d, pl.row_wgs, s);
            return launch_row1_kernel(pl.threads, pl.wB / 64, d, pl.row_wgs, pl.lds_row, s);
        };
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
