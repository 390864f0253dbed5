// hip_runtime.h stand-in for the CPU emulation harness (tests only).
// Runs the engine's kernel source on the host.  Every work-item is a fiber
// (a user-space context with its own stack); a workgroup's fibers run on one
// host thread, switched at barriers, and several host threads each run their
// own workgroups.  __syncthreads and wave barriers are real barriers over the
// fibers, cross-lane reads go through a per-wave buffer.  It exists so kernel
// indexing can be checked against the oracle without a GPU; the product is
// never built against it.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
// workgroup-shared memory: one copy per host thread, which runs one
// workgroup at a time
#define __shared__ thread_local
#define LUCAS_GLOBAL
#define LUCAS_WAVES_PER_SIMD(n)

struct double2 { double x, y; };
struct int2 { int x, y; };
struct alignas(16) int4 { int x, y, z, w; };
// one work-item per host thread: every value is its own first lane
inline int __builtin_amdgcn_readfirstlane(int x) { return x; }
inline double2 make_double2(double x, double y) { return {x, y}; }
inline int2 make_int2(int x, int y) { return {x, y}; }
struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx { unsigned x, y, z; };
extern thread_local emu_idx threadIdx, blockIdx;
extern emu_idx blockDim;

typedef int hipError_t;
typedef void* hipStream_t;
#define hipSuccess 0
#define hipErrorNotSupported 801
#define hipErrorInvalidValue 1
inline hipError_t hipGetLastError() { return 0; }
#define hipFuncAttributeMaxDynamicSharedMemorySize 0
inline hipError_t hipFuncSetAttribute(const void*, int, int) { return 0; }

namespace emu {
extern std::mutex atom_mu;
void run_grid(unsigned nblocks, unsigned nthreads, const std::function<void()>& body);
// barrier over the calling fiber's workgroup, or over its wave
void group_barrier();
void wave_barrier();
// per-workgroup exchange buffers of the calling host thread (one int / float
// per work-item)
int* lane_buf();
float* shfl_buf();
// value of `x` held by lane `src` of the calling work-item's wave
inline int wave_read(int x, int src) {
    const unsigned w = threadIdx.x >> 6;
    int* const buf = lane_buf();
    buf[threadIdx.x] = x;
    wave_barrier();
    const int r = buf[(w << 6) + (src & 63)];
    wave_barrier();
    return r;
}
}
inline void __syncthreads() { emu::group_barrier(); }
// the lanes of a wave run as separate fibers here: a wave barrier is real
inline void __builtin_amdgcn_wave_barrier() { emu::wave_barrier(); }
inline unsigned __lane_id() { return threadIdx.x & 63; }
// DPP: quad_perm (ctrl < 0x100) and row_ror:n (0x121..0x12F) are the forms used
inline int __builtin_amdgcn_update_dpp(int /*old*/, int src, int ctrl, int, int, bool) {
    const int l = threadIdx.x & 63;
    int s = l;
    if (ctrl < 0x100) s = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
    else if (ctrl >= 0x121 && ctrl <= 0x12F) s = (l & ~15) | ((l + (ctrl - 0x120)) & 15);
    return emu::wave_read(src, s);
}
// ds_swizzle, bit-mask mode (offset bit 15 clear) within 32-lane halves
inline int __builtin_amdgcn_ds_swizzle(int src, int pattern) {
    const int l = threadIdx.x & 63;
    const int am = pattern & 0x1F, om = (pattern >> 5) & 0x1F, xm = (pattern >> 10) & 0x1F;
    return emu::wave_read(src, (l & 32) | ((((l & 31) & am) | om) ^ xm));
}
inline int __builtin_amdgcn_ds_bpermute(int addr, int src) { return emu::wave_read(src, (addr >> 2) & 63); }
inline int __builtin_amdgcn_readlane(int x, int lane) { return emu::wave_read(x, lane & 63); }
using std::max;
using std::min;
inline int __double2loint(double d) { unsigned long long u; std::memcpy(&u, &d, 8); return (int)(u & 0xffffffffu); }
inline int __double2hiint(double d) { unsigned long long u; std::memcpy(&u, &d, 8); return (int)(u >> 32); }
inline double __hiloint2double(int hi, int lo) {
    const unsigned long long u = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
    double d;
    std::memcpy(&d, &u, 8);
    return d;
}
inline float __shfl_xor(float v, int mask, int /*width*/) {
    float* const buf = emu::shfl_buf();
    buf[threadIdx.x] = v;
    __syncthreads();
    float r = buf[threadIdx.x ^ mask];
    __syncthreads();
    return r;
}
inline long long __double_as_longlong(double d) { long long u; std::memcpy(&u, &d, 8); return u; }
inline double __longlong_as_double(long long u) { double d; std::memcpy(&d, &u, 8); return d; }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline unsigned __brev(unsigned x) {
    unsigned r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}
inline int __mul24(int a, int b) { return a * b; }
inline unsigned __builtin_amdgcn_alignbit(unsigned hi, unsigned lo, unsigned s) {
    return (unsigned)((((unsigned long long)hi << 32) | lo) >> (s & 31));
}
inline unsigned __umulhi(unsigned a, unsigned b) { return (unsigned)(((unsigned long long)a * b) >> 32); }
inline unsigned __float_as_uint(float f) { unsigned u; std::memcpy(&u, &f, 4); return u; }
#define __HIP_MEMORY_SCOPE_AGENT 3
template <typename T>
inline T __hip_atomic_load(T* p, int, int) {
    std::lock_guard<std::mutex> lk(emu::atom_mu);
    return *p;
}
inline unsigned atomicMax(unsigned* p, unsigned v) {
    std::lock_guard<std::mutex> lk(emu::atom_mu);
    unsigned old = *p;
    if (v > old) *p = v;
    return old;
}
template <typename T>
inline T __builtin_nontemporal_load(const T* p) { return *p; }
template <typename T>
inline void __builtin_nontemporal_store(T v, T* p) { *p = v; }
using std::fma; using std::fmax; using std::fmin; using std::fabs; using std::rint;
inline float fmaxf(float a, float b) { return std::fmax(a, b); }

template <typename F, typename... Args>
void hipLaunchKernelGGL(F kernel, dim3 grid, dim3 block, size_t, hipStream_t, Args... args) {
    emu::run_grid(grid.x, block.x, [&] { kernel(args...); });
}
