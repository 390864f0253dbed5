// hip_runtime.h stand-in for the CPU emulation harness (tests only).
// Runs the engine's kernel source on the host: one workgroup at a time, every
// work-item a host thread, __syncthreads a real barrier, shuffles through a
// shared buffer.  It exists so kernel indexing can be checked against the
// oracle without a GPU; the product is never built against it.
#pragma once
#include <barrier>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__
#define LUCAS_GLOBAL

struct double2 { double x, y; };
struct int2 { int x, y; };
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
inline hipError_t hipGetLastError() { return 0; }
#define hipFuncAttributeMaxDynamicSharedMemorySize 0
inline hipError_t hipFuncSetAttribute(const void*, int, int) { return 0; }

namespace emu {
extern std::barrier<>* bar;
extern std::vector<float> shfl;
extern std::mutex atom_mu;
void run_grid(unsigned nblocks, unsigned nthreads, const std::function<void()>& body);
}
inline void __syncthreads() { emu::bar->arrive_and_wait(); }
inline float __shfl_xor(float v, int mask, int /*width*/) {
    emu::shfl[threadIdx.x] = v;
    __syncthreads();
    float r = emu::shfl[threadIdx.x ^ mask];
    __syncthreads();
    return r;
}
inline unsigned __umulhi(unsigned a, unsigned b) { return (unsigned)(((unsigned long long)a * b) >> 32); }
inline unsigned __float_as_uint(float f) { unsigned u; std::memcpy(&u, &f, 4); return u; }
inline unsigned atomicMax(unsigned* p, unsigned v) {
    std::lock_guard<std::mutex> lk(emu::atom_mu);
    unsigned old = *p;
    if (v > old) *p = v;
    return old;
}
using std::fma; using std::fmax; using std::fabs; using std::rint;
inline float fmaxf(float a, float b) { return std::fmax(a, b); }

template <typename F, typename... Args>
void hipLaunchKernelGGL(F kernel, dim3 grid, dim3 block, size_t, hipStream_t, Args... args) {
    emu::run_grid(grid.x, block.x, [&] { kernel(args...); });
}
