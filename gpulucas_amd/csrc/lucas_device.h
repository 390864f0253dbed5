// lucas_device.h -- device helpers shared by the engine's kernel files:
// global-memory access, complex arithmetic, small DFTs, the rounding and
// packing of convolution outputs, the carry-save digit restoration and the
// wave reductions of the roundoff and product-size diagnostics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lucas_desc.h"

// Pointers read from a job descriptor are generic; loading through a global
// address-space cast lets the compiler emit global (not flat) accesses, which
// do not share the LDS wait counter.  Complex words move as 16-byte vectors.
#ifndef LUCAS_GLOBAL
#define LUCAS_GLOBAL __attribute__((address_space(1)))
#endif
// Register budget of a kernel: compiled to fit n waves per SIMD.
#ifndef LUCAS_WAVES_PER_SIMD
#define LUCAS_WAVES_PER_SIMD(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif
typedef double lucas_v2d __attribute__((vector_size(16)));
typedef int lucas_v2i __attribute__((vector_size(8)));
template <typename T>
__device__ __forceinline__ T gload(const T* p, size_t i) {
    return ((const LUCAS_GLOBAL T*)p)[i];
}
__device__ __forceinline__ double2 gload2(const double2* p, size_t i) {
    const lucas_v2d v = ((const LUCAS_GLOBAL lucas_v2d*)p)[i];
    return make_double2(v[0], v[1]);
}
__device__ __forceinline__ void gstore2(double2* p, size_t i, double2 x) {
    const lucas_v2d v = {x.x, x.y};
    ((LUCAS_GLOBAL lucas_v2d*)p)[i] = v;
}
// Spectrum word moved non-temporally when `nt` (the state outgrows the
// caches: no allocation in L2, no write-back of dirty lines at kernel end).
__device__ __forceinline__ double2 gload2_nt(const double2* p, size_t i, bool nt) {
    if (nt) return make_double2(__builtin_nontemporal_load(&p[i].x), __builtin_nontemporal_load(&p[i].y));
    return gload2(p, i);
}
__device__ __forceinline__ void gstore2_nt(double2* p, size_t i, double2 x, bool nt) {
    if (nt) {
        __builtin_nontemporal_store(x.x, &p[i].x);
        __builtin_nontemporal_store(x.y, &p[i].y);
    } else {
        gstore2(p, i, x);
    }
}
__device__ __forceinline__ int2 gload2i(const int* p, size_t i) {
    const lucas_v2i v = ((const LUCAS_GLOBAL lucas_v2i*)p)[i];
    return make_int2(v[0], v[1]);
}
__device__ __forceinline__ void gstore2i(int* p, size_t i, int2 x) {
    const lucas_v2i v = {x.x, x.y};
    ((LUCAS_GLOBAL lucas_v2i*)p)[i] = v;
}

// Dynamic LDS of the running workgroup (sized per launch).
__device__ __forceinline__ double2* lucas_dynamic_lds() {
    extern __shared__ double2 lucas_lds[];
    return lucas_lds;
}

namespace {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by SIGN*i
template <int SIGN>
__device__ __forceinline__ double2 mul_si(double2 a) {
    return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// Quotient t / d for 0 <= t, d < 2^16 from a host-computed multiply-high
// reciprocal m = ceil(2^32 / d) (d == 1 has no 32-bit reciprocal).
__device__ __forceinline__ int udiv(int t, int d, unsigned int m) {
    return d == 1 ? t : (int)__umulhi((unsigned int)t, m);
}

// ---------------------------------------------------------------------------
// Small DFTs, y_s = sum_r v_r exp(SIGN*2*pi*i*r*s/R), in place.
// ---------------------------------------------------------------------------
template <int R, int SIGN> struct Dft;

template <int SIGN> struct Dft<2, SIGN> {
    __device__ __forceinline__ static void run(double2* v) {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int SIGN> struct Dft<3, SIGN> {
    __device__ __forceinline__ static void run(double2* v) {
        const double S3 = 0.8660254037844386;  // sqrt(3)/2
        double2 t = cadd(v[1], v[2]);
        double2 u = cscale(mul_si<SIGN>(csub(v[1], v[2])), S3);
        double2 m = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, u);
        v[2] = csub(m, u);
    }
};

template <int SIGN> struct Dft<4, SIGN> {
    __device__ __forceinline__ static void run(double2* v) {
        double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        double2 s13 = cadd(v[1], v[3]), d13 = mul_si<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    }
};

template <int SIGN> struct Dft<5, SIGN> {
    __device__ __forceinline__ static void run(double2* v) {
        const double C1 = 0.30901699437494745, C2 = -0.8090169943749475;
        const double S1 = 0.9510565162951535, S2 = 0.5877852522924731;
        double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        double2 a0 = v[0];
        double2 b1 = make_double2(a0.x + C1 * t1.x + C2 * t2.x, a0.y + C1 * t1.y + C2 * t2.y);
        double2 b2 = make_double2(a0.x + C2 * t1.x + C1 * t2.x, a0.y + C2 * t1.y + C1 * t2.y);
        double2 d1 = mul_si<SIGN>(make_double2(S1 * t3.x + S2 * t4.x, S1 * t3.y + S2 * t4.y));
        double2 d2 = mul_si<SIGN>(make_double2(S2 * t3.x - S1 * t4.x, S2 * t3.y - S1 * t4.y));
        v[0] = cadd(a0, cadd(t1, t2));
        v[1] = cadd(b1, d1);
        v[4] = csub(b1, d1);
        v[2] = cadd(b2, d2);
        v[3] = csub(b2, d2);
    }
};

template <int SIGN> struct Dft<7, SIGN> {
    __device__ __forceinline__ static void run(double2* v) {
        const double C1 = 0.6234898018587335, C2 = -0.2225209339563144, C3 = -0.9009688679024191;
        const double S1 = 0.7818314824680298, S2 = 0.9749279121818236, S3 = 0.4338837391175581;
        double2 t1 = cadd(v[1], v[6]), t2 = cadd(v[2], v[5]), t3 = cadd(v[3], v[4]);
        double2 u1 = csub(v[1], v[6]), u2 = csub(v[2], v[5]), u3 = csub(v[3], v[4]);
        double2 a0 = v[0];
        double2 b1 = make_double2(a0.x + C1 * t1.x + C2 * t2.x + C3 * t3.x, a0.y + C1 * t1.y + C2 * t2.y + C3 * t3.y);
        double2 b2 = make_double2(a0.x + C2 * t1.x + C3 * t2.x + C1 * t3.x, a0.y + C2 * t1.y + C3 * t2.y + C1 * t3.y);
        double2 b3 = make_double2(a0.x + C3 * t1.x + C1 * t2.x + C2 * t3.x, a0.y + C3 * t1.y + C1 * t2.y + C2 * t3.y);
        double2 d1 = mul_si<SIGN>(make_double2(S1 * u1.x + S2 * u2.x + S3 * u3.x, S1 * u1.y + S2 * u2.y + S3 * u3.y));
        double2 d2 = mul_si<SIGN>(make_double2(S2 * u1.x - S3 * u2.x - S1 * u3.x, S2 * u1.y - S3 * u2.y - S1 * u3.y));
        double2 d3 = mul_si<SIGN>(make_double2(S3 * u1.x - S1 * u2.x + S2 * u3.x, S3 * u1.y - S1 * u2.y + S2 * u3.y));
        v[0] = cadd(a0, cadd(t1, cadd(t2, t3)));
        v[1] = cadd(b1, d1);
        v[6] = csub(b1, d1);
        v[2] = cadd(b2, d2);
        v[5] = csub(b2, d2);
        v[3] = cadd(b3, d3);
        v[4] = csub(b3, d3);
    }
};

template <int SIGN> struct Dft<8, SIGN> {
    __device__ __forceinline__ static void run(double2* v) {
        const double R2 = 0.7071067811865476;
        double2 e[4] = {v[0], v[2], v[4], v[6]};
        double2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, SIGN>::run(e);
        Dft<4, SIGN>::run(o);
        // twiddles exp(SIGN*2*pi*i*s/8), s = 1..3
        double2 o1 = cmul(o[1], make_double2(R2, SIGN * R2));
        double2 o2 = mul_si<SIGN>(o[2]);
        double2 o3 = cmul(o[3], make_double2(-R2, SIGN * R2));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};


// Maximum over a wave of a non-negative value: DPP quad permutes and row
// rotates take the max within each 16-lane row (no LDS crossbar), four lane
// reads combine the rows.  The result is wave-uniform.
__device__ __forceinline__ unsigned int wave_max_u32(unsigned int x) {
    x = max(x, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));   // lane ^ 1
    x = max(x, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));   // lane ^ 2
    x = max(x, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x124, 0xF, 0xF, false));  // row rotate 4
    x = max(x, (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xF, 0xF, false));  // row rotate 8
    return max(max((unsigned int)__builtin_amdgcn_readlane((int)x, 0), (unsigned int)__builtin_amdgcn_readlane((int)x, 16)),
               max((unsigned int)__builtin_amdgcn_readlane((int)x, 32), (unsigned int)__builtin_amdgcn_readlane((int)x, 48)));
}

// Wave max of the roundoff, folded into the job's maximum.  err >= 0, so
// float order is the order of the bit patterns.  One atomic per wave, issued
// only when the wave beats `seen` (the job's max as read at kernel start):
// after the first iterations a new maximum is rare, so there is no barrier
// and almost no atomic traffic at the end of the kernel.
__device__ __forceinline__ void wave_max_err(float err, unsigned int seen, unsigned int* dst) {
    const unsigned int m = wave_max_u32(__float_as_uint(err));
    if ((threadIdx.x & 63) == 0 && m > seen) atomicMax(dst, m);
}

// Bit length of a non-negative integer-valued double (0 for 0), capped at 64.
__device__ __forceinline__ unsigned int bit_length(double v) {
    const unsigned int e = ((unsigned int)__double2hiint(v) >> 20) & 0x7ff;
    return e ? min(e - 1022u, 64u) : 0u;
}

// Product-size diagnostic of the rounded convolution outputs c_j (the
// reference's computeMaxBitVector, gpuLucas.cu:826-836, reduced by
// findMaxErrorHOST, 811-819): dst[0] = max bit length of |c_j|; dst[1] = max
// over j of the 64-bit leading-zero count of |c_j|, the figure the reference
// prints as "max Bit Vector" (64 minus the smallest bit length).  pmax and
// pmin are the work-item's largest and smallest |c_j|; seen0, seen1 the
// job's values as read at kernel start.
__device__ __forceinline__ void wave_max_bits(double pmax, double pmin, unsigned int seen0, unsigned int seen1,
                                              unsigned int* dst) {
    const unsigned int b = wave_max_u32(bit_length(pmax));
    const unsigned int z = wave_max_u32(64u - bit_length(pmin));
    if ((threadIdx.x & 63) == 0) {
        if (b > seen0) atomicMax(dst, b);
        if (z > seen1) atomicMax(dst + 1, z);
    }
}

// Carry-save restoration of one digit (reference llintToIrrBal<n>,
// IrrBaseBalanced.cu:52-110) at compile-time depth NC.  The window is stored
// top-down: mag[i], neg[i] are |c| and sign of the rounded product of digit
// D-i, w[i] its width, i = 0..NC.  Returns the balanced digit and writes the
// carry into *carry.
template <int NC>
__device__ __forceinline__ int restore_digit(const unsigned long long* mag, const bool* neg, const int* w,
                                             int* carry) {
    const int W = w[0];
    const unsigned int mask = (1u << W) - 1u;
    int sum = 0;
    int shift = 0;
#pragma unroll
    for (int i = 0; i <= NC; ++i) {
        if (i > 0) shift += w[i];
        // Fields at or above bit 64 are zero (the reference's shift is
        // undefined there).  Depths 2 and 3 run at q >= 16 (q <= 20), so their
        // shifts stay below 3*21 = 63 and need no guard.
        const int part = (NC <= 3 || shift < 64) ? (int)((unsigned int)(mag[i] >> shift) & mask) : 0;
        sum += neg[i] ? -part : part;
    }
    // The reference's carry (IrrBaseBalanced.cu:103-109): h = (sum + half) >> W
    // for sum >= half, -((-sum + half) >> W) for sum < -half, else 0; here
    // without branches, as the magnitude's carry with sum's sign.  |sum| =
    // half with sum < 0 is not past the threshold: no carry.
    const int half = 1 << (W - 1);
    const int sg = sum >> 31;         // 0 or -1
    const int a = (sum ^ sg) - sg;    // |sum|
    const int hm = (a + half - (sum == -half ? 1 : 0)) >> W;
    const int h = (hm ^ sg) - sg;
    *carry = h;
    return sum - h * (1 << W);
}

// Rounding of an unweighted convolution output v to c = rint(v) (ties to
// even, as the reference's __double2ll_rn, gpuLucas.cu:791-796), packed as
// the carry restoration reads it: sign in bit 63, bit 62 set iff the digit
// is wide (q+1 bits), |c| in bits 0..51.  For |v| < 2^52, t = |v| + 2^52
// lies in [2^52, 2^53) where doubles are the integers, so the addition rounds
// |v| to |c| (rint is odd-symmetric) and t's mantissa field holds |c|; the
// packing only rewrites t's sign and exponent bits.  *err receives |v - c|,
// and *mag |c|.  |v| >= 2^52 (or NaN) has no meaningful rounding at all: t's
// exponent then leaves that of [2^52, 2^53), the roundoff reported is 1 and
// the gate fires.
__device__ __forceinline__ double round_pack(double v, bool wide, float* err, double* mag) {
    const double av = fabs(v);
    const double t = av + 0x1p52;
    const double c = t - 0x1p52;
    const unsigned int th = (unsigned int)__double2hiint(t);
    *err = th < 0x43400000u ? (float)fabs(av - c) : 1.0f;
    *mag = c;
    const unsigned int hi = (th & 0xFFFFFu) | ((unsigned int)wide << 30) |
                            ((unsigned int)__double2hiint(v) & 0x80000000u);
    return __hiloint2double((int)hi, __double2loint(t));
}

}  // namespace
