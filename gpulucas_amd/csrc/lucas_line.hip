// lucas_line.hip -- line kernels (gfx950): one Lucas-Lehmer squaring
// s <- s^2 - 2 mod 2^p - 1 by the balanced-digit IBDWT (the reference's
// mersenneIter, gpuLucas.cu:845-942) as two launches, each one HBM round trip
// of the 8N-byte state, with every transform held by single waves:
//
//   row kernel    : one row pair (k, A-k) per workgroup, one row per wave
//                   (long rows: two waves per row, joined by a radix-2 step):
//                   outer twiddle, forward row transform, real-signal square
//                   (ComplexPointwiseSqr, gpuLucas.cu:647-665), inverse row
//                   transform, inverse outer twiddle.
//   column kernel : K own columns and H halo columns per workgroup, one column
//                   per wave (1024-word columns: two waves per column, joined
//                   by a radix-2 step): inverse column transform, unweight,
//                   -2, round
//                   (invDWTproductMinus2, gpuLucas.cu:774-805), carry-save
//                   restoration (llintToIrrBal<n>, IrrBaseBalanced.cu:52-110)
//                   and ripple + weight (loadIntToDoubleIBDWT,
//                   gpuLucas.cu:730-766) by one work-item per row, forward
//                   column transform.
//
// A line of length L = 64E lives in its wave's registers as E words per lane
// (element lane + 64e).  A transform is a radix-E step in registers followed
// by the 64-point transforms over the lanes as two radix-8 steps, each
// exchanged through the wave's own LDS region: no workgroup barrier, and
// every shape constant (lengths, strides, divisions) is known at compile
// time.  Workgroup barriers remain only where lines meet: loading and
// storing tiles, the square's row pairs, the carry across columns and the
// radix-2 steps of split lines.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "lucas_desc.h"
#include "lucas_device.h"
#include "lucas_line.h"

namespace {

// Orders one wave's LDS accesses: the LDS executes a wave's accesses in
// issue order, so a wave reading what its own lanes wrote needs no wait,
// only that the compiler keeps the program order.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("" ::: "memory");
}

template <int SIGN> struct Dft<1, SIGN> {
    __device__ __forceinline__ static void run(double2*) {}
};

// exp(SIGN*2*pi*i*j/16) * a for the twiddles inside the radix-16 DFT.
template <int J, int SIGN>
__device__ __forceinline__ double2 mul_w16(double2 a) {
    constexpr double c[10] = {1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 0.0,
                              -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0,
                              -0.9238795325112867};
    constexpr double s[10] = {0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867, 1.0,
                              0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 0.0,
                              -0.3826834323650898};
    if constexpr (J == 4) return mul_si<SIGN>(a);
    else return cmul(a, make_double2(c[J], SIGN * s[J]));
}

// Radix 16 as 4 x 4: x[n1 + 4n2] -> DFT4 over n2 (k2), twiddle
// w16^(n1*k2), DFT4 over n1 (k1); output k2 + 4k1.
template <int SIGN> struct Dft<16, SIGN> {
    __device__ __forceinline__ static void run(double2* v) {
        double2 t[16];
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) {
            double2 a[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
            Dft<4, SIGN>::run(a);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) t[4 * n1 + k2] = a[k2];
        }
        t[5] = mul_w16<1, SIGN>(t[5]);
        t[6] = mul_w16<2, SIGN>(t[6]);
        t[7] = mul_w16<3, SIGN>(t[7]);
        t[9] = mul_w16<2, SIGN>(t[9]);
        t[10] = mul_w16<4, SIGN>(t[10]);
        t[11] = mul_w16<6, SIGN>(t[11]);
        t[13] = mul_w16<3, SIGN>(t[13]);
        t[14] = mul_w16<6, SIGN>(t[14]);
        t[15] = mul_w16<9, SIGN>(t[15]);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            double2 b[4] = {t[k2], t[4 + k2], t[8 + k2], t[12 + k2]};
            Dft<4, SIGN>::run(b);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) v[k2 + 4 * k1] = b[k1];
        }
    }
};

// omega_8^d * a with omega_8 = exp(-2*pi*i/8).
template <int D>
__device__ __forceinline__ double2 mul_w8(double2 a) {
    constexpr double R2 = 0.7071067811865476;
    if constexpr (D == 0) return a;
    else if constexpr (D == 2) return make_double2(a.y, -a.x);
    else if constexpr (D == 4) return make_double2(-a.x, -a.y);
    else if constexpr (D == 6) return make_double2(-a.y, a.x);
    else if constexpr (D == 1) return make_double2(R2 * (a.x + a.y), R2 * (a.y - a.x));
    else if constexpr (D == 3) return make_double2(R2 * (a.y - a.x), -R2 * (a.x + a.y));
    else if constexpr (D == 5) return make_double2(-R2 * (a.x + a.y), R2 * (a.x - a.y));
    else return make_double2(R2 * (a.x - a.y), R2 * (a.x + a.y));
}

// Wave-local LDS positions (16-byte words within a line's region) of the
// 64-point steps.  S1 holds the radix-E outputs y[k1][n] (n < 64), S2 the
// first radix-8 step's outputs u[k1][a][c].  The XOR terms make every
// 16-lane group of each access hit 16 different bank groups: the two k1 of a
// group differ in bit 3, and the a ^ c term spreads S2's column reads.
__device__ __forceinline__ int s1_pos(int k1, int n) { return k1 * 64 + (n ^ ((k1 & 1) << 3)); }
__device__ __forceinline__ int s2_pos(int k1, int c, int a) { return k1 * 64 + ((c ^ (k1 & 1)) << 3) + (a ^ c); }

// Twiddles of a line transform (see Line): read from the workgroup's LDS copy
// of the line table (long lines), or held in registers, loaded once per
// kernel from the global table (lines of up to 8 words per lane, whose
// transforms are short chains: no LDS round trip for a twiddle).
struct TwLds {
    const double2* T1;
    const double2* T2;
    int l;
    __device__ __forceinline__ double2 t1(int k) const { return T1[k * 64 + l]; }
    __device__ __forceinline__ double2 t2(int, int c, int a) const { return T2[c * 8 + a]; }
};
template <int E>
struct TwReg {
    double2 a1[E > 1 ? E - 1 : 1];
    // omega_64^(a*c) depends on a = lane mod 8 only (the same for every
    // butterfly of a lane)
    double2 a2[7];
    // tab: omega_L^(n*k1) at [k1*64 + n], omega_64^(a*c) at [64E + c*8 + a]
    __device__ __forceinline__ void load(const double2* tab, int l) {
#pragma unroll
        for (int k = 1; k < E; ++k) a1[k - 1] = gload2(tab, k * 64 + l);
#pragma unroll
        for (int c = 1; c < 8; ++c) a2[c - 1] = gload2(tab, 64 * E + c * 8 + (l & 7));
    }
    __device__ __forceinline__ double2 t1(int k) const { return a1[k - 1]; }
    __device__ __forceinline__ double2 t2(int, int c, int) const { return a2[c - 1]; }
};

// Transforms of one line of length L = 64E held by one wave.
//   Input distribution (natural): x[e] = element l + 64e of lane l.
//   Output distribution: butterfly beta = l + 64i (i < NB, active iff
//   beta < NA) holds in x[8i + d] the frequency k1 + E*c + 8E*d, with
//   k1 = beta >> 3, c = beta & 7.
// fwd maps natural input to that distribution (forward DFT, unnormalised);
// inv maps it back (inverse DFT, unnormalised: inv(fwd(x)) = L x).
// Twiddles (tw, a TwLds or TwReg): t1(k1) = omega_L^(n*k1) for lane n,
// t2(i, c, a) = omega_64^(a*c).
template <int E>
struct Line {
    static_assert(E < 8 || E % 8 == 0, "line length: 64E with E < 8 or a multiple of 8");
    static constexpr int L = 64 * E;
    static constexpr int NA = 8 * E;                 // radix-8 butterflies per 64-point step
    static constexpr int NB = E >= 8 ? E / 8 : 1;    // butterflies per lane
    static constexpr int NX = E >= 8 ? E : 8;        // words per lane
    static constexpr bool ALL = NA >= 64;            // every lane busy in the radix-8 steps

    __device__ __forceinline__ static int out_index(int beta, int d) {
        return (beta >> 3) + E * (beta & 7) + 8 * E * d;
    }
    // Region position d*NA + beta of frequency k (the output distribution
    // laid out lane-contiguously).
    __device__ __forceinline__ static int out_pos(int k) {
        const int d = k / NA, rem = k - d * NA;
        const int c = rem / E, k1 = rem - c * E;
        return d * NA + 8 * k1 + c;
    }

    template <class TW>
    __device__ __forceinline__ static void fwd(double2 (&x)[NX], double2* R, const TW& tw, int l) {
        Dft<E, -1>::run(x);
#pragma unroll
        for (int k = 1; k < E; ++k) x[k] = cmul(x[k], tw.t1(k));
        wave_sync();
#pragma unroll
        for (int k = 0; k < E; ++k) R[s1_pos(k, l)] = x[k];
        wave_sync();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int beta = l + 64 * i;
            if (ALL || beta < NA) {
                const int k1 = beta >> 3, a = beta & 7;
#pragma unroll
                for (int b = 0; b < 8; ++b) x[8 * i + b] = R[s1_pos(k1, a + 8 * b)];
                Dft<8, -1>::run(&x[8 * i]);
#pragma unroll
                for (int c = 1; c < 8; ++c) x[8 * i + c] = cmul(x[8 * i + c], tw.t2(i, c, a));
            }
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int beta = l + 64 * i;
            if (ALL || beta < NA) {
                const int k1 = beta >> 3, a = beta & 7;
#pragma unroll
                for (int c = 0; c < 8; ++c) R[s2_pos(k1, c, a)] = x[8 * i + c];
            }
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int beta = l + 64 * i;
            if (ALL || beta < NA) {
                const int k1 = beta >> 3, c = beta & 7;
#pragma unroll
                for (int a = 0; a < 8; ++a) x[8 * i + a] = R[s2_pos(k1, c, a)];
                Dft<8, -1>::run(&x[8 * i]);
            }
        }
    }

    template <class TW>
    __device__ __forceinline__ static void inv(double2 (&x)[NX], double2* R, const TW& tw, int l) {
#pragma unroll
        for (int i = 0; i < NB; ++i)
            if (ALL || l + 64 * i < NA) Dft<8, 1>::run(&x[8 * i]);
        wave_sync();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int beta = l + 64 * i;
            if (ALL || beta < NA) {
                const int k1 = beta >> 3, c = beta & 7;
#pragma unroll
                for (int a = 0; a < 8; ++a) R[s2_pos(k1, c, a)] = x[8 * i + a];
            }
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int beta = l + 64 * i;
            if (ALL || beta < NA) {
                const int k1 = beta >> 3, a = beta & 7;
#pragma unroll
                for (int c = 0; c < 8; ++c) x[8 * i + c] = R[s2_pos(k1, c, a)];
#pragma unroll
                for (int c = 1; c < 8; ++c) x[8 * i + c] = cmulc(x[8 * i + c], tw.t2(i, c, a));
                Dft<8, 1>::run(&x[8 * i]);
            }
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int beta = l + 64 * i;
            if (ALL || beta < NA) {
                const int k1 = beta >> 3, a = beta & 7;
#pragma unroll
                for (int b = 0; b < 8; ++b) R[s1_pos(k1, a + 8 * b)] = x[8 * i + b];
            }
        }
        wave_sync();
#pragma unroll
        for (int k = 0; k < E; ++k) x[k] = R[s1_pos(k, l)];
#pragma unroll
        for (int k = 1; k < E; ++k) x[k] = cmulc(x[k], tw.t1(k));
        Dft<E, 1>::run(x);
    }
};

// Phase stamps (profiling library only, LUCAS_PHASE_STAMPS): work-item 0 of
// each of the first LUCAS_LINE_STAMP_WGS workgroups records the 100 MHz
// real-time counter at the phase boundaries of the steady-state column
// launches and of the row launches, six stamps per workgroup.
#define LUCAS_LINE_STAMP_WGS 8192
#ifdef LUCAS_PHASE_STAMPS
__device__ unsigned long long lucas_line_stamps_col[LUCAS_LINE_STAMP_WGS * 8];
__device__ unsigned long long lucas_line_stamps_row[LUCAS_LINE_STAMP_WGS * 8];
#define LINE_STAMP(arr, on, k)                                                       \
    do {                                                                             \
        if ((on) && threadIdx.x == 0 && blockIdx.x < LUCAS_LINE_STAMP_WGS)           \
            arr[blockIdx.x * 8 + (k)] = __builtin_amdgcn_s_memrealtime();            \
    } while (0)
#else
#define LINE_STAMP(arr, on, k) \
    do {                       \
    } while (0)
#endif

// ---------------------------------------------------------------------------
// Row kernel: rows R0, R1 of a pair (wave 0, wave 1).  Workgroup 0 takes rows
// 0 and A/2, each paired with itself.  Frequency K = R + A*kk of the
// half-length transform pairs with M - K: row A - R, column B-1-kk (row 0:
// row 0, column -kk mod B).  With E = Z_K + conj(Z_{M-K}) and
// O = -i (Z_K - conj(Z_{M-K})) the packed spectrum of the squared real signal
// is (E^2 + omega_M^K O^2 + 2i E O) / 4; the 1/4 is in the unweight table.
// ---------------------------------------------------------------------------
template <int EB>
__device__ __forceinline__ void line_row_body(const LucasJobDesc& J, int wg) {
    using F = Line<EB>;
    constexpr int B = 64 * EB, NA = F::NA, NB = F::NB;
    double2* const lds = lucas_dynamic_lds();
    double2* const T2R = lds + 2 * B;  // rt2 rows of the two rows, [wave][e]
    const int A = J.A;
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool self = wg == 0;
    const int R = self ? (w ? A / 2 : 0) : (w ? A - wg : wg);
    double2* const mine = lds + w * B;
    double2* const theirs = self ? mine : lds + (w ^ 1) * B;

    // table loads first: the load counter drains in issue order, so waiting
    // for them never waits for the row
    LINE_STAMP(lucas_line_stamps_row, true, 0);
    // the line twiddles, in registers (a row workgroup has two waves per CU
    // slot, so registers are plentiful; measured faster than an LDS copy)
    TwReg<EB> twr;
    twr.load(J.lnB, l);
    double2 x[F::NX];
#pragma unroll
    for (int e = 0; e < EB; ++e) x[e] = gload2_nt(J.spec, (size_t)R * J.ld + l + 64 * e, J.nt);
    // Every twiddle the kernel needs is loaded here, so no later phase waits
    // on a global load: the outer twiddle omega_M^(R*c), c = l + 64e, is
    // rt1[R][l] * rt2[R][e] (the rt2 row goes to LDS for the store), and the
    // square's omega_M^K = omega_M^R * omega_B^(k1 + EB*c) * omega_8^d.
    const double2 t1 = gload2(J.rt1, (size_t)R * 64 + l);
    // short rows keep the whole factor row (wave-uniform) in registers
    constexpr bool T2REG = EB <= 8;
    double2 t2r[T2REG ? EB : 1];
    double2 t2v = make_double2(0.0, 0.0);
    if constexpr (T2REG) {
#pragma unroll
        for (int e = 0; e < EB; ++e) t2r[e] = gload2(J.rt2, (size_t)R * EB + e);
    } else if (l < EB) {
        t2v = gload2(J.rt2, (size_t)R * EB + l);
    }
    const double2 wR = gload2(J.twMA, R);
    double2 wbv[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int beta = l + 64 * i;
        wbv[i] = gload2(J.twB, (F::ALL || beta < NA) ? (beta >> 3) + EB * (beta & 7) : 0);
    }
    if (!T2REG && l < EB) T2R[w * EB + l] = t2v;
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EB; ++e) {
        if constexpr (T2REG) {
            t2r[e] = cmul(t1, t2r[e]);  // kept for the store
            x[e] = cmul(x[e], t2r[e]);
        } else {
            x[e] = cmul(x[e], cmul(t1, T2R[w * EB + e]));
        }
    }
    LINE_STAMP(lucas_line_stamps_row, true, 1);
    F::fwd(x, mine, twr, l);
    wave_sync();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int beta = l + 64 * i;
        if (F::ALL || beta < NA) {
#pragma unroll
            for (int d = 0; d < 8; ++d) mine[d * NA + beta] = x[8 * i + d];
        }
    }
    __syncthreads();
    LINE_STAMP(lucas_line_stamps_row, true, 2);
    {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int beta = l + 64 * i;
            if (F::ALL || beta < NA) {
                // omega_B^kk = omega_B^(k1 + EB*c) * omega_8^d
                const double2 wb = cmul(wR, wbv[i]);
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    const int kk = F::out_index(beta, d);
                    int pos;
                    if (self && R == 0) {
                        pos = F::out_pos(kk == 0 ? 0 : B - kk);
                    } else {
                        pos = B - 1 - (d * NA + beta);  // frequency B-1-kk
                    }
                    const double2 zk = x[8 * i + d], zm = theirs[pos];
                    double2 wk;
                    switch (d) {
                        case 0: wk = mul_w8<0>(wb); break;
                        case 1: wk = mul_w8<1>(wb); break;
                        case 2: wk = mul_w8<2>(wb); break;
                        case 3: wk = mul_w8<3>(wb); break;
                        case 4: wk = mul_w8<4>(wb); break;
                        case 5: wk = mul_w8<5>(wb); break;
                        case 6: wk = mul_w8<6>(wb); break;
                        default: wk = mul_w8<7>(wb); break;
                    }
                    const double2 Ev = cadd(zk, cconj(zm));
                    const double2 Ov = mul_si<-1>(csub(zk, cconj(zm)));
                    const double2 P = cadd(cmul(Ev, Ev), cmul(wk, cmul(Ov, Ov)));
                    const double2 Q = cmul(Ev, Ov);
                    x[8 * i + d] = make_double2(P.x - 2.0 * Q.y, P.y + 2.0 * Q.x);
                }
            }
        }
    }
    __syncthreads();  // every partner read is done before the regions are reused
    LINE_STAMP(lucas_line_stamps_row, true, 3);
    F::inv(x, mine, twr, l);
    LINE_STAMP(lucas_line_stamps_row, true, 4);
#pragma unroll
    for (int e = 0; e < EB; ++e) {
        const double2 tw = T2REG ? t2r[e] : cmul(t1, T2R[w * EB + e]);
        gstore2_nt(J.spec2, (size_t)R * J.ld + l + 64 * e, cmulc(x[e], tw), J.nt);
    }
    LINE_STAMP(lucas_line_stamps_row, true, 5);
}

// ---------------------------------------------------------------------------
// Row kernel, split rows: each row of the pair is held by two waves, wave
// (r, h) holding elements h*B/2 + n (n < B/2) of row r, so a long row has
// twice the waves of the one-wave-per-row kernel.  A radix-2 step joins the
// halves through LDS: the even frequencies kk = 2k are the half-length
// transform of y0[n] = x[n] + x[n + B/2], the odd ones kk = 2k + 1 that of
// y1[n] = (x[n] - x[n + B/2]) omega_B^n, and wave h transforms y_h.  The
// square's partner of kk = 2k + h, column B-1-kk of the partner row, is
// 2(B/2-1-k) + (1-h): the other half of the partner row at the reversed
// position.  Row 0 pairs kk with -kk mod B, in the same half: k with
// -k mod B/2 for h = 0, with B/2-1-k for h = 1.  The inverse runs the steps
// backwards: half-length inverses u_h, then z[n] = u0 + u1 conj(omega_B^n)
// and z[n + B/2] = u0 - u1 conj(omega_B^n).
// ---------------------------------------------------------------------------
// omega_64^t = exp(-2*pi*i*t/64) for a compile-time t (after unrolling): the
// quarter-wave cosines cos(2*pi*j/64), j = 0..16, and the quadrant symmetries.
__device__ constexpr double kCos64[17] = {1.0, 0.9951847266721969, 0.9807852804032304, 0.9569403357322088,
                                          0.9238795325112867, 0.881921264348355, 0.8314696123025452,
                                          0.773010453362737, 0.7071067811865476, 0.6343932841636455,
                                          0.5555702330196022, 0.47139673682599764, 0.3826834323650898,
                                          0.2902846772544624, 0.19509032201612828, 0.0980171403295606, 0.0};
__device__ __forceinline__ double2 omega64(int t) {
    const int q = (t >> 4) & 3, r = t & 15;
    const double c = kCos64[r], sn = kCos64[16 - r];  // cos, sin of 2*pi*r/64
    if (q == 0) return make_double2(c, -sn);
    if (q == 1) return make_double2(-sn, -c);
    if (q == 2) return make_double2(-c, sn);
    return make_double2(sn, c);
}

template <int EB>
__device__ __forceinline__ void line_row_split_body(const LucasJobDesc& J, int wg) {
    constexpr int EH = EB / 2;
    using F = Line<EH>;
    constexpr int B = 64 * EB, LH = 64 * EH, NA = F::NA, NB = F::NB;
    double2* const lds = lucas_dynamic_lds();
    double2* const T2R = lds + 2 * B;  // each wave's half of its row's rt2 row, [wave][e]
    const int A = J.A;
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int h = w & 1;
    const bool self = wg == 0;
    const int R = self ? ((w >> 1) ? A / 2 : 0) : ((w >> 1) ? A - wg : wg);
    double2* const mine = lds + w * LH;
    double2* const sib = lds + (w ^ 1) * LH;  // the other half of this row
    // the square's partners: the other half of the partner row (w ^ 3), of
    // this row for row A/2, this half for row 0
    double2* const part = self ? (R == 0 ? mine : sib) : lds + (w ^ 3) * LH;
    const bool neg = self && R == 0 && h == 0;  // partner -k mod B/2

    LINE_STAMP(lucas_line_stamps_row, true, 0);
    TwReg<EH> twr;
    twr.load(J.lnB, l);
    double2 x[F::NX];
    const size_t base = (size_t)R * J.ld + (size_t)h * LH;
#pragma unroll
    for (int e = 0; e < EH; ++e) x[e] = gload2_nt(J.spec, base + l + 64 * e, J.nt);
    // outer twiddle omega_M^(R*c), c = h*B/2 + l + 64e: rt1[R][l] * rt2[R][h*EH + e]
    const double2 t1 = gload2(J.rt1, (size_t)R * 64 + l);
    constexpr bool T2REG = EH <= 8;
    double2 t2r[T2REG ? EH : 1];
    double2 t2v = make_double2(0.0, 0.0);
    if constexpr (T2REG) {
#pragma unroll
        for (int e = 0; e < EH; ++e) t2r[e] = gload2(J.rt2, (size_t)R * EB + h * EH + e);
    } else if (l < EH) {
        t2v = gload2(J.rt2, (size_t)R * EB + h * EH + l);
    }
    // the radix-2 step's omega_B^n, n = l + 64e: omega_B^l * omega_B^(64e),
    // the second factor omega_64^(e*64/EB) a constant
    static_assert(64 % EB == 0, "split rows: EB divides 64");
    const double2 wl = gload2(J.twB, l);
    const double2 wR = gload2(J.twMA, R);
    double2 wbv[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int beta = l + 64 * i;
        wbv[i] = gload2(J.twB, (F::ALL || beta < NA) ? h + 2 * ((beta >> 3) + EH * (beta & 7)) : 0);
    }
    if (!T2REG && l < EH) T2R[w * EH + l] = t2v;
    wave_sync();
#pragma unroll
    for (int e = 0; e < EH; ++e) {
        if constexpr (T2REG) {
            t2r[e] = cmul(t1, t2r[e]);  // kept for the store
            x[e] = cmul(x[e], t2r[e]);
        } else {
            x[e] = cmul(x[e], cmul(t1, T2R[w * EH + e]));
        }
        mine[l + 64 * e] = x[e];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EH; ++e) {
        const double2 o = sib[l + 64 * e];
        if (h == 0) x[e] = cadd(x[e], o);
        else x[e] = cmul(csub(o, x[e]), cmul(wl, omega64(e * (64 / EB))));
    }
    __syncthreads();  // the sibling's reads are done before the region is reused
    LINE_STAMP(lucas_line_stamps_row, true, 1);
    F::fwd(x, mine, twr, l);
    wave_sync();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int beta = l + 64 * i;
        if (F::ALL || beta < NA) {
#pragma unroll
            for (int d = 0; d < 8; ++d) mine[d * NA + beta] = x[8 * i + d];
        }
    }
    __syncthreads();
    LINE_STAMP(lucas_line_stamps_row, true, 2);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int beta = l + 64 * i;
        if (F::ALL || beta < NA) {
            // omega_B^kk = omega_B^(h + 2(k1 + EH*c)) * omega_8^d
            const double2 wb = cmul(wR, wbv[i]);
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                const int k = F::out_index(beta, d);
                const int pos = neg ? F::out_pos(k == 0 ? 0 : LH - k) : LH - 1 - (d * NA + beta);
                const double2 zk = x[8 * i + d], zm = part[pos];
                double2 wk;
                switch (d) {
                    case 0: wk = mul_w8<0>(wb); break;
                    case 1: wk = mul_w8<1>(wb); break;
                    case 2: wk = mul_w8<2>(wb); break;
                    case 3: wk = mul_w8<3>(wb); break;
                    case 4: wk = mul_w8<4>(wb); break;
                    case 5: wk = mul_w8<5>(wb); break;
                    case 6: wk = mul_w8<6>(wb); break;
                    default: wk = mul_w8<7>(wb); break;
                }
                const double2 Ev = cadd(zk, cconj(zm));
                const double2 Ov = mul_si<-1>(csub(zk, cconj(zm)));
                const double2 P = cadd(cmul(Ev, Ev), cmul(wk, cmul(Ov, Ov)));
                const double2 Q = cmul(Ev, Ov);
                x[8 * i + d] = make_double2(P.x - 2.0 * Q.y, P.y + 2.0 * Q.x);
            }
        }
    }
    __syncthreads();  // every partner read is done before the regions are reused
    LINE_STAMP(lucas_line_stamps_row, true, 3);
    F::inv(x, mine, twr, l);
    wave_sync();
#pragma unroll
    for (int e = 0; e < EH; ++e) {
        if (h) x[e] = cmulc(x[e], cmul(wl, omega64(e * (64 / EB))));
        mine[l + 64 * e] = x[e];
    }
    __syncthreads();
    LINE_STAMP(lucas_line_stamps_row, true, 4);
#pragma unroll
    for (int e = 0; e < EH; ++e) {
        const double2 o = sib[l + 64 * e];
        const double2 z = h ? csub(o, x[e]) : cadd(x[e], o);
        const double2 tw = T2REG ? t2r[e] : cmul(t1, T2R[w * EH + e]);
        gstore2_nt(J.spec2, base + l + 64 * e, cmulc(z, tw), J.nt);
    }
    LINE_STAMP(lucas_line_stamps_row, true, 5);
}

template <int EB, int S>
__device__ __forceinline__ void line_row_any(const LucasJobDesc& J, int wg) {
    if constexpr (S == 1) line_row_body<EB>(J, wg);
    else line_row_split_body<EB>(J, wg);
}

// ---------------------------------------------------------------------------
// Column kernel: workgroup `tile` owns columns c0 = K*tile .. c0+K-1; local
// column lc is global column c0 - H + lc (mod B).  Local columns 0..H-1 are
// the halo: the restoration of digit j reads the rounded products of digits
// j-NC .. j, and a tile's first digits need the products of the H columns to
// its left.  A halo column left of column 0 wraps to column B - H + lc of the
// previous row (the cyclic wrap of IrrBaseBalanced.cu:64-70).
// S = 1: wave w holds local column w.  S = 2: waves 2lc and 2lc+1 hold the
// halves of local column lc (rows h*A/2 + n of wave h), joined by a radix-2
// step as in the split row kernel: the forward transform of half h is that
// of y0[n] = x[n] + x[n + A/2] (even frequencies) or
// y1[n] = (x[n] - x[n + A/2]) omega_A^n (odd ones); the inverse takes the
// half-length inverses u_h to z[n] = u0 + u1 conj(omega_A^n),
// z[n + A/2] = u0 - u1 conj(omega_A^n).  Frequency r of a column sits in
// half r & 1 at the half-line position of r >> 1.
// ---------------------------------------------------------------------------
template <int EA, int K, int NC, int S>
__device__ __forceinline__ void line_col_body(const LucasJobDesc& J, int tile, int mode) {
    static_assert(NC >= 2 && NC <= 3, "the row restoration assumes widths of at least 16 bits");
    static_assert(S == 1 || (S == 2 && EA % 2 == 0 && EA / 2 <= 8), "split columns: halves of at most 512 words");
    constexpr int EL = EA / S;  // words per lane of a wave's line
    using F = Line<EL>;
    constexpr int A = 64 * EA, LH = 64 * EL, H = lucas_line_halo(NC), NL = K + H, RS = A + 1;
    constexpr int NW = lucas_line_col_waves(K, NC, S), T = 64 * NW;  // waves S*NL.. are helpers
    constexpr int NA = F::NA, NB = F::NB;
    // the line twiddles: in registers for short lines, else an LDS copy
    constexpr bool REG = !lucas_line_col_twl(EA, S);
    double2* const lds = lucas_dynamic_lds();
    double2* const T1 = lds + NL * RS;
    double2* const T2 = T1 + A;
    double2* const roww = REG ? T1 : T2 + 64;                 // {whi, iwhi} per row
    int* const rowr = reinterpret_cast<int*>(roww + A);      // rhi per row
    double2* const locw = roww + A + (A + 3) / 4;            // {wlo, iwlo} per local digit
    int* const locr = reinterpret_cast<int*>(locw + 2 * NL);  // rlo per local digit
    const int B = J.B, N = J.N;
    const int c0 = tile * K;
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool in_digits = (mode & LUCAS_IN_DIGITS) != 0;
    const bool out_digits = (mode & LUCAS_OUT_DIGITS) != 0;
    auto gcol = [&](int lc) {
        const int g = c0 - H + lc;
        return g < 0 ? g + B : g;
    };
    const int lcw = w / S, h = w % S;  // this wave's local column and half
    const bool colw = lcw < NL;        // this wave holds (half) a local column (else a helper)
    const bool own = colw && lcw >= H;  // ... one of the tile's own columns
    double2* const column = lds + (colw ? lcw : 0) * RS;
    double2* const region = column + h * LH;  // the wave's line
    // region position of frequency r of a column
    auto spec_pos = [](int r) { return S == 1 ? F::out_pos(r) : (r & 1) * LH + F::out_pos(r >> 1); };
    const unsigned int seen = __hip_atomic_load(J.maxerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int seen_b = __hip_atomic_load(J.pbits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int seen_z = __hip_atomic_load(J.pbits + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    const bool stamp = mode == 0;
    LINE_STAMP(lucas_line_stamps_col, stamp, 0);
    // Tables (T1, T2; per row; per local digit), loads first, then the tile.
    constexpr int NTAB = 2 * A + 64 + 2 * NL, NTJ = (NTAB + T - 1) / T;
    double2 tv[NTJ];
    int ti[NTJ];
    TwReg<EL> twr;
    if constexpr (REG) twr.load(J.lnA, l);
    // split columns: omega_A^n, n = l + 64e, as omega_A^l * omega_64^(e*64/EA)
    const double2 wa = S == 2 ? gload2(J.twA, l) : make_double2(1.0, 0.0);
    const TwLds twl{T1, T2, l};
#pragma unroll
    for (int j = 0; j < NTJ; ++j) {
        const int t = tid + T * j;
        if (t < A + 64) {
            if (!REG) tv[j] = gload2(J.lnA, t);
        } else if (t < 2 * A + 64) {
            const int r = t - A - 64;
            tv[j] = make_double2(gload(J.whi, r), gload(J.iwhi, r));
            ti[j] = gload(J.rhi, r);
        } else if (t < NTAB) {
            const int u = t - 2 * A - 64;
            const int jj = 2 * gcol(u >> 1) + (u & 1);
            tv[j] = make_double2(gload(J.wlo, jj), gload(J.iwlo, jj));
            ti[j] = gload(J.rlo, jj);
        }
    }
    constexpr int NTILE = (A * NL + T - 1) / T;
    double2 tile_v[NTILE];
    if (!in_digits) {
#pragma unroll
        for (int i = 0; i < NTILE; ++i) {
            const int t = tid + T * i;
            if (t < A * NL) {
                const int r = t / NL, lc = t - r * NL;
                tile_v[i] = gload2_nt(J.spec2, (size_t)r * J.ld + gcol(lc), J.nt);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NTJ; ++j) {
        const int t = tid + T * j;
        if (t < A + 64) {
            if (!REG) T1[t] = tv[j];
        } else if (t < 2 * A + 64) {
            roww[t - A - 64] = tv[j];
            rowr[t - A - 64] = ti[j];
        } else if (t < NTAB) {
            locw[t - 2 * A - 64] = tv[j];
            locr[t - 2 * A - 64] = ti[j];
        }
    }
    if (!in_digits) {
#pragma unroll
        for (int i = 0; i < NTILE; ++i) {
            const int t = tid + T * i;
            if (t < A * NL) {
                const int r = t / NL, lc = t - r * NL;
                lds[lc * RS + spec_pos(r)] = tile_v[i];
            }
        }
    }
    __syncthreads();
    LINE_STAMP(lucas_line_stamps_col, stamp, 1);

    float err = 0.0f;
    double pmax = 0.0, pmin = 0x1p60;  // largest / smallest |c_j| of this work-item
    double2 x[F::NX];
    // weight 2^(r_j/N) of the digits (2m, 2m+1) of element (r, local column lc)
    auto weighted = [&](int lc, int r, int2 d) {
        const int rh = rowr[r];
        const double wh = roww[r].x;
        const double ae = locw[2 * lc].x * wh * (locr[2 * lc] + rh >= N ? 0.5 : 1.0);
        const double ao = locw[2 * lc + 1].x * wh * (locr[2 * lc + 1] + rh >= N ? 0.5 : 1.0);
        return make_double2(ae * (double)d.x, ao * (double)d.y);
    };

    if (in_digits) {
        // the own columns' digits, weighted, straight into the forward transform
        if (own) {
            const int gc = c0 + lcw - H;
            auto dig = [&](int r) { return weighted(lcw, r, gload2i(J.digits, (size_t)gc + (size_t)B * r)); };
#pragma unroll
            for (int e = 0; e < EL; ++e) {
                const int n = l + 64 * e;
                if constexpr (S == 1) {
                    x[e] = dig(n);
                } else {
                    const double2 a = dig(n), b = dig(n + LH);
                    x[e] = h ? cmul(csub(a, b), cmul(wa, omega64(e * (64 / EA)))) : cadd(a, b);
                }
            }
        }
    } else {
        // inverse column transform of every local column
        if (colw) {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int beta = l + 64 * i;
                if (F::ALL || beta < NA) {
#pragma unroll
                    for (int d = 0; d < 8; ++d) x[8 * i + d] = region[d * NA + beta];
                }
            }
            if constexpr (REG) F::inv(x, region, twr, l);
            else F::inv(x, region, twl, l);
            if constexpr (S == 2) {
                wave_sync();
#pragma unroll
                for (int e = 0; e < EL; ++e) {
                    if (h) x[e] = cmulc(x[e], cmul(wa, omega64(e * (64 / EA))));
                    region[l + 64 * e] = x[e];
                }
            }
        }
        if constexpr (S == 2) {
            // the radix-2 step: each half reads the other's u
            __syncthreads();
            if (colw) {
#pragma unroll
                for (int e = 0; e < EL; ++e) {
                    const double2 o = column[(h ^ 1) * LH + l + 64 * e];
                    x[e] = h ? csub(o, x[e]) : cadd(x[e], o);
                }
            }
            __syncthreads();  // every read done before the products overwrite the column
        }
        if (colw) {
            // unweight, -2 on digit 0, round half-even; the rounded products
            // go back as round_pack words (natural row order)
            const int gc = gcol(lcw);
            const int rle = locr[2 * lcw], rlo = locr[2 * lcw + 1];
            const double ile = locw[2 * lcw].y, ilo = locw[2 * lcw + 1].y;
            wave_sync();
#pragma unroll
            for (int e = 0; e < EL; ++e) {
                const int r = h * LH + l + 64 * e;
                const int rh = rowr[r];
                const double iwh = roww[r].y;
                const int rje = rle + rh, rjo = rlo + rh;
                double ve = x[e].x * (ile * iwh * (rje >= N ? 2.0 : 1.0));
                const double vo = x[e].y * (ilo * iwh * (rjo >= N ? 2.0 : 1.0));
                if (gc == 0 && r == 0) ve -= 2.0;
                float ee, eo;
                double me, mo;
                const double pe = round_pack(ve, (rje >= N ? rje - N : rje) < J.b, &ee, &me);
                const double po = round_pack(vo, (rjo >= N ? rjo - N : rjo) < J.b, &eo, &mo);
                err = fmaxf(err, fmaxf(ee, eo));
                pmax = fmax(pmax, fmax(me, mo));
                pmin = fmin(pmin, fmin(me, mo));
                column[r] = make_double2(pe, po);
            }
            // long columns reduce the maxima here, not at the end: held to
            // the end they would keep every element's intermediates live
            // through the carry; short ones defer the reduction off the
            // critical path
            if constexpr (!REG) {
                wave_max_err(err, seen, J.maxerr);
                wave_max_bits(pmax, pmin, seen_b, seen_z, J.pbits);
            }
        }
        __syncthreads();
        LINE_STAMP(lucas_line_stamps_col, stamp, 2);
        // Carry-save restoration + single ripple x_j = d_j + h_{j-1}
        // (llintToIrrBal<n>, IrrBaseBalanced.cu:52-110, and
        // loadIntToDoubleIBDWT, gpuLucas.cu:730-766), one work-item per row
        // over the row's local digits u = 0 .. 2NL-1 (digit 2*(c0-H) + u).
        // Digit t sums, for k = 0..NC, the field of product t-k that starts k
        // digits above it: bits [w_{t-k} + .. + w_{t-1}, + w_t) of |c_{t-k}|.
        // Walking up the row, a window holds the NC+1 products that feed the
        // current digit, each shifted down by the widths of the digits passed
        // (w_t <= 21, so every field is in the low 32 bits).
        const unsigned int mq = (1u << J.q) - 1u, mq1 = (2u << J.q) - 1u;
        // RR rows at once (r0, r0 + T, ...), their chains interleaved
        auto carry_rows = [&](auto rr_c, int r0) {
            constexpr int RR = decltype(rr_c)::value;
            unsigned int hw[RR][2 * NL], lw[RR][2 * NL];
#pragma unroll
            for (int i = 0; i < RR; ++i) {
                const int r = r0 + i * T;
#pragma unroll
                for (int lc = 0; lc < NL; ++lc) {
                    const int rr = (c0 - H + lc < 0) ? (r == 0 ? A - 1 : r - 1) : r;
                    const double2 v = lds[lc * RS + rr];
                    hw[i][2 * lc] = (unsigned int)__double2hiint(v.x);
                    lw[i][2 * lc] = (unsigned int)__double2loint(v.x);
                    hw[i][2 * lc + 1] = (unsigned int)__double2hiint(v.y);
                    lw[i][2 * lc + 1] = (unsigned int)__double2loint(v.y);
                }
            }
            // window slot k holds product t-k: (wh[k]:wl[k]) its remaining
            // magnitude, ws[k] its sign (+1 / -1); wh is 0 once a product has
            // been shifted by one width (its magnitude is below 2^36)
            unsigned int wh[RR][NC + 1], wl[RR][NC + 1];
            int ws[RR][NC + 1];
            int hprev[RR], h0[RR], dlo[RR];
#pragma unroll
            for (int i = 0; i < RR; ++i) {
                hprev[i] = h0[i] = dlo[i] = 0;
#pragma unroll
                for (int k = 0; k <= NC; ++k) wh[i][k] = wl[i][k] = 0u, ws[i][k] = 1;
            }
#pragma unroll
            for (int t = 2 * H - 1 - NC; t < 2 * NL; ++t) {
#pragma unroll
                for (int i = 0; i < RR; ++i) {
                    if (t > 2 * H - 1 - NC) {
                        // pass digit t-1: every product in the window moves
                        // down by its width; the oldest leaves
                        const int wp = J.q + (int)((hw[i][t - 1] >> 30) & 1u);
#pragma unroll
                        for (int k = NC; k >= 1; --k) {
                            wl[i][k] = __builtin_amdgcn_alignbit(wh[i][k - 1], wl[i][k - 1], wp);
                            wh[i][k] = wh[i][k - 1] >> wp;
                            ws[i][k] = ws[i][k - 1];
                        }
                    }
                    wh[i][0] = hw[i][t] & 0xFFFFFu;
                    wl[i][0] = lw[i][t];
                    ws[i][0] = ((int)hw[i][t] >> 31) | 1;
                }
                if (t < 2 * H - 1) continue;  // the window fills below the first digit restored
#pragma unroll
                for (int i = 0; i < RR; ++i) {
                    const bool wide = (hw[i][t] >> 30) & 1u;
                    const unsigned int M = wide ? mq1 : mq;
                    const int W = J.q + (int)wide;
                    int sum = 0;
#pragma unroll
                    for (int k = 0; k <= NC; ++k) sum += __mul24(ws[i][k], (int)(wl[i][k] & M));
                    // the reference's carry: (sum + half) >> W for sum >= half,
                    // -((half - sum) >> W) for sum < -half, else 0
                    const int half = (int)((M + 1u) >> 1);
                    const int h = (sum + half + (sum >> 31) + (sum == -half ? 1 : 0)) >> W;
                    const int d = sum - (h << W);
                    if (t == 2 * H - 1) {
                        hprev[i] = h;  // carry into the first own element
                    } else if ((t & 1) == 0) {
                        dlo[i] = d;
                        h0[i] = h;
                    } else {
                        const int lc = t >> 1, lo = lc - H, r = r0 + i * T;
                        const int2 xv = make_int2(dlo[i] + hprev[i], d + h0[i]);
                        hprev[i] = h;
                        if (out_digits) {
                            gstore2i(J.digits, (size_t)(c0 + lo) + (size_t)B * r, xv);
                        } else {
                            lds[lc * RS + r] = weighted(lc, r, xv);
                        }
                    }
                }
            }
        };
        // two rows per work-item where a second one exists (a wave-uniform
        // choice: A - T is a multiple of 64), so no pass runs a few waves alone
        for (int r = tid; r < A; r += 2 * T) {
            if (r + T < A) carry_rows(std::integral_constant<int, 2>{}, r);
            else carry_rows(std::integral_constant<int, 1>{}, r);
        }
        __syncthreads();
        LINE_STAMP(lucas_line_stamps_col, stamp, 3);
        if (out_digits) {
            if (REG && colw) {
                wave_max_err(err, seen, J.maxerr);
                wave_max_bits(pmax, pmin, seen_b, seen_z, J.pbits);
            }
            return;
        }
        if (own) {
#pragma unroll
            for (int e = 0; e < EL; ++e) {
                const int n = l + 64 * e;
                if constexpr (S == 1) {
                    x[e] = column[n];
                } else {
                    const double2 a = column[n], b = column[n + LH];
                    x[e] = h ? cmul(csub(a, b), cmul(wa, omega64(e * (64 / EA)))) : cadd(a, b);
                }
            }
        }
    }
    // both halves read before the transforms reuse the column
    if constexpr (S == 2) __syncthreads();
    // forward column transform of the own columns, outputs to the region
    if (own) {
        if constexpr (REG) F::fwd(x, region, twr, l);
        else F::fwd(x, region, twl, l);
        wave_sync();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int beta = l + 64 * i;
            if (F::ALL || beta < NA) {
#pragma unroll
                for (int d = 0; d < 8; ++d) region[d * NA + beta] = x[8 * i + d];
            }
        }
    }
    __syncthreads();
    LINE_STAMP(lucas_line_stamps_col, stamp, 4);
    // store row-major: the outer twiddle is the row kernel's
    for (int t = tid; t < A * K; t += T) {
        const int k = t / K, lo = t - k * K;
        gstore2_nt(J.spec, (size_t)k * J.ld + c0 + lo, lds[(H + lo) * RS + spec_pos(k)], J.nt);
    }
    LINE_STAMP(lucas_line_stamps_col, stamp, 5);
    if (REG && colw && !in_digits) {
        wave_max_err(err, seen, J.maxerr);
        wave_max_bits(pmax, pmin, seen_b, seen_z, J.pbits);
    }
}

// Column tile of workgroup `wg` of a single-job launch of n workgroups:
// contiguous tile ranges per XCD (workgroup 8k + x runs tile k of XCD x's
// range), so tiles that share halo lines share an L2.
__device__ __forceinline__ int line_col_tile(int wg, int n) {
    return (n % 8 == 0 && n >= 16) ? (n / 8) * (wg & 7) + (wg >> 3) : wg;
}

}  // namespace

#define LUCAS_LINE_COL_KERNELS(EA, K, NC, S)                                                                   \
    extern "C" __global__ void __launch_bounds__(64 * lucas_line_col_waves(K, NC, S))                           \
            LUCAS_WAVES_PER_SIMD(lucas_line_col_waves_per_simd(EA, K, NC, S))                                     \
            lucas_lcol_##EA##_##K##_##NC##_##S(const LucasJobDesc* __restrict__ jobs,                           \
                                               const int2* __restrict__ wgmap, int mode) {                       \
        const int2 jm = wgmap[blockIdx.x];                                                                       \
        line_col_body<EA, K, NC, S>(jobs[jm.x], jm.y, mode);                                                     \
    }                                                                                                            \
    extern "C" __global__ void __launch_bounds__(64 * lucas_line_col_waves(K, NC, S))                           \
            LUCAS_WAVES_PER_SIMD(lucas_line_col_waves_per_simd(EA, K, NC, S))                                     \
            lucas_lcol1_##EA##_##K##_##NC##_##S(const LucasJobDesc J, int mode) {                                \
        line_col_body<EA, K, NC, S>(J, line_col_tile(blockIdx.x, J.col_wgs), mode);                              \
    }
LUCAS_LINE_COLS(LUCAS_LINE_COL_KERNELS)

#define LUCAS_LINE_ROW_KERNELS(EB, S)                                                                         \
    extern "C" __global__ void __launch_bounds__(lucas_line_row_threads(S))                                   \
            lucas_lrow_##EB##_##S(const LucasJobDesc* __restrict__ jobs, const int2* __restrict__ wgmap) {   \
        const int2 jm = wgmap[blockIdx.x];                                                                    \
        line_row_any<EB, S>(jobs[jm.x], jm.y);                                                                \
    }                                                                                                         \
    extern "C" __global__ void __launch_bounds__(lucas_line_row_threads(S))                                   \
            lucas_lrow1_##EB##_##S(const LucasJobDesc J) {                                                    \
        line_row_any<EB, S>(J, blockIdx.x);                                                                   \
    }
LUCAS_LINE_ROWS(LUCAS_LINE_ROW_KERNELS)

namespace lucas {

namespace {
typedef void (*LineColKernel)(const LucasJobDesc*, const int2*, int);
typedef void (*LineCol1Kernel)(const LucasJobDesc, int);
typedef void (*LineRowKernel)(const LucasJobDesc*, const int2*);
typedef void (*LineRow1Kernel)(const LucasJobDesc);
struct LineCol {
    LineColKernel many;
    LineCol1Kernel one;
    int threads, lds_words;
};
struct LineRow {
    LineRowKernel many;
    LineRow1Kernel one;
    int threads, lds_words;
};
#define LUCAS_LINE_COL_ENTRY(EA, K, NC, S)                                                                      \
    {lucas_lcol_##EA##_##K##_##NC##_##S, lucas_lcol1_##EA##_##K##_##NC##_##S, 64 * lucas_line_col_waves(K, NC, S), \
     lucas_line_col_words(EA, K, NC, S)},
#define LUCAS_LINE_ROW_ENTRY(EB, S) \
    {lucas_lrow_##EB##_##S, lucas_lrow1_##EB##_##S, lucas_line_row_threads(S), lucas_line_row_words(EB)},
const LineCol kLineCols[] = {LUCAS_LINE_COLS(LUCAS_LINE_COL_ENTRY)};
const LineRow kLineRows[] = {LUCAS_LINE_ROWS(LUCAS_LINE_ROW_ENTRY)};
}  // namespace

hipError_t configure_line_kernels(size_t max_lds) {
    for (const LineCol& k : kLineCols) {
        if ((size_t)16 * k.lds_words > max_lds) return hipErrorInvalidValue;
        hipError_t e = hipFuncSetAttribute((const void*)k.many, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           16 * k.lds_words);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k.one, hipFuncAttributeMaxDynamicSharedMemorySize, 16 * k.lds_words);
        if (e != hipSuccess) return e;
    }
    for (const LineRow& k : kLineRows) {
        hipError_t e = hipFuncSetAttribute((const void*)k.many, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           16 * k.lds_words);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k.one, hipFuncAttributeMaxDynamicSharedMemorySize, 16 * k.lds_words);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Copies the line kernels' phase stamps (profiling library only).
hipError_t read_line_phase_stamps(unsigned long long* col, int ncol, unsigned long long* row, int nrow) {
#ifdef LUCAS_PHASE_STAMPS
    if (ncol < 0 || nrow < 0 || ncol > LUCAS_LINE_STAMP_WGS || nrow > LUCAS_LINE_STAMP_WGS) return hipErrorInvalidValue;
    hipError_t e =
            hipMemcpyFromSymbol(col, HIP_SYMBOL(lucas_line_stamps_col), sizeof(unsigned long long) * 8 * ncol);
    if (e == hipSuccess)
        e = hipMemcpyFromSymbol(row, HIP_SYMBOL(lucas_line_stamps_row), sizeof(unsigned long long) * 8 * nrow);
    return e;
#else
    (void)col, (void)ncol, (void)row, (void)nrow;
    return hipErrorNotSupported;
#endif
}

// Column launch of line instantiation `idx` (lucas_line_col_index): the
// batched kernel over (jobs, map) when `one` is null, else the single-job one.
hipError_t launch_line_col(int idx, const LucasJobDesc* jobs, const int2* map, const LucasJobDesc* one, int nwg,
                           int mode, hipStream_t stream) {
    const LineCol& k = kLineCols[idx];
    if (one)
        hipLaunchKernelGGL(k.one, dim3(nwg), dim3(k.threads), 16 * (size_t)k.lds_words, stream, *one, mode);
    else
        hipLaunchKernelGGL(k.many, dim3(nwg), dim3(k.threads), 16 * (size_t)k.lds_words, stream, jobs, map, mode);
    return hipGetLastError();
}

hipError_t launch_line_row(int idx, const LucasJobDesc* jobs, const int2* map, const LucasJobDesc* one, int nwg,
                           hipStream_t stream) {
    const LineRow& k = kLineRows[idx];
    if (one)
        hipLaunchKernelGGL(k.one, dim3(nwg), dim3(k.threads), 16 * (size_t)k.lds_words, stream, *one);
    else
        hipLaunchKernelGGL(k.many, dim3(nwg), dim3(k.threads), 16 * (size_t)k.lds_words, stream, jobs, map);
    return hipGetLastError();
}

}  // namespace lucas
