// lucas_kernels.hip -- CDNA4 (gfx950) kernels for one Lucas-Lehmer squaring
// s <- s^2 - 2 mod 2^p - 1 by Crandall's balanced-digit IBDWT, as computed by
// the reference's mersenneIter (gpuLucas.cu:845-942).  Two kernels make one
// iteration, each one HBM round trip of the 8N-byte state:
//
//   lucas_row : forward row FFT (length B) + real-signal untangle, pointwise
//               square, re-tangle (replaces ComplexPointwiseSqr,
//               gpuLucas.cu:647-665, and the cuFFT D2Z/Z2D pair at
//               gpuLucas.cu:869/877) + inverse row FFT.
//   lucas_col : inverse column FFT (length A) with its outer twiddles,
//               unweight, -2, round (invDWTproductMinus2, gpuLucas.cu:774-805),
//               carry-save radix restoration (llintToIrrBal<n>,
//               IrrBaseBalanced.cu:52-110), single ripple carry and weighting
//               (loadIntToDoubleIBDWT, gpuLucas.cu:730-766), forward column FFT
//               with outer twiddles.
//
// The complex FFT of length M = N/2 is a two-pass (A x B) Stockham
// decomposition; each pass runs from LDS with one workgroup per tile.  The
// restoration needs the products of the n+1 preceding digits, which live in
// the H = ceil((n+1)/2) columns to the left of a column tile: each column
// workgroup recomputes those halo columns instead of exchanging them.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lucas_desc.h"
#include "lucas_device.h"

// Phase stamps (built only with LUCAS_PHASE_STAMPS, into the separate
// profiling library): work-item 0 of each of the first LUCAS_STAMP_WGS
// workgroups records the 100 MHz real-time counter at the phase boundaries of
// the steady-state (mode 0) column launches and of the row launches, six
// stamps per workgroup, into a per-device array the host reads back.  The
// default build compiles the marks to nothing.
#define LUCAS_STAMP_WGS 8192
#ifdef LUCAS_PHASE_STAMPS
__device__ unsigned long long lucas_stamps_col[LUCAS_STAMP_WGS * 8];
__device__ unsigned long long lucas_stamps_row[LUCAS_STAMP_WGS * 8];
#define LUCAS_STAMP_COL(k)                                                                 \
    do {                                                                                   \
        if (mode == 0 && threadIdx.x == 0 && blockIdx.x < LUCAS_STAMP_WGS)                 \
            lucas_stamps_col[blockIdx.x * 8 + (k)] = __builtin_amdgcn_s_memrealtime();     \
    } while (0)
#define LUCAS_STAMP_ROW(k)                                                                 \
    do {                                                                                   \
        if (threadIdx.x == 0 && blockIdx.x < LUCAS_STAMP_WGS)                              \
            lucas_stamps_row[blockIdx.x * 8 + (k)] = __builtin_amdgcn_s_memrealtime();     \
    } while (0)
#else
#define LUCAS_STAMP_COL(k) \
    do {                   \
    } while (0)
#define LUCAS_STAMP_ROW(k) \
    do {                   \
    } while (0)
#endif


namespace {

// LDS word of tile element i (16-byte words).  The low four bits (the bank
// group within a 256-byte row of banks) are XORed with bits 4-7 and 6-9, a
// permutation within each aligned 16-word block: power-of-two strides
// (Stockham writes at stride R*Ns, wave-factor outputs at stride L1*P)
// spread over all bank groups instead of piling onto one.
__device__ __forceinline__ int pidx(int i) { return i ^ ((i >> 4) & 15) ^ ((i >> 6) & 15); }

// ---------------------------------------------------------------------------
// Sequences in the LDS tile.  Sequence q (q < nsq) starts at word
//   base(q) = off + (q mod Q1)*s1 + (q div Q1)*s2     (Q1 = 0: off + q*s1)
// and its element j is word base(q) + j*sIdx (before padding).
// ---------------------------------------------------------------------------
struct SeqGeom {
    int nsq;
    unsigned int msq;  // reciprocal of nsq for udiv
    int Q1;
    unsigned int mQ1;
    int s1, s2, off, sIdx;
};

__device__ __forceinline__ int seq_base(const SeqGeom& g, int q) {
    if (g.Q1 == 0) return g.off + q * g.s1;
    const int hi = udiv(q, g.Q1, g.mQ1);
    return g.off + (q - hi * g.Q1) * g.s1 + hi * g.s2;
}

// Reciprocal of a power of two for udiv (2^k: t / 2^k = umulhi(t, 2^(32-k))).
__device__ __forceinline__ unsigned int pow2_recip(int d) {
    return d <= 1 ? 0u : (1u << (32 - (31 - __clz(d))));
}

// ---------------------------------------------------------------------------
// One Stockham autosort stage over the sequences of g, each of length L.
// Stage with radix R after Ns = product of the previous radices:
//   for j < L/R, k = j mod Ns:
//     v_r = x[j + r*L/R] * w_L^(r*k*L/(Ns*R))
//     y[(j/Ns)*Ns*R + k + s*Ns] = DFT_R(v)_s
// Reads of one stage all happen before its writes (registers, then barrier),
// so one LDS buffer serves all stages.  SEQ_FAST maps consecutive threads to
// consecutive sequences instead of consecutive butterflies.
// ---------------------------------------------------------------------------
template <int T, int R, int SIGN, bool SEQ_FAST>
__device__ __forceinline__ void fft_stage(double2* lds, const LucasStage& st, const SeqGeom& g, const double2* tw,
                                          int tid) {
    constexpr int BM = (LUCAS_EMAX + R - 1) / R;
    const int Ns = st.Ns, nb = st.nb, sIdx = g.sIdx;
    const int total = nb * g.nsq;
    double2 v[BM][R];
    int outb[BM];
#pragma unroll
    for (int i = 0; i < BM; ++i) {
        const int t = tid + i * T;
        if (t < total) {
            int seq, j;
            if (SEQ_FAST) {
                j = udiv(t, g.nsq, g.msq);
                seq = t - j * g.nsq;
            } else {
                seq = udiv(t, nb, st.mnb);
                j = t - seq * nb;
            }
            const int jq = udiv(j, Ns, st.mNs);
            const int k = j - jq * Ns;
            const int base = seq_base(g, seq);
            // this butterfly's R-1 twiddles (LDS)
            const double2* const twk = tw + st.twoff + k * (R - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = lds[pidx(base + (j + r * nb) * sIdx)];
            // each twiddle is read where it is used (fewer live registers); the
            // first stage's are all 1 (a wave-uniform branch skips them)
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[i][r] = SIGN < 0 ? cmul(v[i][r], twk[r - 1]) : cmulc(v[i][r], twk[r - 1]);
            }
            Dft<R, SIGN>::run(v[i]);
            outb[i] = base + (jq * Ns * R + k) * sIdx;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BM; ++i) {
        const int t = tid + i * T;
        if (t < total) {
#pragma unroll
            for (int r = 0; r < R; ++r) lds[pidx(outb[i] + r * Ns * sIdx)] = v[i][r];
        }
    }
    __syncthreads();
}

// Stage s of a table copied into LDS, made wave-uniform (scalar registers)
// so the radix dispatch below is a scalar branch.
__device__ __forceinline__ LucasStage uniform_stage(const LucasStage* lst, int s) {
    const int4* q = reinterpret_cast<const int4*>(lst + s);
    const int4 a = q[0], b = q[1];
    LucasStage st;
    st.R = __builtin_amdgcn_readfirstlane(a.x);
    st.Ns = __builtin_amdgcn_readfirstlane(a.y);
    st.nb = __builtin_amdgcn_readfirstlane(a.z);
    st.twstep = __builtin_amdgcn_readfirstlane(a.w);
    st.mNs = (unsigned int)__builtin_amdgcn_readfirstlane(b.x);
    st.mnb = (unsigned int)__builtin_amdgcn_readfirstlane(b.y);
    st.twoff = __builtin_amdgcn_readfirstlane(b.z);
    st.pad = 0;
    return st;
}

// Forward transforms only: an inverse transform is conj(DFT(conj(x))), so the
// kernels conjugate on the way in and out.
template <int T, int SIGN, bool SEQ_FAST>
__device__ __forceinline__ void fft_lds(double2* lds, const LucasStage* stages, int nst, const SeqGeom& g,
                                        const double2* tw, int tid) {
    for (int s = 0; s < nst; ++s) {
        const LucasStage st = uniform_stage(stages, s);
        switch (st.R) {
            case 2: fft_stage<T, 2, SIGN, SEQ_FAST>(lds, st, g, tw, tid); break;
            case 3: fft_stage<T, 3, SIGN, SEQ_FAST>(lds, st, g, tw, tid); break;
            case 4: fft_stage<T, 4, SIGN, SEQ_FAST>(lds, st, g, tw, tid); break;
            case 5: fft_stage<T, 5, SIGN, SEQ_FAST>(lds, st, g, tw, tid); break;
            case 7: fft_stage<T, 7, SIGN, SEQ_FAST>(lds, st, g, tw, tid); break;
            default: fft_stage<T, 8, SIGN, SEQ_FAST>(lds, st, g, tw, tid); break;
        }
    }
}

// ---------------------------------------------------------------------------
// Wave-register transforms of W = 64P points: one wave64 holds a sequence as
// v[m] = element 64m + lane and runs the radix-2 decimation-in-frequency
// levels in registers (half-sizes >= 64) and across lanes (32 .. 1) with no
// LDS traffic and no workgroup barrier.
// ---------------------------------------------------------------------------

// Value held by lane ^ MASK of the same wave: DPP quad permutes (1, 2), DPP
// row rotate by 8 (8), ds_swizzle bit mode within 32 lanes (4, 16),
// ds_bpermute (32).
template <int MASK>
__device__ __forceinline__ int lane_xor_i(int x) {
    if constexpr (MASK == 1) return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
    else if constexpr (MASK == 2) return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
    else if constexpr (MASK == 8) return __builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false);
    else if constexpr (MASK == 4 || MASK == 16) return __builtin_amdgcn_ds_swizzle(x, 0x1F | (MASK << 10));
    else return __builtin_amdgcn_ds_bpermute((int)((__lane_id() ^ 32u) << 2), x);
}

template <int MASK>
__device__ __forceinline__ double lane_xor_d(double v) {
    const int hi = lane_xor_i<MASK>(__double2hiint(v));
    const int lo = lane_xor_i<MASK>(__double2loint(v));
    return __hiloint2double(hi, lo);
}

// Radix-2 DIF level of half-size H across lanes.  The lane with bit H clear
// holds a and ends with a+b; its partner holds b and ends with (a-b)*w.
// s = +1 / -1 on the low / high lane; the low lane's w is omega^0.
template <int H>
__device__ __forceinline__ double2 xlevel(double2 v, double s, double2 w) {
    const double2 p = make_double2(lane_xor_d<H>(v.x), lane_xor_d<H>(v.y));
    return cmul(make_double2(fma(s, v.x, p.x), fma(s, v.y, p.y)), w);
}

// Bit reversal of m over log2(P) bits.
template <int P>
__device__ __forceinline__ constexpr int brev_slot(int m) {
    return P == 4 ? (((m & 1) << 1) | (m >> 1)) : m;
}

// Cross-lane level of half-size H for all P values of the lane.  tab holds
// omega_2H^t, t < H: the high lane takes omega_2H^(lane mod H), the low lane
// omega^0 = tab[0].  Lanes read one address each, consecutive or shared, so
// the LDS read is free of bank conflicts; it is issued just before use, so
// no level's twiddle stays live across the others.
template <int H, int P>
__device__ __forceinline__ void xlevel_all(double2 (&v)[P], const double2* tab, int lane) {
    const bool hi = (lane & H) != 0;
    const double2 w = tab[hi ? (lane & (H - 1)) : 0];
    const double s = hi ? -1.0 : 1.0;
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = xlevel<H>(v[m], s, w);
}

// LDS twiddle table of a W = 64P point wave factor (lucas_wave_table_words(P)
// entries): omega_W^u for u < W/2 (in-register levels), then for each
// cross-lane half-size H = 1, 2, .., 32 the H entries omega_2H^t at offset
// W/2 + H - 1.  The kernels fill it from their global table of omega_L
// (L = L1*W): entry t is omega_L^(u*L1), u = wave_table_exponent<P>(t).
// Exponent u (of omega_W) of entry t of that table.
template <int P>
__device__ __forceinline__ int wave_table_exponent(int t) {
    if (t < 32 * P) return t;
    const int i = t - 32 * P;                // (H - 1) + tt
    const int H = 1 << (31 - __clz(i + 1));  // largest power of two <= i + 1
    return (i - (H - 1)) * (32 * P / H);     // omega_2H^tt = omega_W^(tt*W/(2H))
}

// W = 64P point DIF of v (element 64m + lane in v[m]), twiddles from the
// wave table twW (see wave_table_exponent).  Afterwards v[m] holds
// X[brev6(lane)*P + brev_slot(m)].
template <int P>
__device__ __forceinline__ void wave_dif(double2 (&v)[P], const double2* twW, int lane) {
    // in-register levels, half-size h = 64*hs: omega_2h^e = omega_W^(e*P/(2hs)), e = (64m + lane) mod h
#pragma unroll
    for (int hs = P / 2; hs >= 1; hs >>= 1) {
#pragma unroll
        for (int m = 0; m < P; ++m) {
            if (m & hs) continue;
            const double2 a = v[m], b = v[m + hs];
            const int e = 64 * (m & (hs - 1)) + lane;
            v[m] = cadd(a, b);
            v[m + hs] = cmul(csub(a, b), twW[e * (P / (2 * hs))]);
        }
    }
    const double2* const xt = twW + 32 * P - 1;  // level H at xt + H
    xlevel_all<32, P>(v, xt + 32, lane);
    xlevel_all<16, P>(v, xt + 16, lane);
    xlevel_all<8, P>(v, xt + 8, lane);
    xlevel_all<4, P>(v, xt + 4, lane);
    xlevel_all<2, P>(v, xt + 2, lane);
    xlevel_all<1, P>(v, xt + 1, lane);
}

// Last factor of a length L = L1*W forward transform of the sequences of g,
// after the length-L1 Stockham passes over the W interleaved sub-sequences
// (n2 = element mod W).  Four-step: sub-transform (q, k1) takes elements
// W*k1 + n2 of sequence q, n2 = 64m + lane, multiplies them by
// omega_L^(n2*k1), runs the W-point DIF in one wave and writes X[k1 + L1*k2]
// in natural order.  Wave w takes sub-transform w (the planner keeps their
// number within the workgroup's waves); all reads precede all writes.
template <int T, int P, typename Post>
__device__ __forceinline__ void fft_wave(double2* lds, const SeqGeom& g, int L1, unsigned int mL1,
                                         const double2* twL, const double2* twW, int tid, Post post) {
    constexpr int W = 64 * P;
    const int lane = tid & 63, sub = tid >> 6;
    const bool active = sub < g.nsq * L1;  // wave-uniform
    double2 v[P];
    int ob = 0, q = 0, k1 = 0;
    if (active) {
        q = udiv(sub, L1, mL1);
        k1 = sub - q * L1;
        const int base = seq_base(g, q);
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int n2 = 64 * m + lane;
            const double2 x = lds[pidx(base + (W * k1 + n2) * g.sIdx)];
            v[m] = L1 > 1 ? cmul(x, twL[n2 * k1]) : x;
        }
        wave_dif<P>(v, twW, lane);
        ob = base + k1 * g.sIdx;
    }
    __syncthreads();
    if (active) {
        const int kb = (int)(__brev((unsigned int)lane) >> 26) * P;
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int k = L1 * (kb + brev_slot<P>(m));  // output element k + k1 of sequence q
            lds[pidx(ob + k * g.sIdx)] = post(q, k + k1, v[m]);
        }
    }
    __syncthreads();
}

// Identity post-operation of a transform's outputs.
struct PostNone {
    __device__ __forceinline__ double2 operator()(int, int, double2 v) const { return v; }
};

// Forward transform of length L of the sequences of g: the length-L1
// Stockham passes (stages, twiddles tws) over the sub-sequences of gA, then,
// when P > 0, the W = 64P wave-register factor, whose outputs pass through
// post(sequence, element, value) on their way to LDS.  P = 0: Stockham only
// over g (post is not applied).
template <int T, int P, typename Post = PostNone>
__device__ __forceinline__ void fft_seq(double2* lds, const LucasStage* stages, int nst, const double2* tws,
                                        const SeqGeom& g, const SeqGeom& gA, int L1, unsigned int mL1,
                                        const double2* twL, const double2* twW, int tid, Post post = Post()) {
    if constexpr (P > 0) {
        fft_lds<T, -1, true>(lds, stages, nst, gA, tws, tid);
        fft_wave<T, P>(lds, g, L1, mL1, twL, twW, tid, post);
    } else {
        fft_lds<T, -1, true>(lds, stages, nst, g, tws, tid);
    }
}


// Window of CNT rounded products (round_pack words) of digits 2m+1-k,
// k = k0 .. k0+CNT-1, for element m = (row r, local column lc) of the tile,
// top-down: |c| in mag, sign in neg, the digit's width in w.  A local column
// left of global column 0 is the previous row's halo (the cyclic wrap of
// IrrBaseBalanced.cu:64-70).
template <int CNT>
__device__ __forceinline__ void load_window(const double2* lds, const LucasJobDesc& J, int r, int lc, int c0mH,
                                            int ncl, int k0, unsigned long long* mag, bool* neg, int* w) {
#pragma unroll
    for (int i = 0; i < CNT; ++i) {
        const int k = k0 + i;
        const int ie = k >> 1;  // element offset back
        const int rr = (c0mH + lc - ie < 0) ? (r == 0 ? J.A - 1 : r - 1) : r;
        const double2 v = lds[pidx(rr * ncl + (lc - ie))];
        const unsigned long long bits = (unsigned long long)__double_as_longlong((k & 1) ? v.x : v.y);
        neg[i] = (bits >> 63) != 0;
        mag[i] = bits & ((1ull << 52) - 1);
        w[i] = J.q + (int)((bits >> 62) & 1);  // width of digit 2m+1-k
    }
}

// Carry-save restoration for element m = (row r, local column lc) of the
// tile: digits 2m+1 and 2m from the products of digits 2m+1 .. 2m-NC.
// Returns (d_2m, d_2m+1 + h_2m) and the carry h_2m+1.
template <int NC>
__device__ __forceinline__ int2 restore_element(const double2* lds, const LucasJobDesc& J, int r, int lc, int c0mH,
                                                int ncl, int* h1) {
    unsigned long long mag[NC + 2];
    bool neg[NC + 2];
    int w[NC + 2];
    load_window<NC + 2>(lds, J, r, lc, c0mH, ncl, 0, mag, neg, w);
    int h0;
    const int d1 = restore_digit<NC>(mag, neg, w, h1);               // digit 2m+1
    const int d0 = restore_digit<NC>(mag + 1, neg + 1, w + 1, &h0);  // digit 2m
    return make_int2(d0, d1 + h0);
}

// Carry h_2m-1 out of digit 2m-1, the digit before element m = (row r,
// local column lc): the ripple into the row's first own element.
template <int NC>
__device__ __forceinline__ int restore_carry_before(const double2* lds, const LucasJobDesc& J, int r, int lc,
                                                    int c0mH, int ncl) {
    unsigned long long mag[NC + 1];
    bool neg[NC + 1];
    int w[NC + 1];
    load_window<NC + 1>(lds, J, r, lc, c0mH, ncl, 2, mag, neg, w);
    int h;
    (void)restore_digit<NC>(mag, neg, w, &h);
    return h;
}

// ---------------------------------------------------------------------------
// Column kernel: one workgroup of T work-items per C-column tile of one job.
// The tile holds C+H columns, local column lc = global column c0-H+lc, element
// (r, lc) at lds[pidx(r*(C+H) + lc)].  Everything that does not depend on the
// data (stage twiddles, weight and r tables of the tile's rows and digits) is
// copied into LDS while the tile's own loads are in flight.
// ---------------------------------------------------------------------------
// DENSE: built for two 1024-item workgroups per CU (64 VGPRs), which loads
// its tile in smaller batches.
template <int T, int PA, bool DENSE = false>
__device__ __forceinline__ void col_kernel_body(const LucasJobDesc& J, int tile, int mode) {
    LUCAS_STAMP_COL(0);
    double2* const lds = lucas_dynamic_lds();
    const int A = J.A, B = J.B, C = J.C, H = J.halo, N = J.N;
    const int n = J.ncarry;
    const int c0 = tile * C;
    const int tid = threadIdx.x;
    const int ncl = C + H;  // local columns: [0,H) halo, [H, H+C) own
    const unsigned int mncl = J.m_ncl, mC = J.m_C;
    // LDS carve-up, in the order of lucas_col_lds_words()
    double2* p = lds + lucas_pad_words(A * ncl);
    double2* const tws = p;  // stage twiddles of the length-A transform
    p += J.ntwsA;
    double2* const roww = p;  // {2^(rhi/N), 2^(-rhi/N)/(4M)} per row
    p += A;
    double2* const colw = p;  // {2^(rlo/N), 2^(-rlo/N)} per local digit 2*lc + e
    p += 2 * ncl;
    int* const rowr = reinterpret_cast<int*>(p);  // rhi per row
    p += (A + 3) / 4;
    int* const colr = reinterpret_cast<int*>(p);  // rlo per local digit
    p += (2 * ncl + 3) / 4;
    LucasStage* const lst = reinterpret_cast<LucasStage*>(p);  // column transform stages
    p += 2 * J.nradA;
    double2* const twL = p;  // omega_A^t, t < A (wave factor only)
    p += PA ? A : 0;
    double2* const twW = p;  // wave-factor twiddles (wave_table_exponent)
    p += lucas_wave_table_words(PA);
    signed char* const hb = reinterpret_cast<signed char*>(p);  // carries h_2m+1 of own elements, [A][C]
    p += (A * C + 15) / 16;
    signed char* const hf = reinterpret_cast<signed char*>(p);  // carry into each row's first own element, [A]
    // the job's maximum so far: a relaxed device-coherent load, first needed
    // by the wave reductions at the end, so no wait follows it here
    const unsigned int seen = __hip_atomic_load(J.maxerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int seen_b = __hip_atomic_load(J.pbits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int seen_z = __hip_atomic_load(J.pbits + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double2* const twA = J.twA;
    const bool in_digits = (mode & LUCAS_IN_DIGITS) != 0;
    float err = 0.0f;
    double pmax = 0.0, pmin = 0x1p60;  // largest / smallest |c_j| of this work-item

    // The prologue issues every global load before it uses any result: a
    // loop that loads, waits and stores one element at a time keeps a single
    // request per work-item in flight, and the kernel then moves its tile at
    // a fraction of the HBM bandwidth.  The tile goes in batches of TB
    // elements per work-item (values and outer-twiddle factors in registers),
    // the first batch in flight while the tables are copied.
    constexpr int TB = DENSE ? 2 : 4;
    constexpr int wA = 64 * PA;
    const int A1 = PA ? A / wA : A;
    const int ntile = in_digits ? 0 : A * ncl;
    double2 tv[TB], ta[TB], tm[TB];
    // tile element t = (row r, local column lc) with the inverse outer
    // twiddle omega_M^(-c*r), conjugated.  omega_M^(c*r) is a product of two
    // table entries: for the dense kernels (M of a million and more) one
    // entry of each of the column's two twiddle rows (ct1, ct2), a few KB
    // per workgroup; below that omega_A^(e div B) * omega_M^(e mod B) with
    // e = c*r, from two tables of A and B entries that stay in L1 / L2.
    const double inv_B = 1.0 / (double)B;
    auto tile_issue = [&](int t0) {
#pragma unroll
        for (int k = 0; k < TB; ++k) {
            const int t = t0 + tid + k * T;
            if (t >= ntile) continue;
            const int r = udiv(t, ncl, mncl), lc = t - r * ncl;
            int gc = c0 - H + lc;
            if (gc < 0) gc += B;
            tv[k] = gload2(J.spec2, (size_t)r * B + gc);
            if constexpr (DENSE) {
                ta[k] = gload2(J.ct1, (size_t)gc * 64 + (r & 63));
                tm[k] = gload2(J.ct2, (size_t)gc * J.ohA + (r >> 6));
            } else {
                // e < M < 2^31; the quotient from the double reciprocal is
                // within one of e div B
                const int e = gc * r;
                int hi = (int)((double)e * inv_B);
                int lo = e - hi * B;
                if (lo < 0) {
                    --hi;
                    lo += B;
                } else if (lo >= B) {
                    ++hi;
                    lo -= B;
                }
                ta[k] = gload2(twA, hi);
                tm[k] = gload2(J.twMB, lo);
            }
        }
    };
    auto tile_store = [&](int t0) {
#pragma unroll
        for (int k = 0; k < TB; ++k) {
            const int t = t0 + tid + k * T;
            if (t < ntile) lds[pidx(t)] = cconj(cmulc(tv[k], cmul(ta[k], tm[k])));
        }
    };
    // The load counter drains in issue order, so the tables' loads go first:
    // waiting for them never waits for the tile's.  Tables, one pass: all
    // loads of an index before its stores; the first index of each
    // work-item is loaded before the tile and stored after it is issued.
    const int n_tws = J.ntwsA;                     // stage twiddles
    const int n_twL = (PA && A1 > 1) ? A : 0;      // omega_A^t (wave factor's outer twiddles)
    const int n_wav = PA ? 32 * PA + 63 : 0;       // wave-factor table
    const int n_col = 2 * ncl;                     // per local digit
    const int n_st = 8 * J.nradA;                  // stage table words
    const int nmax = max(max(max(n_tws, n_st), max(n_twL, n_wav)), max(A, n_col));
    const int* const st_src = reinterpret_cast<const int*>(J.stA);
    int* const st_dst = reinterpret_cast<int*>(lst);
    double2 v_tws, v_twL = make_double2(0.0, 0.0), v_wav = make_double2(0.0, 0.0);
    double v_wh, v_iwh, v_wl, v_iwl;
    int v_rh, v_rl, v_st;
    auto table_issue = [&](int t) {
        if (t < n_st) v_st = st_src[t];
        if (t < n_tws) v_tws = gload2(J.twsA, t);
        if (t < A) {
            if constexpr (PA > 0) v_twL = gload2(twA, t);
            v_wh = gload(J.whi, t);
            v_iwh = gload(J.iwhi, t);
            v_rh = gload(J.rhi, t);
        }
        if constexpr (PA > 0)
            if (t < n_wav) v_wav = gload2(twA, (size_t)wave_table_exponent<PA>(t) * A1);
        if (t < n_col) {
            int gc = c0 - H + (t >> 1);
            if (gc < 0) gc += B;
            const int jj = 2 * gc + (t & 1);
            v_wl = gload(J.wlo, jj);
            v_iwl = gload(J.iwlo, jj);
            v_rl = gload(J.rlo, jj);
        }
    };
    auto table_store = [&](int t) {
        if (t < n_st) st_dst[t] = v_st;
        if (t < n_tws) tws[t] = v_tws;
        if (t < n_twL) twL[t] = v_twL;
        if (t < n_wav) twW[t] = v_wav;
        if (t < A) {
            roww[t] = make_double2(v_wh, v_iwh);
            rowr[t] = v_rh;
        }
        if (t < n_col) {
            colw[t] = make_double2(v_wl, v_iwl);
            colr[t] = v_rl;
        }
    };
    table_issue(tid);
    if (ntile) tile_issue(0);
    table_store(tid);
    for (int t = tid + T; t < nmax; t += T) {
        table_issue(t);
        table_store(t);
    }
    if (ntile) {
        tile_store(0);
        for (int t0 = TB * T; t0 < ntile; t0 += TB * T) {
            tile_issue(t0);
            tile_store(t0);
        }
    }
    __syncthreads();
    LUCAS_STAMP_COL(1);  // tile and tables loaded

    // weight: (y_2m, y_2m+1) = 2^(r_j/N) (x_2m, x_2m+1) for own element (r, lo)
    auto weighted = [&](int r, int lo, int2 x) {
        const int e = 2 * (lo + H);
        const int rh = rowr[r];
        const double wh = roww[r].x;
        const double ae = colw[e].x * wh * (colr[e] + rh >= N ? 0.5 : 1.0);
        const double ao = colw[e + 1].x * wh * (colr[e + 1] + rh >= N ? 0.5 : 1.0);
        return make_double2(ae * (double)x.x, ao * (double)x.y);
    };

    if (in_digits) {
        // the tile's digits, TB elements per work-item per batch, loads first
        for (int t0 = 0; t0 < A * C; t0 += TB * T) {
            int2 x[TB];
#pragma unroll
            for (int k = 0; k < TB; ++k) {
                const int t = t0 + tid + k * T;
                if (t >= A * C) continue;
                const int r = udiv(t, C, mC), lo = t - r * C;
                x[k] = gload2i(J.digits, (size_t)(c0 + lo) + (size_t)B * r);
            }
#pragma unroll
            for (int k = 0; k < TB; ++k) {
                const int t = t0 + tid + k * T;
                if (t >= A * C) continue;
                const int r = udiv(t, C, mC), lo = t - r * C;
                lds[pidx(r * ncl + H + lo)] = weighted(r, lo, x[k]);
            }
        }
    } else {
        // inverse column FFT of the tile and halo
        const SeqGeom g{ncl, mncl, 0, 0u, 1, 0, 0, ncl};
        const SeqGeom gA{ncl * wA, J.m_sqA0, 0, 0u, 1, 0, 0, wA * ncl};
        // unweight, -2 on digit 0, round half-even, roundoff of element
        // (r, lc); the rounded products go back as round_pack words
        auto unweight = [&](int lc, int r, double2 v) -> double2 {
            int gc = c0 - H + lc;
            if (gc < 0) gc += B;
            v = cconj(v);
            const int rh = rowr[r];
            const double iwh = roww[r].y;
            const int e = 2 * lc;
            // r_j = rlo + rhi (mod N): the wrap halves the weight; wide iff r_j < b
            const int rje = colr[e] + rh, rjo = colr[e + 1] + rh;
            const double ie = colw[e].y * iwh * (rje >= N ? 2.0 : 1.0);
            const double io = colw[e + 1].y * iwh * (rjo >= N ? 2.0 : 1.0);
            double ve = v.x * ie;
            const double vo = v.y * io;
            if (gc == 0 && r == 0) ve -= 2.0;
            float ee, eo;
            double me, mo;
            const double pe = round_pack(ve, (rje >= N ? rje - N : rje) < J.b, &ee, &me);
            const double po = round_pack(vo, (rjo >= N ? rjo - N : rjo) < J.b, &eo, &mo);
            err = fmaxf(err, fmaxf(ee, eo));
            pmax = fmax(pmax, fmax(me, mo));
            pmin = fmin(pmin, fmin(me, mo));
            return make_double2(pe, po);
        };
        if constexpr (PA > 0) {
            // applied to the wave factor's outputs as they are written
            fft_seq<T, PA>(lds, lst, J.nradA, tws, g, gA, A1, J.m_A1, twL, twW, tid, unweight);
        } else {
            fft_seq<T, PA>(lds, lst, J.nradA, tws, g, gA, A1, J.m_A1, twL, twW, tid);
            for (int t = tid; t < A * ncl; t += T) {
                const int r = udiv(t, ncl, mncl), lc = t - r * ncl;
                lds[pidx(t)] = unweight(lc, r, lds[pidx(t)]);
            }
            __syncthreads();
        }
        // Carry-save restoration + single ripple x_j = d_j + h_{j-1}
        // (loadIntToDoubleIBDWT, gpuLucas.cu:730-766), one work-item per own
        // element, rows in order.  T is a multiple of C, so a pass holds whole
        // rows: its items read their own row's products and the halo of the
        // row before, which no pass writes.  Stage 1 restores digits 2m, 2m+1
        // and publishes h_2m+1 in hb; after the barrier stage 2 adds h_2m-1
        // from the element to the left (for a row's first element, from hf)
        // and overwrites its own element with the weighted new digits (or
        // stores them).
        LUCAS_STAMP_COL(2);  // inverse FFT and unweight done
        const int c0mH = c0 - H;
        // the carry into each row's first own element comes from its left
        // neighbour in the halo columns, which no pass overwrites: one
        // work-item per row, before the first barrier below
        for (int r = tid; r < A; r += T) {
            int h;
            switch (n) {
                case 2: h = restore_carry_before<2>(lds, J, r, H, c0mH, ncl); break;
                case 3: h = restore_carry_before<3>(lds, J, r, H, c0mH, ncl); break;
                case 4: h = restore_carry_before<4>(lds, J, r, H, c0mH, ncl); break;
                case 5: h = restore_carry_before<5>(lds, J, r, H, c0mH, ncl); break;
                default: h = restore_carry_before<6>(lds, J, r, H, c0mH, ncl); break;
            }
            hf[r] = (signed char)h;
        }
        const int nper = (A * C + T - 1) / T;
        for (int i = 0; i < nper; ++i) {
            const int t = tid + i * T;
            const bool own = t < A * C;
            const int r = udiv(own ? t : 0, C, mC), lo = (own ? t : 0) - r * C;
            int2 x = make_int2(0, 0);
            int h1 = 0;
            if (own) {
                const int lc = lo + H;
                switch (n) {
                    case 2: x = restore_element<2>(lds, J, r, lc, c0mH, ncl, &h1); break;
                    case 3: x = restore_element<3>(lds, J, r, lc, c0mH, ncl, &h1); break;
                    case 4: x = restore_element<4>(lds, J, r, lc, c0mH, ncl, &h1); break;
                    case 5: x = restore_element<5>(lds, J, r, lc, c0mH, ncl, &h1); break;
                    default: x = restore_element<6>(lds, J, r, lc, c0mH, ncl, &h1); break;
                }
                hb[t] = (signed char)h1;
            }
            __syncthreads();
            if (!own) continue;
            x.x += (int)(lo == 0 ? hf[r] : hb[t - 1]);
            if (mode & LUCAS_OUT_DIGITS) {
                gstore2i(J.digits, (size_t)(c0 + lo) + (size_t)B * r, x);
                continue;
            }
            lds[pidx(r * ncl + H + lo)] = weighted(r, lo, x);
        }
        if (mode & LUCAS_OUT_DIGITS) {
            wave_max_err(err, seen, J.maxerr);
            wave_max_bits(pmax, pmin, seen_b, seen_z, J.pbits);
            return;
        }
    }
    __syncthreads();
    LUCAS_STAMP_COL(3);  // carry done
    // forward column FFT of the own columns
    {
        const SeqGeom g{C, mC, 0, 0u, 1, 0, H, ncl};
        const SeqGeom gA{C * wA, J.m_sqA1, C, mC, 1, ncl, H, wA * ncl};
        fft_seq<T, PA>(lds, lst, J.nradA, tws, g, gA, A1, J.m_A1, twL, twW, tid);
    }
    LUCAS_STAMP_COL(4);  // forward FFT done
    // store row-major; the outer twiddle omega_M^(c*k) is applied by the row
    // kernel as it loads
    for (int t = tid; t < A * C; t += T) {
        const int k = udiv(t, C, mC), lo = t - k * C;
        gstore2(J.spec, (size_t)k * B + c0 + lo, lds[pidx(k * ncl + H + lo)]);
    }
    LUCAS_STAMP_COL(5);  // stores issued
    wave_max_err(err, seen, J.maxerr);
    if (!in_digits) wave_max_bits(pmax, pmin, seen_b, seen_z, J.pbits);
}

// ---------------------------------------------------------------------------
// Row kernel: one workgroup per row pair (k, A-k) of one job, 0 < k < A/2.
// Rows 0 and A/2 pair with themselves and share workgroup 0, so a launch has
// ceil(A/2) workgroups.  Frequency K = k + A*kc of the half-length complex
// transform Z pairs with M-K; the real-signal square is formed per pair as
//   E = Z_K + conj(Z_{M-K}),  O = -i (Z_K - conj(Z_{M-K}))   (2x scaled)
//   Z'_K     = E^2 + w O^2 + 2i E O,          w = omega_M^K
//   Z'_{M-K} = conj(E^2 + w O^2) + 2i conj(E O)
// which is 4x the packed spectrum of the squared real signal; the 1/4 is
// folded into the unweight table.
// ---------------------------------------------------------------------------
template <int T, int PB>
__device__ __forceinline__ void row_kernel_body(const LucasJobDesc& J, int r0) {
    LUCAS_STAMP_ROW(0);
    double2* const lds = lucas_dynamic_lds();
    const int A = J.A, B = J.B;
    const int tid = threadIdx.x;
    // workgroup 0: row 0, and row A/2 when A is even (both self-paired);
    // workgroup k > 0: rows k and A-k
    const bool self = r0 == 0;
    const int r1 = self ? A / 2 : A - r0;
    const int nrows = (self && (A & 1)) ? 1 : 2;
    // LDS carve-up, in the order of lucas_row_lds_words()
    double2* const tws = lds + lucas_pad_words(2 * B);  // stage twiddles of the length-B transform
    double2* const twu = tws + J.ntwsB;          // omega_B^t, t < B (untangle)
    LucasStage* const lst = reinterpret_cast<LucasStage*>(twu + B);  // row transform stages
    double2* const twW = twu + B + 2 * J.nradB;  // wave-factor twiddles (wave_table_exponent)

    // Both rows with the column kernel's outer twiddle omega_M^(c*k) (one
    // entry of each of the row's two twiddle rows, rt1 and rt2), and the
    // tables.  Every global load is issued before its result is used (see
    // the column kernel): the rows go in batches of 4 elements per work-item
    // (values and twiddle factors in registers), the first batch in flight
    // while the tables are copied.
    constexpr int TB = 4;
    constexpr int wB = 64 * PB;
    const int B1 = PB ? B / wB : B;
    const int ntile = nrows * B;
    double2 tv[TB], ta[TB], tm[TB];
    auto tile_issue = [&](int t0) {
#pragma unroll
        for (int k = 0; k < TB; ++k) {
            const int t = t0 + tid + k * T;
            if (t >= ntile) continue;
            const int rl = t >= B ? 1 : 0, kc = t - rl * B;
            const int row = rl ? r1 : r0;
            tv[k] = gload2(J.spec, (size_t)row * B + kc);
            ta[k] = gload2(J.rt1, (size_t)row * 64 + (kc & 63));
            tm[k] = gload2(J.rt2, (size_t)row * J.ohB + (kc >> 6));
        }
    };
    auto tile_store = [&](int t0) {
#pragma unroll
        for (int k = 0; k < TB; ++k) {
            const int t = t0 + tid + k * T;
            if (t < ntile) lds[pidx(t)] = cmul(tv[k], cmul(ta[k], tm[k]));
        }
    };
    // The load counter drains in issue order, so the tables' loads go first:
    // waiting for them never waits for the tile's.  Tables, one pass: all
    // loads of an index before its stores; the first index of each
    // work-item is loaded before the tile and stored after it is issued.
    const int n_tws = J.ntwsB;
    const int n_wav = PB ? 32 * PB + 63 : 0;
    const int n_st = 8 * J.nradB;  // stage table words
    const int nmax = max(max(n_tws, n_st), max(B, n_wav));
    const int* const st_src = reinterpret_cast<const int*>(J.stB);
    int* const st_dst = reinterpret_cast<int*>(lst);
    double2 v_tws, v_twu, v_wav = make_double2(0.0, 0.0);
    int v_st;
    auto table_issue = [&](int t) {
        if (t < n_st) v_st = st_src[t];
        if (t < n_tws) v_tws = gload2(J.twsB, t);
        if (t < B) v_twu = gload2(J.twB, t);
        if constexpr (PB > 0)
            if (t < n_wav) v_wav = gload2(J.twB, (size_t)wave_table_exponent<PB>(t) * B1);
    };
    auto table_store = [&](int t) {
        if (t < n_st) st_dst[t] = v_st;
        if (t < n_tws) tws[t] = v_tws;
        if (t < B) twu[t] = v_twu;
        if (t < n_wav) twW[t] = v_wav;
    };
    table_issue(tid);
    const double2 wr0 = gload2(J.twMA, r0), wr1 = gload2(J.twMA, r1);
    tile_issue(0);
    table_store(tid);
    for (int t = tid + T; t < nmax; t += T) {
        table_issue(t);
        table_store(t);
    }
    tile_store(0);
    for (int t0 = TB * T; t0 < ntile; t0 += TB * T) {
        tile_issue(t0);
        tile_store(t0);
    }
    const SeqGeom g{nrows, pow2_recip(nrows), 0, 0u, B, 0, 0, 1};
    const SeqGeom gA{nrows * wB, pow2_recip(nrows * wB), wB, pow2_recip(wB), 1, B, 0, wB};
    // forward row FFT; square; inverse row FFT (conjugated input).  Two call
    // sites: a loop over the two transforms costs registers across the loop.
    __syncthreads();
    LUCAS_STAMP_ROW(1);  // rows and tables loaded
    fft_seq<T, PB>(lds, lst, J.nradB, tws, g, gA, B1, J.m_B1, twu, twW, tid);
    LUCAS_STAMP_ROW(2);  // forward FFT done
    {
        // Pairs (kc in row r0) <-> (pc in row r1) for rows k, A-k: kc < B,
        // pc = B-1-kc.  Workgroup 0 pairs within each row instead: row 0 has
        // kc <= B/2 with pc = -kc mod B, then row A/2 (LDS row 1) has
        // kc <= (B-1)/2 with pc = B-1-kc.
        const int n0 = self ? B / 2 + 1 : B;
        const int npairs = self ? n0 + (nrows == 2 ? (B + 1) / 2 : 0) : B;
        for (int t = tid; t < npairs; t += T) {
            const bool half = self && t >= n0;  // row A/2 of workgroup 0
            const int kc = half ? t - n0 : t;
            int pc = (self && !half) ? B - kc : B - 1 - kc;
            if (pc == B) pc = 0;
            const int ko = half ? B + kc : kc;
            const int po = (self && !half) ? pc : B + pc;
            const double2 zk = lds[pidx(ko)], zm = lds[pidx(po)];
            const double2 w = cmul(half ? wr1 : wr0, twu[kc]);
            const double2 E = cadd(zk, cconj(zm));
            const double2 O = mul_si<-1>(csub(zk, cconj(zm)));
            const double2 P = cadd(cmul(E, E), cmul(w, cmul(O, O)));
            const double2 Q = cmul(E, O);
            // stored conjugated for the inverse transform
            lds[pidx(ko)] = make_double2(P.x - 2.0 * Q.y, -(P.y + 2.0 * Q.x));
            if (po != ko) lds[pidx(po)] = make_double2(P.x + 2.0 * Q.y, P.y - 2.0 * Q.x);
        }
    }
    __syncthreads();
    LUCAS_STAMP_ROW(3);  // square done
    fft_seq<T, PB>(lds, lst, J.nradB, tws, g, gA, B1, J.m_B1, twu, twW, tid);
    LUCAS_STAMP_ROW(4);  // inverse FFT done
    for (int t = tid; t < nrows * B; t += T) {
        const int rl = t >= B ? 1 : 0, kc = t - rl * B;
        gstore2(J.spec2, (size_t)(rl ? r1 : r0) * B + kc, cconj(lds[pidx(t)]));
    }
    LUCAS_STAMP_ROW(5);  // stores issued
}

// The single-job kernels take the descriptor as a kernel argument (4 KiB limit).
static_assert(sizeof(LucasJobDesc) <= 4096, "LucasJobDesc must fit the kernel argument segment");

// Column tile of workgroup `wg` of a single-job launch: the tiles are dealt
// to the 8 XCDs (workgroup 8k + x runs tile k of XCD x's contiguous range)
// so neighbouring tiles, which share halo lines, share an L2.
__device__ __forceinline__ int single_col_tile(int wg, int n) {
    return (n % 8 == 0 && n >= 16) ? (n / 8) * (wg & 7) + (wg >> 3) : wg;
}

}  // namespace

// One kernel per (workgroup size, wave factor P of its transform): each
// holds a single wave-factor variant, which keeps the code of a column/row
// kernel pair within the instruction cache.  The batched kernels read
// (job, tile) from a workgroup map and the job's descriptor from an array;
// the single-job kernels take the descriptor by value (kernel argument
// segment) and derive the tile from the workgroup id, which takes two
// dependent memory reads off the start of every workgroup.
#define LUCAS_KERNEL_PAIR(T, P)                                                                            \
    extern "C" __global__ void __launch_bounds__(T)                                                      \
            lucas_col_kernel_##T##_p##P(const LucasJobDesc* __restrict__ jobs,                            \
                                        const int2* __restrict__ wgmap, int mode) {                       \
        const int2 jm = wgmap[blockIdx.x];                                                               \
        col_kernel_body<T, P>(jobs[jm.x], jm.y, mode);                                                   \
    }                                                                                                    \
    extern "C" __global__ void __launch_bounds__(T)                                                      \
            lucas_row_kernel_##T##_p##P(const LucasJobDesc* __restrict__ jobs,                            \
                                        const int2* __restrict__ wgmap) {                                 \
        const int2 jm = wgmap[blockIdx.x];                                                               \
        row_kernel_body<T, P>(jobs[jm.x], jm.y);                                                         \
    }                                                                                                    \
    extern "C" __global__ void __launch_bounds__(T)                                                      \
            lucas_col1_kernel_##T##_p##P(const LucasJobDesc J, int mode) {                                 \
        col_kernel_body<T, P>(J, single_col_tile(blockIdx.x, J.col_wgs), mode);                          \
    }                                                                                                    \
    extern "C" __global__ void __launch_bounds__(T) lucas_row1_kernel_##T##_p##P(const LucasJobDesc J) {  \
        row_kernel_body<T, P>(J, blockIdx.x);                                                            \
    }
LUCAS_KERNEL_PAIR(256, 0)
LUCAS_KERNEL_PAIR(256, 1)
LUCAS_KERNEL_PAIR(256, 2)
LUCAS_KERNEL_PAIR(256, 4)
LUCAS_KERNEL_PAIR(512, 0)
LUCAS_KERNEL_PAIR(512, 1)
LUCAS_KERNEL_PAIR(512, 2)
LUCAS_KERNEL_PAIR(512, 4)
LUCAS_KERNEL_PAIR(1024, 0)
LUCAS_KERNEL_PAIR(1024, 1)
LUCAS_KERNEL_PAIR(1024, 2)
LUCAS_KERNEL_PAIR(1024, 4)

// Column kernels of 1024 work-items limited to 64 VGPRs (8 waves per SIMD),
// so two workgroups share a CU: for grids with more column workgroups than
// CUs.  The register cap costs a few spilled values.
#define LUCAS_DENSE_COL(P)                                                                                   \
    extern "C" __global__ void __launch_bounds__(1024, 8)                                                  \
            lucas_col_kernel_1024d_p##P(const LucasJobDesc* __restrict__ jobs,                              \
                                        const int2* __restrict__ wgmap, int mode) {                         \
        const int2 jm = wgmap[blockIdx.x];                                                                 \
        col_kernel_body<1024, P, true>(jobs[jm.x], jm.y, mode);                                            \
    }                                                                                                      \
    extern "C" __global__ void __launch_bounds__(1024, 8)                                                  \
            lucas_col1_kernel_1024d_p##P(const LucasJobDesc J, int mode) {                                   \
        col_kernel_body<1024, P, true>(J, single_col_tile(blockIdx.x, J.col_wgs), mode);                    \
    }
LUCAS_DENSE_COL(0)
LUCAS_DENSE_COL(1)
LUCAS_DENSE_COL(2)
LUCAS_DENSE_COL(4)

namespace lucas {

namespace {
typedef void (*ColKernel)(const LucasJobDesc*, const int2*, int);
typedef void (*RowKernel)(const LucasJobDesc*, const int2*);
typedef void (*Col1Kernel)(const LucasJobDesc, int);
typedef void (*Row1Kernel)(const LucasJobDesc);
// [workgroup size 256/512/1024][P 0/1/2/4]
const ColKernel kColKernels[4][4] = {
        {lucas_col_kernel_256_p0, lucas_col_kernel_256_p1, lucas_col_kernel_256_p2, lucas_col_kernel_256_p4},
        {lucas_col_kernel_512_p0, lucas_col_kernel_512_p1, lucas_col_kernel_512_p2, lucas_col_kernel_512_p4},
        {lucas_col_kernel_1024_p0, lucas_col_kernel_1024_p1, lucas_col_kernel_1024_p2, lucas_col_kernel_1024_p4},
        {lucas_col_kernel_1024d_p0, lucas_col_kernel_1024d_p1, lucas_col_kernel_1024d_p2,
         lucas_col_kernel_1024d_p4}};
const RowKernel kRowKernels[3][4] = {
        {lucas_row_kernel_256_p0, lucas_row_kernel_256_p1, lucas_row_kernel_256_p2, lucas_row_kernel_256_p4},
        {lucas_row_kernel_512_p0, lucas_row_kernel_512_p1, lucas_row_kernel_512_p2, lucas_row_kernel_512_p4},
        {lucas_row_kernel_1024_p0, lucas_row_kernel_1024_p1, lucas_row_kernel_1024_p2, lucas_row_kernel_1024_p4}};
const Col1Kernel kCol1Kernels[4][4] = {
        {lucas_col1_kernel_256_p0, lucas_col1_kernel_256_p1, lucas_col1_kernel_256_p2, lucas_col1_kernel_256_p4},
        {lucas_col1_kernel_512_p0, lucas_col1_kernel_512_p1, lucas_col1_kernel_512_p2, lucas_col1_kernel_512_p4},
        {lucas_col1_kernel_1024_p0, lucas_col1_kernel_1024_p1, lucas_col1_kernel_1024_p2,
         lucas_col1_kernel_1024_p4},
        {lucas_col1_kernel_1024d_p0, lucas_col1_kernel_1024d_p1, lucas_col1_kernel_1024d_p2,
         lucas_col1_kernel_1024d_p4}};
const Row1Kernel kRow1Kernels[3][4] = {
        {lucas_row1_kernel_256_p0, lucas_row1_kernel_256_p1, lucas_row1_kernel_256_p2, lucas_row1_kernel_256_p4},
        {lucas_row1_kernel_512_p0, lucas_row1_kernel_512_p1, lucas_row1_kernel_512_p2, lucas_row1_kernel_512_p4},
        {lucas_row1_kernel_1024_p0, lucas_row1_kernel_1024_p1, lucas_row1_kernel_1024_p2,
         lucas_row1_kernel_1024_p4}};
int threads_index(int threads) { return threads == 1024 ? 2 : (threads == 512 ? 1 : 0); }
int p_index(int p) { return p == 4 ? 3 : p; }
}  // namespace

hipError_t configure_kernels(size_t max_lds) {
    for (int t = 0; t < 4; ++t)
        for (int q = 0; q < 4; ++q) {
            hipError_t e = hipFuncSetAttribute((const void*)kColKernels[t][q],
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
            if (e == hipSuccess)
                e = hipFuncSetAttribute((const void*)kCol1Kernels[t][q], hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)max_lds);
            if (e == hipSuccess && t < 3)
                e = hipFuncSetAttribute((const void*)kRowKernels[t][q], hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)max_lds);
            if (e == hipSuccess && t < 3)
                e = hipFuncSetAttribute((const void*)kRow1Kernels[t][q], hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)max_lds);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

bool phase_stamps_built() {
#ifdef LUCAS_PHASE_STAMPS
    return true;
#else
    return false;
#endif
}

// Copies the stamps of the first ncol column and nrow row workgroups of the
// current device (eight words per workgroup, six used) after the device
// has finished the launches that wrote them.
hipError_t read_phase_stamps(unsigned long long* col, int ncol, unsigned long long* row, int nrow) {
#ifdef LUCAS_PHASE_STAMPS
    if (ncol < 0 || nrow < 0 || ncol > LUCAS_STAMP_WGS || nrow > LUCAS_STAMP_WGS) return hipErrorInvalidValue;
    hipError_t e = hipMemcpyFromSymbol(col, HIP_SYMBOL(lucas_stamps_col), sizeof(unsigned long long) * 8 * ncol);
    if (e == hipSuccess)
        e = hipMemcpyFromSymbol(row, HIP_SYMBOL(lucas_stamps_row), sizeof(unsigned long long) * 8 * nrow);
    return e;
#else
    (void)col, (void)ncol, (void)row, (void)nrow;
    return hipErrorNotSupported;
#endif
}

// dense: the 64-VGPR variant (1024 work-items only)
hipError_t launch_col_kernel(int threads, int p, int dense, const LucasJobDesc* jobs, const int2* map, int nwg,
                             size_t lds, int mode, hipStream_t stream) {
    const int ti = threads == 1024 && dense ? 3 : threads_index(threads);
    hipLaunchKernelGGL(kColKernels[ti][p_index(p)], dim3(nwg), dim3(threads), lds, stream, jobs, map, mode);
    return hipGetLastError();
}

hipError_t launch_row_kernel(int threads, int p, const LucasJobDesc* jobs, const int2* map, int nwg, size_t lds,
                             hipStream_t stream) {
    hipLaunchKernelGGL(kRowKernels[threads_index(threads)][p_index(p)], dim3(nwg), dim3(threads), lds, stream, jobs,
                       map);
    return hipGetLastError();
}

// Single-job launches: the descriptor travels as the kernel argument.
hipError_t launch_col1_kernel(int threads, int p, int dense, const LucasJobDesc& job, int nwg, size_t lds, int mode,
                              hipStream_t stream) {
    const int ti = threads == 1024 && dense ? 3 : threads_index(threads);
    hipLaunchKernelGGL(kCol1Kernels[ti][p_index(p)], dim3(nwg), dim3(threads), lds, stream, job, mode);
    return hipGetLastError();
}

hipError_t launch_row1_kernel(int threads, int p, const LucasJobDesc& job, int nwg, size_t lds, hipStream_t stream) {
    hipLaunchKernelGGL(kRow1Kernels[threads_index(threads)][p_index(p)], dim3(nwg), dim3(threads), lds, stream, job);
    return hipGetLastError();
}

}  // namespace lucas
