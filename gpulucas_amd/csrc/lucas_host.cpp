// lucas_host.cpp -- planning, extended-precision tables and residue arithmetic
// (host only).  Replaces the reference's host setup:
//   initConstantSymbols     gpuLucas.cu:222-237   (q, bases, masks)
//   setSliceAndDice         gpuLucas.cu:323-343   (carry depth n from q)
//   computeWeightVectors    gpuLucas.cu:673-686   (QD dd_real weights)
//   computeBitsPerWord(+8)  gpuLucas.cu:688-714   (widths and HI masks)
//   print_residue           gpuLucas.cu:1288-1352 (Res64)
//   rebalanceIrrIntSEQGPU   IrrBaseBalanced.cu:133-169 (final normalisation)
// Widths and weights derive from the exact integer r_j = (-j p) mod N:
// ceil(jp/N) - jp/N = r_j / N, so no cancellation-prone real arithmetic (the
// reason the reference needed double-double) is left; only 2^(r/N) and the
// twiddles use long double, rounded once to double.
#include "lucas_host.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <limits>

#include "lucas_desc.h"
#include "lucas_line.h"

namespace lucas {

static const long double kPi = 3.141592653589793238462643383279502884L;
// LDS one workgroup may allocate (column tile + digit buffer, or two rows):
// all of a gfx950 CU's 160 KB.
static const size_t kMaxLds = kMaxLdsBytes;
static const int kCUs = 256;  // compute units of an MI355X

bool valid_length(long long N) {
    if (N < 4 || (N & 1)) return false;
    for (int f : {2, 3, 5, 7})
        while (N % f == 0) N /= f;
    return N == 1;
}

int carry_depth(int q) {
    if (q >= 21) return -1;
    if (q >= 18) return 2;
    if (q >= 16) return 3;
    if (q >= 14) return 4;
    if (q >= 12) return 5;
    return 6;
}

std::vector<int> radix_plan(int L, int pref) {
    std::vector<int> r;
    int twos = 0;
    while (L % 2 == 0) { L /= 2; ++twos; }
    if (pref == 4) {
        // 2^k as 4s; an odd power takes one 8 (or a 2 when k == 1)
        int fours = twos / 2;
        if (twos % 2 && twos >= 3) {
            r.push_back(8);
            fours = (twos - 3) / 2;
        } else if (twos % 2) {
            r.push_back(2);
        }
        for (int i = 0; i < fours; ++i) r.push_back(4);
    } else {
        // 2^k as 8s, with a 4x4 instead of 8x2 when k mod 3 == 1.
        int eights = twos / 3, rem = twos % 3;
        if (rem == 1 && eights > 0) { --eights; rem = 4; }
        for (int i = 0; i < eights; ++i) r.push_back(8);
        if (rem == 1) r.push_back(2);
        if (rem == 2) r.push_back(4);
        if (rem == 4) { r.push_back(4); r.push_back(4); }
    }
    for (int f : {7, 5, 3})
        while (L % f == 0) { L /= f; r.push_back(f); }
    if (L != 1) r.clear();
    return r;
}

int stage_twiddle_count(int L, const std::vector<int>& rad) {
    (void)L;
    int n = 0, Ns = 1;
    for (int R : rad) {
        n += Ns * (R - 1);
        Ns *= R;
    }
    return n;
}

static long long stage_cost(const std::vector<int>& rad, long long elems, int threads) {
    // per-work-item element passes through LDS for one FFT of the tile
    long long per = (elems + threads - 1) / threads;
    return (long long)rad.size() * per;
}

static std::string base_plan(int p, int N, Plan* pl) {
    if (p < 7) return "p must be >= 7";
    if (!valid_length(N)) return "N must be even and of the form 2^a 3^b 5^c 7^d";
    if ((long long)N > (long long)p) return "N must not exceed p";
    pl->p = p;
    pl->N = N;
    pl->M = N / 2;
    pl->q = p / N;
    pl->b = p % N;
    pl->ncarry = carry_depth(pl->q);
    if (pl->ncarry < 0) return "p/N >= 21 bits per word is out of range";
    pl->halo = (pl->ncarry + 2) / 2;  // ceil((n+1)/2)
    // The restoration keeps product bits below the sum of n+1 digit widths;
    // a product has at most 2(q+1) + log2(N) + 1 bits, and only products
    // below 2^53 are exact FP64 integers in the first place.
    int coverage = (pl->ncarry + 1) * pl->q;
    int need = std::min(53, 2 * (pl->q + 1) + (int)ceil(log2((double)N)) + 1);
    if (coverage < need) return "word size too small for the carry-save restoration";
    return "";
}

// Modelled cost of a tiling (lower is better): per-work-item LDS passes on
// the critical path of each kernel, serialised when the grid exceeds about
// two workgroups per CU, plus a small term for redundant halo work.
static long long model_cost(const Plan& pl) {
    const int T = pl.threads;
    long long col_t = stage_cost(pl.radA, (long long)pl.A * (pl.C + pl.halo), T) +
                      stage_cost(pl.radA, (long long)pl.A * pl.C, T) + 8;
    long long row_t = 2 * stage_cost(pl.radB, 2LL * pl.B, T) + 8;
    col_t *= (pl.col_wgs + 511) / 512;
    row_t *= (pl.row_wgs + 511) / 512;
    // narrow tiles read short row segments: penalise C < 4 (uncoalesced)
    long long coal = pl.C >= 4 ? 0 : (4 / pl.C) * 4;
    long long total = (long long)pl.A * (2 * pl.C + pl.halo) * pl.col_wgs / 256 + 4LL * pl.B * pl.row_wgs / 256;
    return 64 * (col_t + row_t + coal) + total / 64;
}

// Wave-register factors W (256, 128, 64) a length-L transform over nseq
// sequences with T work-items admits: W | L, and the nseq*L/W sub-transforms
// fit one per wave of the workgroup; largest first, then 0 (none).  A larger
// W leaves fewer Stockham stages; a smaller one spreads the transform over
// more waves with shorter per-wave chains.
static std::vector<int> wave_factors(int L, int nseq, int T) {
    std::vector<int> out;
    for (int W : {256, 128, 64})
        if (L % W == 0 && (long long)nseq * (L / W) <= T / 64) out.push_back(W);
    out.push_back(0);
    return out;
}

// Words per work-item a transform with these Stockham stages can hold: a
// stage of radix R runs ceil(EMAX/R) butterflies of R words per work-item.
static int stage_capacity(const std::vector<int>& rad) {
    int cap = 8;
    for (int R : rad) cap = std::min(cap, (LUCAS_EMAX + R - 1) / R * R);
    return cap;
}

std::vector<Plan> plan_candidates(int p, int N, bool wide) {
    std::vector<Plan> out;
    Plan pl;
    if (!base_plan(p, N, &pl).empty()) return out;
    // every tile below holds at most 8 words per item of a 1024-item
    // workgroup, and C + H >= 3: A <= 8192/3 and B = M/A <= 4096
    const int a_lo = std::max(2, (pl.M + 4095) / 4096), a_hi = std::min(pl.M, 8192 / 3);
    for (int A = a_lo; A <= a_hi; ++A) {
        if (pl.M % A) continue;
        const int B = pl.M / A;
        if (B < pl.halo) continue;
        if (radix_plan(A, 8).empty() || radix_plan(B, 8).empty()) continue;
        for (int C : {1, 2, 4, 8, 16, 32, 64}) {
            if (B % C) continue;
            const long long tile = std::max<long long>((long long)A * (C + pl.halo), 2LL * B);
            for (int T : {LUCAS_THREADS_SMALL, LUCAS_THREADS_LARGE, LUCAS_THREADS_MAX}) {
                if (tile > (long long)(wide ? 8 : LUCAS_EMAX) * T) continue;
                if (wide && tile <= (long long)LUCAS_EMAX * T) continue;  // ordinary
                // a workgroup more than twice the tile would leave most work-items idle
                if (T > LUCAS_THREADS_SMALL && tile < T / 2) continue;
                const std::vector<int> wAs = wave_factors(A, C + pl.halo, T), wBs = wave_factors(B, 2, T);
                for (size_t ia = 0; ia < wAs.size(); ++ia)
                for (size_t ib = 0; ib < wBs.size(); ++ib)
                for (int pref : {8, 4}) {
                    const int wA = wAs[ia], wB = wBs[ib];
                    const int A1 = wA ? A / wA : A, B1 = wB ? B / wB : B;
                    std::vector<int> ra = radix_plan(A1, pref), rb = radix_plan(B1, pref);
                    if (pref == 4 && ra == radix_plan(A1, 8) && rb == radix_plan(B1, 8)) continue;
                    if ((ra.empty() && A1 > 1) || (rb.empty() && B1 > 1)) continue;
                    if ((int)ra.size() > LUCAS_MAX_RADICES || (int)rb.size() > LUCAS_MAX_RADICES) continue;
                    if (wide && ((long long)A * (C + pl.halo) > (long long)stage_capacity(ra) * T ||
                                 2LL * B > (long long)stage_capacity(rb) * T))
                        continue;
                    const size_t lds_col = 16 * (size_t)lucas_col_lds_words(
                            A, C, pl.halo, stage_twiddle_count(A1, ra), (int)ra.size(), wA / 64);
                    const size_t lds_row =
                            16 * (size_t)lucas_row_lds_words(B, stage_twiddle_count(B1, rb), (int)rb.size(), wB / 64);
                    if (lds_col > kMaxLds || lds_row > kMaxLds) continue;
                    Plan c = pl;
                    c.A = A;
                    c.B = B;
                    c.C = C;
                    c.threads = T;
                    c.wA = wA;
                    c.wB = wB;
                    c.radA = ra;
                    c.radB = rb;
                    c.col_wgs = B / C;
                    c.row_wgs = (A + 1) / 2;  // row pairs; rows 0 and A/2 share one
                    c.lds_col = lds_col;
                    c.lds_row = lds_row;
                    c.radix_pref = pref;
                    // more column workgroups than CUs: two per CU if their LDS fits
                    c.col_dense = T == LUCAS_THREADS_MAX && c.col_wgs > kCUs && 2 * lds_col <= kMaxLds;
                    out.push_back(c);
                }
            }
        }
    }
    std::stable_sort(out.begin(), out.end(), [](const Plan& a, const Plan& b) { return model_cost(a) < model_cost(b); });
    return out;
}

std::vector<Plan> line_candidates(int p, int N) {
    std::vector<Plan> out;
    Plan pl;
    if (!base_plan(p, N, &pl).empty()) return out;
    for (int EA = 1; 64 * EA <= pl.M; ++EA) {
        const int A = 64 * EA;
        if (pl.M % A || (pl.M / A) % 64) continue;
        const int B = pl.M / A, EB = B / 64;
        for (int S = 1; S <= 2; ++S) {
            const int ir = lucas_line_row_index(EB, S);
            if (ir < 0) continue;
            for (int CS = 1; CS <= 2; ++CS)
            for (int K = 1; K <= 16; ++K) {
                const int ic = lucas_line_col_index(EA, K, pl.ncarry, CS);
                if (ic < 0 || B % K) continue;
                Plan c = pl;
                c.A = A;
                c.B = B;
                c.C = K;
                c.threads = 64 * lucas_line_col_waves(K, pl.ncarry, CS);
                c.col_wgs = B / K;
                c.row_wgs = A / 2;
                c.line = 1;
                c.line_col = ic;
                c.line_row = ir;
                c.line_split = S - 1;
                c.line_col_split = CS - 1;
                c.ld = A >= 1024 && B % 256 == 0 ? B + 8 : B;
                c.lds_col = 16 * (size_t)lucas_line_col_words(EA, K, pl.ncarry, CS);
                c.lds_row = 16 * (size_t)lucas_line_row_words(EB);
                if (c.lds_col > kMaxLds || c.lds_row > kMaxLds) continue;
                out.push_back(c);
            }
        }
    }
    return out;
}

std::string make_line_plan(int p, int N, int A, int K, Plan* out, int nt, int split, int col_split) {
    Plan pl;
    std::string why = base_plan(p, N, &pl);
    if (!why.empty()) return why;
    for (const Plan& c : line_candidates(p, N)) {
        if ((A && c.A != A) || (K && c.C != K) || (split >= 0 && c.line_split != split) ||
            (col_split >= 0 && c.line_col_split != col_split))
            continue;
        *out = c;
        out->nt = nt ? 1 : 0;
        return "";
    }
    return "no line-kernel shape for this N and carry depth";
}

int line_kind(const Plan& pl) {
    return pl.line ? 1 + pl.nt + 2 * pl.line_split + 4 * pl.line_col_split : 0;
}

std::string make_line_plan_kind(int p, int N, int A, int K, int kind, Plan* out) {
    if (kind < 1 || kind > 8) return "line kind out of range";
    return make_line_plan(p, N, A, K, out, (kind - 1) & 1, ((kind - 1) >> 1) & 1, (kind - 1) >> 2);
}

// Measured fastest tilings per signal length (gpulucas_amd/tune.py); line:
// the line kind (line_kind: 1 for the line kernels with A rows and C own
// columns per column workgroup, +1 with non-temporal spectrum loads and
// stores, +2 with rows split over two waves, +4 with columns split).
struct TunedShape {
    int N, A, C, threads, pref, wA, wB, line;
};
static const TunedShape kTuned[] = {
#include "tuned_shapes.inc"
    {0, 0, 0, 0, 0, 0, 0, 0}};

std::string make_plan_exact(int p, int N, int A, int C, int threads, int pref, Plan* out, int wA, int wB,
                            bool wide) {
    Plan pl;
    std::string why = base_plan(p, N, &pl);
    if (!why.empty()) return why;
    for (const Plan& c : plan_candidates(p, N, wide)) {
        if (A && c.A != A) continue;
        if (C && c.C != C) continue;
        if (threads && c.threads != threads) continue;
        if (pref && c.radix_pref != pref) continue;
        if (wA >= 0 && c.wA != wA) continue;
        if (wB >= 0 && c.wB != wB) continue;
        *out = c;
        return "";
    }
    return "no admissible (A, B, C) tiling for this N";
}

std::string make_plan(int p, int N, int forceA, int forceC, Plan* out, bool shared) {
    if (!forceA && !forceC) {
        for (const TunedShape* t = kTuned; t->N; ++t) {
            if (t->N != N) continue;
            if (t->line) {
                // a single exponent's latency plan (batches keep the tiled kernels)
                if (!shared && make_line_plan_kind(p, N, t->A, t->C, t->line, out).empty()) return "";
                continue;
            }
            if (shared) {
                // Throughput over latency: a batch fills the device with
                // many jobs' workgroups, so the smallest workgroups come
                // first -- four 256-item workgroups per CU hide one
                // another's load and barrier latency, where a 1024-item one
                // runs alone on its CU -- each with the widest tile (least
                // halo recomputation) at the tuned split.  MI355X, 64
                // exponents: 1.65x the rate of 1024-item plans at p in
                // [1e6, 5e6], 2.1x at p in [1e5, 1.2e6]; 8 exponents near
                // 9e6: 1.09x; 4 near 3.4e7 (no 256-item tiling): the same.
                for (int T : {LUCAS_THREADS_SMALL, LUCAS_THREADS_LARGE})
                    for (int C : {8, 4, 2, 1})
                        if (C >= t->C && make_plan_exact(p, N, t->A, C, T, t->pref, out).empty()) return "";
                if (N > 65536) {
                    for (int C : {8, 4})
                        if (C > t->C && make_plan_exact(p, N, t->A, C, 1024, t->pref, out).empty()) return "";
                    if (make_plan_exact(p, N, t->A, t->C, 1024, t->pref, out).empty()) return "";
                }
                for (int C : {8, 4})
                    if (C > t->C && make_plan_exact(p, N, t->A, C, t->threads, t->pref, out).empty()) return "";
                if (make_plan_exact(p, N, t->A, t->C, t->threads, t->pref, out).empty()) return "";
            } else if (make_plan_exact(p, N, t->A, t->C, t->threads, t->pref, out, t->wA, t->wB).empty()) {
                return "";
            }
        }
    }
    const std::string why = make_plan_exact(p, N, forceA, forceC, 0, 0, out);
    if (why.empty() || make_plan_exact(p, N, forceA, forceC, 0, 0, out, -1, -1, true).empty()) return "";
    return why;
}

// Bits per word the default length rule allows at length N.  Calibrated on
// MI355X: max roundoff over 600 iterations from a random residue stays near
// 0.12 on this line (0.08 at 19.2 bits for N = 65536, 0.03 at 18.2 bits for
// N = 163840, 0.19 at 18.5 bits for N = 2160900).
double max_bits_per_word(int N) {
    double bits = 19.6 - 0.35 * log2((double)N / 65536.0);
    return std::min(20.5, std::max(16.0, bits));
}

// Relative cost of length N: words times weighted transform stages (each
// Stockham stage is one pass over the data; measured on MI355X, radix-7 and
// radix-5 passes cost more per element than radix 8/4, radix 2 and 3 less).
// Lengths whose complex length has fewer than 2^4 leave odd row lengths and
// one-column tiles: costed double.
static double length_cost(long long N) {
    double w = 0.0;
    for (int R : radix_plan((int)(N / 2), 8)) w += R == 7 ? 1.6 : R == 5 ? 1.2 : R == 2 ? 0.6 : R == 3 ? 0.9 : 1.0;
    long long m = N / 2;
    int twos = 0;
    while (m % 2 == 0 && twos < 4) {
        m /= 2;
        ++twos;
    }
    if (twos < 4 && N >= 4096) w *= 2.0;
    return (double)N * w;
}

// Default length among the lengths with an ordinary tiling, or (`wide`)
// among those with a wide one; 0 if none.
static int default_length_of(int p, bool wide) {
    auto admissible = [&](long long N) {
        Plan pl;
        return make_plan_exact(p, (int)N, 0, 0, 0, 0, &pl, -1, -1, wide).empty();
    };
    // No tiling exists past M = A*B with 2B and 3A (C + H >= 3) within the
    // words a 1024-item workgroup holds (4 per item, 8 for wide tilings).
    const long long words = (wide ? 8LL : (long long)LUCAS_EMAX) * LUCAS_THREADS_MAX;
    const long long max_M = (words / 2) * (words / 3);
    // smallest admissible length under the bits rule
    long long nmin = 0;
    for (long long N = 4; N <= (long long)p && N / 2 <= max_M; N += 2) {
        if (!valid_length(N)) continue;
        if ((double)p / (double)N > max_bits_per_word((int)N)) continue;
        if (admissible(N)) {
            nmin = N;
            break;
        }
    }
    if (!nmin) return 0;
    // cheapest transform among admissible lengths up to 15% longer
    long long best = nmin;
    double best_cost = length_cost(nmin);
    for (long long N = nmin + 2; N <= nmin + nmin * 15 / 100 && N <= (long long)p; N += 2) {
        if (!valid_length(N)) continue;
        const double c = length_cost(N);
        if (c >= best_cost) continue;
        if (!admissible(N)) continue;
        best = N;
        best_cost = c;
    }
    return (int)best;
}

int default_length(int p) {
    if (p < 7) return 0;
    const int N = default_length_of(p, false);
    return N ? N : default_length_of(p, true);
}

unsigned int magic_u32(unsigned int d) {
    if (d <= 1) return 0;
    return (unsigned int)(((1ULL << 32) + d - 1) / d);
}

void stage_table(int L, const std::vector<int>& rad, LucasStage* out) {
    int Ns = 1, off = 0;
    for (size_t s = 0; s < rad.size(); ++s) {
        LucasStage& st = out[s];
        st.twoff = off;
        st.pad = 0;
        off += Ns * (rad[s] - 1);
        st.R = rad[s];
        st.Ns = Ns;
        st.nb = L / rad[s];
        st.twstep = L / (Ns * rad[s]);
        st.mNs = magic_u32((unsigned)Ns);
        st.mnb = magic_u32((unsigned)st.nb);
        Ns *= rad[s];
    }
}

void fill_desc_shape(const Plan& plan, LucasJobDesc* d) {
    d->p = plan.p;
    d->N = plan.N;
    d->M = plan.M;
    d->A = plan.A;
    d->B = plan.B;
    d->q = plan.q;
    d->b = plan.b;
    d->ncarry = plan.ncarry;
    d->halo = plan.halo;
    d->C = plan.C;
    d->col_wgs = plan.col_wgs;
    d->row_wgs = plan.row_wgs;
    d->ld = plan.ld ? plan.ld : plan.B;
    d->threads = plan.threads;
    const int A1 = plan.wA ? plan.A / plan.wA : plan.A, B1 = plan.wB ? plan.B / plan.wB : plan.B;
    d->nradA = (int)plan.radA.size();
    d->nradB = (int)plan.radB.size();
    d->ntwsA = stage_twiddle_count(A1, plan.radA);
    d->ntwsB = stage_twiddle_count(B1, plan.radB);
    d->pA = plan.wA / 64;
    d->pB = plan.wB / 64;
    d->ohA = outer_hi_count(plan.A);
    d->ohB = outer_hi_count(plan.B);
    d->nt = plan.nt;
    stage_table(A1, plan.radA, d->stA);
    stage_table(B1, plan.radB, d->stB);
    d->m_A1 = magic_u32((unsigned)A1);
    d->m_B1 = magic_u32((unsigned)B1);
    d->m_sqA0 = magic_u32((unsigned)((plan.C + plan.halo) * std::max(1, plan.wA)));
    d->m_sqA1 = magic_u32((unsigned)(plan.C * std::max(1, plan.wA)));
    d->m_ncl = magic_u32((unsigned)(plan.C + plan.halo));
    d->m_C = magic_u32((unsigned)plan.C);
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

TableLayout table_layout(const Plan& pl) {
    TableLayout l;
    size_t off = 0;
    auto take = [&](size_t n) {
        const size_t o = off;
        off += align256(n);
        return o;
    };
    const size_t c2 = sizeof(double) * 2;  // one complex entry
    const int A1 = pl.wA ? pl.A / pl.wA : pl.A, B1 = pl.wB ? pl.B / pl.wB : pl.B;
    l.twA = take(c2 * pl.A);
    l.twMA = take(c2 * pl.A);
    l.twB = take(c2 * pl.B);
    // the column kernel's outer twiddles: two-factor rows for the dense
    // kernels, omega_M^t (t < B) below that
    l.twMB = take(pl.col_dense ? 0 : c2 * pl.B);
    l.ct1 = take(pl.col_dense ? c2 * pl.B * kOuterLo : 0);
    l.ct2 = take(pl.col_dense ? c2 * pl.B * outer_hi_count(pl.A) : 0);
    l.rt1 = take(c2 * pl.A * kOuterLo);
    l.rt2 = take(c2 * pl.A * outer_hi_count(pl.B));
    l.rlo = take(sizeof(int) * 2 * pl.B);
    l.rhi = take(sizeof(int) * pl.A);
    l.wlo = take(sizeof(double) * 2 * pl.B);
    l.iwlo = take(sizeof(double) * 2 * pl.B);
    l.whi = take(sizeof(double) * pl.A);
    l.iwhi = take(sizeof(double) * pl.A);
    l.twsA = take(c2 * std::max(1, stage_twiddle_count(A1, pl.radA)));
    l.twsB = take(c2 * std::max(1, stage_twiddle_count(B1, pl.radB)));
    l.lnA = take(pl.line ? c2 * (pl.A + 64) : 0);
    l.lnB = take(pl.line ? c2 * (pl.B + 64) : 0);
    l.acc = take(256);  // roundoff maximum and product-size diagnostic
    l.bytes = off;
    return l;
}

void table_image(const Plan& pl, const Tables& t, const TableLayout& l, char* img) {
    (void)pl;
    memset(img, 0, l.bytes);
    auto put = [&](size_t o, const void* src, size_t n) {
        if (n) memcpy(img + o, src, n);
    };
    put(l.twA, t.twA.data(), sizeof(double) * t.twA.size());
    put(l.twMA, t.twMA.data(), sizeof(double) * t.twMA.size());
    put(l.twB, t.twB.data(), sizeof(double) * t.twB.size());
    put(l.twMB, t.twMB.data(), sizeof(double) * t.twMB.size());
    put(l.ct1, t.ct1.data(), sizeof(double) * t.ct1.size());
    put(l.ct2, t.ct2.data(), sizeof(double) * t.ct2.size());
    put(l.rt1, t.rt1.data(), sizeof(double) * t.rt1.size());
    put(l.rt2, t.rt2.data(), sizeof(double) * t.rt2.size());
    put(l.rlo, t.rlo.data(), sizeof(int) * t.rlo.size());
    put(l.rhi, t.rhi.data(), sizeof(int) * t.rhi.size());
    put(l.wlo, t.wlo.data(), sizeof(double) * t.wlo.size());
    put(l.iwlo, t.iwlo.data(), sizeof(double) * t.iwlo.size());
    put(l.whi, t.whi.data(), sizeof(double) * t.whi.size());
    put(l.iwhi, t.iwhi.data(), sizeof(double) * t.iwhi.size());
    put(l.twsA, t.twsA.data(), sizeof(double) * t.twsA.size());
    put(l.twsB, t.twsB.data(), sizeof(double) * t.twsB.size());
    put(l.lnA, t.lnA.data(), sizeof(double) * t.lnA.size());
    put(l.lnB, t.lnB.data(), sizeof(double) * t.lnB.size());
}

void desc_tables(const TableLayout& l, char* m, LucasJobDesc* d) {
    d->twA = (const double2*)(m + l.twA);
    d->twMA = (const double2*)(m + l.twMA);
    d->twB = (const double2*)(m + l.twB);
    d->twMB = (const double2*)(m + l.twMB);
    d->ct1 = (const double2*)(m + l.ct1);
    d->ct2 = (const double2*)(m + l.ct2);
    d->rt1 = (const double2*)(m + l.rt1);
    d->rt2 = (const double2*)(m + l.rt2);
    d->rlo = (const int*)(m + l.rlo);
    d->rhi = (const int*)(m + l.rhi);
    d->wlo = (const double*)(m + l.wlo);
    d->iwlo = (const double*)(m + l.iwlo);
    d->whi = (const double*)(m + l.whi);
    d->iwhi = (const double*)(m + l.iwhi);
    d->twsA = (const double2*)(m + l.twsA);
    d->twsB = (const double2*)(m + l.twsB);
    d->lnA = (const double2*)(m + l.lnA);
    d->lnB = (const double2*)(m + l.lnB);
    d->maxerr = (unsigned int*)(m + l.acc);
    d->pbits = d->maxerr + 1;
}

size_t spec_bytes(const Plan& pl) {
    return sizeof(double) * 2 * (size_t)pl.A * (size_t)(pl.ld ? pl.ld : pl.B);
}

size_t job_bytes(const Plan& pl) {
    size_t s = 0;
    s += align256(spec_bytes(pl)) * 2;  // spec, spec2
    s += align256(sizeof(int) * (size_t)pl.N);             // digits
    return s + table_layout(pl).bytes;
}

static void root_table(long long count, long long denom, std::vector<double>* out) {
    out->resize(2 * (size_t)count);
    for (long long t = 0; t < count; ++t) {
        long double ang = 2.0L * kPi * (long double)t / (long double)denom;
        (*out)[2 * t] = (double)cosl(ang);
        (*out)[2 * t + 1] = (double)(-sinl(ang));
    }
}

// Stage twiddles of a length-L Stockham transform: stage s (radix R after
// Ns) holds w_L^(r*k*L/(Ns*R)) at [twoff + k*(R-1) + r-1], k < Ns, 1 <= r < R.
static void stage_twiddles(int L, const std::vector<int>& rad, std::vector<double>* out) {
    out->assign(2 * (size_t)std::max(1, stage_twiddle_count(L, rad)), 0.0);
    int Ns = 1;
    size_t off = 0;
    for (int R : rad) {
        const long long twstep = L / (Ns * R);
        for (int k = 0; k < Ns; ++k)
            for (int r = 1; r < R; ++r) {
                const long long e = ((long long)r * k * twstep) % L;
                long double ang = 2.0L * kPi * (long double)e / (long double)L;
                (*out)[2 * off] = (double)cosl(ang);
                (*out)[2 * off + 1] = (double)(-sinl(ang));
                ++off;
            }
        Ns *= R;
    }
}

void make_tables(const Plan& pl, Tables* t) {
    stage_twiddles(pl.wA ? pl.A / pl.wA : pl.A, pl.radA, &t->twsA);
    stage_twiddles(pl.wB ? pl.B / pl.wB : pl.B, pl.radB, &t->twsB);
    if (pl.line) {
        // line twiddles of a length-L transform: omega_L^(n*k1) at [k1*64 + n],
        // then omega_64^(a*c) at [L + c*8 + a]
        auto line_tw = [](int L, std::vector<double>* out) {
            out->resize(2 * (size_t)(L + 64));
            auto put = [&](size_t i, long long e, long long den) {
                const long double ang = 2.0L * kPi * (long double)e / (long double)den;
                (*out)[2 * i] = (double)cosl(ang);
                (*out)[2 * i + 1] = (double)(-sinl(ang));
            };
            for (int t = 0; t < L; ++t) put(t, (long long)(t & 63) * (t >> 6), L);
            for (int u = 0; u < 64; ++u) put(L + u, (long long)(u & 7) * (u >> 3), 64);
        };
        line_tw(pl.A >> pl.line_col_split, &t->lnA);  // split columns: the half-length transform
        line_tw(pl.B >> pl.line_split, &t->lnB);  // split rows: the half-length transform
    }
    root_table(pl.A, pl.A, &t->twA);
    root_table(pl.B, pl.B, &t->twB);
    root_table(pl.A, pl.M, &t->twMA);
    // outer twiddles omega_M^e, e = x*y mod M, in two factors per index
    auto outer = [&](std::vector<double>* out, int n_outer, int n_inner, long long step) {
        out->resize(2 * (size_t)n_outer * n_inner);
        for (long long x = 0; x < n_outer; ++x)
            for (long long y = 0; y < n_inner; ++y) {
                const long long e = (x * y * step) % (long long)pl.M;
                const long double ang = 2.0L * kPi * (long double)e / (long double)pl.M;
                (*out)[2 * (x * n_inner + y)] = (double)cosl(ang);
                (*out)[2 * (x * n_inner + y) + 1] = (double)(-sinl(ang));
            }
    };
    if (pl.col_dense) {
        outer(&t->ct1, pl.B, kOuterLo, 1);
        outer(&t->ct2, pl.B, outer_hi_count(pl.A), kOuterLo);
    } else {
        root_table(pl.B, pl.M, &t->twMB);
    }
    outer(&t->rt1, pl.A, kOuterLo, 1);
    outer(&t->rt2, pl.A, outer_hi_count(pl.B), kOuterLo);
    const long long N = pl.N, p = pl.p;
    t->rlo.resize(2 * (size_t)pl.B);
    t->wlo.resize(2 * (size_t)pl.B);
    t->iwlo.resize(2 * (size_t)pl.B);
    for (long long J = 0; J < 2LL * pl.B; ++J) {
        long long r = r_of(J, p, N);
        t->rlo[J] = (int)r;
        t->wlo[J] = (double)exp2l((long double)r / (long double)N);
        t->iwlo[J] = (double)exp2l(-(long double)r / (long double)N);
    }
    t->rhi.resize(pl.A);
    t->whi.resize(pl.A);
    t->iwhi.resize(pl.A);
    for (long long r = 0; r < pl.A; ++r) {
        long long rr = r_of(2LL * pl.B * r, p, N);
        t->rhi[r] = (int)rr;
        t->whi[r] = (double)exp2l((long double)rr / (long double)N);
        // inverse transform is unnormalised and the square step is 4x scaled
        t->iwhi[r] = (double)(exp2l(-(long double)rr / (long double)N) / (4.0L * (long double)pl.M));
    }
}

void digit_widths(int p, int N, int* widths) {
    const int q = p / N, b = p % N;
    for (long long j = 0; j < N; ++j) widths[j] = q + (r_of(j, p, N) < b ? 1 : 0);
}

void normalize_digits(int p, int N, const int32_t* digits, int32_t* out, uint64_t* res64, int* is_zero) {
    std::vector<int> w(N);
    digit_widths(p, N, w.data());
    std::vector<long long> d(digits, digits + N);
    long long carry = 0;
    for (int pass = 0; pass < 64; ++pass) {
        for (int j = 0; j < N; ++j) {
            long long t = d[j] + carry;
            long long v = t & ((1LL << w[j]) - 1);
            carry = (t - v) >> w[j];
            d[j] = v;
        }
        if (carry == 0) break;  // otherwise 2^p == 1: wrap into digit 0
    }
    bool allzero = true, allones = true;
    for (int j = 0; j < N; ++j) {
        if (d[j] != 0) allzero = false;
        if (d[j] != (1LL << w[j]) - 1) allones = false;
    }
    const bool zero = allzero || allones;
    uint64_t r = 0;
    if (!zero) {
        int off = 0;
        for (int j = 0; j < N && off < 64; ++j) {
            r |= (uint64_t)d[j] << off;
            off += w[j];
        }
    }
    if (out) {
        for (int j = 0; j < N; ++j) out[j] = zero ? 0 : (int32_t)d[j];
    }
    if (res64) *res64 = r;
    if (is_zero) *is_zero = zero ? 1 : 0;
}

}  // namespace lucas
