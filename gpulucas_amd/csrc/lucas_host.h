// lucas_host.h -- host-side planning, extended-precision tables and residue
// arithmetic of the engine (no HIP calls; unit-testable without a GPU).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "lucas_desc.h"

namespace lucas {

// LDS a workgroup may use on gfx950 (160 KB per CU, all of it per workgroup).
static const size_t kMaxLdsBytes = 163840;

struct Plan {
    int p = 0, N = 0, M = 0;
    int A = 0, B = 0, C = 0;
    int q = 0, b = 0;       // q = p / N, b = p % N
    int ncarry = 0;         // reference carry depth n
    int halo = 0;           // ceil((n+1)/2)
    int col_wgs = 0, row_wgs = 0;
    int threads = 0;        // work-items per workgroup (256, 512 or 1024)
    int radix_pref = 8;     // preferred power-of-two radix (8 or 4)
    int wA = 0, wB = 0;     // wave-register factor of each transform (64, 128, 256; 0: none)
    int col_dense = 0;      // column kernel at 8 waves per SIMD (two 1024-item workgroups per CU)
    // Line kernels (lucas_line.h): one wave per row or column, shapes compiled
    // in; line_col / line_row index their instantiations.  C is then the
    // column workgroup's own columns and threads its work-items.
    int line = 0, line_col = -1, line_row = -1;
    int nt = 0;  // line kernels: spectra moved with non-temporal loads and stores
    int line_split = 0;      // line kernels: each row held by two waves (a radix-2 step joins them)
    int line_col_split = 0;  // line kernels: each column held by two waves
    // Row stride of the spectra in words (0: B).  Line plans of 1024 or more
    // rows pad rows whose length is a multiple of 4 KB by 128 bytes: a
    // column tile's reads, 64-96 bytes of each of 1024 rows, then rotate
    // over the memory channels instead of all landing at one offset modulo
    // 4 KB (MI355X, N = 2^21: 33.2 -> 32.4 us/iteration; at 512 and 256 rows
    // the padding measured 0.6-0.8 % slower, so it is not used there).
    int ld = 0;
    std::vector<int> radA, radB;  // Stockham radices of the remaining factors A/wA, B/wB
    size_t lds_col = 0, lds_row = 0;  // bytes of LDS per workgroup
};

// N = 2^a 3^b 5^c 7^d and even.
bool valid_length(long long N);
// Reference carry depth rule (setSliceAndDice, gpuLucas.cu:323-343); -1 if q >= 21.
int carry_depth(int q);
// Radix sequence for a Stockham FFT of length L (radices 8, 4, 2, 3, 5, 7),
// powers of two as 8s (pref 8) or 4s (pref 4).
std::vector<int> radix_plan(int L, int pref = 8);
// Entries of the stage twiddle table of a length-L transform.
int stage_twiddle_count(int L, const std::vector<int>& rad);
// Full plan for (p, N) with optional forced A / C (0 = choose).  Returns an
// empty string on success or the reason for refusal.
// `shared`: the plan of a job in a batch, where throughput counts rather
// than one job's latency: the tuned split and workgroup size with the widest
// admissible tile (C up to 8, the least halo recomputation) and the largest
// wave factors.
std::string make_plan(int p, int N, int A, int C, Plan* out, bool shared = false);
// Every admissible (A, C) tiling of (p, N), cheapest modelled first.
// Ordinary tilings hold at most LUCAS_EMAX*T words per workgroup.  `wide`
// lists instead the larger tilings only transforms of radix 3, 5, 7 and 8
// stages can run (a stage of radix R holds ceil(EMAX/R)*R words per
// work-item, up to 8); they are offered where no ordinary tiling exists,
// which extends the longest lengths.
std::vector<Plan> plan_candidates(int p, int N, bool wide = false);
// Plan matching a full description (A, C, threads, pref 0 = any; wave
// factors wA, wB < 0 = any) among the ordinary (or with `wide`, the wide)
// tilings; "" on success.
std::string make_plan_exact(int p, int N, int A, int C, int threads, int pref, Plan* out, int wA = -1, int wB = -1,
                            bool wide = false);
// Line-kernel plan for (p, N) with A rows and K own columns per column
// workgroup (A = 0 or K = 0: the first instantiated choice); "" on success.
// `nt`: spectra moved with non-temporal loads and stores (faster once the
// state outgrows the caches).
// split / col_split: 0 one wave per row / column, 1 two, -1 either (the
// first candidate).
std::string make_line_plan(int p, int N, int A, int K, Plan* out, int nt = 0, int split = -1, int col_split = -1);
// The line kind of a plan (lucas_plan_info.line): 1 + non-temporal +
// 2 * split rows + 4 * split columns; 0 for the tiled kernels.
int line_kind(const Plan& pl);
// Plan of line kind `kind` (1..8) with A rows and K own columns.
std::string make_line_plan_kind(int p, int N, int A, int K, int kind, Plan* out);
// Every line-kernel plan of (p, N).
std::vector<Plan> line_candidates(int p, int N);
// Default N for p: smallest admissible length under the bits-per-word rule.
int default_length(int p);
// Bits per word allowed at length N by the default rule.
double max_bits_per_word(int N);

// Multiply-high reciprocal of d: umulhi(t, m) == t / d for t, d < 2^16.
unsigned int magic_u32(unsigned int d);
// Stage table of a length-L transform with radix sequence `rad`.
void stage_table(int L, const std::vector<int>& rad, LucasStage* out);
// Fill every descriptor field that depends only on the plan (not on memory).
void fill_desc_shape(const Plan& plan, LucasJobDesc* d);

// Device footprint of one job under `plan` (state + tables), bytes.
// Bytes of one spectrum (A rows of the plan's row stride).
size_t spec_bytes(const Plan& plan);
size_t job_bytes(const Plan& plan);

// Outer twiddles omega_M^(c*r) (c < B, r < A) as the product of two table
// entries: split a row index r = 64*rh + rl (column kernel) or a column
// index c = 64*ch + cl (row kernel).
static const int kOuterLo = 64;
inline int outer_hi_count(int L) { return (L + kOuterLo - 1) / kOuterLo; }

// Tables (host copies).
struct Tables {
    std::vector<double> twA, twB, twMA, twMB;  // interleaved re,im
    std::vector<double> ct1, ct2;              // [c][rl]: omega_M^(c*rl); [c][rh]: omega_M^(c*64*rh)
    std::vector<double> rt1, rt2;              // [r][cl]: omega_M^(r*cl); [r][ch]: omega_M^(r*64*ch)
    std::vector<double> twsA, twsB;            // stage twiddles, interleaved
    std::vector<double> lnA, lnB;              // line-kernel twiddles (line plans), interleaved
    std::vector<int> rlo, rhi;
    std::vector<double> wlo, whi, iwlo, iwhi;
};
void make_tables(const Plan& plan, Tables* t);

// Device layout of a job's tables and accumulators: byte offsets (from the
// start of the table block, each 256-aligned) and the block's size.
struct TableLayout {
    size_t twA, twMA, twB, twMB, ct1, ct2, rt1, rt2, rlo, rhi, wlo, iwlo, whi, iwhi, twsA, twsB, lnA, lnB, acc;
    size_t bytes;
};
TableLayout table_layout(const Plan& plan);
// Host image of the table block (layout.bytes, accumulators zero).
void table_image(const Plan& plan, const Tables& t, const TableLayout& lay, char* img);
// Points a descriptor's table and accumulator fields into the block at `base`.
void desc_tables(const TableLayout& lay, char* base, LucasJobDesc* d);

// r_j = (-j p) mod N.
inline long long r_of(long long j, long long p, long long N) {
    long long r = (-(j % N) * (p % N)) % N;
    return r < 0 ? r + N : r;
}
// Reference digit widths and weights for comparison (computeBitsPerWord,
// computeWeightVectors), computed exactly / in long double.
void digit_widths(int p, int N, int* widths);

// Normalise a pseudo-balanced digit vector mod 2^p - 1 into digits in
// [0, 2^w_j); returns Res64 and whether the residue is 0.
void normalize_digits(int p, int N, const int32_t* digits, int32_t* out, uint64_t* res64, int* is_zero);

}  // namespace lucas
