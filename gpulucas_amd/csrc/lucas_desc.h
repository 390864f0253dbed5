// lucas_desc.h -- per-exponent job descriptor shared by the host engine and the
// HIP kernels.  One descriptor describes one Mersenne exponent p transformed at
// length N; a kernel launch covers a list of descriptors (one exponent in the
// plain path, many in the batched and served paths).
//
// Data layout in HBM for one job (M = N/2 complex FP64 words, M = A*B):
//   element m = c + B*r packs digits (2m, 2m+1) as one complex word; the M
//   words form a row-major A x B matrix (row r, column c).
//   * spec  : written by the column kernel, read by the row kernel.
//   * spec2 : written by the row kernel, read by the column kernel.
//   * digits: N int32 pseudo-balanced digits (state between step() calls).
// Per-digit metadata is regenerated from small tables: digit j has
// r_j = (-j*p) mod N, width q+1 iff r_j < (p mod N), IBDWT weight 2^(r_j/N).
#pragma once

#include <hip/hip_runtime.h>

#define LUCAS_MAX_RADICES 16
// Elements (complex words) each work-item may hold while an FFT stage is in
// flight; a tile of a workgroup with T work-items holds at most EMAX*T words.
#define LUCAS_EMAX 4
// Work-items per workgroup: kernels are built for 256, 512 and 1024.
#define LUCAS_THREADS_SMALL 256
#define LUCAS_THREADS_LARGE 512
#define LUCAS_THREADS_MAX 1024
// Deepest carry-save restoration supported (reference carry depth n <= 6).
#define LUCAS_NMAX 6

// Column-kernel modes.
#define LUCAS_IN_DIGITS 1   // start from int32 digits instead of spec2
#define LUCAS_OUT_DIGITS 2  // stop after the carry and write int32 digits

// One Stockham stage of a length-L transform (L = product of the radices):
// radix R after Ns = product of the previous radices; nb = L/R butterflies
// per sequence; twiddle index step L/(Ns*R).  mNs and mnb are multiply-high
// reciprocals: q = umulhi(t, m) = t / d for t, d < 2^16 (m = ceil(2^32/d),
// unused when d == 1).
struct LucasStage {
    int R, Ns, nb, twstep;
    unsigned int mNs, mnb;
    int twoff;  // this stage's block in the stage twiddle table: [k][r-1], k < Ns
    int pad;
};

struct LucasJobDesc {
    int p, N, M, A, B;      // exponent, signal length, complex length, M = A*B
    int q, b;               // q = p / N (low word width), b = p % N (# wide words)
    int ncarry;             // carry-save depth n (products spread over n+1 digits)
    int halo;               // halo columns needed by the carry: ceil((n+1)/2)
    int C;                  // columns per column workgroup (B % C == 0)
    int col_wgs, row_wgs;   // workgroups per launch of each kernel
    int threads;            // work-items per workgroup (256, 512 or 1024)
    int nradA, nradB;
    int ntwsA, ntwsB;       // entries of the two stage twiddle tables
    // Each transform of length L = L1 * W runs as Stockham passes of length
    // L1 in LDS and a W = 64P point wave-register factor (P = 0: no wave
    // factor, L1 = L).
    int pA, pB;             // P of the column (length A) and row (length B) transforms
    int ohA, ohB;           // ceil(A/64), ceil(B/64): row lengths of ct2 and rt2
    int nt;                 // line kernels: spectra loaded and stored non-temporally
    int ld;                 // line kernels: row stride of the spectra, in words (>= B)
    LucasStage stA[LUCAS_MAX_RADICES];  // column transform Stockham stages (length A1)
    LucasStage stB[LUCAS_MAX_RADICES];  // row transform Stockham stages (length B1)
    unsigned int m_ncl, m_C;             // reciprocals of C+halo and C
    unsigned int m_A1, m_B1;             // reciprocals of A1 and B1
    unsigned int m_sqA0, m_sqA1;         // reciprocals of (C+halo)*64pA and C*64pA
    // Twiddle tables (forward sign, exp(-2*pi*i*t/L)), rounded from long double.
    const double2* twA;     // omega_A^t, t < A
    const double2* twB;     // omega_B^t, t < B
    const double2* twMA;    // omega_M^t, t < A
    const double2* twMB;    // omega_M^t, t < B (column kernels below the dense size)
    // Outer twiddles omega_M^(c*r) = ct1[c][r mod 64] * ct2[c][r div 64]
    // (column kernel) = rt1[r][c mod 64] * rt2[r][c div 64] (row kernel).
    const double2* ct1;     // [B][64]: omega_M^(c*rl)
    const double2* ct2;     // [B][ceil(A/64)]: omega_M^(64*c*rh)
    const double2* rt1;     // [A][64]: omega_M^(r*cl)
    const double2* rt2;     // [A][ceil(B/64)]: omega_M^(64*r*ch)
    // Line-kernel twiddles (line plans only): omega_L^(n*k1) at [k1*64 + n]
    // (L entries, L = A or B), then omega_64^(a*c) at [L + c*8 + a].
    const double2* lnA;
    const double2* lnB;
    const double2* twsA;    // stage twiddles of the length-A transform (see LucasStage)
    const double2* twsB;    // stage twiddles of the length-B transform
    // Weight tables: j = Jlo + 2B*r;  r_j = (rlo[Jlo] + rhi[r]) mod N.
    const int* rlo;         // (-Jlo*p) mod N, Jlo < 2B
    const int* rhi;         // (-2B*r*p) mod N, r < A
    const double* wlo;      // 2^(rlo/N)
    const double* whi;      // 2^(rhi/N)
    const double* iwlo;     // 2^(-rlo/N)
    const double* iwhi;     // 2^(-rhi/N) / (4M)   (inverse-FFT scale folded in)
    // State.
    double2* spec;
    double2* spec2;
    int* digits;
    unsigned int* maxerr;   // max |x - rint(x)| since last reset, as float bits
    // product-size diagnostic since last reset: [0] max bit length of the
    // rounded products |c_j|, [1] max 64-bit leading-zero count of |c_j|
    unsigned int* pbits;
};

// Dynamic LDS of the two kernels for one job, in 16-byte words.  Each kernel
// carves its LDS in the order listed.
//   column kernel: (C+H) x A tile; stage twiddles; per-row
//     {2^(rhi/N), 2^(-rhi/N)/(4M)}; per local digit {2^(rlo/N), 2^(-rlo/N)};
//     per-row rhi; per local digit rlo; stage table;
//     omega_A^t, t < A, and the wave-factor table (with a wave factor only);
//     int8 carries of the own elements, A x C; int8 carry into each row's
//     first own element, A
//   row kernel: two rows; stage twiddles; omega_B^t, t < B; stage table;
//     the wave-factor table (with a wave factor only)
// A tile of n words occupies whole 16-word blocks (its word index is permuted
// within each block).
__host__ __device__ inline int lucas_pad_words(int n) { return (n + 15) & ~15; }
// Twiddle table of a W = 64P point wave-register factor: omega_W^u, u < W/2,
// then omega_2H^t, t < H, for H = 1 .. 32 (63 entries, one pad word).
__host__ __device__ constexpr int lucas_wave_table_words(int P) { return P ? 32 * P + 64 : 0; }
__host__ __device__ inline int lucas_col_lds_words(int A, int C, int H, int ntws, int nst, int pA) {
    const int ncl = C + H;
    return lucas_pad_words(A * ncl) + ntws + A + 2 * ncl + (A + 3) / 4 + (2 * ncl + 3) / 4 + 2 * nst +
           (pA ? A : 0) + lucas_wave_table_words(pA) + (A * C + 15) / 16 + (A + 15) / 16;
}
__host__ __device__ inline int lucas_row_lds_words(int B, int ntws, int nst, int pB) {
    return lucas_pad_words(2 * B) + ntws + B + 2 * nst + lucas_wave_table_words(pB);
}
