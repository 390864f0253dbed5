// lucas_line.h -- the line kernels' instantiations and LDS sizes, shared by
// the host planner and the kernel file.
//
// Line kernels give every wave one whole line of the A x B matrix of complex
// words: a row (row kernel) or a column (column kernel) of length L = 64*E.
// Lane l holds elements l + 64*e (e < E) in registers; the transform runs a
// radix-E step in registers and the remaining 64 points as two radix-8 steps
// exchanged through the wave's own LDS region (no workgroup barrier).  The
// shapes are compile-time: the column kernel is instantiated per (EA, K, NC)
// -- A = 64*EA rows, K own columns per workgroup, carry depth NC, halo
// H = ceil((NC+1)/2) columns -- and the row kernel per EB (B = 64*EB).
#pragma once

// X(EA, K, NC, S): S waves per column (S = 2: each wave holds one half, EA/2
// words per lane, and a radix-2 step joins the halves)
#define LUCAS_LINE_COLS(X)                                                                                  \
    X(1, 2, 2, 1) X(1, 2, 3, 1) X(1, 4, 2, 1) X(1, 4, 3, 1)                                                 \
    X(2, 2, 2, 1) X(2, 2, 3, 1) X(2, 4, 2, 1) X(2, 4, 3, 1)                                                 \
    X(3, 2, 2, 1) X(3, 2, 3, 1) X(3, 4, 2, 1) X(3, 4, 3, 1)                                                 \
    X(4, 1, 2, 1) X(4, 2, 2, 1) X(4, 2, 3, 1) X(4, 4, 2, 1) X(4, 4, 3, 1)                                   \
    X(5, 2, 2, 1) X(5, 2, 3, 1) X(5, 4, 2, 1) X(5, 4, 3, 1)                                                 \
    X(7, 2, 2, 1) X(7, 2, 3, 1) X(7, 4, 2, 1) X(7, 4, 3, 1)                                                 \
    X(8, 2, 2, 1) X(8, 2, 3, 1) X(8, 4, 2, 1) X(8, 4, 3, 1)                                                 \
    X(16, 2, 2, 1) X(16, 2, 3, 1) X(16, 4, 2, 1) X(16, 4, 3, 1)                                             \
    X(16, 2, 2, 2) X(16, 2, 3, 2) X(16, 4, 2, 2) X(16, 4, 3, 2)
// X(EB, S): rows of B = 64*EB words, S waves per row (S = 2: each wave holds
// one half, EB/2 words per lane, and a radix-2 step joins the halves)
#define LUCAS_LINE_ROWS(X) \
    X(1, 1) X(2, 1) X(3, 1) X(4, 1) X(5, 1) X(7, 1) X(8, 1) X(16, 1) X(8, 2) X(16, 2) X(32, 2)

// Halo columns of carry depth n: the restoration of a column's first digits
// reads the products of the n+1 digits before them.
inline constexpr int lucas_line_halo(int nc) { return (nc + 2) / 2; }

// Waves of a column workgroup: S per local column, rounded up to a multiple
// of the CU's 4 SIMDs (the extra waves take a share of the row-wise carry
// and of the tile loads and stores, so every SIMD carries the same).
inline constexpr int lucas_line_col_waves(int K, int NC, int S) {
    return (S * (K + lucas_line_halo(NC)) + 3) / 4 * 4;
}
// The column transform's twiddles: an LDS copy for lines of more than 8
// words per lane, else registers.
inline constexpr bool lucas_line_col_twl(int EA, int S) { return EA / S > 8; }
// LDS of a column workgroup, in 16-byte words: the NL = K + H column regions
// (A + 1 words each, the pad keeping neighbouring columns in different bank
// groups); the column transform's twiddles omega_L^(l*k1) [k1][l] and
// omega_64^t (LDS twiddles only); per row {2^(rhi/N), 2^(-rhi/N)/(4M)} and
// rhi; per local digit {2^(rlo/N), 2^(-rlo/N)} and rlo.
inline constexpr int lucas_line_col_words(int EA, int K, int NC, int S) {
    const int A = 64 * EA, NL = K + lucas_line_halo(NC);
    return NL * (A + 1) + (lucas_line_col_twl(EA, S) ? A + 64 : 0) + A + (A + 3) / 4 + 2 * NL + (2 * NL + 3) / 4;
}
// LDS of a row workgroup (one row pair, 2S waves): one region per wave (the
// two rows) and the outer twiddle factors rt2 of both rows (the transform's
// twiddles are held in registers).
inline constexpr int lucas_line_row_words(int EB) { return 2 * 64 * EB + 2 * EB; }
// Work-items of a row workgroup: S waves for each of the pair's two rows.
inline constexpr int lucas_line_row_threads(int S) { return 128 * S; }

// Waves per SIMD a column kernel is compiled for (its register budget): two
// workgroups per CU where the LDS holds two, so that one workgroup's loads
// and stores overlap the other's transforms; else one.
inline constexpr int lucas_line_col_waves_per_simd(int EA, int K, int NC, int S) {
    return (2 * 16 * lucas_line_col_words(EA, K, NC, S) <= 160 * 1024 ? 2 : 1) * lucas_line_col_waves(K, NC, S) / 4;
}

// Index of an instantiation in the lists above, or -1.
inline int lucas_line_col_index(int EA, int K, int NC, int S) {
    int i = 0;
#define LUCAS_LINE_X(a, k, n, s)                          \
    if (EA == (a) && K == (k) && NC == (n) && S == (s)) \
        return i;                                       \
    ++i;
    LUCAS_LINE_COLS(LUCAS_LINE_X)
#undef LUCAS_LINE_X
    return -1;
}
inline int lucas_line_row_index(int EB, int S) {
    int i = 0;
#define LUCAS_LINE_X(b, s)          \
    if (EB == (b) && S == (s)) \
        return i;                 \
    ++i;
    LUCAS_LINE_ROWS(LUCAS_LINE_X)
#undef LUCAS_LINE_X
    return -1;
}
