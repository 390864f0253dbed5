/*
 * gmp_ll.c -- ORACLE (test infrastructure only, never the product path).
 *
 * Lucas-Lehmer sequence s_0 = 4, s_{k+1} = s_k^2 - 2 mod M_p (M_p = 2^p - 1)
 * computed with GMP big integers.  This is the arithmetic that the
 * reference's IBDWT iteration must reproduce exactly:
 *   - s_0 = 4 and p-2 squarings: gpuLucas.cu:1368-1375 and the loop at
 *     gpuLucas.cu:1399 (iter runs from 2 to p-1);
 *   - one squaring minus 2 per iteration: gpuLucas.cu:845-942
 *     (mersenneIter) with the -2 applied in invDWTproductMinus2,
 *     gpuLucas.cu:774-805;
 *   - Res64 = low 64 bits of the canonical residue in [0, M_p):
 *     print_residue, gpuLucas.cu:1288-1352; the prime test is
 *     "residue == 0" (gpuLucas.cu:1506-1512, done canonically here).
 *
 * Reduction mod M_p uses the Mersenne fold s = (s >> p) + (s & M_p), which
 * is exact because 2^p == 1 (mod M_p).
 *
 * Exposed as a shared library (ctypes) for the tests and the CPU baseline in
 * bench.py, and as a CLI for the fixture generator under tests/golden/.
 */
#include <gmp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
    int p;
    mpz_t s, mp, t;
} gll_state;

static void gll_init(gll_state *st, int p) {
    st->p = p;
    mpz_init_set_ui(st->s, 4);
    mpz_init(st->mp);
    mpz_init(st->t);
    mpz_ui_pow_ui(st->mp, 2, (unsigned long)p);
    mpz_sub_ui(st->mp, st->mp, 1);
    if (mpz_cmp(st->s, st->mp) >= 0) mpz_mod(st->s, st->s, st->mp);
}

static void gll_clear(gll_state *st) {
    mpz_clear(st->s);
    mpz_clear(st->mp);
    mpz_clear(st->t);
}

/* s <- s^2 - 2 mod M_p, with s kept canonical in [0, M_p). */
static void gll_step(gll_state *st) {
    mpz_mul(st->s, st->s, st->s);
    if (mpz_cmp_ui(st->s, 2) < 0) {
        /* s^2 in {0, 1}: s^2 - 2 is negative, add M_p once. */
        mpz_add(st->s, st->s, st->mp);
    }
    mpz_sub_ui(st->s, st->s, 2);
    while (mpz_sizeinbase(st->s, 2) > (size_t)st->p) {
        mpz_tdiv_q_2exp(st->t, st->s, (mp_bitcnt_t)st->p); /* high part */
        mpz_tdiv_r_2exp(st->s, st->s, (mp_bitcnt_t)st->p); /* low part  */
        mpz_add(st->s, st->s, st->t);
    }
    if (mpz_cmp(st->s, st->mp) == 0) mpz_set_ui(st->s, 0);
}

static uint64_t gll_res64(const gll_state *st) {
    /* low 64 bits of the canonical residue */
    uint64_t lo = 0;
    size_t n = mpz_size(st->s);
    if (n > 0) lo = (uint64_t)mpz_getlimbn(st->s, 0);
    return lo;
}

/*
 * Res64 of s_k for each k in `iters` (ascending).  Returns 0 on success,
 * -1 on bad arguments.  `is_zero[i]` (optional) is 1 when s_k == 0 mod M_p.
 */
int gmp_ll_res64(int p, const long long *iters, int n, uint64_t *out, int *is_zero) {
    if (p < 2 || n < 0) return -1;
    for (int i = 1; i < n; i++)
        if (iters[i] < iters[i - 1]) return -1;
    gll_state st;
    gll_init(&st, p);
    long long k = 0;
    for (int i = 0; i < n; i++) {
        if (iters[i] < 0) { gll_clear(&st); return -1; }
        while (k < iters[i]) { gll_step(&st); k++; }
        out[i] = gll_res64(&st);
        if (is_zero) is_zero[i] = (mpz_sgn(st.s) == 0);
    }
    gll_clear(&st);
    return 0;
}

/*
 * Canonical residue s_k as little-endian 32-bit words (ceil(p/32) words).
 * Used by the tests to compare whole residues, not only 64 bits.
 */
int gmp_ll_residue_words(int p, long long k, uint32_t *words, int nwords) {
    if (p < 2 || k < 0 || nwords < (p + 31) / 32) return -1;
    gll_state st;
    gll_init(&st, p);
    for (long long i = 0; i < k; i++) gll_step(&st);
    memset(words, 0, sizeof(uint32_t) * (size_t)nwords);
    size_t count = 0;
    mpz_export(words, &count, -1, sizeof(uint32_t), 0, 0, st.s);
    gll_clear(&st);
    return (int)count;
}

/*
 * Reduce an arbitrary signed digit vector sum_j d_j 2^{e_j} mod M_p to its
 * canonical Res64.  e[j] are the bit offsets of the digits.  Used by tests
 * to check digit-vector invariants independently of any IBDWT code.
 */
int gmp_digits_res64(int p, const int64_t *digits, const int64_t *offsets, int n,
                     uint64_t *out, int *is_zero) {
    if (p < 2 || n < 0) return -1;
    mpz_t acc, term, mp, t;
    mpz_init(acc); mpz_init(term); mpz_init(mp); mpz_init(t);
    mpz_ui_pow_ui(mp, 2, (unsigned long)p);
    mpz_sub_ui(mp, mp, 1);
    for (int j = 0; j < n; j++) {
        if (digits[j] == 0) continue;
        int64_t d = digits[j];
        uint64_t mag = d < 0 ? (uint64_t)(-(d + 1)) + 1u : (uint64_t)d;
        mpz_set_ui(term, (unsigned long)mag);
        mpz_mul_2exp(term, term, (mp_bitcnt_t)offsets[j]);
        if (d < 0) mpz_sub(acc, acc, term); else mpz_add(acc, acc, term);
    }
    mpz_mod(acc, acc, mp); /* canonical, non-negative */
    *out = mpz_size(acc) ? (uint64_t)mpz_getlimbn(acc, 0) : 0;
    if (is_zero) *is_zero = mpz_sgn(acc) == 0;
    mpz_clear(acc); mpz_clear(term); mpz_clear(mp); mpz_clear(t);
    return 0;
}

/*
 * Steady-state single-thread timing: `warm` untimed iterations, then `iters`
 * timed ones.  Returns milliseconds per iteration.
 */
double gmp_ll_time(int p, int warm, int iters) {
    gll_state st;
    gll_init(&st, p);
    for (int i = 0; i < warm; i++) gll_step(&st);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < iters; i++) gll_step(&st);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    gll_clear(&st);
    double s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return iters > 0 ? 1e3 * s / iters : 0.0;
}

const char *gmp_ll_version(void) { return gmp_version; }

#ifdef GMP_LL_MAIN
/* CLI: gmp_ll p k1 [k2 ...]  -> one line "k res64hex zero" per k, flushed.
 *      gmp_ll --time p warm iters -> "ms_per_iter" (gmp_ll_time), one line. */
int main(int argc, char **argv) {
    if (argc == 5 && strcmp(argv[1], "--time") == 0) {
        printf("%.6f\n", gmp_ll_time(atoi(argv[2]), atoi(argv[3]), atoi(argv[4])));
        return 0;
    }
    if (argc < 3) {
        fprintf(stderr, "usage: %s p k1 [k2 ...]\n", argv[0]);
        return 2;
    }
    int p = atoi(argv[1]);
    int n = argc - 2;
    gll_state st;
    gll_init(&st, p);
    long long k = 0;
    for (int i = 0; i < n; i++) {
        long long target = atoll(argv[2 + i]);
        if (target < k) { fprintf(stderr, "iterations must ascend\n"); return 2; }
        while (k < target) { gll_step(&st); k++; }
        printf("%lld %016llx %d\n", k, (unsigned long long)gll_res64(&st), mpz_sgn(st.s) == 0);
        fflush(stdout);
    }
    gll_clear(&st);
    return 0;
}
#endif
