/*
 * ibdwt_oracle.c -- ORACLE (test infrastructure only; never linked into, or
 * called by, the product library).
 *
 * Plain-C restatement of the reference's Lucas-Lehmer IBDWT iteration
 * (ah42/gpuLucas), function by function, on the CPU:
 *
 *   widths         computeBitsPerWord        gpuLucas.cu:688-694
 *   HI masks       computeBitsPerWordVectors gpuLucas.cu:700-714
 *   q, bases       initConstantSymbols       gpuLucas.cu:222-237
 *   carry depth    setSliceAndDice           gpuLucas.cu:323-343
 *   weights        computeWeightVectors      gpuLucas.cu:673-686
 *   s_0 = 4        loadValue4ToFFTarray      gpuLucas.cu:718-726
 *   convolution    cufftExecD2Z / ComplexPointwiseSqr / cufftExecZ2D
 *                                            gpuLucas.cu:857-879, 647-665
 *   unweight, -2   invDWTproductMinus2<e>    gpuLucas.cu:774-805
 *   restoration    llintToIrrBal<n>          IrrBaseBalanced.cu:52-110
 *   ripple + load  loadIntToDoubleIBDWT      gpuLucas.cu:730-766
 *   final carry    addPseudoBalanced, rebalanceIrrIntSEQGPU
 *                                            IrrBaseBalanced.cu:117-169
 *   Res64          print_residue             gpuLucas.cu:1288-1352
 *
 * Differences from the reference that do not change any integer result:
 *   - widths use exact integer ceil(j*p/N) (the reference uses double
 *     ceil(PoverN*j), equal at every shape the tests use);
 *   - weights are 2^(r_j/N) with exact r_j = N*ceil(jp/N) - jp, in long
 *     double (the reference uses QD dd_real);
 *   - the convolution is a long-double complex FFT (cuFFT is closed and not
 *     available), or, for small N, the exact weighted cyclic convolution in
 *     __int128 (ibdwt_conv_exact), which pins the FFT path;
 *   - shifts of 64 or more in the restoration give 0 (undefined in the
 *     reference's llintToIrrBal<6> at q <= 10);
 *   - the Res64 is taken from a canonical residue; the reference's zero test
 *     (gpuLucas.cu:1506-1512) only looks for positive digits.
 *
 * `mutation` plants one deliberate mistake (tests check the GMP comparison
 * catches each): 1 drop the -2, 2 one carry term too few, 3 carry sign from
 * the current digit for every term, 4 every digit given width q, 5 ripple
 * carry taken from digit j+1, 6 weight exponent sign flipped.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef long double complex cld;

static const long double PI_L = 3.141592653589793238462643383279502884L;

typedef struct {
    int p, N, q, b, n;
    int *w;          /* width of digit j */
    int64_t *e;      /* e_j = ceil(j p / N), j = 0..N */
    long double *a;  /* 2^(e_j - jp/N) */
    long double *ainv;
    uint8_t *bpw8;   /* bit i set <=> digit j-i is wide */
} plan_t;

/* setSliceAndDice, gpuLucas.cu:323-343 */
int ibdwt_carry_depth(int q) {
    if (q >= 21) return -1;
    if (q >= 18) return 2;
    if (q >= 16) return 3;
    if (q >= 14) return 4;
    if (q >= 12) return 5;
    return 6;
}

static int plan_init(plan_t *P, int p, int N, int mutation) {
    memset(P, 0, sizeof(*P));
    if (N <= 0 || p <= 0) return -1;
    P->p = p;
    P->N = N;
    P->q = p / N;          /* h_LO_BITS, gpuLucas.cu:224 */
    P->b = p % N;
    P->n = ibdwt_carry_depth(P->q);
    if (P->n < 0) return -1;
    P->w = (int *)malloc(sizeof(int) * N);
    P->e = (int64_t *)malloc(sizeof(int64_t) * (N + 1));
    P->a = (long double *)malloc(sizeof(long double) * N);
    P->ainv = (long double *)malloc(sizeof(long double) * N);
    P->bpw8 = (uint8_t *)malloc(N);
    for (int64_t j = 0; j <= N; j++) P->e[j] = (j * (int64_t)p + N - 1) / N;
    for (int j = 0; j < N; j++) {
        P->w[j] = (int)(P->e[j + 1] - P->e[j]);
        if (mutation == 4) P->w[j] = P->q;
    }
    /* computeBitsPerWordVectors, gpuLucas.cu:700-714 */
    for (int j = 0; j < N; j++) {
        uint8_t m = 0;
        for (int bit = 0; bit < 8; bit++) {
            int k = ((j - bit) % N + N) % N;
            m |= (uint8_t)((P->w[k] == P->q ? 0 : 1) << bit);
        }
        P->bpw8[j] = m;
    }
    /* computeWeightVectors, gpuLucas.cu:673-686: a_j = 2^(ceil(jp/N) - jp/N) */
    for (int64_t j = 0; j < N; j++) {
        int64_t r = P->e[j] * N - j * (int64_t)p; /* in [0, N) */
        long double ex = (long double)r / (long double)N;
        if (mutation == 6) ex = -ex;
        P->a[j] = exp2l(ex);
        P->ainv[j] = 1.0L / (P->a[j] * (long double)N);
    }
    return 0;
}

static void plan_free(plan_t *P) {
    free(P->w);
    free(P->e);
    free(P->a);
    free(P->ainv);
    free(P->bpw8);
}

/* ---- long-double mixed-radix FFT: y_k = sum_j x_j exp(sign 2 pi i jk / n) */
static void fft_rec(const cld *x, int n, int s, cld *y, int sign, cld *scratch) {
    if (n == 1) {
        y[0] = x[0];
        return;
    }
    int f = 2;
    while (n % f) f++;
    int m = n / f;
    for (int r = 0; r < f; r++) fft_rec(x + (size_t)r * s, m, s * f, scratch + (size_t)r * m, sign, scratch + n);
    for (int k = 0; k < m; k++) {
        for (int qq = 0; qq < f; qq++) {
            cld acc = 0;
            int kk = k + qq * m;
            for (int r = 0; r < f; r++) {
                long long t = ((long long)r * kk) % n;
                long double ang = sign * 2.0L * PI_L * (long double)t / (long double)n;
                acc += scratch[(size_t)r * m + k] * (cosl(ang) + I * sinl(ang));
            }
            y[kk] = acc;
        }
    }
}

static void fft_ld(cld *data, int n, int sign) {
    /* scratch sized for every recursion level: sum of n over levels <= 2n */
    cld *scratch = (cld *)malloc(sizeof(cld) * (size_t)n * 4);
    cld *out = (cld *)malloc(sizeof(cld) * (size_t)n);
    fft_rec(data, n, 1, out, sign, scratch);
    memcpy(data, out, sizeof(cld) * (size_t)n);
    free(scratch);
    free(out);
}

/*
 * One weighted squaring, unweighted and rounded, with the -2 on digit 0:
 * c_j = rint(IDFT(DFT(a x)^2)_j * ainv_j - 2[j=0])  (gpuLucas.cu:774-805).
 * Returns the max roundoff |x - rint(x)|.
 */
static double conv_fft(const plan_t *P, const int32_t *x, int64_t *c, int mutation) {
    int N = P->N;
    cld *z = (cld *)malloc(sizeof(cld) * (size_t)N);
    for (int j = 0; j < N; j++) z[j] = P->a[j] * (long double)x[j];
    fft_ld(z, N, -1);
    for (int j = 0; j < N; j++) z[j] = z[j] * z[j]; /* ComplexPointwiseSqr */
    fft_ld(z, N, +1);                               /* unnormalised; 1/N in ainv */
    double err = 0;
    for (int j = 0; j < N; j++) {
        long double v = creall(z[j]) * P->ainv[j];
        if (j == 0 && mutation != 1) v -= 2.0L;
        long double r = rintl(v);
        double d = (double)fabsl(v - r);
        if (d > err) err = d;
        c[j] = (int64_t)r;
    }
    free(z);
    return err;
}

/*
 * Exact weighted cyclic convolution: since e_i + e_k - e_{i+k} (or - p after
 * a wrap) is 0 or 1, c_j = sum_{i+k = j mod N} x_i x_k 2^delta exactly.
 */
int ibdwt_conv_exact(int p, int N, const int32_t *x, int64_t *c) {
    plan_t P;
    if (plan_init(&P, p, N, 0)) return -1;
    for (int j = 0; j < N; j++) {
        __int128 acc = 0;
        for (int i = 0; i < N; i++) {
            int k = j - i;
            int64_t shift;
            if (k >= 0) {
                shift = P.e[i] + P.e[k] - P.e[j];
            } else {
                k += N;
                shift = P.e[i] + P.e[k] - P.e[j] - p;
            }
            acc += ((__int128)x[i] * x[k]) << shift;
        }
        if (j == 0) acc -= 2;
        c[j] = (int64_t)acc;
    }
    plan_free(&P);
    return 0;
}

int ibdwt_conv_fft(int p, int N, const int32_t *x, int64_t *c, double *maxerr) {
    plan_t P;
    if (plan_init(&P, p, N, 0)) return -1;
    double e = conv_fft(&P, x, c, 0);
    if (maxerr) *maxerr = e;
    plan_free(&P);
    return 0;
}

/*
 * llintToIrrBal<n>, IrrBaseBalanced.cu:52-110: carry-save radix restoration
 * of the rounded products c into a balanced digit d_j and carry h_j.
 */
static void restore(const plan_t *P, const int64_t *c, int32_t *d, int8_t *h, int mutation) {
    int N = P->N, n = P->n;
    if (mutation == 2) n = n - 1;
    for (int j = 0; j < N; j++) {
        uint8_t bperW8 = P->bpw8[j];
        int isHi = bperW8 & 1;
        int BITS = P->q + isHi;
        if (mutation == 4) BITS = P->q;
        int64_t myBase = (int64_t)1 << BITS;
        int64_t myMask = myBase - 1;
        int64_t baseOver2 = myBase >> 1;
        int sgn_j = c[j] < 0 ? -1 : 1;
        uint64_t mag = (uint64_t)(c[j] < 0 ? -c[j] : c[j]);
        int64_t sum = sgn_j * (int64_t)(mag & (uint64_t)myMask);
        int shiftBits = 0;
        for (int i = 1; i <= n; i++) {
            int k = ((j - i) % N + N) % N;
            shiftBits += P->q + ((bperW8 >> i) & 1);
            int sg = c[k] < 0 ? -1 : 1;
            if (mutation == 3) sg = sgn_j;
            uint64_t mk = (uint64_t)(c[k] < 0 ? -c[k] : c[k]);
            int64_t part = shiftBits < 64 ? (int64_t)((mk >> shiftBits) & (uint64_t)myMask) : 0;
            sum += sg * part;
        }
        int64_t hival = 0;
        if (sum < -baseOver2)
            hival = -((-sum + baseOver2) >> BITS);
        else if (sum >= baseOver2)
            hival = (sum + baseOver2) >> BITS;
        d[j] = (int32_t)(sum - hival * myBase);
        h[j] = (int8_t)hival;
    }
}

/* Exposed restoration (mutation-free) for unit tests. */
int ibdwt_restore(int p, int N, const int64_t *c, int32_t *d, int8_t *h) {
    plan_t P;
    if (plan_init(&P, p, N, 0)) return -1;
    restore(&P, c, d, h, 0);
    plan_free(&P);
    return 0;
}

/*
 * loadIntToDoubleIBDWT ripple (gpuLucas.cu:730-766): the next FFT input is
 * a_j (d_j + h_{j-1}), h_{-1} = h_{N-1}.  The oracle keeps the integer part
 * x_j = d_j + h_{j-1} as its state.
 */
static void ripple(const plan_t *P, const int32_t *d, const int8_t *h, int32_t *x, int mutation) {
    int N = P->N;
    for (int j = 0; j < N; j++) {
        int k = mutation == 5 ? (j + 1) % N : (j == 0 ? N - 1 : j - 1);
        x[j] = d[j] + h[k];
    }
}

/*
 * `iters` LL iterations.  x (N digits) holds the state: on entry s_k in
 * pseudo-balanced digits (NULL-equivalent start: x = [4, 0, ...]), on exit
 * s_{k+iters}.  use_exact selects the __int128 convolution.
 */
int ibdwt_iterate(int p, int N, int32_t *x, long long iters, int use_exact, int mutation, double *maxerr) {
    plan_t P;
    if (plan_init(&P, p, N, mutation)) return -1;
    int64_t *c = (int64_t *)malloc(sizeof(int64_t) * N);
    int32_t *d = (int32_t *)malloc(sizeof(int32_t) * N);
    int8_t *h = (int8_t *)malloc(N);
    double err = 0;
    for (long long it = 0; it < iters; it++) {
        if (use_exact) {
            plan_t Q = P;
            (void)Q;
            ibdwt_conv_exact(p, N, x, c);
            if (mutation == 1) c[0] += 2;
        } else {
            double e = conv_fft(&P, x, c, mutation);
            if (e > err) err = e;
        }
        restore(&P, c, d, h, mutation);
        ripple(&P, d, h, x, mutation);
    }
    if (maxerr) *maxerr = err;
    free(c);
    free(d);
    free(h);
    plan_free(&P);
    return 0;
}

/*
 * addPseudoBalanced + rebalanceIrrIntSEQGPU (IrrBaseBalanced.cu:117-169)
 * applied to the (d, h) form is equivalent to a sequential balanced carry over
 * x; then print_residue's borrow pass (gpuLucas.cu:1302-1317) makes digits
 * non-negative and the low 64 bits are packed (gpuLucas.cu:1325-1344).  Here
 * the residue is made canonical in [0, M_p) before packing.
 */
int ibdwt_res64(int p, int N, const int32_t *x, uint64_t *res64, int *is_zero) {
    plan_t P;
    if (plan_init(&P, p, N, 0)) return -1;
    int64_t *v = (int64_t *)malloc(sizeof(int64_t) * N);
    for (int j = 0; j < N; j++) v[j] = x[j];
    /* rebalanceIrrIntSEQGPU: balanced sequential carry, final carry to digit 0 */
    int64_t carry = 0;
    for (int j = 0; j < N; j++) {
        int64_t base = (int64_t)1 << P.w[j], half = base >> 1;
        int64_t t = v[j] + carry;
        carry = 0;
        while (t >= half) { t -= base; carry++; }
        while (t < -half) { t += base; carry--; }
        v[j] = t;
    }
    v[0] += carry;
    /* canonical non-negative digits with wrap-around borrows (2^p == 1) */
    for (int pass = 0; pass < 64; pass++) {
        carry = 0;
        for (int j = 0; j < N; j++) {
            int64_t t = v[j] + carry;
            int64_t m = t & (((int64_t)1 << P.w[j]) - 1);
            carry = (t - m) >> P.w[j];
            v[j] = m;
        }
        if (!carry) break;
        v[0] += carry;
    }
    int allzero = 1, allones = 1;
    for (int j = 0; j < N; j++) {
        if (v[j]) allzero = 0;
        if (v[j] != ((int64_t)1 << P.w[j]) - 1) allones = 0;
    }
    uint64_t r = 0;
    if (!(allzero || allones)) {
        int off = 0;
        for (int j = 0; j < N && off < 64; j++) {
            r |= (uint64_t)v[j] << off;
            off += P.w[j];
        }
    }
    *res64 = r;
    if (is_zero) *is_zero = allzero || allones;
    free(v);
    plan_free(&P);
    return 0;
}

/* Reference-formula widths (double ceil, gpuLucas.cu:688-694) for comparison. */
void ibdwt_widths_double(int p, int N, int *w) {
    double PoverN = p / (double)N;
    for (int j = 1; j <= N; j++) w[j - 1] = (int)(ceil(PoverN * j) - ceil(PoverN * (j - 1)));
}

/* Exact widths, HI masks and weights (double) as used by the oracle. */
int ibdwt_tables(int p, int N, int *w, uint8_t *bpw8, double *a, double *ainv) {
    plan_t P;
    if (plan_init(&P, p, N, 0)) return -1;
    for (int j = 0; j < N; j++) {
        if (w) w[j] = P.w[j];
        if (bpw8) bpw8[j] = P.bpw8[j];
        if (a) a[j] = (double)P.a[j];
        if (ainv) ainv[j] = (double)P.ainv[j];
    }
    plan_free(&P);
    return 0;
}

/*
 * computeMaxBitVector (gpuLucas.cu:826-836) reduced by findMaxErrorHOST
 * (gpuLucas.cu:811-819): per product the 64-bit leading-zero count of |c_j|
 * as a float (__clzll(0) is 64), maximum over j -- the reference's "max Bit
 * Vector".  Also the largest bit length of |c_j|, the figure that bounds the
 * products against the 53-bit FP64 mantissa.
 */
void ibdwt_product_bits(int N, const int64_t *c, int *max_bits, int *max_clz) {
    float mz = 0.0f;
    int mb = 0;
    for (int j = 0; j < N; j++) {
        uint64_t m = c[j] >= 0 ? (uint64_t)c[j] : (uint64_t)0 - (uint64_t)c[j];
        int z = m ? __builtin_clzll(m) : 64;
        if ((float)z > mz) mz = (float)z;
        if (64 - z > mb) mb = 64 - z;
    }
    *max_bits = mb;
    *max_clz = (int)mz;
}
