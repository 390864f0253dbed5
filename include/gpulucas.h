/*
 * gpulucas.h -- C ABI of the MI355X Lucas-Lehmer engine (libgpulucas.so).
 *
 * The reference (ah42/gpuLucas) has no library boundary: its documented API is
 * errorTest(p, N) / mersenneTest(p, N) (gpuLucas.cu:69-73, README:92-96),
 * implemented as static functions over globals (errorTrial gpuLucas.cu:1192,
 * mersenneTest gpuLucas.cu:1358) and called only from main.  This header is
 * the boundary those entry points become: plain C types, status codes instead
 * of exit(), an opaque context that owns all device state.
 *
 * Conventions (as in the reference):
 *   - s_0 = 4; "iteration k" is s_k after k squarings (gpuLucas.cu:1368-1375,
 *     1399); a full test is p-2 iterations and M_p is prime iff s_{p-2} == 0.
 *   - Res64 = low 64 bits of the canonical residue in [0, M_p)
 *     (print_residue, gpuLucas.cu:1288-1352).
 *   - N = 0 asks the engine to choose the signal length (findSignalSize,
 *     gpuLucas.cu:1064-1187); otherwise N must be even and 2^a 3^b 5^c 7^d.
 *   - Roundoff = max |x - rint(x)| over every unweighted convolution output
 *     (invDWTproductMinus2<1>, gpuLucas.cu:774-805), kept on the device.
 */
#ifndef GPULUCAS_H
#define GPULUCAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the reference printed and called exit(-1): cuda_safecalls.h:45-98). */
#define LUCAS_OK 0
#define LUCAS_ERR_INVALID_ARG (-1)     /* bad p / N / iteration count / handle */
#define LUCAS_ERR_OOM (-2)             /* device memory short (gpuLucas.cu:956-976) */
#define LUCAS_ERR_ROUNDOFF (-3)        /* roundoff >= limit (gpuLucas.cu:904-907, 1239-1242) */
#define LUCAS_ERR_DEVICE (-4)          /* HIP runtime error */
#define LUCAS_ERR_IO (-5)              /* checkpoint read/write failure (gpuLucas.cu:1569-1665) */
#define LUCAS_ERR_NO_DEVICE (-6)       /* no usable GPU */
#define LUCAS_ERR_BUSY (-7)            /* server queue full / admission refused */
#define LUCAS_ERR_UNSUPPORTED (-8)     /* not available in this build of the library */

/* Reference ERROR_LIMIT (gpuLucas.cu:176). */
#define LUCAS_DEFAULT_ERROR_LIMIT 0.30

typedef struct lucas_ctx lucas_ctx;
typedef struct lucas_batch lucas_batch;
typedef struct lucas_server lucas_server;

/* Transform shape chosen for (p, N); filled by lucas_plan. */
typedef struct {
    int p, N;          /* exponent and signal length (real words) */
    int q;             /* low word width floor(p/N); wide words have q+1 bits */
    int n_wide;        /* number of (q+1)-bit words = p mod N */
    int carry_depth;   /* reference carry depth n (setSliceAndDice, gpuLucas.cu:323-343) */
    int A, B;          /* complex length N/2 = A*B: column and row FFT lengths */
    int C;             /* columns per column workgroup */
    int halo;          /* halo columns recomputed per column workgroup */
    int col_wgs, row_wgs;
    int threads;       /* work-items per workgroup */
    int radix_pref;    /* power-of-two radix preference of the FFT stages (8 or 4) */
    int wave_A, wave_B; /* wave-register factor of the column / row transform (64, 128, 256; 0: none) */
    int col_dense;     /* 1: column kernel capped at 64 VGPRs, two 1024-item workgroups per CU */
    int line;          /* 1: line kernels (one wave per row or column, shapes compiled in);
                          C = own columns and threads = work-items of a column workgroup;
                          2: the same with non-temporal spectrum loads and stores;
                          +2: every row held by two waves; +4: every column
                          held by two waves (values 1..8) */
    size_t device_bytes; /* device footprint of one job */
} lucas_plan_info;

typedef struct {
    int p, N;
    int is_prime;          /* 1 iff the final residue is 0 */
    uint64_t res64;        /* Res64 of s_{iterations} */
    long long iterations;  /* iterations performed in this call (from s_0 unless resumed) */
    double max_roundoff;   /* over every iteration of the call */
    double ms_per_iter;    /* wall-clock mean */
    /* Product-size diagnostic over every iteration of the call
     * (computeMaxBitVector, gpuLucas.cu:826-836, printed at 913-916): the
     * largest bit length of a rounded convolution output |c_j|, and the
     * reference's "max Bit Vector", the largest 64-bit leading-zero count of
     * |c_j| (64 minus the smallest bit length). */
    int max_product_bits;
    int max_product_clz;
} lucas_result;

typedef struct {
    int device;                 /* HIP device ordinal */
    double error_limit;         /* roundoff gate; <= 0 means LUCAS_DEFAULT_ERROR_LIMIT */
    long long checkpoint_every; /* 0: no checkpoints (reference -c, gpuLucas.cu:204) */
    const char* checkpoint_path;/* file; ".bak" rotation as gpuLucas.cu:1641-1665 */
    int resume;                 /* 1: resume from checkpoint_path if it is valid */
} lucas_options;

/* ---- library ---------------------------------------------------------- */
const char* lucas_version(void);
const char* lucas_status_string(int status);
int lucas_device_count(int* count);

/* ---- planning (host only, no GPU needed) ------------------------------ */
/* Default signal length for p (replaces findSignalSize, gpuLucas.cu:1064-1187,
 * with a table rule; lucas_autotune_N measures on the GPU). */
int lucas_default_N(int p);
/* Bits per word the default rule allows at length N: the calibrated line on
 * which the roundoff from a random residue stays near 0.12 (no reference
 * counterpart; the reference's window was p/20 .. p/14, gpuLucas.cu:1072-1077). */
double lucas_max_bits_per_word(int N);
/* Shape of (p, N); N == 0 chooses lucas_default_N(p).  Replaces the setup of
 * initConstantSymbols / setSliceAndDice (gpuLucas.cu:222-237, 323-343). */
int lucas_plan(int p, int N, lucas_plan_info* out);
/* Same with an explicit column length A and tile width C (0 = automatic). */
int lucas_plan_shape(int p, int N, int A, int C, lucas_plan_info* out);
/* Every admissible (A, C) tiling of (p, N), cheapest modelled first; fills up
 * to `cap` entries and writes the total into *count. */
int lucas_plan_list(int p, int N, lucas_plan_info* out, int cap, int* count);
/* The line-kernel plans of (p, N) (line = 1, +2 for split rows, +4 for
 * split columns; none where no instantiated shape fits N and the carry
 * depth): *count = how many, the first `cap` in out. */
int lucas_line_plan_list(int p, int N, lucas_plan_info* out, int cap, int* count);
/* Per-digit widths and weights as the reference computes them
 * (computeBitsPerWord gpuLucas.cu:688-694, computeWeightVectors 673-686):
 * widths[N], weights[N] = 2^(ceil(jp/N) - jp/N); any pointer may be NULL. */
int lucas_digit_tables(int p, int N, int* widths, double* weights, double* inv_weights);
/* Canonical Res64 of a pseudo-balanced digit vector (host O(N) carry
 * normalisation; replaces addPseudoBalanced + rebalanceIrrIntSEQGPU +
 * print_residue, IrrBaseBalanced.cu:117-169, gpuLucas.cu:1288-1352).
 * `normalized` (optional, N entries) receives the digits in [0, 2^w). */
int lucas_digits_res64(int p, int N, const int32_t* digits, uint64_t* res64, int* is_zero,
                       int32_t* normalized);

/* Device-to-device copy of `bytes` repeated `reps` times on `device`; mean
 * microseconds per copy.  A known-size stream for bandwidth and HBM-counter
 * calibration (no reference counterpart). */
int lucas_device_copy_bench(size_t bytes, int reps, int device, double* us_per_copy);

/* ---- single-exponent context ------------------------------------------ */
/* Allocates device state for (p, N) and sets s_0 = 4 (loadValue4ToFFTarray,
 * gpuLucas.cu:718-726; mallocArrays gpuLucas.cu:948-1029). */
int lucas_ctx_create(int p, int N, int device, lucas_ctx** out);
int lucas_ctx_create_shape(int p, int N, int A, int C, int device, lucas_ctx** out);
/* Context from a full plan description (p, N, A, C, threads, radix_pref as
 * returned by lucas_plan_list / lucas_autotune_shape). */
int lucas_ctx_create_plan(const lucas_plan_info* plan, int device, lucas_ctx** out);
int lucas_ctx_destroy(lucas_ctx* ctx);                 /* freeArrays, gpuLucas.cu:1031-1057 */
int lucas_ctx_info(const lucas_ctx* ctx, lucas_plan_info* out);
/* Advance `iters` LL iterations (mersenneIter x iters, gpuLucas.cu:845-942). */
int lucas_ctx_step(lucas_ctx* ctx, long long iters);
long long lucas_ctx_iteration(const lucas_ctx* ctx);
/* Current pseudo-balanced digits (N int32), D2H. */
int lucas_ctx_get_digits(lucas_ctx* ctx, int32_t* digits);
/* Device address of the context's digit vector (N int32, pseudo-balanced,
 * valid between steps), for device-side interop such as
 * lucas_step_digits_noalloc.  No reference counterpart (its state was a
 * global, gpuLucas.cu:179-257). */
int lucas_ctx_device_digits(lucas_ctx* ctx, void** ptr);
/* Replace the state by a digit vector (|d_j| < 2^(w_j+1)) at iteration `iter`. */
int lucas_ctx_set_digits(lucas_ctx* ctx, const int32_t* digits, long long iter);
int lucas_ctx_res64(lucas_ctx* ctx, uint64_t* res64, int* is_zero);
/* Max roundoff since the last reset (findMaxErrorHOST, gpuLucas.cu:811-819). */
int lucas_ctx_max_roundoff(lucas_ctx* ctx, int reset, double* out);
/* Product-size diagnostic since the last reset, over every iteration
 * (computeMaxBitVector, gpuLucas.cu:826-836, reduced by findMaxErrorHOST,
 * 811-819): *max_bits = largest bit length of a rounded convolution output
 * |c_j| (the margin to the 53-bit FP64 mantissa); *max_clz = largest 64-bit
 * leading-zero count of |c_j|, the value the reference prints as "max Bit
 * Vector".  Either pointer may be NULL. */
int lucas_ctx_product_bits(lucas_ctx* ctx, int reset, int* max_bits, int* max_clz);
/* Makes a step of `iters` iterations launch-ready (no reference
 * counterpart: the reference launched every kernel from the host,
 * gpuLucas.cu:845-942): builds and uploads its replay graphs and runs the
 * step once, then restores the state (digits, roundoff and product-size
 * accumulators, iteration count), so that the next lucas_ctx_step of that
 * length pays no graph instantiation or first-launch cost. */
int lucas_ctx_prepare(lucas_ctx* ctx, long long iters);
/* Mean time per kernel launch over `iters` iterations launched as in
 * lucas_ctx_step (row kernel, column kernel, alternating), from HIP events
 * recorded on the context's stream after every launch (microseconds; each
 * interval includes that launch's dispatch gap).  Max iters 4096. */
int lucas_ctx_time_kernels(lucas_ctx* ctx, int iters, double* col_us, double* row_us);
/* Phase profile (no counterpart in the reference).  Only the profiling build
 * libgpulucas_phases.so implements it; libgpulucas.so returns
 * LUCAS_ERR_UNSUPPORTED.  Runs `iters` iterations in the lucas_ctx_step order
 * (digits are not touched) and copies, for each workgroup of the last column
 * and row launch, 8 words of which the first 6 are 100 MHz real-time stamps:
 * column: start, tile loaded, inverse FFT + unweight done, carry done,
 * forward FFT done, stores issued; row: start, rows loaded, forward FFT done,
 * square done, inverse FFT done, stores issued.  *n_col, *n_row receive the
 * workgroup counts (at most max_workgroups each, else LUCAS_ERR_INVALID_ARG).
 * Single-exponent contexts only. */
int lucas_ctx_phase_stamps(lucas_ctx* ctx, int iters, unsigned long long* col_stamps,
                           unsigned long long* row_stamps, int max_workgroups, int* n_col, int* n_row);
/* Mean wall time of `iters` iterations (graph-launched), in ms/iteration. */
int lucas_ctx_time_iterations(lucas_ctx* ctx, long long iters, double* ms_per_iter);
/* Checkpoint (own versioned format; rotates path -> path.bak like
 * writeCheckpoint / readCheckpoint, gpuLucas.cu:1569-1665).  lucas_ctx_load
 * tries path, then path.bak, each in this format or in the reference's
 * (program name, version, p, N, loop index, weighted double signal), so a
 * run checkpointed by gpuLucas resumes here.  Unreadable or inconsistent
 * files return LUCAS_ERR_IO before any device work. */
int lucas_ctx_save(lucas_ctx* ctx, const char* path);
int lucas_ctx_load(const char* path, int device, lucas_ctx** out);

/* ---- no device allocation --------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the computation of lucas_ctx_set_digits +
 * lucas_ctx_step(iters) + lucas_ctx_get_digits on a context of (p, N), with
 * the pseudo-balanced digits (N int32) read and written in place in
 * caller-owned device memory `digits`, launched on `stream` (a hipStream_t;
 * NULL: the null stream).  Every intermediate (spectra, weight and twiddle
 * tables, roundoff accumulator) lives in a static device array of
 * LUCAS_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends; the call makes no hipMalloc / hipFree.  N above
 * LUCAS_NOALLOC_MAX_N, or a plan whose tables do not fit the array, returns
 * LUCAS_ERR_INVALID_ARG.  Calls on one GPU share its array and overwrite it:
 * calls that overlap on one GPU race.  max_roundoff (optional; non-NULL
 * synchronises the stream) receives the call's max roundoff. */
#define LUCAS_NOALLOC_MAX_N 524288
#define LUCAS_NOALLOC_SCRATCH_BYTES (10 * 1024 * 1024)
int lucas_step_digits_noalloc(int p, int N, int32_t* digits, long long iters, int device, void* stream,
                              double* max_roundoff);

/* ---- reference entry points ------------------------------------------- */
/* Full test: p-2 iterations from s_0 = 4 (or from a checkpoint), roundoff
 * tracked on every iteration, Res64 + prime flag (mersenneTest,
 * gpuLucas.cu:1358-1551).  The roundoff gate is read after every chunk of
 * max(1000, min(65536, p/1000)) iterations (the reference read it every p/1000
 * and only printed a breach, gpuLucas.cu:1401-1407): a breach stops the run
 * and returns LUCAS_ERR_ROUNDOFF with is_prime = 0 and `iterations` the
 * count reached; no checkpoint is written past a breach.  Otherwise
 * LUCAS_OK.  `out` is filled in both cases. */
int lucas_mersenne_test(int p, int N, const lucas_options* opt, lucas_result* out);
/* Timing + roundoff trial of `iterations` (clamped to >= 100) from s_0 = 4
 * (errorTrial, gpuLucas.cu:1192-1282).  out->ms_per_iter is the mean time
 * of the iterations alone (the replay graphs are built before the timed
 * region); returns LUCAS_ERR_ROUNDOFF when max roundoff >= error_limit (the
 * reference returned -1).  iterations <= 0 uses clamp(p/10000, 100, 2500)
 * (gpuLucas.cu:529-541). */
int lucas_error_test(int p, int N, int iterations, const lucas_options* opt, lucas_result* out);
/* Measured choice of N for p (findSignalSize, gpuLucas.cu:1064-1187).  The
 * candidates are the reference's window, N from p/20 to the smallest power of
 * two >= p/16 but at most p/14 (1072-1077), restricted to lengths with a
 * tiling whose bits per word stay at or under the calibrated line, plus
 * lucas_default_N(p).  Each runs max(iters, 300) iterations from
 * pseudo-random balanced digits (graphs built beforehand); the fastest whose
 * roundoff stays below the limit wins.  When none passes, the window widens
 * past its top, up to 10 times (1166-1181).  Writes the chosen N and its
 * plan. */
int lucas_autotune_N(int p, int iters, const lucas_options* opt, lucas_plan_info* out);
/* The first window's candidate lengths of lucas_autotune_N (host only):
 * fills up to `cap` entries, ascending, and writes the total into *count. */
int lucas_autotune_candidates(int p, int* out, int cap, int* count);

/* Measured choice of shape for (p, N): times `iters` iterations of up to
 * `max_candidates` plans -- the line-kernel shapes first (from N = 2^19 each
 * also with non-temporal spectra), then the tilings (radix split A x B, tile
 * width C, workgroup size) -- and keeps the fastest (replaces the hand-set
 * signal-size and slice constants, gpuLucas.cu:159-173, 323-343).
 * *best_ms = ms/iteration. */
int lucas_autotune_shape(int p, int N, int iters, int max_candidates, const lucas_options* opt,
                         lucas_plan_info* out, double* best_ms);

/* ---- batched exponents (no reference counterpart: the reference runs one
 * exponent per process).  All jobs advance together, one launch of each
 * kernel per iteration covering every job. */
int lucas_batch_create(const int* ps, const int* Ns, int njobs, int device, lucas_batch** out);
int lucas_batch_destroy(lucas_batch* b);
int lucas_batch_step(lucas_batch* b, long long iters);
int lucas_batch_res64(lucas_batch* b, uint64_t* res64, int* is_zero, double* max_roundoff);
int lucas_batch_get_digits(lucas_batch* b, int job, int32_t* digits);
int lucas_batch_time_iterations(lucas_batch* b, long long iters, double* ms_per_iter);

/* ---- serving: continuous batching of (p, iterations) requests ---------- */
typedef struct {
    int device;
    size_t hbm_budget_bytes;  /* admission limit on resident job state (0 = 80% of free) */
    int max_jobs;             /* slots (0 = 256) */
    int chunk_iters;          /* iterations per scheduling tick (0 = 256) */
    const char* checkpoint_dir; /* NULL: no checkpoints; else <dir>/<p>_<id>.chk per job */
    long long checkpoint_every;
    double error_limit;       /* roundoff gate per job; <= 0: LUCAS_DEFAULT_ERROR_LIMIT */
} lucas_server_options;

typedef struct {
    long long id;
    int p, N;
    long long iterations;
    int done;               /* 1 when finished */
    int status;             /* LUCAS_OK, LUCAS_ERR_ROUNDOFF (max_roundoff reached
                               the limit: the residue is not trustworthy), or error */
    uint64_t res64;
    int is_zero;
    double max_roundoff;
    double submit_s, start_s, finish_s; /* seconds since server start */
} lucas_job_result;

int lucas_server_start(const lucas_server_options* opt, lucas_server** out);
/* Enqueue a job; N = 0 chooses the default length.  Thread-safe. */
int lucas_server_submit(lucas_server* s, int p, int N, long long iterations, long long* id);
int lucas_server_poll(lucas_server* s, long long id, lucas_job_result* out);
/* Block until every submitted job is done (timeout_s <= 0: no limit). */
int lucas_server_drain(lucas_server* s, double timeout_s);
/* Aggregate counters: iterations done, busy seconds, peak jobs resident. */
int lucas_server_stats(lucas_server* s, long long* iterations, double* busy_s, int* peak_jobs);
/* Stop the worker and free the server.  Jobs still queued or resident are
 * dropped with it.  No other call on `s` may run concurrently with it or
 * follow it (a thread blocked in lucas_server_drain must have returned). */
int lucas_server_stop(lucas_server* s);

#ifdef __cplusplus
}
#endif
#endif /* GPULUCAS_H */
