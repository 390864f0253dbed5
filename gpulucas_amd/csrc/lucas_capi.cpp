// lucas_capi.cpp -- extern "C" boundary of libgpulucas (declared in
// include/gpulucas.h): contexts, the reference's mersenneTest / errorTest
// entry points, checkpoints, batches and the serving loop.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gpulucas.h"
#include "lucas_engine.h"
#include "lucas_host.h"

using namespace lucas;

struct lucas_ctx {
    Engine* engine = nullptr;
    Job* job = nullptr;
};

struct lucas_batch {
    Engine* engine = nullptr;
    std::vector<Job*> jobs;
};

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

void fill_info(const Plan& pl, lucas_plan_info* out) {
    memset(out, 0, sizeof(*out));
    out->p = pl.p;
    out->N = pl.N;
    out->q = pl.q;
    out->n_wide = pl.b;
    out->carry_depth = pl.ncarry;
    out->A = pl.A;
    out->B = pl.B;
    out->C = pl.C;
    out->halo = pl.halo;
    out->col_wgs = pl.col_wgs;
    out->row_wgs = pl.row_wgs;
    out->threads = pl.threads;
    out->radix_pref = pl.radix_pref;
    out->wave_A = pl.wA;
    out->wave_B = pl.wB;
    out->col_dense = pl.col_dense;
    out->line = line_kind(pl);
    out->device_bytes = job_bytes(pl);
}

int plan_for(int p, int N, int A, int C, Plan* pl, bool shared = false) {
    if (p < 7) return LUCAS_ERR_INVALID_ARG;
    if (N == 0) N = default_length(p);
    if (N <= 0) return LUCAS_ERR_INVALID_ARG;
    return make_plan(p, N, A, C, pl, shared).empty() ? LUCAS_OK : LUCAS_ERR_INVALID_ARG;
}

int ctx_new(const Plan& pl, int device, lucas_ctx** out) {
    lucas_ctx* c = new lucas_ctx();
    c->engine = new Engine(device);
    int rc = c->engine->init();
    if (rc == LUCAS_OK) rc = job_create(pl, device, c->engine->stream(), &c->job);
    if (rc != LUCAS_OK) {
        delete c->engine;
        delete c;
        return rc;
    }
    c->engine->set_jobs({c->job});
    *out = c;
    return LUCAS_OK;
}

// The lengths findSignalSize tries (gpuLucas.cu:1072-1077): from p/20 up to
// the smallest power of two >= p/16, but not past p/14.
void autotune_window(int p, long long* lo, long long* hi) {
    *lo = ((long long)p + 19) / 20;
    long long pow2 = 1;
    while (pow2 < (long long)p / 16) pow2 <<= 1;
    *hi = std::min(pow2, (long long)p / 14);
    if (*hi < *lo) *hi = *lo;
}

// Candidate lengths in [lo, hi]: every admissible N (2^a 3^b 5^c 7^d with a
// tiling) at or under the calibrated bits-per-word line, plus the default
// length of p, ascending.
std::vector<int> autotune_candidates(int p, long long lo, long long hi) {
    std::vector<int> c;
    for (long long N = std::max(4LL, lo + (lo & 1)); N <= hi && N <= (long long)p; N += 2) {
        if (!valid_length(N)) continue;
        if ((double)p / (double)N > max_bits_per_word((int)N)) continue;
        Plan pl;
        if (make_plan(p, (int)N, 0, 0, &pl).empty()) c.push_back((int)N);
    }
    const int dn = default_length(p);
    if (dn > 0 && std::find(c.begin(), c.end(), dn) == c.end()) {
        c.push_back(dn);
        std::sort(c.begin(), c.end());
    }
    return c;
}

double error_limit_of(const lucas_options* opt) {
    return (opt && opt->error_limit > 0) ? opt->error_limit : LUCAS_DEFAULT_ERROR_LIMIT;
}

// ---- checkpoint format -----------------------------------------------------
// 48-byte header + int32 digits[N]:
//   char magic[8] = "GLUAMD01"; int32 p, N; int64 iter; uint64 fnv1a(digits);
//   uint64 reserved[2]
struct CkptHeader {
    char magic[8];
    int32_t p, N;
    int64_t iter;
    uint64_t checksum;
    uint64_t reserved[2];
};
static_assert(sizeof(CkptHeader) == 48, "checkpoint header layout");

uint64_t fnv1a(const void* data, size_t n) {
    const unsigned char* b = (const unsigned char*)data;
    uint64_t h = 1469598103934665603ULL;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}

int write_checkpoint(const char* path, int p, int N, long long iter, const int32_t* digits) {
    std::string tmp = std::string(path) + ".tmp", bak = std::string(path) + ".bak";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return LUCAS_ERR_IO;
    CkptHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, "GLUAMD01", 8);
    h.p = p;
    h.N = N;
    h.iter = iter;
    h.checksum = fnv1a(digits, sizeof(int32_t) * (size_t)N);
    bool ok = fwrite(&h, sizeof(h), 1, f) == 1 && fwrite(digits, sizeof(int32_t), (size_t)N, f) == (size_t)N;
    ok = (fflush(f) == 0) && ok;
    if (ok) fsync(fileno(f));
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        unlink(tmp.c_str());
        return LUCAS_ERR_IO;
    }
    // rotate: path -> path.bak, tmp -> path (writeCheckpoint, gpuLucas.cu:1641-1665)
    unlink(bak.c_str());
    rename(path, bak.c_str());
    if (rename(tmp.c_str(), path) != 0) return LUCAS_ERR_IO;
    return LUCAS_OK;
}

int read_checkpoint_file(const char* path, int* p, int* N, long long* iter, std::vector<int32_t>* digits) {
    FILE* f = fopen(path, "rb");
    if (!f) return LUCAS_ERR_IO;
    CkptHeader h;
    int rc = LUCAS_ERR_IO;
    if (fread(&h, sizeof(h), 1, f) == 1 && memcmp(h.magic, "GLUAMD01", 8) == 0 && h.N > 0 && h.N <= (1 << 28) &&
        h.p > 0 && h.iter >= 0) {
        digits->resize((size_t)h.N);
        if (fread(digits->data(), sizeof(int32_t), (size_t)h.N, f) == (size_t)h.N &&
            fnv1a(digits->data(), sizeof(int32_t) * (size_t)h.N) == h.checksum) {
            *p = h.p;
            *N = h.N;
            *iter = h.iter;
            rc = LUCAS_OK;
        }
    }
    fclose(f);
    return rc;
}

// A checkpoint written by the reference (writeCheckpoint, gpuLucas.cu:1641-1665):
//   char name[16] = "gpuLucas"; char version[16]; int32 p, N, iter;
//   double signal[N]
// The signal is the next squaring's FFT input, the weighted digits
// x_j * 2^(r_j/N) (loadIntToDoubleIBDWT, gpuLucas.cu:730-766), and holds
// s_(iter-1): the reference resumes at loop index iter + 1
// (readCheckpoint, gpuLucas.cu:1632-1634; one squaring per index from
// index 2 = s_1).  Each digit is recovered by dividing out its weight and
// rounding; a value further than 1/4 from an integer, or a digit past its
// width, marks the file as not a valid state.
int read_reference_checkpoint_file(const char* path, int* p, int* N, long long* iter,
                                   std::vector<int32_t>* digits) {
    FILE* f = fopen(path, "rb");
    if (!f) return LUCAS_ERR_IO;
    char name[16], vers[16];
    int32_t hp = 0, hN = 0, hiter = 0;
    int rc = LUCAS_ERR_IO;
    if (fread(name, 1, 16, f) == 16 && fread(vers, 1, 16, f) == 16 && memcmp(name, "gpuLucas", 9) == 0 &&
        fread(&hp, 4, 1, f) == 1 && fread(&hN, 4, 1, f) == 1 && fread(&hiter, 4, 1, f) == 1 && hp >= 7 &&
        hN <= (1 << 28) && valid_length(hN) && hN <= hp && hiter >= 2 && hiter <= hp) {
        std::vector<double> sig((size_t)hN);
        if (fread(sig.data(), sizeof(double), (size_t)hN, f) == (size_t)hN) {
            std::vector<int> w((size_t)hN);
            digit_widths(hp, hN, w.data());
            digits->resize((size_t)hN);
            rc = LUCAS_OK;
            for (int j = 0; j < hN && rc == LUCAS_OK; ++j) {
                const double a = (double)exp2l((long double)r_of(j, hp, hN) / (long double)hN);
                const double x = sig[(size_t)j] / a, xr = rint(x);
                if (!(fabs(x - xr) <= 0.25) || !(fabs(xr) < ldexp(1.0, w[(size_t)j] + 1)))
                    rc = LUCAS_ERR_IO;
                else
                    (*digits)[(size_t)j] = (int32_t)xr;
            }
            if (rc == LUCAS_OK) {
                *p = hp;
                *N = hN;
                *iter = hiter - 1;
            }
        }
    }
    fclose(f);
    return rc;
}

// primary, then ".bak" (readCheckpoint, gpuLucas.cu:1569-1591); each in this
// library's format or the reference's
int read_checkpoint(const char* path, int* p, int* N, long long* iter, std::vector<int32_t>* digits) {
    const std::string bak = std::string(path) + ".bak";
    for (const char* f : {path, bak.c_str()}) {
        if (read_checkpoint_file(f, p, N, iter, digits) == LUCAS_OK) return LUCAS_OK;
        if (read_reference_checkpoint_file(f, p, N, iter, digits) == LUCAS_OK) return LUCAS_OK;
    }
    return LUCAS_ERR_IO;
}

}  // namespace

extern "C" {

const char* lucas_version(void) { return "gpulucas-amd 0.1.0 (gfx950)"; }

const char* lucas_status_string(int status) {
    switch (status) {
        case LUCAS_OK: return "ok";
        case LUCAS_ERR_INVALID_ARG: return "invalid argument";
        case LUCAS_ERR_OOM: return "out of device memory";
        case LUCAS_ERR_ROUNDOFF: return "roundoff limit exceeded";
        case LUCAS_ERR_DEVICE: return "HIP device error";
        case LUCAS_ERR_IO: return "checkpoint I/O error";
        case LUCAS_ERR_NO_DEVICE: return "no usable GPU";
        case LUCAS_ERR_BUSY: return "busy";
        case LUCAS_ERR_UNSUPPORTED: return "not supported by this build";
        default: return "unknown status";
    }
}

int lucas_device_count(int* count) {
    if (!count) return LUCAS_ERR_INVALID_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        return LUCAS_ERR_NO_DEVICE;
    }
    *count = c;
    return LUCAS_OK;
}

int lucas_device_copy_bench(size_t bytes, int reps, int device, double* us_per_copy) {
    if (!us_per_copy || bytes == 0 || reps <= 0) return LUCAS_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return LUCAS_ERR_NO_DEVICE;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return LUCAS_ERR_OOM;
    if (hipMalloc(&b, bytes) != hipSuccess) {
        hipFree(a);
        return LUCAS_ERR_OOM;
    }
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LUCAS_OK;
    float ms = 0.f;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 1, bytes, s) != hipSuccess ||
        hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess || hipEventRecord(e0, s) != hipSuccess)
        rc = LUCAS_ERR_DEVICE;
    for (int i = 0; i < reps && rc == LUCAS_OK; ++i)
        if (hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) rc = LUCAS_ERR_DEVICE;
    if (rc == LUCAS_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = LUCAS_ERR_DEVICE;
    if (rc == LUCAS_OK) *us_per_copy = 1000.0 * ms / reps;
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    hipFree(a);
    hipFree(b);
    return rc;
}

int lucas_default_N(int p) { return default_length(p); }

double lucas_max_bits_per_word(int N) { return N > 0 ? max_bits_per_word(N) : 0.0; }

int lucas_plan(int p, int N, lucas_plan_info* out) { return lucas_plan_shape(p, N, 0, 0, out); }

int lucas_plan_shape(int p, int N, int A, int C, lucas_plan_info* out) {
    if (!out) return LUCAS_ERR_INVALID_ARG;
    Plan pl;
    int rc = plan_for(p, N, A, C, &pl);
    if (rc != LUCAS_OK) return rc;
    fill_info(pl, out);
    return LUCAS_OK;
}

int lucas_plan_list(int p, int N, lucas_plan_info* out, int cap, int* count) {
    if (!count || cap < 0 || (cap > 0 && !out)) return LUCAS_ERR_INVALID_ARG;
    std::vector<Plan> c = plan_candidates(p, N);
    if (c.empty()) c = plan_candidates(p, N, true);
    if (c.empty()) return LUCAS_ERR_INVALID_ARG;
    *count = (int)c.size();
    for (int i = 0; i < cap && i < (int)c.size(); ++i) fill_info(c[i], &out[i]);
    return LUCAS_OK;
}

int lucas_line_plan_list(int p, int N, lucas_plan_info* out, int cap, int* count) {
    if (!count || cap < 0 || (cap > 0 && !out)) return LUCAS_ERR_INVALID_ARG;
    std::vector<Plan> c = line_candidates(p, N);
    *count = (int)c.size();
    for (int i = 0; i < cap && i < (int)c.size(); ++i) fill_info(c[i], &out[i]);
    return LUCAS_OK;
}

int lucas_digit_tables(int p, int N, int* widths, double* weights, double* inv_weights) {
    if (p < 7 || !valid_length(N) || N > p) return LUCAS_ERR_INVALID_ARG;
    if (widths) digit_widths(p, N, widths);
    for (long long j = 0; j < N; ++j) {
        const long long r = r_of(j, p, N);
        const long double a = exp2l((long double)r / (long double)N);
        if (weights) weights[j] = (double)a;
        if (inv_weights) inv_weights[j] = (double)(1.0L / (a * (long double)N));
    }
    return LUCAS_OK;
}

int lucas_digits_res64(int p, int N, const int32_t* digits, uint64_t* res64, int* is_zero, int32_t* normalized) {
    if (p < 7 || !valid_length(N) || N > p || !digits) return LUCAS_ERR_INVALID_ARG;
    normalize_digits(p, N, digits, normalized, res64, is_zero);
    return LUCAS_OK;
}

int lucas_ctx_create(int p, int N, int device, lucas_ctx** out) {
    return lucas_ctx_create_shape(p, N, 0, 0, device, out);
}

int lucas_ctx_create_shape(int p, int N, int A, int C, int device, lucas_ctx** out) {
    if (!out) return LUCAS_ERR_INVALID_ARG;
    Plan pl;
    int rc = plan_for(p, N, A, C, &pl);
    if (rc != LUCAS_OK) return rc;
    return ctx_new(pl, device, out);
}

int lucas_ctx_create_plan(const lucas_plan_info* plan, int device, lucas_ctx** out) {
    if (!plan || !out) return LUCAS_ERR_INVALID_ARG;
    Plan pl;
    if (plan->line) {
        if (!make_line_plan_kind(plan->p, plan->N, plan->A, plan->C, plan->line, &pl).empty())
            return LUCAS_ERR_INVALID_ARG;
        return ctx_new(pl, device, out);
    }
    const bool ordinary = make_plan_exact(plan->p, plan->N, plan->A, plan->C, plan->threads, plan->radix_pref, &pl,
                                          plan->wave_A, plan->wave_B).empty();
    if (!ordinary && !make_plan_exact(plan->p, plan->N, plan->A, plan->C, plan->threads, plan->radix_pref, &pl,
                                      plan->wave_A, plan->wave_B, true).empty())
        return LUCAS_ERR_INVALID_ARG;
    return ctx_new(pl, device, out);
}

int lucas_ctx_destroy(lucas_ctx* ctx) {
    if (!ctx) return LUCAS_ERR_INVALID_ARG;
    if (ctx->engine) ctx->engine->sync();
    job_destroy(ctx->job);
    delete ctx->engine;
    delete ctx;
    return LUCAS_OK;
}

int lucas_ctx_info(const lucas_ctx* ctx, lucas_plan_info* out) {
    if (!ctx || !out) return LUCAS_ERR_INVALID_ARG;
    fill_info(ctx->job->plan, out);
    return LUCAS_OK;
}

int lucas_ctx_step(lucas_ctx* ctx, long long iters) {
    if (!ctx || iters < 0) return LUCAS_ERR_INVALID_ARG;
    int rc = ctx->engine->step(iters);
    if (rc == LUCAS_OK) rc = ctx->engine->sync();
    return rc;
}

long long lucas_ctx_iteration(const lucas_ctx* ctx) { return ctx ? ctx->job->iter : -1; }

int lucas_ctx_get_digits(lucas_ctx* ctx, int32_t* digits) {
    if (!ctx || !digits) return LUCAS_ERR_INVALID_ARG;
    return job_get_digits(ctx->job, ctx->engine->stream(), digits);
}

int lucas_ctx_device_digits(lucas_ctx* ctx, void** ptr) {
    if (!ctx || !ptr) return LUCAS_ERR_INVALID_ARG;
    *ptr = ctx->job->desc.digits;
    return LUCAS_OK;
}

int lucas_ctx_set_digits(lucas_ctx* ctx, const int32_t* digits, long long iter) {
    if (!ctx || !digits || iter < 0) return LUCAS_ERR_INVALID_ARG;
    std::vector<int> w(ctx->job->plan.N);
    digit_widths(ctx->job->plan.p, ctx->job->plan.N, w.data());
    for (int j = 0; j < ctx->job->plan.N; ++j) {
        const long long lim = 1LL << (w[j] + 1);
        if (digits[j] <= -lim || digits[j] >= lim) return LUCAS_ERR_INVALID_ARG;
    }
    int rc = job_set_digits(ctx->job, ctx->engine->stream(), digits);
    if (rc == LUCAS_OK) ctx->job->iter = iter;
    return rc;
}

int lucas_ctx_res64(lucas_ctx* ctx, uint64_t* res64, int* is_zero) {
    if (!ctx) return LUCAS_ERR_INVALID_ARG;
    std::vector<int32_t> d(ctx->job->plan.N);
    int rc = job_get_digits(ctx->job, ctx->engine->stream(), d.data());
    if (rc != LUCAS_OK) return rc;
    normalize_digits(ctx->job->plan.p, ctx->job->plan.N, d.data(), nullptr, res64, is_zero);
    return LUCAS_OK;
}

int lucas_ctx_max_roundoff(lucas_ctx* ctx, int reset, double* out) {
    if (!ctx || !out) return LUCAS_ERR_INVALID_ARG;
    return job_read_maxerr(ctx->job, ctx->engine->stream(), reset, out);
}

int lucas_ctx_product_bits(lucas_ctx* ctx, int reset, int* max_bits, int* max_clz) {
    if (!ctx) return LUCAS_ERR_INVALID_ARG;
    return job_read_pbits(ctx->job, ctx->engine->stream(), reset, max_bits, max_clz);
}

int lucas_ctx_prepare(lucas_ctx* ctx, long long iters) {
    if (!ctx || iters < 0) return LUCAS_ERR_INVALID_ARG;
    int rc = ctx->engine->prepare_graphs(iters);
    if (rc == LUCAS_OK) rc = ctx->engine->sync();
    if (rc != LUCAS_OK || iters == 0) return rc;
    // Rehearsal: the first launch of a new graph costs tens of microseconds
    // more than the next ones, so the step is run once here and the state
    // (digits, roundoff and product-size accumulators, iteration count) put
    // back as it was.
    Job* j = ctx->job;
    hipStream_t st = ctx->engine->stream();
    const size_t dbytes = sizeof(int32_t) * (size_t)j->plan.N;
    void* save = nullptr;
    unsigned int acc[3] = {0, 0, 0};
    if (hipMalloc(&save, dbytes) != hipSuccess) {
        (void)hipGetLastError();
        return LUCAS_ERR_OOM;
    }
    hipError_t e = hipMemcpyAsync(save, j->desc.digits, dbytes, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(acc, j->desc.maxerr, sizeof(acc), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    rc = e == hipSuccess ? ctx->engine->step(iters) : LUCAS_ERR_DEVICE;
    if (rc == LUCAS_OK) {
        j->iter -= iters;
        e = hipMemcpyAsync(j->desc.digits, save, dbytes, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(j->desc.maxerr, acc, sizeof(acc), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = LUCAS_ERR_DEVICE;
    }
    hipFree(save);
    return rc;
}

int lucas_ctx_time_kernels(lucas_ctx* ctx, int iters, double* col_us, double* row_us) {
    if (!ctx || !col_us || !row_us) return LUCAS_ERR_INVALID_ARG;
    return ctx->engine->time_kernels(iters, col_us, row_us);
}

int lucas_ctx_phase_stamps(lucas_ctx* ctx, int iters, unsigned long long* col_stamps,
                           unsigned long long* row_stamps, int max_workgroups, int* n_col, int* n_row) {
    if (!ctx || !col_stamps || !row_stamps || !n_col || !n_row || max_workgroups <= 0) return LUCAS_ERR_INVALID_ARG;
    return ctx->engine->phase_stamps(iters, col_stamps, row_stamps, max_workgroups, n_col, n_row);
}

int lucas_ctx_time_iterations(lucas_ctx* ctx, long long iters, double* ms_per_iter) {
    if (!ctx || iters <= 0 || !ms_per_iter) return LUCAS_ERR_INVALID_ARG;
    int rc = ctx->engine->prepare_graphs(iters);  // graph builds stay outside the timed region
    if (rc == LUCAS_OK) rc = ctx->engine->sync();
    if (rc != LUCAS_OK) return rc;
    const double t0 = now_s();
    rc = ctx->engine->step(iters);
    if (rc == LUCAS_OK) rc = ctx->engine->sync();
    *ms_per_iter = 1e3 * (now_s() - t0) / (double)iters;
    return rc;
}

int lucas_ctx_save(lucas_ctx* ctx, const char* path) {
    if (!ctx || !path) return LUCAS_ERR_INVALID_ARG;
    std::vector<int32_t> d(ctx->job->plan.N);
    int rc = job_get_digits(ctx->job, ctx->engine->stream(), d.data());
    if (rc != LUCAS_OK) return rc;
    return write_checkpoint(path, ctx->job->plan.p, ctx->job->plan.N, ctx->job->iter, d.data());
}

int lucas_ctx_load(const char* path, int device, lucas_ctx** out) {
    if (!path || !out) return LUCAS_ERR_INVALID_ARG;
    int p = 0, N = 0;
    long long iter = 0;
    std::vector<int32_t> d;
    int rc = read_checkpoint(path, &p, &N, &iter, &d);
    if (rc != LUCAS_OK) return rc;
    lucas_ctx* c = nullptr;
    rc = lucas_ctx_create(p, N, device, &c);
    if (rc != LUCAS_OK) return rc;
    rc = lucas_ctx_set_digits(c, d.data(), iter);
    if (rc != LUCAS_OK) {
        lucas_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return LUCAS_OK;
}

int lucas_mersenne_test(int p, int N, const lucas_options* opt, lucas_result* out) {
    if (!out || p < 7) return LUCAS_ERR_INVALID_ARG;
    memset(out, 0, sizeof(*out));
    const int device = opt ? opt->device : 0;
    const double limit = error_limit_of(opt);
    lucas_ctx* ctx = nullptr;
    int rc = LUCAS_ERR_IO;
    if (opt && opt->resume && opt->checkpoint_path) {
        rc = lucas_ctx_load(opt->checkpoint_path, device, &ctx);
        if (rc == LUCAS_OK && ctx->job->plan.p != p) {
            lucas_ctx_destroy(ctx);
            return LUCAS_ERR_INVALID_ARG;  // checkpoint of another exponent (gpuLucas.cu:1616-1620)
        }
    }
    if (rc != LUCAS_OK) {
        rc = lucas_ctx_create(p, N, device, &ctx);
        if (rc != LUCAS_OK) return rc;
    }
    const long long total = (long long)p - 2;
    const long long start = ctx->job->iter;
    const long long every = (opt && opt->checkpoint_every > 0 && opt->checkpoint_path) ? opt->checkpoint_every : 0;
    // The roundoff gate is read after every chunk (the reference checked every
    // p/1000 iterations and only printed a breach, gpuLucas.cu:1401-1407): a
    // breach stops the run, and no checkpoint is written past it.
    const long long gate = std::max(1000LL, std::min(1LL << 16, (long long)p / 1000));
    double err = 0.0;
    bool breach = false;
    const double t0 = now_s();
    while (rc == LUCAS_OK && ctx->job->iter < total) {
        long long chunk = std::min(total - ctx->job->iter, gate - ctx->job->iter % gate);
        if (every) chunk = std::min(chunk, every - (ctx->job->iter % every));
        rc = lucas_ctx_step(ctx, chunk);
        if (rc == LUCAS_OK) rc = lucas_ctx_max_roundoff(ctx, 0, &err);
        if (rc == LUCAS_OK && err >= limit) {
            breach = true;
            break;
        }
        if (rc == LUCAS_OK && every && ctx->job->iter % every == 0 && ctx->job->iter < total)
            rc = lucas_ctx_save(ctx, opt->checkpoint_path);
    }
    const double t1 = now_s();
    if (rc == LUCAS_OK) rc = lucas_ctx_max_roundoff(ctx, 0, &err);
    uint64_t r = 0;
    int zero = 0;
    if (rc == LUCAS_OK) rc = lucas_ctx_res64(ctx, &r, &zero);
    int bits = 0, clz = 0;
    if (rc == LUCAS_OK) rc = lucas_ctx_product_bits(ctx, 0, &bits, &clz);
    out->p = p;
    out->N = ctx->job->plan.N;
    out->res64 = r;
    out->is_prime = breach ? 0 : zero;
    out->iterations = ctx->job->iter - start;
    out->max_roundoff = err;
    out->ms_per_iter = out->iterations > 0 ? 1e3 * (t1 - t0) / (double)out->iterations : 0.0;
    out->max_product_bits = bits;
    out->max_product_clz = clz;
    lucas_ctx_destroy(ctx);
    if (rc == LUCAS_OK && (breach || err >= limit)) rc = LUCAS_ERR_ROUNDOFF;
    return rc;
}

int lucas_error_test(int p, int N, int iterations, const lucas_options* opt, lucas_result* out) {
    if (!out || p < 7) return LUCAS_ERR_INVALID_ARG;
    memset(out, 0, sizeof(*out));
    if (iterations <= 0) iterations = std::min(2500, std::max(100, p / 10000));  // gpuLucas.cu:529-541
    if (iterations < 100) iterations = 100;                                   // gpuLucas.cu:1195-1196
    lucas_ctx* ctx = nullptr;
    int rc = lucas_ctx_create(p, N, opt ? opt->device : 0, &ctx);
    if (rc != LUCAS_OK) return rc;
    double ms = 0.0, err = 0.0;
    // the replay graphs are built before the timed region (time_iterations),
    // so ms/iteration covers the iterations only
    rc = lucas_ctx_time_iterations(ctx, iterations, &ms);
    if (rc == LUCAS_OK) rc = lucas_ctx_max_roundoff(ctx, 0, &err);
    uint64_t r = 0;
    int zero = 0;
    if (rc == LUCAS_OK) rc = lucas_ctx_res64(ctx, &r, &zero);
    int bits = 0, clz = 0;
    if (rc == LUCAS_OK) rc = lucas_ctx_product_bits(ctx, 0, &bits, &clz);
    out->p = p;
    out->N = ctx->job->plan.N;
    out->res64 = r;
    out->is_prime = 0;
    out->iterations = iterations;
    out->max_roundoff = err;
    out->ms_per_iter = ms;
    out->max_product_bits = bits;
    out->max_product_clz = clz;
    lucas_ctx_destroy(ctx);
    if (rc == LUCAS_OK && err >= error_limit_of(opt)) rc = LUCAS_ERR_ROUNDOFF;
    return rc;
}

int lucas_autotune_N(int p, int iters, const lucas_options* opt, lucas_plan_info* out) {
    if (!out || p < 7) return LUCAS_ERR_INVALID_ARG;
    if (iters < 300) iters = 300;
    const double limit = error_limit_of(opt);
    const int device = opt ? opt->device : 0;
    long long lo = 0, hi = 0;
    autotune_window(p, &lo, &hi);
    double best_ms = 1e30;
    Plan best;
    int last_rc = LUCAS_ERR_ROUNDOFF;
    // findSignalSize widens the window by 8 x 256 words up to 10 times when
    // no length passes (gpuLucas.cu:1166-1181); here by an eighth.
    for (int retry = 0; retry <= 10 && best.N == 0; ++retry) {
        for (int N : autotune_candidates(p, lo, hi)) {
            Plan pl;
            if (!make_plan(p, N, 0, 0, &pl).empty()) continue;
            lucas_ctx* ctx = nullptr;
            int rc = ctx_new(pl, device, &ctx);
            if (rc != LUCAS_OK) {
                last_rc = rc;
                continue;
            }
            // steady-state roundoff from a pseudo-random balanced start
            std::vector<int> w(N);
            digit_widths(p, N, w.data());
            std::vector<int32_t> d(N);
            uint64_t s = 0x9E3779B97F4A7C15ULL ^ (uint64_t)p ^ ((uint64_t)N << 32);
            for (int j = 0; j < N; ++j) {
                s ^= s << 13;
                s ^= s >> 7;
                s ^= s << 17;
                const int32_t half = 1 << (w[j] - 1);
                d[j] = (int32_t)(s % (uint64_t)(2 * half)) - half;
            }
            double ms = 0, err = 0;
            rc = lucas_ctx_set_digits(ctx, d.data(), 0);
            if (rc == LUCAS_OK) rc = lucas_ctx_time_iterations(ctx, iters, &ms);
            if (rc == LUCAS_OK) rc = lucas_ctx_max_roundoff(ctx, 0, &err);
            lucas_ctx_destroy(ctx);
            if (rc != LUCAS_OK) {
                last_rc = rc;
                continue;
            }
            if (err < limit && ms < best_ms) {
                best_ms = ms;
                best = pl;
            }
        }
        lo = hi + 1;
        hi += std::max(8LL * 256, hi / 8);
    }
    if (best.N == 0) return last_rc;
    fill_info(best, out);
    return LUCAS_OK;
}

int lucas_autotune_candidates(int p, int* out, int cap, int* count) {
    if (p < 7 || !count || cap < 0 || (cap > 0 && !out)) return LUCAS_ERR_INVALID_ARG;
    long long lo = 0, hi = 0;
    autotune_window(p, &lo, &hi);
    const std::vector<int> c = autotune_candidates(p, lo, hi);
    *count = (int)c.size();
    for (int i = 0; i < cap && i < (int)c.size(); ++i) out[i] = c[i];
    return LUCAS_OK;
}

int lucas_autotune_shape(int p, int N, int iters, int max_candidates, const lucas_options* opt,
                         lucas_plan_info* out, double* best_ms) {
    if (!out || p < 7) return LUCAS_ERR_INVALID_ARG;
    if (N == 0) N = default_length(p);
    // the line-kernel shapes first (few, and usually the fastest), each also
    // with non-temporal spectrum loads and stores where the state outgrows
    // the L2s (N >= 2^19), then the tilings of the general kernels in
    // modelled order
    std::vector<Plan> cands;
    for (const Plan& c : line_candidates(p, N)) {
        cands.push_back(c);
        if (N >= 524288) {
            Plan nt = c;
            nt.nt = 1;
            cands.push_back(nt);
        }
    }
    std::vector<Plan> tiled = plan_candidates(p, N);
    if (tiled.empty()) tiled = plan_candidates(p, N, true);
    cands.insert(cands.end(), tiled.begin(), tiled.end());
    if (cands.empty()) return LUCAS_ERR_INVALID_ARG;
    if (max_candidates > 0 && (int)cands.size() > max_candidates) cands.resize(max_candidates);
    if (iters < 10) iters = 10;
    const int device = opt ? opt->device : 0;
    double best = 1e30;
    Plan bestp;
    int last_rc = LUCAS_OK;
    for (const Plan& pl : cands) {
        lucas_ctx* ctx = nullptr;
        int rc = ctx_new(pl, device, &ctx);
        double ms = 0;
        if (rc == LUCAS_OK) rc = lucas_ctx_time_iterations(ctx, 40, &ms);  // warm-up, graph build
        if (rc == LUCAS_OK) rc = lucas_ctx_time_iterations(ctx, iters, &ms);
        if (ctx) lucas_ctx_destroy(ctx);
        if (rc != LUCAS_OK) {
            last_rc = rc;
            continue;
        }
        if (ms < best) {
            best = ms;
            bestp = pl;
        }
    }
    if (bestp.N == 0) return last_rc != LUCAS_OK ? last_rc : LUCAS_ERR_DEVICE;
    fill_info(bestp, out);
    if (best_ms) *best_ms = best;
    return LUCAS_OK;
}

// ---- batches ----------------------------------------------------------------

int lucas_batch_create(const int* ps, const int* Ns, int njobs, int device, lucas_batch** out) {
    if (!ps || njobs <= 0 || !out) return LUCAS_ERR_INVALID_ARG;
    std::vector<Plan> plans(njobs);
    for (int i = 0; i < njobs; ++i) {
        int rc = plan_for(ps[i], Ns ? Ns[i] : 0, 0, 0, &plans[i], true);
        if (rc != LUCAS_OK) return rc;
    }
    lucas_batch* b = new lucas_batch();
    b->engine = new Engine(device);
    int rc = b->engine->init();
    for (int i = 0; i < njobs && rc == LUCAS_OK; ++i) {
        Job* j = nullptr;
        rc = job_create(plans[i], device, b->engine->stream(), &j);
        if (rc == LUCAS_OK) b->jobs.push_back(j);
    }
    if (rc != LUCAS_OK) {
        for (Job* j : b->jobs) job_destroy(j);
        delete b->engine;
        delete b;
        return rc;
    }
    b->engine->set_jobs(b->jobs);
    *out = b;
    return LUCAS_OK;
}

int lucas_batch_destroy(lucas_batch* b) {
    if (!b) return LUCAS_ERR_INVALID_ARG;
    b->engine->sync();
    for (Job* j : b->jobs) job_destroy(j);
    delete b->engine;
    delete b;
    return LUCAS_OK;
}

int lucas_batch_step(lucas_batch* b, long long iters) {
    if (!b || iters < 0) return LUCAS_ERR_INVALID_ARG;
    int rc = b->engine->step(iters);
    if (rc == LUCAS_OK) rc = b->engine->sync();
    return rc;
}

int lucas_batch_res64(lucas_batch* b, uint64_t* res64, int* is_zero, double* max_roundoff) {
    if (!b) return LUCAS_ERR_INVALID_ARG;
    for (size_t i = 0; i < b->jobs.size(); ++i) {
        Job* j = b->jobs[i];
        std::vector<int32_t> d(j->plan.N);
        int rc = job_get_digits(j, b->engine->stream(), d.data());
        if (rc != LUCAS_OK) return rc;
        uint64_t r = 0;
        int z = 0;
        normalize_digits(j->plan.p, j->plan.N, d.data(), nullptr, &r, &z);
        if (res64) res64[i] = r;
        if (is_zero) is_zero[i] = z;
        if (max_roundoff) {
            rc = job_read_maxerr(j, b->engine->stream(), 0, &max_roundoff[i]);
            if (rc != LUCAS_OK) return rc;
        }
    }
    return LUCAS_OK;
}

int lucas_batch_get_digits(lucas_batch* b, int job, int32_t* digits) {
    if (!b || job < 0 || job >= (int)b->jobs.size() || !digits) return LUCAS_ERR_INVALID_ARG;
    return job_get_digits(b->jobs[job], b->engine->stream(), digits);
}

int lucas_batch_time_iterations(lucas_batch* b, long long iters, double* ms_per_iter) {
    if (!b || iters <= 0 || !ms_per_iter) return LUCAS_ERR_INVALID_ARG;
    int rc = b->engine->prepare_graphs(iters);  // graph builds stay outside the timed region
    if (rc == LUCAS_OK) rc = b->engine->sync();
    if (rc != LUCAS_OK) return rc;
    const double t0 = now_s();
    rc = b->engine->step(iters);
    if (rc == LUCAS_OK) rc = b->engine->sync();
    *ms_per_iter = 1e3 * (now_s() - t0) / (double)iters;
    return rc;
}

}  // extern "C"

// ---- serving ----------------------------------------------------------------
//
// Continuous batching over three threads that share the device:
//   * the admitter takes queued requests in order while job slots and the
//     HBM budget allow, builds each job's tables and uploads them on its own
//     stream (device memory from a pool of freed job blocks), and hands the
//     resident job to the worker;
//   * the worker owns the launch engine: each tick it adds the jobs that
//     became resident, advances every resident job by one chunk in a single
//     batched launch sequence, copies out the digits of the jobs that reached
//     their iteration count and of those due a checkpoint, and drops the
//     finished ones from the batch;
//   * the retirer normalises finished residues (host O(N)), applies the
//     roundoff gate, writes checkpoint files, and returns device memory and
//     budget to the admitter.
// So admission, retirement and checkpoint I/O overlap the device work.  A
// job's arithmetic is independent of its neighbours in the batch, so its
// residue is the same alone or in any mix.

struct lucas_server {
    struct Pending {
        long long id;
        Plan plan;
        long long iters;
    };
    struct Active {
        long long id;
        Job* job;
        long long target;
    };
    // work for the retirer: a finished job (its digits and roundoff already
    // read), or a checkpoint of a running one
    struct Retire {
        long long id = 0;
        Job* job = nullptr;       // finished job to release, or nullptr
        int status = LUCAS_OK;
        int p = 0, N = 0;
        long long iter = 0;
        double max_roundoff = 0;
        std::vector<int32_t> digits;
        std::string checkpoint;   // non-empty: write digits there
    };
    lucas_server_options opt{};
    std::string ckdir;
    double limit = LUCAS_DEFAULT_ERROR_LIMIT;
    Engine* engine = nullptr;
    DevicePool* pool = nullptr;
    hipStream_t admit_stream = nullptr;
    std::thread worker, admitter, retirer;
    std::mutex mu;
    std::condition_variable cv;         // worker: jobs ready / stop
    std::condition_variable admit_cv;   // admitter: queue or budget changed / stop
    std::condition_variable retire_cv;  // retirer: work queued / stop
    std::condition_variable done_cv;    // a job finished
    std::deque<Pending> queue;
    std::deque<Active> ready;           // resident, not yet in the batch
    std::deque<Retire> retiring;
    std::map<long long, lucas_job_result> records;
    size_t budget = 0, used = 0;
    int resident = 0;                   // jobs admitted and not yet retired
    long long next_id = 1;
    long long outstanding = 0;
    bool stopping = false;
    double t0 = 0;
    long long iters_done = 0;
    double busy_s = 0;
    int peak_jobs = 0;
    int worker_rc = LUCAS_OK;

    void run();
    void admit();
    void retire();
    // <dir>/<p>_<id>.chk: two resident jobs of one exponent keep separate files
    std::string checkpoint_path(int p, long long id) const {
        return ckdir + "/" + std::to_string(p) + "_" + std::to_string(id) + ".chk";
    }
    int max_jobs() const { return opt.max_jobs > 0 ? opt.max_jobs : 256; }
};

void lucas_server::admit() {
    hipSetDevice(opt.device);
    for (;;) {
        Pending pnd;
        {
            std::unique_lock<std::mutex> lk(mu);
            admit_cv.wait(lk, [&] {
                if (stopping) return true;
                if (queue.empty() || resident >= max_jobs()) return false;
                // a job larger than what is free waits for memory, unless the
                // device holds nothing (the submit check bounds it by the budget)
                return used + job_bytes(queue.front().plan) <= budget || resident == 0;
            });
            if (stopping) return;
            pnd = queue.front();
            queue.pop_front();
            used += job_bytes(pnd.plan);
            ++resident;
        }
        Job* j = nullptr;
        const int rc = job_create(pnd.plan, opt.device, admit_stream, &j, pool);
        std::lock_guard<std::mutex> lk(mu);
        lucas_job_result& rec = records[pnd.id];
        rec.start_s = now_s() - t0;
        if (rc != LUCAS_OK) {
            used -= job_bytes(pnd.plan);
            --resident;
            rec.done = 1;
            rec.status = rc;
            rec.finish_s = rec.start_s;
            --outstanding;
            done_cv.notify_all();
            admit_cv.notify_all();
            continue;
        }
        ready.push_back({pnd.id, j, pnd.iters});
        cv.notify_all();
    }
}

void lucas_server::retire() {
    for (;;) {
        Retire w;
        {
            std::unique_lock<std::mutex> lk(mu);
            retire_cv.wait(lk, [&] { return stopping || !retiring.empty(); });
            if (retiring.empty()) return;  // stopping with nothing left
            w = std::move(retiring.front());
            retiring.pop_front();
        }
        if (!w.checkpoint.empty()) {
            write_checkpoint(w.checkpoint.c_str(), w.p, w.N, w.iter, w.digits.data());
            continue;
        }
        lucas_job_result res;
        memset(&res, 0, sizeof(res));
        res.status = w.status;
        if (res.status == LUCAS_OK) {
            normalize_digits(w.p, w.N, w.digits.data(), nullptr, &res.res64, &res.is_zero);
            // the gate of errorTrial / mersenneTest (gpuLucas.cu:904-907, 1239-1242)
            if (w.max_roundoff >= limit) res.status = LUCAS_ERR_ROUNDOFF;
        }
        const size_t bytes = w.job ? job_bytes(w.job->plan) : 0;
        job_destroy(w.job, pool);
        std::lock_guard<std::mutex> lk(mu);
        lucas_job_result& rec = records[w.id];
        rec.done = 1;
        rec.status = res.status;
        rec.res64 = res.res64;
        rec.is_zero = res.is_zero;
        rec.max_roundoff = w.max_roundoff;
        rec.finish_s = now_s() - t0;
        used -= bytes;
        --resident;
        --outstanding;
        done_cv.notify_all();
        admit_cv.notify_all();
    }
}

void lucas_server::run() {
    hipSetDevice(opt.device);
    const long long chunk_max = opt.chunk_iters > 0 ? opt.chunk_iters : 256;
    std::vector<Active> active;
    for (;;) {
        bool changed = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stopping || !ready.empty() || !active.empty(); });
            if (stopping) break;
            while (!ready.empty()) {
                active.push_back(ready.front());
                ready.pop_front();
                changed = true;
            }
        }
        if (changed) {
            std::vector<Job*> js;
            for (const Active& a : active) js.push_back(a.job);
            engine->set_jobs(js);
        }
        long long chunk = chunk_max;
        for (const Active& a : active) chunk = std::min(chunk, a.target - a.job->iter);
        const double tb = now_s();
        int rc = chunk > 0 ? engine->step(chunk) : LUCAS_OK;
        if (rc == LUCAS_OK) rc = engine->sync();
        {
            std::lock_guard<std::mutex> lk(mu);
            busy_s += now_s() - tb;
            if (rc != LUCAS_OK) worker_rc = rc;
            iters_done += chunk * (long long)active.size();
            peak_jobs = std::max(peak_jobs, (int)active.size());
        }
        // finished jobs and due checkpoints: digits out, the rest to the retirer
        std::vector<Active> keep;
        std::vector<Retire> out;
        for (Active& a : active) {
            const bool finished = a.job->iter >= a.target || rc != LUCAS_OK;
            const bool ckpt = !finished && !ckdir.empty() && opt.checkpoint_every > 0 &&
                              a.job->iter % opt.checkpoint_every < chunk;
            if (!finished && !ckpt) {
                keep.push_back(a);
                continue;
            }
            Retire w;
            w.id = a.id;
            w.p = a.job->plan.p;
            w.N = a.job->plan.N;
            w.iter = a.job->iter;
            w.status = rc;
            w.digits.resize(a.job->plan.N);
            if (w.status == LUCAS_OK) w.status = job_get_digits(a.job, engine->stream(), w.digits.data());
            if (finished) {
                if (w.status == LUCAS_OK) w.status = job_read_maxerr(a.job, engine->stream(), 0, &w.max_roundoff);
                w.job = a.job;
            } else {
                keep.push_back(a);
                if (w.status != LUCAS_OK) continue;
                w.checkpoint = checkpoint_path(w.p, a.id);
            }
            out.push_back(std::move(w));
        }
        if (keep.size() != active.size()) {
            active.swap(keep);
            std::vector<Job*> js;
            for (const Active& a : active) js.push_back(a.job);
            engine->set_jobs(js);
        }
        if (!out.empty()) {
            std::lock_guard<std::mutex> lk(mu);
            for (Retire& w : out) retiring.push_back(std::move(w));
            retire_cv.notify_all();
        }
    }
    // shut down: release resident jobs
    for (Active& a : active) job_destroy(a.job, pool);
    active.clear();
}

extern "C" {

int lucas_server_start(const lucas_server_options* opt, lucas_server** out) {
    if (!out) return LUCAS_ERR_INVALID_ARG;
    lucas_server* s = new lucas_server();
    if (opt) s->opt = *opt;
    if (opt && opt->checkpoint_dir) s->ckdir = opt->checkpoint_dir;
    s->opt.checkpoint_dir = nullptr;
    s->limit = s->opt.error_limit > 0 ? s->opt.error_limit : LUCAS_DEFAULT_ERROR_LIMIT;
    s->engine = new Engine(s->opt.device);
    int rc = s->engine->init();
    if (rc == LUCAS_OK && hipStreamCreateWithFlags(&s->admit_stream, hipStreamNonBlocking) != hipSuccess)
        rc = LUCAS_ERR_DEVICE;
    if (rc != LUCAS_OK) {
        delete s->engine;
        delete s;
        return rc;
    }
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) freeb = 0;
    s->budget = s->opt.hbm_budget_bytes ? s->opt.hbm_budget_bytes : (size_t)(0.8 * (double)freeb);
    // freed job blocks kept for reuse: up to a quarter of the budget
    s->pool = new DevicePool(s->opt.device, s->budget / 4);
    s->t0 = now_s();
    s->worker = std::thread([s] { s->run(); });
    s->admitter = std::thread([s] { s->admit(); });
    s->retirer = std::thread([s] { s->retire(); });
    *out = s;
    return LUCAS_OK;
}

static const long long kServerBusyJobs = 4;

int lucas_server_submit(lucas_server* s, int p, int N, long long iterations, long long* id) {
    if (!s || iterations <= 0) return LUCAS_ERR_INVALID_ARG;
    Plan pl;
    // Lightly loaded, a served job's latency is what counts: the single-job
    // plan.  With kServerBusyJobs or more jobs already waiting or running the
    // device is shared anyway, and the batch plan's small workgroups and wide
    // tiles raise the aggregate rate (MI355X: four exponents near 3e6 as one
    // batch already match their single-job plans run back to back).
    long long outstanding = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        outstanding = s->outstanding;
    }
    int rc = plan_for(p, N, 0, 0, &pl, outstanding >= kServerBusyJobs);
    if (rc != LUCAS_OK) return rc;
    if (job_bytes(pl) > s->budget) return LUCAS_ERR_BUSY;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return LUCAS_ERR_BUSY;
    const long long jid = s->next_id++;
    lucas_job_result& rec = s->records[jid];
    memset(&rec, 0, sizeof(rec));
    rec.id = jid;
    rec.p = p;
    rec.N = pl.N;
    rec.iterations = iterations;
    rec.submit_s = now_s() - s->t0;
    s->queue.push_back({jid, pl, iterations});
    ++s->outstanding;
    s->admit_cv.notify_all();
    if (id) *id = jid;
    return LUCAS_OK;
}

int lucas_server_poll(lucas_server* s, long long id, lucas_job_result* out) {
    if (!s || !out) return LUCAS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->records.find(id);
    if (it == s->records.end()) return LUCAS_ERR_INVALID_ARG;
    *out = it->second;
    return LUCAS_OK;
}

int lucas_server_drain(lucas_server* s, double timeout_s) {
    if (!s) return LUCAS_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto pred = [&] { return s->outstanding == 0; };
    if (timeout_s > 0) {
        if (!s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return LUCAS_ERR_BUSY;
    } else {
        s->done_cv.wait(lk, pred);
    }
    return s->worker_rc;
}

int lucas_server_stats(lucas_server* s, long long* iterations, double* busy_s, int* peak_jobs) {
    if (!s) return LUCAS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (iterations) *iterations = s->iters_done;
    if (busy_s) *busy_s = s->busy_s;
    if (peak_jobs) *peak_jobs = s->peak_jobs;
    return LUCAS_OK;
}

int lucas_server_stop(lucas_server* s) {
    if (!s) return LUCAS_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv.notify_all();
        s->admit_cv.notify_all();
        s->retire_cv.notify_all();
    }
    if (s->admitter.joinable()) s->admitter.join();
    if (s->worker.joinable()) s->worker.join();
    if (s->retirer.joinable()) s->retirer.join();
    // jobs admitted but never picked up, and retire work left over
    for (lucas_server::Active& a : s->ready) job_destroy(a.job, s->pool);
    for (lucas_server::Retire& w : s->retiring) job_destroy(w.job, s->pool);
    delete s->engine;
    if (s->admit_stream) hipStreamDestroy(s->admit_stream);
    delete s->pool;
    delete s;
    return LUCAS_OK;
}

}  // extern "C"
