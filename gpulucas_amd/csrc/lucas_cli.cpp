// lucas_cli.cpp -- command-line driver over the C ABI, with the reference's
// flags and run sequence (main, gpuLucas.cu:387-639):
//   gpulucas [-h] [-v] [-q] [-f N] [-d device] [-c checkpoint_freq] [-n] p
//   1. resume from <p>.chk (or <p>.chk.bak) when present (gpuLucas.cu:499-500)
//   2. N from -f, from the checkpoint, or measured (findSignalSize, 502-506)
//   3. error trial of clamp(p/10000, 100, 2500) iterations; refuse N when the
//      roundoff reaches the limit (529-568)
//   4. full test with periodic checkpoints and progress lines (1399-1484);
//      SIGINT/SIGQUIT/SIGTERM/SIGTSTP checkpoint and exit (365-369, 598-606);
//      the roundoff gate is read after every slice: a breach prints "Error
//      limit exceeded" (also under -q), writes no result and exits 4 (the
//      reference only printed it, 1404-1407)
//   5. result line on stdout and appended to results.txt (1514-1547)
#include <getopt.h>
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <string>

#include "gpulucas.h"

static volatile sig_atomic_t g_stop = 0;
static void on_signal(int) { g_stop = 1; }

static void usage(const char* prog) {
    printf("Usage: %s [-h] [-v] [-q] [-f N] [-d device] [-c checkpoint_freq] [-n] p\n", prog);
    printf("  -f N   signal length (2^a 3^b 5^c 7^d, even); default: measured\n");
    printf("  -d k   GPU ordinal (default 0)\n");
    printf("  -c k   checkpoint every k iterations (default 50000)\n");
    printf("  -v     verbose: per-checkpoint residues\n");
    printf("  -q     quiet\n");
    printf("  -n p   exponent (also accepted positionally)\n");
}

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static void friendly(char* buf, size_t n, double secs) {
    long t = (long)secs;
    if (t >= 86400)
        snprintf(buf, n, "%ldd %02ld:%02ld:%02ld", t / 86400, (t % 86400) / 3600, (t % 3600) / 60, t % 60);
    else
        snprintf(buf, n, "%02ld:%02ld:%02ld", t / 3600, (t % 3600) / 60, t % 60);
}

// An exponent argument: decimal digits only, at most 2^31 - 1 (the
// reference's check, gpuLucas.cu:437-441).  0 = ok, 1 = not a number,
// 2 = too large.
static int parse_exponent(const char* s, int* p) {
    if (!*s || strlen(s) != strspn(s, "0123456789")) return 1;
    if (strlen(s) > 10 || strtoll(s, nullptr, 10) > 2147483647LL) return 2;
    *p = (int)strtoll(s, nullptr, 10);
    return 0;
}

int main(int argc, char** argv) {
    int p = 0, N = 0, device = 0, verbose = 0, quiet = 0;
    long long ckfreq = 50000;
    int opt;
    while ((opt = getopt(argc, argv, "hvqf:n:d:c:")) != -1) {
        switch (opt) {
            case 'h': usage(argv[0]); return 0;
            case 'v': verbose = 1; break;
            case 'q': quiet = 1; break;
            case 'f': N = atoi(optarg); break;
            case 'n':
                if (parse_exponent(optarg, &p) == 2) {
                    fprintf(stderr, "testPrime too large, aborting\n");
                    return 2;
                }
                break;
            case 'd': device = atoi(optarg); break;
            case 'c': ckfreq = atoll(optarg); break;
            default: usage(argv[0]); return 2;
        }
    }
    if (optind < argc) {
        const int e = parse_exponent(argv[optind], &p);
        if (e == 2) {
            fprintf(stderr, "testPrime too large, aborting\n");
            return 2;
        }
        if (e == 1) {
            usage(argv[0]);
            fprintf(stderr, "Non-option argument %s\n", argv[optind]);
            return 2;
        }
    }
    if (p < 7 || ckfreq <= 0) {
        usage(argv[0]);
        return 2;
    }
    const std::string ck = std::to_string(p) + ".chk";

    lucas_ctx* ctx = nullptr;
    int rc = lucas_ctx_load(ck.c_str(), device, &ctx);
    if (rc == LUCAS_OK) {
        lucas_plan_info info;
        lucas_ctx_info(ctx, &info);
        if (info.p != p) {
            fprintf(stderr, "checkpoint %s belongs to another exponent\n", ck.c_str());
            lucas_ctx_destroy(ctx);
            return 1;
        }
        if (!quiet) printf("Resuming M%d from %s at iteration %lld (N = %d)\n", p, ck.c_str(), lucas_ctx_iteration(ctx), info.N);
    } else {
        ctx = nullptr;
        if (N == 0) {
            lucas_plan_info best;
            lucas_options o;
            memset(&o, 0, sizeof(o));
            o.device = device;
            rc = lucas_autotune_N(p, 300, &o, &best);
            N = rc == LUCAS_OK ? best.N : lucas_default_N(p);
            if (!quiet) printf("Signal length N = %d (%s)\n", N, rc == LUCAS_OK ? "measured" : "default rule");
        }
        // error trial before the full run (errorTrial, gpuLucas.cu:1192-1282)
        const int trial = std::min(2500, std::max(100, p / 10000));
        lucas_options o;
        memset(&o, 0, sizeof(o));
        o.device = device;
        lucas_result tr;
        rc = lucas_error_test(p, N, trial, &o, &tr);
        if (rc == LUCAS_ERR_ROUNDOFF) {
            fprintf(stderr, "roundoff %.4f at N = %d exceeds %.2f; choose a longer -f\n", tr.max_roundoff, N,
                    LUCAS_DEFAULT_ERROR_LIMIT);
            return 1;
        }
        if (rc != LUCAS_OK) {
            fprintf(stderr, "error test: %s\n", lucas_status_string(rc));
            return 1;
        }
        if (!quiet) {
            char eta[64];
            friendly(eta, sizeof(eta), tr.ms_per_iter * 1e-3 * (double)(p - 2));
            printf("Error trial: %d iterations, %.4f ms/iter, max roundoff %.4f; estimated run time %s\n", trial,
                   tr.ms_per_iter, tr.max_roundoff, eta);
        }
        rc = lucas_ctx_create(p, N, device, &ctx);
        if (rc != LUCAS_OK) {
            fprintf(stderr, "context: %s\n", lucas_status_string(rc));
            return 1;
        }
    }

    struct sigaction sa;
    memset(&sa, 0, sizeof(sa));
    sa.sa_handler = on_signal;
    for (int s : {SIGINT, SIGQUIT, SIGTERM, SIGTSTP}) sigaction(s, &sa, nullptr);

    lucas_plan_info info;
    lucas_ctx_info(ctx, &info);
    const long long total = (long long)p - 2;
    const long long slice = 4096;  // signals are checked between slices
    double t_last = now_s();
    long long it_last = lucas_ctx_iteration(ctx);
    while (lucas_ctx_iteration(ctx) < total && !g_stop) {
        const long long it = lucas_ctx_iteration(ctx);
        long long next = std::min(total, (it / ckfreq + 1) * ckfreq);
        next = std::min(next, it + slice);
        rc = lucas_ctx_step(ctx, next - it);
        double gate_err = 0;
        if (rc == LUCAS_OK) rc = lucas_ctx_max_roundoff(ctx, 0, &gate_err);
        if (rc != LUCAS_OK) {
            fprintf(stderr, "step: %s\n", lucas_status_string(rc));
            lucas_ctx_destroy(ctx);
            return 1;
        }
        const long long now = lucas_ctx_iteration(ctx);
        if (gate_err >= LUCAS_DEFAULT_ERROR_LIMIT) {
            // no result line and no checkpoint past the breach
            fprintf(stderr, "Error limit exceeded: max roundoff %.4f >= %.2f at iteration %lld (N = %d); "
                    "no result written, rerun with a longer -f\n", gate_err, LUCAS_DEFAULT_ERROR_LIMIT, now, info.N);
            lucas_ctx_destroy(ctx);
            return 4;
        }
        if (now % ckfreq == 0 && now < total) {
            if (lucas_ctx_save(ctx, ck.c_str()) != LUCAS_OK) fprintf(stderr, "warning: checkpoint write failed\n");
            if (!quiet) {
                const double t = now_s();
                const double ms = 1e3 * (t - t_last) / (double)std::max(1LL, now - it_last);
                double err = 0;
                lucas_ctx_max_roundoff(ctx, 0, &err);
                char eta[64];
                friendly(eta, sizeof(eta), ms * 1e-3 * (double)(total - now));
                printf("[%4.1f%%] Iter %lld: %.4f ms/iter, roundoff %.4f, ETA %s", 100.0 * (double)now / (double)p,
                       now, ms, err, eta);
                if (verbose) {
                    // the reference's verbose "max Bit Vector" (gpuLucas.cu:911-916)
                    uint64_t r = 0;
                    int z = 0, bits = 0, clz = 0;
                    lucas_ctx_res64(ctx, &r, &z);
                    lucas_ctx_product_bits(ctx, 0, &bits, &clz);
                    printf(", 0x%016llx, product bits %d, max Bit Vector %d", (unsigned long long)r, bits, clz);
                }
                printf("\n");
                fflush(stdout);
                t_last = t;
                it_last = now;
            }
        }
    }
    if (g_stop) {
        rc = lucas_ctx_save(ctx, ck.c_str());
        if (!quiet)
            printf("Interrupted at iteration %lld; state saved to %s (%s)\n", lucas_ctx_iteration(ctx), ck.c_str(),
                   lucas_status_string(rc));
        lucas_ctx_destroy(ctx);
        return rc == LUCAS_OK ? 3 : 1;
    }
    uint64_t res = 0;
    int zero = 0;
    double err = 0;
    lucas_ctx_res64(ctx, &res, &zero);
    lucas_ctx_max_roundoff(ctx, 0, &err);
    lucas_ctx_destroy(ctx);
    char line[160];
    if (zero)
        snprintf(line, sizeof(line), "M( %d )P, n = %d, %s", p, info.N, lucas_version());
    else
        snprintf(line, sizeof(line), "M( %d )C, 0x%016llx, n = %d, %s", p, (unsigned long long)res, info.N,
                 lucas_version());
    printf("%s\n", line);
    FILE* f = fopen("results.txt", "a");
    if (f) {
        fprintf(f, "%s\n", line);
        fclose(f);
    }
    unlink(ck.c_str());
    unlink((ck + ".bak").c_str());
    return 0;
}
