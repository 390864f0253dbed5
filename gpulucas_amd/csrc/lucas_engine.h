// lucas_engine.h -- device-side job state and the launch engine (HIP runtime).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <vector>

#include "lucas_desc.h"
#include "lucas_host.h"

namespace lucas {

// One exponent's device-resident state: residue digits, two spectrum
// buffers, weight/twiddle tables and the roundoff accumulator, carved from a
// single allocation.
struct Job {
    Plan plan;
    int device = 0;
    char* mem = nullptr;
    size_t bytes = 0;
    LucasJobDesc desc{};
    long long iter = 0;
};

// Device allocations of jobs that come and go (a server's): freed blocks
// are kept by exact size, up to `cap` bytes in all, and handed to the next
// job of that size instead of going through hipFree / hipMalloc.
class DevicePool {
public:
    DevicePool(int device, size_t cap) : device_(device), cap_(cap) {}
    ~DevicePool();
    // A block of exactly `bytes` (cached or new); hipErrorOutOfMemory when
    // neither is possible.
    hipError_t acquire(size_t bytes, char** ptr);
    // Returns a block; kept while the cache stays within cap, else freed.
    void release(char* ptr, size_t bytes);
    size_t cached() const { return cached_; }

private:
    int device_;
    size_t cap_, cached_ = 0;
    std::mutex mu_;
    std::multimap<size_t, char*> free_;
};

// Creates a job with s_0 = 4 (digits = [4, 0, ...]) and uploads its tables on
// `stream`; device memory from `pool` when given.  Returns a LUCAS_* code.
int job_create(const Plan& plan, int device, hipStream_t stream, Job** out, DevicePool* pool = nullptr);
// Releases a job's device memory (to `pool` when given) and the job.
void job_destroy(Job* job, DevicePool* pool = nullptr);
int job_get_digits(Job* job, hipStream_t stream, int32_t* digits);
int job_set_digits(Job* job, hipStream_t stream, const int32_t* digits);
int job_read_maxerr(Job* job, hipStream_t stream, int reset, double* out);
// Product-size diagnostic since the last reset (see LucasJobDesc::pbits).
int job_read_pbits(Job* job, hipStream_t stream, int reset, int* max_bits, int* max_clz);

// A set of jobs advanced together: one launch of each kernel per iteration
// covers every job (workgroup -> (job, tile) maps built on the host).
class Engine {
public:
    explicit Engine(int device);
    ~Engine();
    int init();
    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }

    void set_jobs(const std::vector<Job*>& jobs);  // composition change
    const std::vector<Job*>& jobs() const { return jobs_; }

    // Advances every job by `iters` (>= 1) iterations; digits in, digits out.
    // Every launch is a replay-graph launch.
    int step(long long iters);
    // Builds (without running) every replay graph a step of `iters`
    // iterations launches, so that a later timed step builds nothing.
    int prepare_graphs(long long iters);
    // Graph slots: 0 .. kLevels-1 hold 2^slot steady-state iterations.
    static const int kHead = 8, kTail = 9, kGraphSlots = 10;
    // Mean microseconds of each kernel launch over `iters` iterations in the
    // step() order, from events after every launch (digits are not touched).
    int time_kernels(int iters, double* col_us, double* row_us);
    // One job only, profiling library only: runs `iters` iterations in the
    // step() order (digits are not touched) and copies the phase stamps of
    // the last column and row launch (8 words per workgroup).
    int phase_stamps(int iters, unsigned long long* col, unsigned long long* row, int max_wgs, int* ncol,
                     int* nrow);
    int sync();

private:
    int prepare();
    int launch_col(int mode);
    int launch_row();
    int build_graph(int slot);

    // Jobs whose kernel has the same workgroup size and wave factor share
    // one launch; column and row launches are grouped separately.  Groups
    // 0..15 are the tiled kernels (3 workgroup sizes and the dense 1024
    // column variant x 4 wave factors P = 0, 1, 2, 4); group kLineBase + i
    // is line instantiation i.
    static const int kLineBase = 16;
    struct Group {
        int threads = 0, p = 0, dense = 0;
        int line = -1;  // line instantiation, -1: tiled kernels
        int2* d_map = nullptr;
        size_t cap = 0;
        int n = 0;
        size_t lds = 0;
    };

    int device_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    std::vector<Job*> jobs_;
    bool dirty_ = true;
    LucasJobDesc* d_desc_ = nullptr;
    size_t cap_desc_ = 0;
    bool single_ = false;  // one job: launch the by-value-descriptor kernels with one_
    LucasJobDesc one_;
    std::map<int, Group> colg_, rowg_;
    // replay graphs by slot (see kHead, kTail), built on first use; after a
    // change of jobs a slot's executable graph is kept as stale and updated
    // in place from the new capture where the kernel sequence allows
    hipGraph_t graph_[kGraphSlots] = {};
    hipGraphExec_t gexec_[kGraphSlots] = {};
    bool stale_[kGraphSlots] = {};
    // whole-step graphs of a single-job engine: one graph per short step
    // length k <= kWholeMax (head, k - 1 pairs, tail), the most recent
    // kWholeCache lengths kept
    static const int kWholeMax = 64, kWholeCache = 8;
    struct Whole {
        long long k = 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Whole> whole_;
    int whole_graph(long long k, hipGraphExec_t* out);
    void destroy_graphs();
};

}  // namespace lucas
