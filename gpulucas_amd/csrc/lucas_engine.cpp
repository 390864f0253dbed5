// lucas_engine.cpp -- device memory, launch sequencing and HIP graphs.
//
// Iteration order for `iters` squarings of every job (digits in, digits out):
//   col(IN_DIGITS)                      weight + forward column FFT
//   [row ; col] x (iters - 1)           one LL iteration each
//   row ; col(OUT_DIGITS)               last iteration, stop after the carry
// The steady-state [row ; col] pairs are replayed from a captured hipGraph.
// This replaces the reference's per-iteration launch train (2 cuFFT execs + 4
// kernels with a cutilCheckMsg after each, gpuLucas.cu:845-942).
#include "lucas_engine.h"

#include <string.h>

#include <algorithm>
#include <vector>

#include "gpulucas.h"

namespace lucas {

static const int kXcds = 8;  // accelerator complex dies of an MI355X

hipError_t launch_col_kernel(int threads, int p, int dense, const LucasJobDesc* jobs, const int2* map, int nwg,
                             size_t lds, int mode, hipStream_t stream);
hipError_t launch_row_kernel(int threads, int p, const LucasJobDesc* jobs, const int2* map, int nwg, size_t lds,
                             hipStream_t stream);
hipError_t launch_col1_kernel(int threads, int p, int dense, const LucasJobDesc& job, int nwg, size_t lds, int mode,
                              hipStream_t stream);
hipError_t launch_row1_kernel(int threads, int p, const LucasJobDesc& job, int nwg, size_t lds, hipStream_t stream);
hipError_t configure_kernels(size_t max_lds);
hipError_t configure_line_kernels(size_t max_lds);
hipError_t launch_line_col(int idx, const LucasJobDesc* jobs, const int2* map, const LucasJobDesc* one, int nwg,
                           int mode, hipStream_t stream);
hipError_t launch_line_row(int idx, const LucasJobDesc* jobs, const int2* map, const LucasJobDesc* one, int nwg,
                           hipStream_t stream);
bool phase_stamps_built();
hipError_t read_phase_stamps(unsigned long long* col, int ncol, unsigned long long* row, int nrow);
hipError_t read_line_phase_stamps(unsigned long long* col, int ncol, unsigned long long* row, int nrow);

// Replay graphs.  A step of k iterations is the head graph (column kernel,
// digits in), k - 1 steady-state [row ; col] pairs and the tail graph (row,
// column kernel with digits out).  The pairs run as graphs of 2^l pairs,
// l = 0 .. kLevels - 1: graphs of 128 while at least 128 remain in a step of
// kLongStep or more (measured 0.8 % faster per iteration than graphs of 32 at
// N = 163840), graphs of 32 in a shorter step, then the binary digits of the
// remainder, largest first.
// So every iteration of every step is graph-launched; a step of k < 128 pairs
// launches at most 2 + log2(k) graphs.  A single-job engine runs a step of
// at most kWholeMax iterations as one whole-step graph instead (each graph
// launch costs host time the GPU waits for at the start of a step).  Graphs
// are built on first use, or ahead of a timed region by prepare_graphs().
static const int kLevels = 8;  // 1, 2, 4, .., 128 pairs
static const long long kLongStep = 512;

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

DevicePool::~DevicePool() {
    hipSetDevice(device_);
    for (auto& kv : free_) hipFree(kv.second);
}

hipError_t DevicePool::acquire(size_t bytes, char** ptr) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.find(bytes);
        if (it != free_.end()) {
            *ptr = it->second;
            free_.erase(it);
            cached_ -= bytes;
            return hipSuccess;
        }
    }
    hipSetDevice(device_);
    hipError_t e = hipMalloc((void**)ptr, bytes);
    if (e == hipErrorOutOfMemory) {
        // give the cache back to the device and try once more
        std::multimap<size_t, char*> drop;
        {
            std::lock_guard<std::mutex> lk(mu_);
            drop.swap(free_);
            cached_ = 0;
        }
        (void)hipGetLastError();
        for (auto& kv : drop) hipFree(kv.second);
        e = hipMalloc((void**)ptr, bytes);
    }
    return e;
}

void DevicePool::release(char* ptr, size_t bytes) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (cached_ + bytes <= cap_) {
            free_.emplace(bytes, ptr);
            cached_ += bytes;
            return;
        }
    }
    hipSetDevice(device_);
    hipFree(ptr);
}

int job_create(const Plan& plan, int device, hipStream_t stream, Job** out, DevicePool* pool) {
    if (hipSetDevice(device) != hipSuccess) return LUCAS_ERR_NO_DEVICE;
    Job* job = new Job();
    job->plan = plan;
    job->device = device;
    job->bytes = job_bytes(plan);
    hipError_t e = pool ? pool->acquire(job->bytes, &job->mem) : hipMalloc((void**)&job->mem, job->bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        delete job;
        return e == hipErrorOutOfMemory ? LUCAS_ERR_OOM : LUCAS_ERR_DEVICE;
    }
    auto give_back = [&]() {
        if (pool)
            pool->release(job->mem, job->bytes);
        else
            hipFree(job->mem);
    };
    Tables t;
    make_tables(plan, &t);
    // spectra, digits, then the table block (layout mirrors job_bytes)
    const size_t o_spec = 0;
    const size_t o_spec2 = align256(spec_bytes(plan));
    const size_t o_dig = 2 * o_spec2;
    const size_t o_tab = o_dig + align256(sizeof(int) * plan.N);
    const TableLayout lay = table_layout(plan);
    if (o_tab + lay.bytes > job->bytes) {
        give_back();
        delete job;
        return LUCAS_ERR_INVALID_ARG;
    }
    // Host image of everything from the digits on (spectra need no init).
    std::vector<char> host(job->bytes - o_dig, 0);
    ((int*)host.data())[0] = 4;  // s_0 = 4 with a_0 = 1 (loadValue4ToFFTarray, gpuLucas.cu:718-726)
    table_image(plan, t, lay, host.data() + (o_tab - o_dig));
    // the accumulators, and a reused block's old state, are overwritten here
    e = hipMemcpyAsync(job->mem + o_dig, host.data(), host.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        give_back();
        delete job;
        return LUCAS_ERR_DEVICE;
    }
    LucasJobDesc& d = job->desc;
    fill_desc_shape(plan, &d);
    char* m = job->mem;
    d.spec = (double2*)(m + o_spec);
    d.spec2 = (double2*)(m + o_spec2);
    d.digits = (int*)(m + o_dig);
    desc_tables(lay, m + o_tab, &d);
    job->iter = 0;
    *out = job;
    return LUCAS_OK;
}

void job_destroy(Job* job, DevicePool* pool) {
    if (!job) return;
    if (job->mem) {
        if (pool) {
            pool->release(job->mem, job->bytes);
        } else {
            hipSetDevice(job->device);
            hipFree(job->mem);
        }
    }
    delete job;
}

int job_get_digits(Job* job, hipStream_t stream, int32_t* digits) {
    hipError_t e = hipMemcpyAsync(digits, job->desc.digits, sizeof(int) * job->plan.N, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e == hipSuccess ? LUCAS_OK : LUCAS_ERR_DEVICE;
}

int job_set_digits(Job* job, hipStream_t stream, const int32_t* digits) {
    hipError_t e = hipMemcpyAsync(job->desc.digits, digits, sizeof(int) * job->plan.N, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e == hipSuccess ? LUCAS_OK : LUCAS_ERR_DEVICE;
}

int job_read_pbits(Job* job, hipStream_t stream, int reset, int* max_bits, int* max_clz) {
    unsigned int v[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(v, job->desc.pbits, sizeof(v), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess && reset) {
        const unsigned int zero[2] = {0, 0};
        e = hipMemcpyAsync(job->desc.pbits, zero, sizeof(zero), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
    }
    if (e != hipSuccess) return LUCAS_ERR_DEVICE;
    if (max_bits) *max_bits = (int)v[0];
    if (max_clz) *max_clz = (int)v[1];
    return LUCAS_OK;
}

int job_read_maxerr(Job* job, hipStream_t stream, int reset, double* out) {
    unsigned int bits = 0;
    hipError_t e = hipMemcpyAsync(&bits, job->desc.maxerr, sizeof(bits), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return LUCAS_ERR_DEVICE;
    float f;
    memcpy(&f, &bits, sizeof(f));
    *out = (double)f;
    if (reset) {
        const unsigned int zero = 0;
        e = hipMemcpyAsync(job->desc.maxerr, &zero, sizeof(zero), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return LUCAS_ERR_DEVICE;
    }
    return LUCAS_OK;
}

Engine::Engine(int device) : device_(device) {}

Engine::~Engine() {
    hipSetDevice(device_);
    destroy_graphs();
    if (d_desc_) hipFree(d_desc_);
    for (auto* gs : {&colg_, &rowg_})
        for (auto& kv : *gs)
            if (kv.second.d_map) hipFree(kv.second.d_map);
    if (ev0_) hipEventDestroy(ev0_);
    if (ev1_) hipEventDestroy(ev1_);
    if (stream_) hipStreamDestroy(stream_);
}

int Engine::init() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return LUCAS_ERR_NO_DEVICE;
    }
    if (device_ < 0 || device_ >= count) return LUCAS_ERR_INVALID_ARG;
    if (hipSetDevice(device_) != hipSuccess) return LUCAS_ERR_NO_DEVICE;
    if (configure_kernels(kMaxLdsBytes) != hipSuccess || configure_line_kernels(kMaxLdsBytes) != hipSuccess)
        return LUCAS_ERR_DEVICE;
    if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) return LUCAS_ERR_DEVICE;
    if (hipEventCreate(&ev0_) != hipSuccess || hipEventCreate(&ev1_) != hipSuccess) return LUCAS_ERR_DEVICE;
    return LUCAS_OK;
}

void Engine::set_jobs(const std::vector<Job*>& jobs) {
    jobs_ = jobs;
    dirty_ = true;
}

int Engine::prepare() {
    if (!dirty_) return LUCAS_OK;
    hipSetDevice(device_);
    // the slot graphs become stale (updated in place when next used); the
    // whole-step graphs of single jobs are rebuilt
    for (int l = 0; l < kGraphSlots; ++l) stale_[l] = gexec_[l] != nullptr;
    for (Whole& w : whole_) {
        if (w.exec) hipGraphExecDestroy(w.exec);
        if (w.graph) hipGraphDestroy(w.graph);
    }
    whole_.clear();
    std::vector<LucasJobDesc> desc;
    std::map<int, std::vector<int2>> cmap, rmap;
    static const int kThreads[4] = {LUCAS_THREADS_SMALL, LUCAS_THREADS_LARGE, LUCAS_THREADS_MAX, LUCAS_THREADS_MAX};
    static const int kP[4] = {0, 1, 2, 4};
    // groups keep their device maps across composition changes; a group no
    // job uses any more launches nothing
    for (auto* gs : {&colg_, &rowg_})
        for (auto& kv : *gs) kv.second.n = 0;
    auto setup = [&](Group& g, int key, bool col) {
        if (key >= kLineBase) {
            g.line = key - kLineBase;
            g.threads = g.p = g.dense = 0;
            g.lds = 0;
        } else {
            g.line = -1;
            g.threads = kThreads[key / 4];
            g.p = kP[key % 4];
            g.dense = col && key / 4 == 3;
        }
    };
    auto group_of = [](int threads, int p, int dense) {
        const int ti = threads == LUCAS_THREADS_MAX ? (dense ? 3 : 2) : (threads == LUCAS_THREADS_LARGE ? 1 : 0);
        return 4 * ti + (p == 4 ? 3 : p);
    };
    for (size_t i = 0; i < jobs_.size(); ++i) {
        const Job* j = jobs_[i];
        desc.push_back(j->desc);
        const int gc = j->plan.line ? kLineBase + j->plan.line_col : group_of(j->plan.threads, j->plan.wA / 64, j->plan.col_dense);
        const int gr = j->plan.line ? kLineBase + j->plan.line_row : group_of(j->plan.threads, j->plan.wB / 64, 0);
        for (int w = 0; w < j->plan.col_wgs; ++w) cmap[gc].push_back(make_int2((int)i, w));
        for (int w = 0; w < j->plan.row_wgs; ++w) rmap[gr].push_back(make_int2((int)i, w));
        if (!colg_.count(gc)) colg_[gc].lds = 0;
        if (!rowg_.count(gr)) rowg_[gr].lds = 0;
        setup(colg_[gc], gc, true);
        setup(rowg_[gr], gr, false);
        if (!j->plan.line) {
            colg_[gc].lds = std::max(colg_[gc].lds, j->plan.lds_col);
            rowg_[gr].lds = std::max(rowg_[gr].lds, j->plan.lds_row);
        }
    }
    // Workgroups are dealt to the 8 XCDs round-robin by workgroup id, and each
    // XCD has its own L2.  Neighbouring column tiles share halo columns and
    // cache lines, so each XCD gets a contiguous range of tiles: workgroup
    // 8k + x runs tile k of XCD x's range.
    for (auto& kv : cmap) {
        std::vector<int2>& m = kv.second;
        const size_t n = m.size();
        if (n < 16) continue;
        std::vector<int2> out;
        out.reserve(n);
        size_t start[kXcds + 1];
        for (int x = 0; x <= kXcds; ++x) start[x] = n * x / kXcds;
        for (size_t k = 0; out.size() < n; ++k)
            for (int x = 0; x < kXcds; ++x)
                if (start[x] + k < start[x + 1]) out.push_back(m[start[x] + k]);
        m.swap(out);
    }
    auto grow = [&](void** ptr, size_t* cap, size_t need) -> bool {
        if (need <= *cap) return true;
        if (*ptr) hipFree(*ptr);
        *ptr = nullptr;
        *cap = 0;
        size_t n = need * 2;
        if (hipMalloc(ptr, n) != hipSuccess) return false;
        *cap = n;
        return true;
    };
    if (!grow((void**)&d_desc_, &cap_desc_, sizeof(LucasJobDesc) * (desc.size() + 1))) return LUCAS_ERR_OOM;
    hipError_t e = hipSuccess;
    if (!desc.empty())
        e = hipMemcpyAsync(d_desc_, desc.data(), sizeof(LucasJobDesc) * desc.size(), hipMemcpyHostToDevice, stream_);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        for (auto& kv : k == 0 ? cmap : rmap) {
            Group& g = k == 0 ? colg_[kv.first] : rowg_[kv.first];
            const std::vector<int2>& m = kv.second;
            if (!grow((void**)&g.d_map, &g.cap, sizeof(int2) * (m.size() + 1))) return LUCAS_ERR_OOM;
            g.n = (int)m.size();
            if (g.n) e = hipMemcpyAsync(g.d_map, m.data(), sizeof(int2) * g.n, hipMemcpyHostToDevice, stream_);
            if (e != hipSuccess) break;
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return LUCAS_ERR_DEVICE;
    // one job: its descriptor goes by value with every launch
    single_ = desc.size() == 1;
    if (single_) one_ = desc[0];
    dirty_ = false;
    return LUCAS_OK;
}

int Engine::launch_col(int mode) {
    for (const auto& kv : colg_) {
        const Group& g = kv.second;
        if (!g.n) continue;
        hipError_t e;
        if (g.line >= 0)
            e = launch_line_col(g.line, d_desc_, g.d_map, single_ ? &one_ : nullptr, g.n, mode, stream_);
        else
            e = single_ ? launch_col1_kernel(g.threads, g.p, g.dense, one_, g.n, g.lds, mode, stream_)
                        : launch_col_kernel(g.threads, g.p, g.dense, d_desc_, g.d_map, g.n, g.lds, mode, stream_);
        if (e != hipSuccess) return LUCAS_ERR_DEVICE;
    }
    return LUCAS_OK;
}

int Engine::launch_row() {
    for (const auto& kv : rowg_) {
        const Group& g = kv.second;
        if (!g.n) continue;
        hipError_t e;
        if (g.line >= 0)
            e = launch_line_row(g.line, d_desc_, g.d_map, single_ ? &one_ : nullptr, g.n, stream_);
        else
            e = single_ ? launch_row1_kernel(g.threads, g.p, one_, g.n, g.lds, stream_)
                        : launch_row_kernel(g.threads, g.p, d_desc_, g.d_map, g.n, g.lds, stream_);
        if (e != hipSuccess) return LUCAS_ERR_DEVICE;
    }
    return LUCAS_OK;
}

void Engine::destroy_graphs() {
    for (int l = 0; l < kGraphSlots; ++l) {
        if (gexec_[l]) hipGraphExecDestroy(gexec_[l]);
        if (graph_[l]) hipGraphDestroy(graph_[l]);
        gexec_[l] = nullptr;
        graph_[l] = nullptr;
        stale_[l] = false;
    }
    for (Whole& w : whole_) {
        if (w.exec) hipGraphExecDestroy(w.exec);
        if (w.graph) hipGraphDestroy(w.graph);
    }
    whole_.clear();
}

// The whole step of k iterations as one graph (head, k - 1 pairs, tail): a
// short step then costs one graph launch.  Cached per k, most recent first.
int Engine::whole_graph(long long k, hipGraphExec_t* out) {
    for (size_t i = 0; i < whole_.size(); ++i)
        if (whole_[i].k == k) {
            std::rotate(whole_.begin(), whole_.begin() + i, whole_.begin() + i + 1);
            *out = whole_[0].exec;
            return LUCAS_OK;
        }
    Whole w;
    w.k = k;
    hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return LUCAS_ERR_DEVICE;
    int rc = launch_col(LUCAS_IN_DIGITS);
    for (long long i = 1; i < k && rc == LUCAS_OK; ++i) {
        rc = launch_row();
        if (rc == LUCAS_OK) rc = launch_col(0);
    }
    if (rc == LUCAS_OK) rc = launch_row();
    if (rc == LUCAS_OK) rc = launch_col(LUCAS_OUT_DIGITS);
    e = hipStreamEndCapture(stream_, &w.graph);
    if (rc != LUCAS_OK || e != hipSuccess) {
        if (w.graph) hipGraphDestroy(w.graph);
        (void)hipGetLastError();
        return LUCAS_ERR_DEVICE;
    }
    // uploaded at once: the first launch then skips the upload
    if (hipGraphInstantiate(&w.exec, w.graph, nullptr, nullptr, 0) != hipSuccess ||
        hipGraphUpload(w.exec, stream_) != hipSuccess) {
        hipGraphDestroy(w.graph);
        return LUCAS_ERR_DEVICE;
    }
    if ((int)whole_.size() == kWholeCache) {
        hipGraphExecDestroy(whole_.back().exec);
        hipGraphDestroy(whole_.back().graph);
        whole_.pop_back();
    }
    whole_.insert(whole_.begin(), w);
    *out = w.exec;
    return LUCAS_OK;
}

// Slot l < kLevels: 2^l steady-state [row ; col] pairs; kHead: the column
// kernel reading digits; kTail: row + column kernel writing digits.
int Engine::build_graph(int slot) {
    if (gexec_[slot] && !stale_[slot]) return LUCAS_OK;
    hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return LUCAS_ERR_DEVICE;
    int rc = LUCAS_OK;
    if (slot == kHead) {
        rc = launch_col(LUCAS_IN_DIGITS);
    } else if (slot == kTail) {
        rc = launch_row();
        if (rc == LUCAS_OK) rc = launch_col(LUCAS_OUT_DIGITS);
    } else {
        for (int i = 0; i < (1 << slot) && rc == LUCAS_OK; ++i) {
            rc = launch_row();
            if (rc == LUCAS_OK) rc = launch_col(0);
        }
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(stream_, &g);
    if (rc != LUCAS_OK || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        (void)hipGetLastError();
        return LUCAS_ERR_DEVICE;
    }
    if (gexec_[slot]) {
        // a stale executable graph: the same kernels with new grid sizes and
        // arguments update in place; another kernel sequence is rebuilt
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        const hipError_t u = hipGraphExecUpdate(gexec_[slot], g, &bad, &res);
        stale_[slot] = false;
        if (u == hipSuccess && res == hipGraphExecUpdateSuccess) {
            if (graph_[slot]) hipGraphDestroy(graph_[slot]);
            graph_[slot] = g;
            return LUCAS_OK;
        }
        (void)hipGetLastError();
        hipGraphExecDestroy(gexec_[slot]);
        gexec_[slot] = nullptr;
    }
    if (graph_[slot]) hipGraphDestroy(graph_[slot]);
    graph_[slot] = g;
    e = hipGraphInstantiate(&gexec_[slot], graph_[slot], nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipGraphUpload(gexec_[slot], stream_);
    if (e != hipSuccess) {
        hipGraphDestroy(graph_[slot]);
        graph_[slot] = nullptr;
        gexec_[slot] = nullptr;
        return LUCAS_ERR_DEVICE;
    }
    return LUCAS_OK;
}

// Graph slots a step of `iters` iterations launches, in order.
static void step_plan(long long iters, std::vector<int>* slots) {
    slots->clear();
    slots->push_back(Engine::kHead);
    long long rem = iters - 1;
    // the largest graph a step replays: 128 pairs in a long step, 32 in a
    // shorter one (a server's chunks, whose job mix changes often, so that a
    // mix change rebuilds graphs of at most 63 iterations in all)
    const int top = iters >= kLongStep ? kLevels - 1 : 5;
    for (; rem >= (1LL << top); rem -= 1LL << top) slots->push_back(top);
    for (int l = top - 1; l >= 0; --l)
        if (rem >= (1LL << l)) {
            slots->push_back(l);
            rem -= 1LL << l;
        }
    slots->push_back(Engine::kTail);
}

int Engine::prepare_graphs(long long iters) {
    if (iters <= 0) return iters == 0 ? LUCAS_OK : LUCAS_ERR_INVALID_ARG;
    if (jobs_.empty()) return LUCAS_OK;
    hipSetDevice(device_);
    int rc = prepare();
    if (rc == LUCAS_OK && single_ && iters <= kWholeMax) {
        hipGraphExec_t x = nullptr;
        return whole_graph(iters, &x);
    }
    std::vector<int> slots;
    step_plan(iters, &slots);
    for (size_t i = 0; i < slots.size() && rc == LUCAS_OK; ++i) rc = build_graph(slots[i]);
    return rc;
}

int Engine::step(long long iters) {
    if (iters <= 0) return iters == 0 ? LUCAS_OK : LUCAS_ERR_INVALID_ARG;
    if (jobs_.empty()) return LUCAS_OK;
    hipSetDevice(device_);
    int rc = prepare();
    if (rc == LUCAS_OK && single_ && iters <= kWholeMax) {
        // one job, short step: the whole step is one graph launch
        hipGraphExec_t x = nullptr;
        rc = whole_graph(iters, &x);
        if (rc == LUCAS_OK && hipGraphLaunch(x, stream_) != hipSuccess) rc = LUCAS_ERR_DEVICE;
        if (rc == LUCAS_OK)
            for (Job* j : jobs_) j->iter += iters;
        return rc;
    }
    std::vector<int> slots;
    step_plan(iters, &slots);
    for (size_t i = 0; i < slots.size() && rc == LUCAS_OK; ++i) {
        rc = build_graph(slots[i]);
        if (rc == LUCAS_OK && hipGraphLaunch(gexec_[slots[i]], stream_) != hipSuccess) rc = LUCAS_ERR_DEVICE;
    }
    if (rc == LUCAS_OK)
        for (Job* j : jobs_) j->iter += iters;
    return rc;
}

int Engine::sync() {
    hipSetDevice(device_);
    return hipStreamSynchronize(stream_) == hipSuccess ? LUCAS_OK : LUCAS_ERR_DEVICE;
}

int Engine::time_kernels(int iters, double* col_us, double* row_us) {
    if (iters <= 0 || iters > 4096 || jobs_.empty()) return LUCAS_ERR_INVALID_ARG;
    hipSetDevice(device_);
    int rc = prepare();
    if (rc != LUCAS_OK) return rc;
    // The iteration sequence row, col, row, col, ... (each kernel reads what
    // the other just wrote, as in step()), with an event after every launch:
    // interval 2i -> 2i+1 is row launch i, 2i+1 -> 2i+2 col launch i.
    std::vector<hipEvent_t> ev(2 * (size_t)iters + 1, nullptr);
    auto cleanup = [&]() {
        for (hipEvent_t e : ev)
            if (e) hipEventDestroy(e);
    };
    for (hipEvent_t& e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            cleanup();
            return LUCAS_ERR_DEVICE;
        }
    rc = launch_col(LUCAS_IN_DIGITS);  // valid spectra; digits untouched
    if (rc == LUCAS_OK && hipEventRecord(ev[0], stream_) != hipSuccess) rc = LUCAS_ERR_DEVICE;
    for (int i = 0; i < iters && rc == LUCAS_OK; ++i) {
        rc = launch_row();
        if (rc == LUCAS_OK && hipEventRecord(ev[2 * i + 1], stream_) != hipSuccess) rc = LUCAS_ERR_DEVICE;
        if (rc == LUCAS_OK) rc = launch_col(0);
        if (rc == LUCAS_OK && hipEventRecord(ev[2 * i + 2], stream_) != hipSuccess) rc = LUCAS_ERR_DEVICE;
    }
    if (rc == LUCAS_OK && hipEventSynchronize(ev[2 * (size_t)iters]) != hipSuccess) rc = LUCAS_ERR_DEVICE;
    double sum_row = 0.0, sum_col = 0.0;
    for (int i = 0; i < iters && rc == LUCAS_OK; ++i) {
        float a = 0.f, b = 0.f;
        if (hipEventElapsedTime(&a, ev[2 * i], ev[2 * i + 1]) != hipSuccess ||
            hipEventElapsedTime(&b, ev[2 * i + 1], ev[2 * i + 2]) != hipSuccess)
            rc = LUCAS_ERR_DEVICE;
        sum_row += a;
        sum_col += b;
    }
    cleanup();
    if (rc != LUCAS_OK) return rc;
    *row_us = 1000.0 * sum_row / iters;
    *col_us = 1000.0 * sum_col / iters;
    return LUCAS_OK;
}

int Engine::phase_stamps(int iters, unsigned long long* col, unsigned long long* row, int max_wgs, int* ncol,
                         int* nrow) {
    if (!phase_stamps_built()) return LUCAS_ERR_UNSUPPORTED;
    if (iters <= 0 || iters > 4096 || jobs_.size() != 1) return LUCAS_ERR_INVALID_ARG;
    hipSetDevice(device_);
    int rc = prepare();
    if (rc != LUCAS_OK) return rc;
    int nc = 0, nr = 0;
    for (const auto& kv : colg_) nc += kv.second.n;
    for (const auto& kv : rowg_) nr += kv.second.n;
    if (nc > max_wgs || nr > max_wgs) return LUCAS_ERR_INVALID_ARG;
    rc = launch_col(LUCAS_IN_DIGITS);  // valid spectra; digits untouched
    for (int i = 0; i < iters && rc == LUCAS_OK; ++i) {
        rc = launch_row();
        if (rc == LUCAS_OK) rc = launch_col(0);
    }
    if (rc == LUCAS_OK && hipStreamSynchronize(stream_) != hipSuccess) rc = LUCAS_ERR_DEVICE;
    if (rc == LUCAS_OK && (jobs_[0]->plan.line ? read_line_phase_stamps(col, nc, row, nr)
                                               : read_phase_stamps(col, nc, row, nr)) != hipSuccess)
        rc = LUCAS_ERR_DEVICE;
    if (rc != LUCAS_OK) return rc;
    *ncol = nc;
    *nrow = nr;
    return LUCAS_OK;
}

}  // namespace lucas
