// server.cpp -- serving loop: continuous batching of imaging jobs.
//
// One scheduler thread per server.  Each iteration it
//   1. admits queued jobs in arrival order while a slot is free and the
//      jobs' device working memory stays under the HBM cap (set-up: device
//      accumulators, grid axes, R1 tables of the job's first K pulses);
//   2. builds a round: admitted jobs in order of least remaining work, each
//      contributing the next pulse window of every (pulse group, 256-voxel
//      tile) of its image as work items, until the round's device-time budget
//      is spent;
//   3. replays the captured graph of the round's buffer [work-list copy from
//      pinned memory -> one persistent kernel per variant present], then the
//      final group sums of jobs whose last window is in the round;
//   4. waits for the previous round, copies finished images (already in
//      pinned memory: their DMA follows the round on its stream) to the
//      callers' buffers and releases their slots and memory.
// Rounds alternate between two list buffers, so the host builds round r + 1
// while the device runs round r.  Windows of one job run in pulse order on
// one stream, and every voxel of a job sums its pulses in the order of that
// job's own group split (plan_bp), so an image is the same alone or in a mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "api_internal.h"
#include "bsar_mixed.h"
#include "bsar_serve.h"

namespace bsar {
namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int64_t id = 0;
    bsar_dataset* ds = nullptr;
    int K = 0, precision = 0;
    int nx = 0, ny = 0, nz = 0;
    int64_t V = 0;
    std::vector<double> axes;     // gx | gy | gz
    double* out = nullptr;
    BpPlan pl;
    int woff = 0;                 // pulses of every group already scheduled
    int64_t footprint = 0;
    int slot = -1;
    int state = BSAR_JOB_QUEUED;
    int status = BSAR_OK;
    double t_submit = 0, t_admit = 0, t_done = 0;
    double r1_exact_frac = 0;
    void* dmem = nullptr;         // accumulators [groups][V] (16 B each) | axes
    size_t dcap = 0;
    void* stage = nullptr;        // pinned copy of the finished image
    size_t scap = 0;
    int variant = 0;
    double cost_vp = 0;           // lane instructions per voxel-pulse (estimate)
    int64_t tiles = 0;
};

// Device blocks reused by size; freed when the server is destroyed or when an
// allocation fails.
struct Pool {
    std::multimap<size_t, void*> free_;
    ~Pool() { trim(); }
    void trim()
    {
        for (auto& kv : free_) (void)hipFree(kv.second);
        free_.clear();
    }
    hipError_t get(size_t bytes, void** p, size_t* cap)
    {
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes + (1u << 20)) {
            *p = it->second;
            *cap = it->first;
            free_.erase(it);
            return hipSuccess;
        }
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            trim();
            e = hipMalloc(p, bytes);
        }
        *cap = bytes;
        return e;
    }
    void put(void* p, size_t cap) { free_.emplace(cap, p); }
};

// Pinned host blocks reused by size (finished images land here by DMA in
// stream order; the host then copies them to the callers' buffers).
struct HostPool {
    std::multimap<size_t, void*> free_;
    ~HostPool() { trim(); }
    void trim()
    {
        for (auto& kv : free_) (void)hipHostFree(kv.second);
        free_.clear();
    }
    hipError_t get(size_t bytes, void** p, size_t* cap)
    {
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes + (1u << 20)) {
            *p = it->second;
            *cap = it->first;
            free_.erase(it);
            return hipSuccess;
        }
        *cap = bytes;
        return hipHostMalloc(p, bytes, hipHostMallocDefault);
    }
    void put(void* p, size_t cap) { free_.emplace(cap, p); }
};

struct Round {
    bool valid = false;
    int buf = 0;
    std::vector<Job*> finishing;
    double units = 0;
    int64_t items = 0;
};

constexpr size_t kAxesMax = 3 * 4096 + 1;   // doubles of staged axes per slot

size_t axes_bytes(const Job* j) { return (size_t)(j->nx + j->ny + j->nz + 1) * sizeof(double); }
size_t acc_bytes(const Job* j) { return ((size_t)j->pl.groups * j->V * 16 + 255) / 256 * 256; }

}  // namespace
}  // namespace bsar

using namespace bsar;

struct bsar_server {
    int device = 0;
    bsar_server_config cfg{};
    bsar_ctx* ctx = nullptr;            // set-up work: tables, uploads, result copies
    hipStream_t cs = nullptr;           // rounds
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stopping = false;
    bool fatal = false;
    std::string err;
    std::deque<Job*> queue;             // submitted, not admitted
    std::unordered_map<int64_t, Job*> jobs;   // every uncollected job
    int64_t next_id = 1;
    int n_active = 0;
    Clock::time_point t0;
    bsar_server_stats st{};
    // scheduler-thread state
    std::vector<Job*> active;
    std::vector<int> free_slots;
    SrvJobDev* h_jobs = nullptr;        // pinned [max_jobs]
    double* h_axes = nullptr;           // pinned [max_jobs][kAxesMax]: grid axes staged for upload
    SrvJobDev* d_jobs = nullptr;
    unsigned long long* d_counts = nullptr;
    SrvList* h_list[2] = {nullptr, nullptr};
    SrvList* d_list[2] = {nullptr, nullptr};
    size_t list_cap = 0;                // items per list
    hipEvent_t ev_beg[2] = {nullptr, nullptr}, ev_end[2] = {nullptr, nullptr};
    std::map<int, hipGraphExec_t> graphs[2];   // by variant mask
    int blocks[kSrvVariants] = {0, 0, 0, 0};
    Pool pool;
    HostPool hpool;
    unsigned long long* h_counts = nullptr;   // pinned [max_jobs]: exact-R1 counts of finished jobs
    double sec_per_unit = 1.0 / 4.5e13; // calibrated from measured rounds
    std::vector<std::vector<SrvItem>> vitems = std::vector<std::vector<SrvItem>>(kSrvVariants);

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

void drop_graphs(bsar_server* s)
{
    for (auto& g : s->graphs)
        for (auto& kv : g) (void)hipGraphExecDestroy(kv.second);
    for (auto& g : s->graphs) g.clear();
}

hipError_t alloc_lists(bsar_server* s, size_t cap)
{
    drop_graphs(s);
    for (int b = 0; b < 2; ++b) {
        if (s->h_list[b]) { (void)hipHostFree(s->h_list[b]); s->h_list[b] = nullptr; }
        if (s->d_list[b]) { (void)hipFree(s->d_list[b]); s->d_list[b] = nullptr; }
    }
    const size_t bytes = sizeof(SrvList) + (cap - 1) * sizeof(SrvItem);
    for (int b = 0; b < 2; ++b) {
        hipError_t e = hipHostMalloc((void**)&s->h_list[b], bytes, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        std::memset(s->h_list[b], 0, bytes);
        e = hipMalloc((void**)&s->d_list[b], bytes);
        if (e != hipSuccess) return e;
    }
    s->list_cap = cap;
    return hipSuccess;
}

// Captures [list copy -> persistent kernel per variant in mask] for buffer b.
hipError_t get_graph(bsar_server* s, int b, int mask, hipGraphExec_t* out)
{
    auto it = s->graphs[b].find(mask);
    if (it != s->graphs[b].end()) { *out = it->second; return hipSuccess; }
    const size_t bytes = sizeof(SrvList) + (s->list_cap - 1) * sizeof(SrvItem);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(s->cs, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return e;
    hipError_t e1 = hipMemcpyAsync(s->d_list[b], s->h_list[b], bytes, hipMemcpyHostToDevice, s->cs);
    for (int v = 0; v < kSrvVariants && e1 == hipSuccess; ++v)
        if (mask & (1 << v)) e1 = launch_serve(v, s->blocks[v], s->d_jobs, s->d_list[b], s->cs);
    e = hipStreamEndCapture(s->cs, &g);
    if (e1 != hipSuccess) { if (g) (void)hipGraphDestroy(g); return e1; }
    if (e != hipSuccess) return e;
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiateWithFlags(&x, g, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return e;
    s->graphs[b][mask] = x;
    *out = x;
    return hipSuccess;
}

// Publishes a job's end.  Admitted jobs (slot >= 0) give back their share of
// the cap.
void notify_done(bsar_server* s, Job* j, int state, int status)
{
    std::lock_guard<std::mutex> lk(s->mu);
    j->state = state;
    j->status = status;
    j->t_done = s->now();
    if (j->slot >= 0) {
        --s->n_active;
        s->st.hbm_in_use -= j->footprint;
    }
    if (state == BSAR_JOB_DONE) {
        ++s->st.completed;
        s->st.voxel_pulses += (double)j->V * j->K;
    } else {
        ++s->st.failed;
    }
    s->cv_done.notify_all();
}

void release(bsar_server* s, Job* j)
{
    if (j->dmem) s->pool.put(j->dmem, j->dcap);
    j->dmem = nullptr;
    if (j->stage) s->hpool.put(j->stage, j->scap);
    j->stage = nullptr;
    if (j->slot >= 0) s->free_slots.push_back(j->slot);
    s->active.erase(std::remove(s->active.begin(), s->active.end(), j), s->active.end());
}

// j may be collected (deleted) by a waiter as soon as notify_done returns:
// release first, touch nothing after.
void fail_job(bsar_server* s, Job* j, int status, const std::string& msg)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = msg;
    }
    release(s, j);
    notify_done(s, j, BSAR_JOB_FAILED, status);
}

// Device set-up of an admitted job.  Returns a status; on failure the caller
// fails the job.
int setup(bsar_server* s, Job* j)
{
    bsar_ctx* c = s->ctx;
    const size_t ab = acc_bytes(j);
    BSAR_CHECK_HIP(c, s->pool.get(ab + axes_bytes(j), &j->dmem, &j->dcap));
    double* dg = (double*)((char*)j->dmem + ab);
    if (j->axes.size() <= kAxesMax) {
        // through the slot's pinned staging area, in round order: no host wait
        double* h = s->h_axes + (size_t)j->slot * kAxesMax;
        std::memcpy(h, j->axes.data(), j->axes.size() * 8);
        BSAR_CHECK_HIP(c, hipMemcpyAsync(dg, h, j->axes.size() * 8, hipMemcpyHostToDevice, s->cs));
    } else {
        BSAR_CHECK_HIP(c, hipMemcpyAsync(dg, j->axes.data(), j->axes.size() * 8, hipMemcpyHostToDevice, c->stream));
        BSAR_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    }
    const bool mixed = j->precision == BSAR_MIXED;
    const int64_t nacc = (int64_t)j->pl.groups * j->V;
    if (mixed) BSAR_CHECK_HIP(c, fill_zero4((float4*)j->dmem, nacc, s->cs));
    else BSAR_CHECK_HIP(c, fill_zero((double2*)j->dmem, nacc, s->cs));
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s->cs, s->d_counts + j->slot, 0ull);
    BSAR_CHECK_HIP(c, hipGetLastError());
    SrvJobDev J{};
    J.a = make_bp_args(j->ds, j->K, dg, j->nx, j->ny, j->nz, j->pl);
    if (mixed) {
        int st = r1_tables(c, j->ds, j->K, j->axes.data(), j->nx, j->axes.data() + j->nx, j->ny,
                           j->axes.data() + j->nx + j->ny, j->nz, &J.m.tab, &J.m.ftx, &J.m.frx);
        if (st) return st;
        J.m.acc = (float4*)j->dmem;
        J.m.exact_count = s->d_counts + j->slot;
        J.m.two_pi = 2.0 * M_PI;
        J.m.inv_two_pi = 1.0 / (2.0 * M_PI);
    } else {
        J.a.acc = (double2*)j->dmem;
    }
    j->variant = 2 * (mixed ? 1 : 0) + (dem_lds_slots(j->ds->dem) <= kSrvLdsDem ? 1 : 0);
    j->cost_vp = mixed ? bp_cost_mixed() : bp_cost_fp64(j->K);
    j->tiles = (j->V + 255) / 256;
    s->h_jobs[j->slot] = J;
    BSAR_CHECK_HIP(c, hipMemcpyAsync(s->d_jobs + j->slot, s->h_jobs + j->slot, sizeof(SrvJobDev),
                                     hipMemcpyHostToDevice, s->cs));
    return BSAR_OK;
}

// Items of one window of job j (all its non-empty groups x tiles).
int64_t window_items(const Job* j)
{
    int64_t g_nonempty = 0;
    for (int g = 0; g < j->pl.groups; ++g)
        if ((int64_t)g * j->pl.glen + j->woff < std::min<int64_t>((int64_t)(g + 1) * j->pl.glen, j->K)) ++g_nonempty;
    return g_nonempty * j->tiles;
}

// Admission.  Returns false when the thread should exit.
bool admit(bsar_server* s, bool in_flight)
{
    std::vector<Job*> fresh;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        for (;;) {
            if (!s->queue.empty() && !s->free_slots.empty()) break;
            if (!s->active.empty() || in_flight) break;
            if (s->stopping && s->queue.empty()) return false;
            s->cv_work.wait(lk);
        }
        const int64_t cap = s->cfg.hbm_cap_bytes;
        while (!s->queue.empty() && !s->free_slots.empty()) {
            Job* j = s->queue.front();
            if (s->st.hbm_in_use + j->footprint > cap) break;    // arrival order: no overtaking
            s->queue.pop_front();
            j->slot = s->free_slots.back();
            s->free_slots.pop_back();
            j->state = BSAR_JOB_RUNNING;
            j->t_admit = s->now();
            s->st.hbm_in_use += j->footprint;
            s->st.hbm_peak = std::max(s->st.hbm_peak, s->st.hbm_in_use);
            ++s->n_active;
            fresh.push_back(j);
        }
    }
    for (Job* j : fresh) {
        int st = setup(s, j);
        if (st) {
            fail_job(s, j, st, s->ctx->err);
            continue;
        }
        s->active.push_back(j);
    }
    return true;
}

void build_round(bsar_server* s, Round& r)
{
    for (auto& v : s->vitems) v.clear();
    std::vector<Job*> cand;
    for (Job* j : s->active)
        if (j->woff < j->pl.glen) cand.push_back(j);
    auto rem = [](const Job* j) { return (double)(j->pl.glen - j->woff) * j->pl.groups * j->pl.nvox_pad * j->cost_vp; };
    std::sort(cand.begin(), cand.end(), [&](const Job* a, const Job* b) {
        const double ra = rem(a), rb = rem(b);
        return ra != rb ? ra < rb : a->id < b->id;
    });
    double budget = s->cfg.round_seconds / s->sec_per_unit;
    int64_t nitems = 0;
    for (Job* j : cand) {
        const double per_pulse = (double)j->pl.groups * j->pl.nvox_pad * j->cost_vp;
        if (nitems > 0 && budget < per_pulse) continue;
        const int64_t need = window_items(j);
        if (nitems + need > (int64_t)s->list_cap) continue;
        // items of about 1/8 of a round each keep the end-of-round tail short
        const double item_units = 256.0 * j->cost_vp * s->blocks[j->variant];
        const int wmax = (int)std::max(1.0, std::floor(budget > 0 ? s->cfg.round_seconds / (8.0 * item_units * s->sec_per_unit) : 1.0));
        int w = std::min(j->pl.glen - j->woff, wmax);
        w = std::min(w, (int)std::max(1.0, std::floor(budget / per_pulse)));
        auto& out = s->vitems[j->variant];
        for (int g = 0; g < j->pl.groups; ++g) {
            const int gend = (int)std::min<int64_t>((int64_t)(g + 1) * j->pl.glen, j->K);
            const int p0 = g * j->pl.glen + j->woff;
            const int p1 = std::min(p0 + w, gend);
            if (p0 >= p1) continue;
            for (int64_t t = 0; t < j->tiles; ++t) {
                SrvItem it{};
                it.slot = j->slot;
                it.g = g;
                it.vox0 = t * 256;
                it.nvox = (int)std::min<int64_t>(256, j->V - t * 256);
                it.p0 = p0;
                it.p1 = p1;
                out.push_back(it);
            }
        }
        nitems += need;
        j->woff += w;
        budget -= w * per_pulse;
        r.units += w * per_pulse;
        if (j->woff >= j->pl.glen) r.finishing.push_back(j);
    }
    r.items = nitems;
}

int launch_round(bsar_server* s, Round& r)
{
    bsar_ctx* c = s->ctx;
    SrvList* L = s->h_list[r.buf];
    unsigned pos = 0;
    int mask = 0;
    for (int v = 0; v < kSrvVariants; ++v) {
        const auto& items = s->vitems[v];
        L->begin[v] = pos;
        L->next[v] = pos;
        if (!items.empty()) {
            std::memcpy(&L->items[pos], items.data(), items.size() * sizeof(SrvItem));
            mask |= 1 << v;
        }
        pos += (unsigned)items.size();
        L->end[v] = pos;
    }
    hipGraphExec_t g = nullptr;
    BSAR_CHECK_HIP(c, get_graph(s, r.buf, mask, &g));
    BSAR_CHECK_HIP(c, hipEventRecord(s->ev_beg[r.buf], s->cs));
    BSAR_CHECK_HIP(c, hipGraphLaunch(g, s->cs));
    // finished images: group sums, then DMA of the image and its exact-R1
    // count to pinned memory, in stream order after the round
    for (Job* j : r.finishing) {
        if (j->precision == BSAR_MIXED)
            BSAR_CHECK_HIP(c, finish_mixed((const float4*)j->dmem, j->pl.groups, j->V, (double2*)j->dmem, s->cs));
        else
            BSAR_CHECK_HIP(c, finish_fp64((const double2*)j->dmem, j->pl.groups, j->V, (double2*)j->dmem, s->cs));
        BSAR_CHECK_HIP(c, s->hpool.get((size_t)j->V * 16, &j->stage, &j->scap));
        BSAR_CHECK_HIP(c, hipMemcpyAsync(j->stage, j->dmem, (size_t)j->V * 16, hipMemcpyDeviceToHost, s->cs));
        BSAR_CHECK_HIP(c, hipMemcpyAsync(s->h_counts + j->slot, s->d_counts + j->slot, sizeof(unsigned long long),
                                         hipMemcpyDeviceToHost, s->cs));
    }
    BSAR_CHECK_HIP(c, hipEventRecord(s->ev_end[r.buf], s->cs));
    return BSAR_OK;
}

int complete_round(bsar_server* s, Round& r)
{
    bsar_ctx* c = s->ctx;
    BSAR_CHECK_HIP(c, hipEventSynchronize(s->ev_end[r.buf]));
    float ms = 0.f;
    BSAR_CHECK_HIP(c, hipEventElapsedTime(&ms, s->ev_beg[r.buf], s->ev_end[r.buf]));
    const double sec = ms * 1e-3;
    if (sec > 2e-3 && r.units > 0) {
        const double x = sec / r.units;
        s->sec_per_unit = 0.8 * s->sec_per_unit + 0.2 * x;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->st.rounds += 1;
        s->st.items += r.items;
        s->st.device_seconds += sec;
    }
    for (Job* j : r.finishing) {
        const unsigned long long exact = s->h_counts[j->slot];
        if (j->V > 0) std::memcpy(j->out, j->stage, (size_t)j->V * 16);
        s->hpool.put(j->stage, j->scap);
        j->stage = nullptr;
        j->r1_exact_frac = j->precision == BSAR_MIXED ? (double)exact / (2.0 * (double)j->V * j->K) : 1.0;
        release(s, j);
        notify_done(s, j, BSAR_JOB_DONE, BSAR_OK);
    }
    return BSAR_OK;
}

// Makes the work lists hold one window of every admitted job.
int fit_lists(bsar_server* s, bool in_flight_prev)
{
    int64_t need = 0;
    for (Job* j : s->active) need = std::max<int64_t>(need, window_items(j));
    if (need <= (int64_t)s->list_cap) return BSAR_OK;
    (void)in_flight_prev;
    BSAR_CHECK_HIP(s->ctx, hipStreamSynchronize(s->cs));   // lists are not in use
    size_t cap = s->list_cap;
    while ((int64_t)cap < need) cap *= 2;
    BSAR_CHECK_HIP(s->ctx, alloc_lists(s, cap));
    return BSAR_OK;
}

void fail_everything(bsar_server* s, const std::string& msg)
{
    std::vector<Job*> all(s->active);
    for (Job* j : all) fail_job(s, j, BSAR_ERR_DEVICE, msg);
    std::deque<Job*> q;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->fatal = true;
        s->err = msg;
        q.swap(s->queue);
    }
    for (Job* j : q) notify_done(s, j, BSAR_JOB_FAILED, BSAR_ERR_DEVICE);
}

void run(bsar_server* s)
{
    (void)hipSetDevice(s->device);
    Round prev;
    int r = 0;
    for (;;) {
        if (!admit(s, prev.valid)) break;
        int st = fit_lists(s, prev.valid);
        Round cur;
        cur.buf = r & 1;
        if (st == BSAR_OK) {
            build_round(s, cur);
            if (cur.items > 0) {
                st = launch_round(s, cur);
                cur.valid = true;
                ++r;
            }
        }
        if (st == BSAR_OK && prev.valid) st = complete_round(s, prev);
        if (st != BSAR_OK) {
            fail_everything(s, s->ctx->err);
            break;
        }
        prev = cur;
    }
    // drain what is left on the device before the thread ends
    (void)hipStreamSynchronize(s->cs);
    bool left = false;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        left = !s->queue.empty();
    }
    if (left || !s->active.empty()) fail_everything(s, "server stopped");
}

void fill_info(const Job* j, bsar_job_info* info)
{
    if (!info) return;
    std::memset(info, 0, sizeof *info);
    info->id = j->id;
    info->state = j->state;
    info->status = j->status;
    info->t_submit = j->t_submit;
    info->t_admit = j->t_admit;
    info->t_done = j->t_done;
    info->voxel_pulses = (double)j->V * j->K;
    info->r1_exact_frac = j->r1_exact_frac;
    info->footprint_bytes = j->footprint;
}

void destroy_server(bsar_server* s)
{
    (void)hipSetDevice(s->device);
    if (s->cs) (void)hipStreamSynchronize(s->cs);
    drop_graphs(s);
    for (int b = 0; b < 2; ++b) {
        if (s->h_list[b]) (void)hipHostFree(s->h_list[b]);
        if (s->d_list[b]) (void)hipFree(s->d_list[b]);
        if (s->ev_beg[b]) (void)hipEventDestroy(s->ev_beg[b]);
        if (s->ev_end[b]) (void)hipEventDestroy(s->ev_end[b]);
    }
    if (s->h_jobs) (void)hipHostFree(s->h_jobs);
    if (s->h_axes) (void)hipHostFree(s->h_axes);
    if (s->d_jobs) (void)hipFree(s->d_jobs);
    if (s->d_counts) (void)hipFree(s->d_counts);
    if (s->h_counts) (void)hipHostFree(s->h_counts);
    s->pool.trim();
    s->hpool.trim();
    if (s->cs) (void)hipStreamDestroy(s->cs);
    if (s->ctx) bsar_ctx_destroy(s->ctx);
    for (auto& kv : s->jobs) delete kv.second;
    delete s;
}

}  // namespace

extern "C" {

int bsar_server_create(int device, const bsar_server_config* cfg, bsar_server** out)
{
    if (!out) return BSAR_ERR_INVALID;
    *out = nullptr;
    bsar_server* s = new (std::nothrow) bsar_server();
    if (!s) return BSAR_ERR_NOMEM;
    s->t0 = Clock::now();
    if (cfg) s->cfg = *cfg;
    if (s->cfg.max_jobs <= 0) s->cfg.max_jobs = 256;
    if (s->cfg.hbm_cap_bytes <= 0) s->cfg.hbm_cap_bytes = (int64_t)200e9;
    if (!(s->cfg.round_seconds > 0)) s->cfg.round_seconds = 0.04;
    s->device = device;
    int st = bsar_ctx_create(device, &s->ctx);
    if (st) { delete s; return st; }
    auto fail = [&](hipError_t e) {
        const int code = e == hipErrorOutOfMemory ? BSAR_ERR_NOMEM : BSAR_ERR_DEVICE;
        destroy_server(s);
        return code;
    };
    hipError_t e = hipStreamCreateWithFlags(&s->cs, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(e);
    for (int b = 0; b < 2; ++b) {
        if ((e = hipEventCreate(&s->ev_beg[b])) != hipSuccess) return fail(e);
        if ((e = hipEventCreate(&s->ev_end[b])) != hipSuccess) return fail(e);
    }
    const int M = s->cfg.max_jobs;
    if ((e = hipHostMalloc((void**)&s->h_jobs, (size_t)M * sizeof(SrvJobDev), hipHostMallocDefault)) != hipSuccess)
        return fail(e);
    std::memset((void*)s->h_jobs, 0, (size_t)M * sizeof(SrvJobDev));
    if ((e = hipHostMalloc((void**)&s->h_axes, (size_t)M * kAxesMax * sizeof(double), hipHostMallocDefault)) !=
        hipSuccess)
        return fail(e);
    if ((e = hipMalloc((void**)&s->d_jobs, (size_t)M * sizeof(SrvJobDev))) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void**)&s->d_counts, (size_t)M * sizeof(unsigned long long))) != hipSuccess) return fail(e);
    if ((e = hipHostMalloc((void**)&s->h_counts, (size_t)M * sizeof(unsigned long long), hipHostMallocDefault)) !=
        hipSuccess)
        return fail(e);
    if ((e = alloc_lists(s, kSrvListCap)) != hipSuccess) return fail(e);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    for (int v = 0; v < kSrvVariants; ++v) {
        int nb = 0;
        const size_t lds = (v & 1) ? (size_t)kSrvLdsDem * sizeof(double) : 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, serve_kernel(v), 256, lds) != hipSuccess || nb <= 0) {
            (void)hipGetLastError();
            nb = 2;
        }
        s->blocks[v] = std::min(nb, 8) * cus;
    }
    for (int k = M - 1; k >= 0; --k) s->free_slots.push_back(k);
    try {
        s->th = std::thread(run, s);
    } catch (...) {
        destroy_server(s);
        return BSAR_ERR_NOMEM;
    }
    *out = s;
    return BSAR_OK;
}

void bsar_server_destroy(bsar_server* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    destroy_server(s);
}

const char* bsar_server_last_error(const bsar_server* s) { return s ? s->err.c_str() : "null server"; }

double bsar_server_clock(const bsar_server* s) { return s ? s->now() : 0.0; }

int bsar_server_submit(bsar_server* s, const bsar_job_desc* d, int64_t* id)
{
    if (!s || !d || !id) return BSAR_ERR_INVALID;
    *id = 0;
    auto bad = [&](int code, const char* msg) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = msg;
        if (code == BSAR_ERR_CAPACITY) ++s->st.rejected;
        return code;
    };
    const bsar_dataset* ds = d->ds;
    if (!ds) return bad(BSAR_ERR_INVALID, "job has no dataset");
    if (ds->device != s->device) return bad(BSAR_ERR_INVALID, "dataset lives on another device");
    if (!rcomp_ready(ds)) return bad(BSAR_ERR_STATE, "no range-compressed data loaded");
    if (d->precision != BSAR_FP64 && d->precision != BSAR_MIXED) return bad(BSAR_ERR_INVALID, "unknown precision");
    const int K = d->pulses == 0 ? ds->n : d->pulses;
    if (K < 1 || K > ds->n) return bad(BSAR_ERR_INVALID, "pulses must be in [1, dataset pulses]");
    if (d->nx < 0 || d->ny < 0 || d->nz < 0) return bad(BSAR_ERR_INVALID, "negative grid size");
    if ((d->nx && !d->gx) || (d->ny && !d->gy) || (d->nz && !d->gz)) return bad(BSAR_ERR_INVALID, "grid axis is null");
    const int64_t V = (int64_t)d->nx * d->ny * d->nz;
    if (V > 0 && !d->out) return bad(BSAR_ERR_INVALID, "out is null");
    if (V > (int64_t(1) << 31)) return bad(BSAR_ERR_INVALID, "grid too large");
    std::unique_ptr<Job> j(new (std::nothrow) Job());
    if (!j) return bad(BSAR_ERR_NOMEM, "host allocation");
    j->ds = const_cast<bsar_dataset*>(ds);
    j->K = K;
    j->precision = d->precision;
    j->nx = d->nx; j->ny = d->ny; j->nz = d->nz;
    j->V = V;
    j->axes.reserve((size_t)d->nx + d->ny + d->nz + 1);
    j->axes.insert(j->axes.end(), d->gx, d->gx + d->nx);
    j->axes.insert(j->axes.end(), d->gy, d->gy + d->ny);
    j->axes.insert(j->axes.end(), d->gz, d->gz + d->nz);
    j->axes.push_back(0.0);
    j->out = d->out;
    j->pl = plan_bp(V, K, d->precision == BSAR_MIXED ? bp_cost_mixed() : bp_cost_fp64(K));
    j->footprint = (int64_t)(acc_bytes(j.get()) + axes_bytes(j.get()));
    if (j->footprint > s->cfg.hbm_cap_bytes) return bad(BSAR_ERR_CAPACITY, "job working memory exceeds the HBM cap");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal || s->stopping) {
        s->err = "server is not accepting jobs";
        return BSAR_ERR_STATE;
    }
    j->id = s->next_id++;
    j->t_submit = s->now();
    ++s->st.submitted;
    Job* p = j.release();
    s->jobs[p->id] = p;
    *id = p->id;
    if (V == 0) {                       // nothing to image
        p->state = BSAR_JOB_DONE;
        p->t_admit = p->t_done = p->t_submit;
        ++s->st.completed;
        s->cv_done.notify_all();
        return BSAR_OK;
    }
    s->queue.push_back(p);
    s->cv_work.notify_one();
    return BSAR_OK;
}

int bsar_server_poll(bsar_server* s, int64_t id, bsar_job_info* info)
{
    if (!s) return BSAR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) { s->err = "unknown job"; return BSAR_ERR_INVALID; }
    fill_info(it->second, info);
    return BSAR_OK;
}

int bsar_server_wait(bsar_server* s, int64_t id, double timeout_s, bsar_job_info* info)
{
    if (!s) return BSAR_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) { s->err = "unknown job"; return BSAR_ERR_INVALID; }
    Job* j = it->second;
    auto done = [&] { return j->state == BSAR_JOB_DONE || j->state == BSAR_JOB_FAILED; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        fill_info(j, info);
        return BSAR_ERR_STATE;
    }
    fill_info(j, info);
    const int status = j->status;
    s->jobs.erase(id);
    delete j;
    return status;
}

int bsar_server_wait_any(bsar_server* s, double timeout_s, bsar_job_info* info)
{
    if (!s || !info) return BSAR_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->jobs.empty()) { s->err = "no uncollected job"; return BSAR_ERR_INVALID; }
    Job* found = nullptr;
    auto any = [&] {
        for (auto& kv : s->jobs)
            if (kv.second->state == BSAR_JOB_DONE || kv.second->state == BSAR_JOB_FAILED) {
                if (!found || kv.second->t_done < found->t_done) found = kv.second;
            }
        return found != nullptr;
    };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, any);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), any)) {
        std::memset(info, 0, sizeof *info);
        return BSAR_ERR_STATE;
    }
    fill_info(found, info);
    const int status = found->status;
    s->jobs.erase(found->id);
    delete found;
    return status;
}

int bsar_server_drain(bsar_server* s)
{
    if (!s) return BSAR_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->n_active == 0; });
    return s->fatal ? BSAR_ERR_DEVICE : BSAR_OK;
}

int bsar_server_get_stats(bsar_server* s, bsar_server_stats* st)
{
    if (!s || !st) return BSAR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->st;
    st->active = s->n_active;
    st->queued = (int)s->queue.size();
    return BSAR_OK;
}

}  // extern "C"
