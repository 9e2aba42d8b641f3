// hfm_server.hip -- continuous-batching solve server (C ABI hfm_server_*).
//
// Requests (Riemann2 / Riemann3 solves of any size) are admitted as HBM budget and resident
// slots allow.  A setup thread builds each admitted solve's stencils (FP64 Selling) and tile
// dependency graph on its own stream while a compute thread runs rounds of the band-ordered block
// fast-iterative method over the union of all resident solves: one multi-solve scheduling kernel
// and one multi-solve update kernel per round, so tiles of every resident solve share each
// launch.  A solve joins at the next batch boundary once its setup is done and leaves when a
// round lists none of its tiles; its values are exported and copied to the caller, and its
// memory is returned to the budget.  Each solve keeps its own keys, band bound and pending
// minimum, so its fixed point does not depend on the other solves in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hfm_internal.h"
#include "hfm_kernels.h"
#include "hfm_tile.h"
#include "selling.h"

namespace hfm {

constexpr int SOLVE_SHIFT = 20;  // list entry = solve << SOLVE_SHIFT | tile
constexpr int MAX_SLOTS = 512;   // descriptor table size (hard cap on resident solves)

struct SolveDesc {
    TileCtx<float> ctx;
    int model;
    int active;
    float delta;                  // initial band width
    int target;                   // tiles per round the band width is steered to
    unsigned *pend;               // 2 * PEND_SLOTS smallest pending keys (alternating rounds)
    float *bound;                 // 2 band bounds + 2 band widths
    int *count;                   // 2 tiles listed per round
    unsigned long long *stats;    // PEND_SLOTS pairs (visits, sweeps)
};

// Scheduling round over every resident solve: grid.y = descriptor index, grid.x = tile blocks.
__global__ __launch_bounds__(256) void k_mschedule(const SolveDesc *__restrict__ descs, int parity,
                                                   int *__restrict__ list, int *__restrict__ list_count) {
    const int s = blockIdx.y;
    const SolveDesc &d = descs[s];
    if (!d.active) return;
    const int nt = d.ctx.g.ntiles;
    if ((int)(blockIdx.x * blockDim.x) >= nt) return;
    float delta = d.delta;
    if (delta > 0.f)
        delta = adapt_delta(d.bound[2 + (parity ^ 1)], d.count[parity ^ 1], d.target, 0.5f * d.delta, 64.f * d.delta);
    const float B = band_bound(key_value(read_pend(d.pend + parity * PEND_SLOTS)), d.bound[parity ^ 1], delta);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d.bound[parity] = B;
        d.bound[2 + parity] = delta;
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    unsigned k = KEY_NONE;
    if (t < nt) k = d.ctx.keys[t];
    const bool take = band_take(k, B);
    const unsigned left = (!take && k != KEY_NONE) ? k : KEY_NONE;
    if (take) d.ctx.keys[t] = KEY_NONE;
    const unsigned lmin = wave_min_u32(left);
    if (lane == 0 && lmin != KEY_NONE) atomicMin(&d.pend[(parity ^ 1) * PEND_SLOTS + ((t >> 6) & (PEND_SLOTS - 1))], lmin);
    const unsigned long long mask = __ballot(take);
    if (mask == 0) return;
    int base = 0;
    if (lane == 0) {
        base = atomicAdd(list_count, __popcll(mask));
        atomicAdd(&d.count[parity], __popcll(mask));
    }
    base = __shfl(base, 0);
    if (take) {
        const unsigned long long below = lane == 0 ? 0ull : (mask & (~0ull >> (64 - lane)));
        list[base + __popcll(below)] = (s << SOLVE_SHIFT) | t;
    }
}

// Update round over the packed list of (solve, tile) entries of every resident solve.
__global__ __launch_bounds__(256) void k_mfim(const SolveDesc *__restrict__ descs, int ndesc, int parity,
                                              const int *__restrict__ list, const int *__restrict__ list_count,
                                              int *__restrict__ list_count_reset) {
    __shared__ float su[4][SU_ROW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n = *list_count;
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) *list_count_reset = 0;
        for (int i = threadIdx.x; i < ndesc; i += blockDim.x)
            if (descs[i].active) descs[i].count[parity ^ 1] = 0;
        for (int i = threadIdx.x; i < ndesc * PEND_SLOTS; i += blockDim.x) {
            const SolveDesc &d = descs[i / PEND_SLOTS];
            if (d.active) d.pend[parity * PEND_SLOTS + (i % PEND_SLOTS)] = KEY_NONE;
        }
    }
    const int gw = blockIdx.x * 4 + wv;
    const int slot = gw & (PEND_SLOTS - 1);
    for (int i = gw; i < n; i += gridDim.x * 4) {
        const int e = __builtin_amdgcn_readfirstlane(list[i]);
        const int s = e >> SOLVE_SHIFT;
        const int tile = e & ((1 << SOLVE_SHIFT) - 1);
        const SolveDesc &d = descs[s];
        unsigned wm = KEY_NONE;
        int sw;
        if (d.model == MODEL_RIEMANN3) sw = fim_tile<MODEL_RIEMANN3, float, 32>(d.ctx, tile, lane, su[wv], wm);
        else sw = fim_tile<MODEL_RIEMANN2, float, 32>(d.ctx, tile, lane, su[wv], wm);
        if (lane == 0) {
            if (wm != KEY_NONE) atomicMin(&d.pend[(parity ^ 1) * PEND_SLOTS + slot], wm);
            atomicAdd(&d.stats[2 * slot], 1ull);
            atomicAdd(&d.stats[2 * slot + 1], (unsigned long long)sw);
        }
    }
}

}  // namespace hfm

using namespace hfm;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct RequestRec {
    int64_t ticket = 0;
    int model = 0;
    Geom g{};
    long long np = 0, N = 0;
    int K = 0;
    double h = 1;
    std::vector<double> metric_host;  // owned copy when the metric is on the host
    const double *metric_dev = nullptr;
    int metric_is_field = 1, dual = 0;
    std::vector<long long> seed_tp;
    std::vector<double> seed_v;
    double *values_out = nullptr;
    size_t bytes = 0;
    Clock::time_point t_submit, t_admit, t_ready, t_done;
    bool done = false;
    hfm_request_result res{};
};

struct Slot {
    std::shared_ptr<RequestRec> rec;
    int idx = -1;  // descriptor index
    char *arena = nullptr;
    int *deps = nullptr;
    double *metric_stage = nullptr;
    float *u = nullptr;
    uint32_t *offs = nullptr;
    float *w = nullptr;
    unsigned long long *fixedmask = nullptr;
    unsigned *keys = nullptr, *tmax = nullptr;
    int *dep_cnt = nullptr, *dep_start = nullptr, *dep_fill = nullptr;
    long long *seed_p = nullptr;
    double *seed_v = nullptr;
    double *values_dev = nullptr;
    int *status = nullptr;
    float delta = 0;
    int64_t rounds = 0;
    hipEvent_t export_done = nullptr;
};

template <typename T> T *carve(char *&p, size_t n) {
    T *r = reinterpret_cast<T *>(p);
    p += ((n * sizeof(T) + 255) / 256) * 256;
    return r;
}

size_t arena_bytes(const RequestRec &r) {
    const size_t np = r.np, nt = r.g.ntiles, K = r.K;
    size_t b = 0;
    auto add = [&](size_t n) { b += ((n + 255) / 256) * 256; };
    add(np * 4);            // u
    add(np * K * 4);        // offs
    add(np * K * 4);        // w
    add(nt * 8);            // fixedmask
    add(nt * 4);            // keys
    add(nt * 4);            // tmax
    add(nt * 4);            // dep_cnt
    add((nt + 1) * 4);      // dep_start
    add(nt * 4);            // dep_fill
    add(r.seed_tp.size() * 8);
    add(r.seed_v.size() * 8);
    add(r.N * 8);           // values_dev
    add(16);                // status
    return b;
}

}  // namespace

struct hfm_server {
    double budget = 0;
    int max_resident = 0;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done, cv_ready;
    std::deque<std::shared_ptr<RequestRec>> queue;   // submitted, not admitted
    std::deque<Slot *> ready;                         // setup done, waiting to join
    std::map<int64_t, std::shared_ptr<RequestRec>> recs;
    int64_t next_ticket = 1;
    double in_use = 0;
    int resident = 0;  // admitted and not yet released
    bool stop = false;
    hfm_server_stats stats{};

    static constexpr int NSETUP = 4;  // setup threads, one stream each
    hipStream_t s_setup[NSETUP] = {}, s_comp = nullptr;
    RedSlot *d_red[NSETUP] = {};  // per setup thread: stencil reductions (mean one-step time)
    std::thread th_setup[NSETUP], th_comp;

    // compute-thread state
    SolveDesc *d_desc = nullptr;
    SolveDesc *h_desc = nullptr;  // pinned
    int *d_list = nullptr;
    long long list_cap = 0;
    int *d_lc = nullptr;          // 2 alternating list counters
    int *d_count = nullptr;       // MAX_SLOTS * 2
    unsigned *d_pend = nullptr;   // MAX_SLOTS * 2 * PEND_SLOTS
    float *d_bound = nullptr;     // MAX_SLOTS * 2
    unsigned long long *d_stats = nullptr;  // MAX_SLOTS * 2 * PEND_SLOTS
    int *h_count = nullptr;       // pinned MAX_SLOTS * 2
    unsigned long long *h_stats = nullptr;  // pinned
    std::vector<Slot *> running;
    std::vector<Slot *> exporting;
    std::vector<int> free_idx;
    int ndesc = 0;
    long long round = 0;
};

namespace {

void fail_request(hfm_server *S, const std::shared_ptr<RequestRec> &r, const std::string &msg) {
    std::lock_guard<std::mutex> lk(S->mu);
    r->res.status = 1;
    std::snprintf(r->res.error, sizeof(r->res.error), "%s", msg.c_str());
    r->t_done = Clock::now();
    r->res.latency_ms = ms_since(r->t_submit, r->t_done);
    r->done = true;
    S->stats.failed++;
    S->cv_done.notify_all();
}

void release_budget(hfm_server *S, size_t bytes) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->in_use -= (double)bytes;
    S->resident--;
    S->cv_submit.notify_all();
}

// Build stencils and tile graph of one admitted request on the setup stream.
Slot *setup_solve(hfm_server *S, const std::shared_ptr<RequestRec> &r, hipStream_t s, RedSlot *red) {
    Slot *sl = new Slot;
    sl->rec = r;
    const size_t ab = arena_bytes(*r);
    HIPCHECK(hipMallocAsync((void **)&sl->arena, ab, s));
    char *p = sl->arena;
    sl->u = carve<float>(p, r->np);
    sl->offs = carve<uint32_t>(p, r->np * r->K);
    sl->w = carve<float>(p, r->np * r->K);
    sl->fixedmask = carve<unsigned long long>(p, r->g.ntiles);
    sl->keys = carve<unsigned>(p, r->g.ntiles);
    sl->tmax = carve<unsigned>(p, r->g.ntiles);
    sl->dep_cnt = carve<int>(p, r->g.ntiles);
    sl->dep_start = carve<int>(p, r->g.ntiles + 1);
    sl->dep_fill = carve<int>(p, r->g.ntiles);
    sl->seed_p = carve<long long>(p, r->seed_tp.size());
    sl->seed_v = carve<double>(p, r->seed_v.size());
    sl->values_dev = carve<double>(p, r->N);
    sl->status = carve<int>(p, 4);
    const double *metric = r->metric_dev;
    const size_t mcount = r->metric_is_field ? (size_t)r->N * r->K : (size_t)r->K;
    if (!metric) {
        HIPCHECK(hipMallocAsync((void **)&sl->metric_stage, mcount * sizeof(double), s));
        HIPCHECK(hipMemcpyAsync(sl->metric_stage, r->metric_host.data(), mcount * sizeof(double),
                                hipMemcpyHostToDevice, s));
        metric = sl->metric_stage;
    }
    HIPCHECK(hipMemcpyAsync(sl->seed_p, r->seed_tp.data(), r->seed_tp.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(sl->seed_v, r->seed_v.data(), r->seed_v.size() * 8, hipMemcpyHostToDevice, s));
    int init[2] = {0, 0x7fffffff};
    HIPCHECK(hipMemcpyAsync(sl->status, init, sizeof(init), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemsetAsync(red, 0, RED_SLOTS * sizeof(RedSlot), s));
    InvHH ih{};
    const int D = r->g.dim;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j <= i; ++j) ih.v[i * (i + 1) / 2 + j] = 1. / (r->h * r->h);
    HIPCHECK(launch_riemann_stencil<float>(r->g, metric, r->metric_is_field, r->dual, ih, sl->offs, sl->w, r->np,
                                           sl->status, red, s));
    StencilData<float> sd{sl->offs, sl->w, nullptr, nullptr, nullptr, r->np};
    HIPCHECK(hipMemsetAsync(sl->dep_cnt, 0, sizeof(int) * r->g.ntiles, s));
    HIPCHECK(launch_tile_deps<float>(r->model, r->g, sd, 0, sl->dep_cnt, nullptr, nullptr, nullptr, nullptr, s));
    HIPCHECK(launch_exclusive_scan(sl->dep_cnt, sl->dep_start, r->g.ntiles, s));
    int hs[2];
    std::vector<RedSlot> h_red(RED_SLOTS);
    int total = 0;
    HIPCHECK(hipMemcpyAsync(hs, sl->status, sizeof(hs), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(h_red.data(), red, RED_SLOTS * sizeof(RedSlot), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(&total, sl->dep_start + r->g.ntiles, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (hs[0] != SELLING_OK) {
        if (sl->metric_stage) HIPCHECK(hipFreeAsync(sl->metric_stage, s));
        HIPCHECK(hipFreeAsync(sl->arena, s));
        HIPCHECK(hipStreamSynchronize(s));
        delete sl;
        throw HipFail{hs[0] == SELLING_NOT_SPD ? "ObtuseSuperbase error : input must be positive definite"
                                              : "ObtuseSuperbase error : did not terminate in 200 iterations"};
    }
    HIPCHECK(hipMallocAsync((void **)&sl->deps, sizeof(int) * std::max(total, 1), s));
    HIPCHECK(hipMemsetAsync(sl->dep_fill, 0, sizeof(int) * r->g.ntiles, s));
    HIPCHECK(launch_tile_deps<float>(r->model, r->g, sd, 1, sl->dep_cnt, sl->dep_start, sl->dep_fill, sl->deps,
                                     nullptr, s));
    if (sl->metric_stage) {
        HIPCHECK(hipFreeAsync(sl->metric_stage, s));
        sl->metric_stage = nullptr;
    }
    HIPCHECK(hipStreamSynchronize(s));
    const RedTotals tot = red_total(h_red.data());
    const double tau = tot.count > 0 ? tot.sum / tot.count : 0.;
    sl->delta = (float)(8. * tau);
    return sl;
}

void setup_loop(hfm_server *S, int which) {
    if (hipSetDevice(0) != hipSuccess) return;
    hipStream_t stream = S->s_setup[which];
    while (true) {
        std::shared_ptr<RequestRec> r;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_submit.wait(lk, [&] {
                if (S->stop) return true;
                if (S->queue.empty()) return false;
                const auto &h = S->queue.front();
                return S->resident < S->max_resident && S->in_use + (double)h->bytes <= S->budget;
            });
            if (S->stop) return;
            r = S->queue.front();
            S->queue.pop_front();
            S->in_use += (double)r->bytes;
            S->resident++;
            S->stats.max_resident_seen = std::max<int64_t>(S->stats.max_resident_seen, S->resident);
        }
        r->t_admit = Clock::now();
        r->res.queue_ms = ms_since(r->t_submit, r->t_admit);
        Slot *sl = nullptr;
        try {
            sl = setup_solve(S, r, stream, S->d_red[which]);
        } catch (const HipFail &e) {
            release_budget(S, r->bytes);
            fail_request(S, r, e.msg);
            continue;
        }
        r->t_ready = Clock::now();
        r->res.setup_ms = ms_since(r->t_admit, r->t_ready);
        std::lock_guard<std::mutex> lk(S->mu);
        S->ready.push_back(sl);
        S->cv_ready.notify_all();
    }
}

void join_solve(hfm_server *S, Slot *sl) {
    const RequestRec &r = *sl->rec;
    hipStream_t s = S->s_comp;
    int idx;
    if (!S->free_idx.empty()) {
        idx = S->free_idx.back();
        S->free_idx.pop_back();
    } else {
        idx = S->ndesc++;
    }
    sl->idx = idx;
    const int parity = (int)(S->round & 1);
    unsigned *pend = S->d_pend + (size_t)idx * 2 * PEND_SLOTS;
    float *bound = S->d_bound + (size_t)idx * 4;
    int *count = S->d_count + (size_t)idx * 2;
    unsigned long long *st = S->d_stats + (size_t)idx * 2 * PEND_SLOTS;
    HIPCHECK(launch_init<float>(r.g, sl->u, sl->fixedmask, sl->keys, sl->tmax, r.np, sl->seed_p, sl->seed_v,
                                (int)r.seed_tp.size(), s));
    HIPCHECK(hipMemsetAsync(pend, 0xff, 2 * PEND_SLOTS * sizeof(unsigned), s));
    HIPCHECK(hipMemsetAsync(count, 0, 2 * sizeof(int), s));
    HIPCHECK(hipMemsetAsync(st, 0, 2 * PEND_SLOTS * sizeof(unsigned long long), s));
    const float binit[4] = {-__builtin_huge_valf(), -__builtin_huge_valf(), sl->delta, sl->delta};
    HIPCHECK(hipMemcpyAsync(bound, binit, sizeof(binit), hipMemcpyHostToDevice, s));
    HIPCHECK(launch_seed_activate(sl->seed_p, sl->seed_v, (int)r.seed_tp.size(), sl->dep_start, sl->deps, sl->keys,
                                  sl->tmax, pend + parity * PEND_SLOTS, s));
    SolveDesc &d = S->h_desc[idx];
    d.ctx = TileCtx<float>{r.g, StencilData<float>{sl->offs, sl->w, nullptr, nullptr, nullptr, r.np}, sl->u,
                           sl->fixedmask, sl->dep_start, sl->deps, sl->keys, sl->tmax};
    d.model = r.model;
    d.active = 1;
    d.delta = sl->delta;
    d.target = std::max(64, std::min(256, r.g.ntiles / 4));
    d.pend = pend;
    d.bound = bound;
    d.count = count;
    d.stats = st;
    S->running.push_back(sl);
}

void finish_export(hfm_server *S, Slot *sl) {
    RequestRec &r = *sl->rec;
    unsigned long long v = 0, sw = 0;
    const unsigned long long *hs = S->h_stats + (size_t)sl->idx * 2 * PEND_SLOTS;
    for (int i = 0; i < PEND_SLOTS; ++i) {
        v += hs[2 * i];
        sw += hs[2 * i + 1];
    }
    (void)sw;
    r.res.tile_visits = (int64_t)v;
    r.res.rounds = sl->rounds;
}

void compute_loop(hfm_server *S) {
    if (hipSetDevice(0) != hipSuccess) return;
    const int batch = 8;
    while (true) {
        // joins
        std::vector<Slot *> joining;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->running.empty() && S->exporting.empty())
                S->cv_ready.wait(lk, [&] { return S->stop || !S->ready.empty(); });
            if (S->stop && S->running.empty() && S->exporting.empty() && S->ready.empty()) return;
            while (!S->ready.empty()) {
                joining.push_back(S->ready.front());
                S->ready.pop_front();
            }
        }
        try {
            for (Slot *sl : joining) join_solve(S, sl);
            if (!S->running.empty()) {
                // list capacity: total tiles of the resident solves
                long long tiles = 0;
                int maxblocks = 0;
                for (Slot *sl : S->running) {
                    tiles += sl->rec->g.ntiles;
                    maxblocks = std::max(maxblocks, (sl->rec->g.ntiles + 255) / 256);
                }
                if (tiles > S->list_cap) {
                    if (S->d_list) HIPCHECK(hipFreeAsync(S->d_list, S->s_comp));
                    S->list_cap = std::max<long long>(tiles, 2 * S->list_cap);
                    HIPCHECK(hipMallocAsync((void **)&S->d_list, S->list_cap * sizeof(int), S->s_comp));
                }
                HIPCHECK(hipMemcpyAsync(S->d_desc, S->h_desc, sizeof(SolveDesc) * S->ndesc, hipMemcpyHostToDevice,
                                        S->s_comp));
                const unsigned grid = (unsigned)std::min<long long>((tiles + 3) / 4, 2048);
                for (int b = 0; b < batch; ++b) {
                    const int parity = (int)(S->round & 1);
                    hipLaunchKernelGGL(k_mschedule, dim3(maxblocks, S->ndesc), dim3(256), 0, S->s_comp, S->d_desc,
                                       parity, S->d_list, &S->d_lc[parity]);
                    HIPCHECK(hipGetLastError());
                    hipLaunchKernelGGL(k_mfim, dim3(grid), dim3(256), 0, S->s_comp, S->d_desc, S->ndesc, parity,
                                       S->d_list, &S->d_lc[parity], &S->d_lc[parity ^ 1]);
                    HIPCHECK(hipGetLastError());
                    S->round++;
                }
                for (Slot *sl : S->running) sl->rounds += batch;
                HIPCHECK(hipMemcpyAsync(S->h_count, S->d_count, sizeof(int) * 2 * S->ndesc, hipMemcpyDeviceToHost,
                                        S->s_comp));
                HIPCHECK(hipMemcpyAsync(S->h_stats, S->d_stats,
                                        sizeof(unsigned long long) * 2 * PEND_SLOTS * S->ndesc,
                                        hipMemcpyDeviceToHost, S->s_comp));
                HIPCHECK(hipStreamSynchronize(S->s_comp));
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    S->stats.batches++;
                    S->stats.rounds += batch;
                }
                const int last = (int)((S->round - 1) & 1);
                std::vector<Slot *> still;
                for (Slot *sl : S->running) {
                    const bool finished = S->h_count[2 * sl->idx + last] == 0;
                    const bool runaway = sl->rounds > 8LL * sl->rec->g.ntiles + 65536;
                    if (!finished && !runaway) {
                        still.push_back(sl);
                        continue;
                    }
                    S->h_desc[sl->idx].active = 0;
                    finish_export(S, sl);
                    if (runaway) {
                        sl->rec->res.status = 2;
                        std::snprintf(sl->rec->res.error, sizeof(sl->rec->res.error),
                                      "block fast-iterative method did not converge");
                    }
                    HIPCHECK(launch_export_values<float>(sl->rec->g, sl->u, sl->values_dev, sl->rec->np, S->s_comp));
                    if (sl->rec->values_out)
                        HIPCHECK(hipMemcpyAsync(sl->rec->values_out, sl->values_dev, sl->rec->N * sizeof(double),
                                                hipMemcpyDeviceToHost, S->s_comp));
                    HIPCHECK(hipEventCreateWithFlags(&sl->export_done, hipEventDisableTiming));
                    HIPCHECK(hipEventRecord(sl->export_done, S->s_comp));
                    S->exporting.push_back(sl);
                }
                S->running.swap(still);
            }
            // completed exports
            std::vector<Slot *> pending;
            for (Slot *sl : S->exporting) {
                if (hipEventQuery(sl->export_done) != hipSuccess) {
                    pending.push_back(sl);
                    continue;
                }
                (void)hipEventDestroy(sl->export_done);
                HIPCHECK(hipFreeAsync(sl->deps, S->s_comp));
                HIPCHECK(hipFreeAsync(sl->arena, S->s_comp));
                S->free_idx.push_back(sl->idx);
                auto r = sl->rec;
                const size_t bytes = r->bytes;
                delete sl;
                release_budget(S, bytes);
                std::lock_guard<std::mutex> lk(S->mu);
                r->t_done = Clock::now();
                r->res.latency_ms = ms_since(r->t_submit, r->t_done);
                r->done = true;
                if (r->res.status == 0) S->stats.completed++;
                else S->stats.failed++;
                S->cv_done.notify_all();
            }
            S->exporting.swap(pending);
            if (S->running.empty() && !S->exporting.empty()) std::this_thread::sleep_for(std::chrono::microseconds(50));
        } catch (const HipFail &e) {
            // A device error poisons every resident solve: fail them all and stop serving.
            std::vector<std::shared_ptr<RequestRec>> victims;
            for (Slot *sl : S->running) victims.push_back(sl->rec);
            for (Slot *sl : S->exporting) victims.push_back(sl->rec);
            S->running.clear();
            S->exporting.clear();
            for (auto &r : victims) fail_request(S, r, e.msg);
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
            S->cv_submit.notify_all();
            return;
        }
    }
}

}  // namespace

extern "C" {

int hfm_server_create(hfm_server **out, double hbm_budget_bytes, int max_resident) {
    if (!out || !(hbm_budget_bytes > 0) || max_resident <= 0) {
        set_error("hfm_server_create: invalid argument");
        return 1;
    }
    hfm_server *S = new hfm_server;
    try {
        S->budget = hbm_budget_bytes;
        S->max_resident = std::min(max_resident, MAX_SLOTS);
        S->stats.hbm_budget = hbm_budget_bytes;
        for (int i = 0; i < hfm_server::NSETUP; ++i) {
            HIPCHECK(hipStreamCreateWithFlags(&S->s_setup[i], hipStreamNonBlocking));
            S->d_red[i] = dalloc<RedSlot>(RED_SLOTS);
        }
        HIPCHECK(hipStreamCreateWithFlags(&S->s_comp, hipStreamNonBlocking));
        S->d_desc = dalloc<SolveDesc>(MAX_SLOTS);
        HIPCHECK(hipHostMalloc((void **)&S->h_desc, sizeof(SolveDesc) * MAX_SLOTS, hipHostMallocDefault));
        std::memset(S->h_desc, 0, sizeof(SolveDesc) * MAX_SLOTS);
        S->d_lc = dalloc<int>(2);
        HIPCHECK(hipMemset(S->d_lc, 0, 2 * sizeof(int)));
        S->d_count = dalloc<int>(MAX_SLOTS * 2);
        S->d_pend = dalloc<unsigned>(MAX_SLOTS * 2 * PEND_SLOTS);
        S->d_bound = dalloc<float>(MAX_SLOTS * 4);
        S->d_stats = dalloc<unsigned long long>(MAX_SLOTS * 2 * PEND_SLOTS);
        HIPCHECK(hipHostMalloc((void **)&S->h_count, sizeof(int) * MAX_SLOTS * 2, hipHostMallocDefault));
        HIPCHECK(hipHostMalloc((void **)&S->h_stats, sizeof(unsigned long long) * MAX_SLOTS * 2 * PEND_SLOTS,
                               hipHostMallocDefault));
        for (int i = 0; i < hfm_server::NSETUP; ++i) S->th_setup[i] = std::thread(setup_loop, S, i);
        S->th_comp = std::thread(compute_loop, S);
    } catch (const HipFail &e) {
        set_error(e.msg);
        delete S;
        return 1;
    }
    *out = S;
    return 0;
}

int hfm_server_submit(hfm_server *S, const hfm_request *q, int64_t *ticket) {
    if (!S || !q || !ticket || !q->model || !q->metric || (q->nseeds > 0 && !q->seeds)) {
        set_error("hfm_server_submit: null argument");
        return 1;
    }
    auto r = std::make_shared<RequestRec>();
    const std::string m = q->model;
    if (m == "Riemann2") r->model = MODEL_RIEMANN2;
    else if (m == "Riemann3") r->model = MODEL_RIEMANN3;
    else {
        set_error("hfm_server_submit: the server solves Riemann2 / Riemann3 requests (got " + m + ")");
        return 1;
    }
    const int dim = r->model == MODEL_RIEMANN2 ? 2 : 3;
    if (!(q->gridScale > 0)) {
        set_error("Parametrization error : gridscale should be positive");
        return 1;
    }
    Geom &g = r->g;
    g.dim = dim;
    g.n[2] = 1;
    r->N = 1;
    for (int i = 0; i < dim; ++i) {
        if (q->dims[i] <= 0 || q->dims[i] > (1 << 16)) {
            set_error("hfm_server_submit: grid extent out of range");
            return 1;
        }
        g.n[i] = (int)q->dims[i];
        r->N *= q->dims[i];
    }
    if (dim == 2) { g.sh[0] = 3; g.sh[1] = 3; g.sh[2] = 0; }
    else { g.sh[0] = 2; g.sh[1] = 2; g.sh[2] = 2; }
    long long nt = 1;
    for (int i = 0; i < 3; ++i) {
        g.nt[i] = (g.n[i] + (1 << g.sh[i]) - 1) >> g.sh[i];
        nt *= g.nt[i];
    }
    if (nt >= (1LL << SOLVE_SHIFT)) {
        set_error("hfm_server_submit: grid too large for the server (use a plan)");
        return 1;
    }
    g.ntiles = (int)nt;
    g.periodic = 0;
    r->np = nt * 64;
    r->K = dim == 2 ? 3 : 6;
    r->h = q->gridScale;
    r->metric_is_field = q->metric_is_field ? 1 : 0;
    r->dual = q->dual ? 1 : 0;
    if (q->metric_on_device) {
        r->metric_dev = q->metric;
    } else {
        const size_t mc = r->metric_is_field ? (size_t)r->N * r->K : (size_t)r->K;
        r->metric_host.assign(q->metric, q->metric + mc);
    }
    // seeds: ADim + floor, out-of-range dropped, first value kept (StaticFactoring.hxx:242-393)
    std::map<long long, std::pair<long long, double>> bylin;
    for (int s = 0; s < q->nseeds; ++s) {
        int c[3] = {0, 0, 0};
        bool ok = true;
        for (int i = 0; i < dim; ++i) {
            const double f = std::floor(q->seeds[s * dim + i] / q->gridScale);
            if (!(f >= 0 && f < g.n[i])) { ok = false; break; }
            c[i] = (int)f;
        }
        if (!ok) continue;
        const long long lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
        bylin.insert({lin, {(long long)tile_of(g, c) * 64 + local_of(g, c), q->seed_values ? q->seed_values[s] : 0.}});
    }
    if (bylin.empty()) {
        set_error("Error : seeds incorrectly set");
        return 1;
    }
    for (auto &kv : bylin) {
        r->seed_tp.push_back(kv.second.first);
        r->seed_v.push_back(kv.second.second);
    }
    r->values_out = q->values_out;
    r->bytes = arena_bytes(*r) + (size_t)nt * 64 * 4 + (q->metric_on_device ? 0 : r->metric_host.size() * 8);
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
        set_error("hfm_server_submit: server stopped");
        return 1;
    }
    if ((double)r->bytes > S->budget) {
        set_error("hfm_server_submit: request needs more HBM than the server budget");
        return 1;
    }
    r->ticket = S->next_ticket++;
    r->t_submit = Clock::now();
    S->recs[r->ticket] = r;
    S->queue.push_back(r);
    S->stats.submitted++;
    *ticket = r->ticket;
    S->cv_submit.notify_all();
    return 0;
}

int hfm_server_poll(hfm_server *S, int64_t ticket, int *done) {
    if (!S || !done) return 1;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->recs.find(ticket);
    if (it == S->recs.end()) {
        set_error("hfm_server_poll: unknown ticket");
        return 1;
    }
    *done = it->second->done ? 1 : 0;
    return 0;
}

int hfm_server_wait(hfm_server *S, int64_t ticket, double timeout_ms, hfm_request_result *res) {
    if (!S) return 1;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->recs.find(ticket);
    if (it == S->recs.end()) {
        set_error("hfm_server_wait: unknown ticket");
        return 1;
    }
    auto r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_ms < 0) S->cv_done.wait(lk, pred);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred)) {
        set_error("hfm_server_wait: timeout");
        return 1;
    }
    if (res) *res = r->res;
    S->recs.erase(ticket);
    return 0;
}

int hfm_server_get_stats(hfm_server *S, hfm_server_stats *st) {
    if (!S || !st) return 1;
    std::lock_guard<std::mutex> lk(S->mu);
    *st = S->stats;
    st->resident = S->resident;
    st->hbm_in_use = S->in_use;
    return 0;
}

void hfm_server_destroy(hfm_server *S) {
    if (!S) return;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        // let admitted work drain
        S->cv_done.wait_for(lk, std::chrono::seconds(60), [&] { return S->resident == 0 && S->queue.empty(); });
        S->stop = true;
        S->cv_submit.notify_all();
        S->cv_ready.notify_all();
    }
    for (int i = 0; i < hfm_server::NSETUP; ++i)
        if (S->th_setup[i].joinable()) S->th_setup[i].join();
    if (S->th_comp.joinable()) S->th_comp.join();
    if (S->s_comp) (void)hipStreamSynchronize(S->s_comp);
    for (int i = 0; i < hfm_server::NSETUP; ++i)
        if (S->s_setup[i]) (void)hipStreamSynchronize(S->s_setup[i]);
    if (S->d_list) (void)hipFree(S->d_list);
    for (int i = 0; i < hfm_server::NSETUP; ++i) dfree(S->d_red[i]);
    dfree(S->d_desc);
    dfree(S->d_lc);
    dfree(S->d_count);
    dfree(S->d_pend);
    dfree(S->d_bound);
    dfree(S->d_stats);
    if (S->h_desc) (void)hipHostFree(S->h_desc);
    if (S->h_count) (void)hipHostFree(S->h_count);
    if (S->h_stats) (void)hipHostFree(S->h_stats);
    if (S->s_comp) (void)hipStreamDestroy(S->s_comp);
    for (int i = 0; i < hfm_server::NSETUP; ++i)
        if (S->s_setup[i]) (void)hipStreamDestroy(S->s_setup[i]);
    delete S;
}

}  // extern "C"
