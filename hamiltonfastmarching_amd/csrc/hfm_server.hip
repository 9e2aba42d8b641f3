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
// Tiles per round shared by the resident solves (each solve's band is steered to
// ROUND_TILES / running, clamped to [16, 256]).  Mixed stream, 128 in flight: 16384 against
// 4096 gave about 7 % more sustained solves/s and a 15 % lower p99, at 8 sweeps per visit.
constexpr int ROUND_TILES = 16384;

struct SolveDesc {
    TileCtx<float> ctx;
    int model;
    int active;
    float delta;                  // initial band width
    int target;                   // tiles per round the band width is steered to
    unsigned *pend;               // 2 * PEND_WORDS: smallest pending keys of alternating rounds, PEND_STRIDE apart
    float *bound;                 // 2 band bounds + 2 band widths
    int *count;                   // 2 tiles listed per round
    unsigned long long *visits;   // tile visits of the solve so far (tiles listed by its rounds)
};

// Jacobi sweeps per tile visit in the serving update kernel (an unconverged tile goes back to
// its solve's schedule).
constexpr int SERVE_SWEEPS = 8;

// Scheduling round over every running solve.  Block b covers tiles [bmap[b].y, bmap[b].y +
// SCHED_BLOCK) of solve bmap[b].x: the grid has one block per SCHED_BLOCK tiles of each running solve.
__global__ __launch_bounds__(SCHED_BLOCK) void k_mschedule(const SolveDesc *__restrict__ descs,
                                                           const int2 *__restrict__ bmap, int parity,
                                                           int *__restrict__ list, int *__restrict__ list_count) {
    const int2 bm = bmap[blockIdx.x];
    const int s = bm.x;
    const SolveDesc &d = descs[s];
    const int nt = d.ctx.g.ntiles;
    float delta = d.delta;
    if (delta > 0.f)
        delta = adapt_delta(d.bound[2 + (parity ^ 1)], d.count[parity ^ 1], d.target, d.delta * (1.f / 16.f),
                            64.f * d.delta);
    const float B =
        band_bound(key_value(read_pend(d.pend + parity * PEND_WORDS, PEND_STRIDE)), d.bound[parity ^ 1], delta);
    if (bm.y == 0 && threadIdx.x == 0) {
        d.bound[parity] = B;
        d.bound[2 + parity] = delta;
    }
    const int t = bm.y + threadIdx.x;
    unsigned k = KEY_NONE;
    if (t < nt) k = d.ctx.keys[t];
    const bool take = band_take(k, B);
    const unsigned left = (!take && k != KEY_NONE) ? k : KEY_NONE;
    if (take) d.ctx.keys[t] = KEY_NONE;
    const int pos = block_take(take, left, list_count, &d.count[parity],
                               &d.pend[(parity ^ 1) * PEND_WORDS + (blockIdx.x & (PEND_SLOTS - 1)) * PEND_STRIDE],
                               d.visits);
    if (take) list[pos] = (s << SOLVE_SHIFT) | t;
}

// Update round over the packed list of (solve, tile) entries of every resident solve.
__global__ __launch_bounds__(256) void k_mfim(const SolveDesc *__restrict__ descs, int ndesc, int parity,
                                              const int *__restrict__ list, long long list_cap,
                                              const int *__restrict__ list_count,
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
            if (d.active) d.pend[parity * PEND_WORDS + (i % PEND_SLOTS) * PEND_STRIDE] = KEY_NONE;
        }
    }
    const int gw = blockIdx.x * 4 + wv;
    const int slot = gw & (PEND_SLOTS - 1);
    const int stride = gridDim.x * 4;
    // read this wave's first entry together with the count (entries past it are never used)
    int next = gw < list_cap ? list[gw] : 0;
    for (int i = gw; i < n; i += stride) {
        const int e = __builtin_amdgcn_readfirstlane(next);
        next = i + stride < n ? list[i + stride] : 0;
        const int s = e >> SOLVE_SHIFT;
        const int tile = e & ((1 << SOLVE_SHIFT) - 1);
        const SolveDesc &d = descs[s];
        unsigned wm = KEY_NONE;
        // at most SERVE_SWEEPS sweeps per visit; an unconverged tile goes back to its solve's schedule
        if (d.model == MODEL_RIEMANN3) (void)fim_tile<MODEL_RIEMANN3, float, SERVE_SWEEPS>(d.ctx, tile, lane, su[wv], wm);
        else (void)fim_tile<MODEL_RIEMANN2, float, SERVE_SWEEPS>(d.ctx, tile, lane, su[wv], wm);
        // each slot on its own 128-byte line: same-line atomics serialise
        if (lane == 0 && wm != KEY_NONE) atomicMin(&d.pend[(parity ^ 1) * PEND_WORDS + slot * PEND_STRIDE], wm);
    }
}

// Setup-time clears of one solve in one launch: the stencil status words, the reduction slots,
// the tile-graph counters and the seed key word.
__global__ __launch_bounds__(256) void k_setup_clear(int *__restrict__ status, RedSlot *__restrict__ red,
                                                     int *__restrict__ dep_cnt, int *__restrict__ dep_fill,
                                                     unsigned *__restrict__ seed_key, int ntiles) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (i0 == 0) {
        status[0] = 0;
        status[1] = 0x7fffffff;
        *seed_key = KEY_NONE;
    }
    int *rw = reinterpret_cast<int *>(red);
    for (int i = i0; i < RED_SLOTS * (int)(sizeof(RedSlot) / sizeof(int)); i += stride) rw[i] = 0;
    for (int i = i0; i < ntiles; i += stride) {
        dep_cnt[i] = 0;
        dep_fill[i] = 0;
    }
}

// Scheduling state of a solve joining the running set at a round of the given parity: the seed
// key (smallest key the seed marks left pending) becomes the pending minimum of that round.
__global__ __launch_bounds__(64) void k_join_state(unsigned *__restrict__ pend, int *__restrict__ count,
                                                   unsigned long long *__restrict__ visits, float *__restrict__ bound,
                                                   const unsigned *__restrict__ seed_key, int parity, float delta) {
    const int i = threadIdx.x;
    for (int j = i; j < 2 * PEND_SLOTS; j += 64)
        pend[(j / PEND_SLOTS) * PEND_WORDS + (j % PEND_SLOTS) * PEND_STRIDE] =
            (j == parity * PEND_SLOTS) ? *seed_key : KEY_NONE;
    if (i == 0) *visits = 0ull;
    if (i < 2) count[i] = 0;
    if (i < 4) bound[i] = i < 2 ? -__builtin_huge_valf() : delta;
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
    std::vector<long long> seed_blob;  // seed_tp then the bits of seed_v: one host-to-device copy
    double *values_out = nullptr;
    size_t bytes = 0;
    Clock::time_point t_submit, t_admit, t_ready, t_done;
    bool done = false;
    hfm_request_result res{};
};

struct Slot {
    std::shared_ptr<RequestRec> rec;
    int idx = -1;  // descriptor index
    long long first_batch = 0;  // first batch of rounds that includes the solve
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
    unsigned *seed_key = nullptr;  // smallest key the seed marks left pending
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
    add(r.seed_tp.size() * 16);  // seed points and values, one block

    add(r.N * 8);           // values_dev
    add(16);                // status
    add(4);                 // seed key
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

    // Setup threads, one stream each.  With the compute stream and the caller's stream this
    // stays within the 4 hardware queues HIP creates by default, so no setup stream shares a
    // queue with the compute stream.
    static constexpr int NSETUP = 2;
    static constexpr int SETUP_GROUP = 8;  // requests set up per pass (two synchronisations)
    hipStream_t s_setup[NSETUP] = {}, s_comp = nullptr;
    RedSlot *d_red[NSETUP] = {};           // per thread: SETUP_GROUP x RED_SLOTS stencil reductions
    RedSlot *h_red[NSETUP] = {};           // pinned copies
    int *h_status[NSETUP] = {};            // pinned SETUP_GROUP x 2
    int *h_total[NSETUP] = {};             // pinned SETUP_GROUP dependency totals
    std::thread th_setup[NSETUP], th_comp;

    // compute-thread state.  One batch of rounds stays in flight while the host reads the
    // counters of the previous one, so the host mirrors the device reads from or writes to
    // are double-buffered by batch parity.
    SolveDesc *d_desc = nullptr;
    SolveDesc *h_desc = nullptr;         // host master copy of the descriptor table
    SolveDesc *h_desc_stage[2] = {};     // pinned, source of the per-batch upload
    int2 *d_bmap = nullptr;              // k_mschedule block table
    int2 *h_bmap_stage[2] = {};          // pinned
    long long bmap_cap = 0;
    int *d_list = nullptr;
    long long list_cap = 0;
    int *d_lc = nullptr;          // 2 alternating list counters
    int *d_count = nullptr;       // MAX_SLOTS * 2
    unsigned *d_pend = nullptr;   // MAX_SLOTS * 2 * PEND_WORDS
    float *d_bound = nullptr;     // MAX_SLOTS * 2
    unsigned long long *d_stats = nullptr;  // MAX_SLOTS tile-visit counters
    int *h_count[2] = {};                   // pinned MAX_SLOTS * 2, per batch parity
    unsigned long long *h_stats[2] = {};    // pinned, per batch parity
    hipEvent_t batch_done[2] = {};
    std::vector<Slot *> running;
    std::vector<Slot *> exporting;
    std::vector<int> free_idx;
    int ndesc = 0;
    long long round = 0;
    long long nlaunched = 0;  // batches launched
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

// Setup of a group of admitted requests on one setup stream: stencils, tile graph, initial
// values and seed marks.  Phase 1 launches every request's stencil, dependency count and scan
// and copies back its status, reductions and dependency total; one synchronisation; phase 2
// allocates and fills the dependency lists, initialises values and marks the seeds; one
// synchronisation.  Failed requests are reported in `err` (empty: ready).
struct SetupItem {
    std::shared_ptr<RequestRec> r;
    Slot *sl = nullptr;
    std::string err;
};

void setup_group(hfm_server *S, std::vector<SetupItem> &items, int which) {
    hipStream_t s = S->s_setup[which];
    const int n = (int)items.size();
    for (int g = 0; g < n; ++g) {
        const std::shared_ptr<RequestRec> &r = items[g].r;
        Slot *sl = new Slot;
        items[g].sl = sl;
        sl->rec = r;
        HIPCHECK(hipMallocAsync((void **)&sl->arena, arena_bytes(*r), s));
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
        const size_t ns = r->seed_tp.size();
        sl->seed_p = carve<long long>(p, 2 * ns);
        sl->seed_v = reinterpret_cast<double *>(sl->seed_p + ns);
        sl->values_dev = carve<double>(p, r->N);
        sl->status = carve<int>(p, 4);
        sl->seed_key = carve<unsigned>(p, 1);
        const double *metric = r->metric_dev;
        if (!metric) {
            const size_t mcount = r->metric_is_field ? (size_t)r->N * r->K : (size_t)r->K;
            HIPCHECK(hipMallocAsync((void **)&sl->metric_stage, mcount * sizeof(double), s));
            HIPCHECK(hipMemcpyAsync(sl->metric_stage, r->metric_host.data(), mcount * sizeof(double),
                                    hipMemcpyHostToDevice, s));
            metric = sl->metric_stage;
        }
        HIPCHECK(hipMemcpyAsync(sl->seed_p, r->seed_blob.data(), r->seed_blob.size() * 8, hipMemcpyHostToDevice, s));
        RedSlot *red = S->d_red[which] + (size_t)g * RED_SLOTS;
        hipLaunchKernelGGL(k_setup_clear, dim3((unsigned)std::min<long long>((r->g.ntiles + 255) / 256 + 16, 1024)),
                           dim3(256), 0, s, sl->status, red, sl->dep_cnt, sl->dep_fill, sl->seed_key, r->g.ntiles);
        HIPCHECK(hipGetLastError());
        InvHH ih{};
        const int D = r->g.dim;
        for (int i = 0; i < D; ++i)
            for (int j = 0; j <= i; ++j) ih.v[i * (i + 1) / 2 + j] = 1. / (r->h * r->h);
        HIPCHECK(launch_riemann_stencil<float>(r->g, metric, r->metric_is_field, r->dual, ih, sl->offs, sl->w, r->np,
                                               sl->status, red, s));
        StencilData<float> sd{sl->offs, sl->w, nullptr, nullptr, nullptr, r->np};
        // The tile graph's first pass keeps each tile's sources (<= 64: 64 ints a tile) in `u` and
        // their counts in `tmax`, and the multi-block scan works in `fixedmask`: all three are
        // written by k_init only after the second pass has read them.
        HIPCHECK(launch_tile_deps<float>(r->model, r->g, sd, 0, sl->dep_cnt, nullptr, nullptr, nullptr,
                                         reinterpret_cast<int *>(sl->u), reinterpret_cast<int *>(sl->tmax), s));
        HIPCHECK(launch_exclusive_scan(sl->dep_cnt, sl->dep_start, r->g.ntiles, s,
                                       reinterpret_cast<int *>(sl->fixedmask)));
        HIPCHECK(hipMemcpyAsync(S->h_status[which] + 2 * g, sl->status, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(S->h_red[which] + (size_t)g * RED_SLOTS, red, RED_SLOTS * sizeof(RedSlot),
                                hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(S->h_total[which] + g, sl->dep_start + r->g.ntiles, sizeof(int),
                                hipMemcpyDeviceToHost, s));
    }
    HIPCHECK(hipStreamSynchronize(s));
    for (int g = 0; g < n; ++g) {
        const std::shared_ptr<RequestRec> &r = items[g].r;
        Slot *sl = items[g].sl;
        const int st = S->h_status[which][2 * g];
        if (st != SELLING_OK) {
            items[g].err = st == SELLING_NOT_SPD ? "ObtuseSuperbase error : input must be positive definite"
                                                 : "ObtuseSuperbase error : did not terminate in 200 iterations";
            if (sl->metric_stage) HIPCHECK(hipFreeAsync(sl->metric_stage, s));
            HIPCHECK(hipFreeAsync(sl->arena, s));
            delete sl;
            items[g].sl = nullptr;
            continue;
        }
        const int total = S->h_total[which][g];
        HIPCHECK(hipMallocAsync((void **)&sl->deps, sizeof(int) * std::max(total, 1), s));
        StencilData<float> sd{sl->offs, sl->w, nullptr, nullptr, nullptr, r->np};
        HIPCHECK(launch_tile_deps<float>(r->model, r->g, sd, 1, sl->dep_cnt, sl->dep_start, sl->dep_fill, sl->deps,
                                         reinterpret_cast<int *>(sl->u), reinterpret_cast<int *>(sl->tmax), s));
        const int nseeds = (int)r->seed_tp.size();
        HIPCHECK(launch_init<float>(r->g, sl->u, sl->fixedmask, sl->keys, sl->tmax, r->np, sl->seed_p, sl->seed_v,
                                    nseeds, s));
        HIPCHECK(launch_seed_activate(sl->seed_p, sl->seed_v, nseeds, sl->dep_start, sl->deps, sl->keys, sl->tmax,
                                      sl->seed_key, s));
        if (sl->metric_stage) {
            HIPCHECK(hipFreeAsync(sl->metric_stage, s));
            sl->metric_stage = nullptr;
        }
        const RedTotals tot = red_total(S->h_red[which] + (size_t)g * RED_SLOTS);
        const double tau = tot.count > 0 ? tot.sum / tot.count : 0.;
        sl->delta = (float)(8. * tau);
    }
    HIPCHECK(hipStreamSynchronize(s));
}

void setup_loop(hfm_server *S, int which) {
    if (hipSetDevice(0) != hipSuccess) return;
    while (true) {
        std::vector<SetupItem> items;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            auto admissible = [&] {
                if (S->queue.empty()) return false;
                const auto &h = S->queue.front();
                return S->resident < S->max_resident && S->in_use + (double)h->bytes <= S->budget;
            };
            S->cv_submit.wait(lk, [&] { return S->stop || admissible(); });
            if (S->stop) return;
            while ((int)items.size() < hfm_server::SETUP_GROUP && admissible()) {
                SetupItem it;
                it.r = S->queue.front();
                S->queue.pop_front();
                S->in_use += (double)it.r->bytes;
                S->resident++;
                S->stats.max_resident_seen = std::max<int64_t>(S->stats.max_resident_seen, S->resident);
                items.push_back(std::move(it));
            }
        }
        const auto t_admit = Clock::now();
        for (auto &it : items) {
            it.r->t_admit = t_admit;
            it.r->res.queue_ms = ms_since(it.r->t_submit, t_admit);
        }
        std::string group_err;
        try {
            setup_group(S, items, which);
        } catch (const HipFail &e) {
            group_err = e.msg;
        }
        const auto t_ready = Clock::now();
        std::vector<Slot *> ok;
        for (auto &it : items) {
            if (!group_err.empty() || !it.err.empty() || !it.sl) {
                release_budget(S, it.r->bytes);
                fail_request(S, it.r, group_err.empty() ? it.err : group_err);
                continue;
            }
            it.r->t_ready = t_ready;
            it.r->res.setup_ms = ms_since(it.r->t_admit, t_ready);
            ok.push_back(it.sl);
        }
        std::lock_guard<std::mutex> lk(S->mu);
        for (Slot *sl : ok) S->ready.push_back(sl);
        if (!ok.empty()) S->cv_ready.notify_all();
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
    sl->first_batch = S->nlaunched;
    const int parity = (int)(S->round & 1);
    unsigned *pend = S->d_pend + (size_t)idx * 2 * PEND_WORDS;
    float *bound = S->d_bound + (size_t)idx * 4;
    int *count = S->d_count + (size_t)idx * 2;
    unsigned long long *st = S->d_stats + idx;
    hipLaunchKernelGGL(k_join_state, dim3(1), dim3(64), 0, s, pend, count, st, bound, sl->seed_key, parity, sl->delta);
    HIPCHECK(hipGetLastError());
    SolveDesc &d = S->h_desc[idx];
    d.ctx = TileCtx<float>{r.g, StencilData<float>{sl->offs, sl->w, nullptr, nullptr, nullptr, r.np}, sl->u,
                           sl->fixedmask, sl->dep_start, sl->deps, sl->keys, sl->tmax};
    d.model = r.model;
    d.active = 1;
    d.delta = sl->delta;
    d.target = 256;  // set per batch from the number of running solves
    d.pend = pend;
    d.bound = bound;
    d.count = count;
    d.visits = st;
    S->running.push_back(sl);
}

void finish_export(Slot *sl, const unsigned long long *h_stats) {
    RequestRec &r = *sl->rec;
    r.res.tile_visits = (int64_t)h_stats[sl->idx];
    r.res.rounds = sl->rounds;
}

// Launch one batch of `batch` rounds over the running solves, then the copies of their
// counters into the host mirrors of parity q = batch index & 1, and record batch_done[q].
void launch_batch(hfm_server *S, long long q, int batch) {
    hipStream_t s = S->s_comp;
    const int par = (int)(q & 1);
    long long tiles = 0;
    std::vector<int2> bm;
    for (Slot *sl : S->running) {
        const int nt = sl->rec->g.ntiles;
        tiles += nt;
        for (int b = 0; b < nt; b += SCHED_BLOCK) bm.push_back(int2{sl->idx, b});
    }
    if (tiles > S->list_cap) {
        if (S->d_list) HIPCHECK(hipFreeAsync(S->d_list, s));
        S->list_cap = std::max<long long>(tiles, 2 * S->list_cap);
        HIPCHECK(hipMallocAsync((void **)&S->d_list, S->list_cap * sizeof(int), s));
    }
    if ((long long)bm.size() > S->bmap_cap) {
        // The staging buffer of the other parity may still feed an in-flight copy: wait for it.
        HIPCHECK(hipStreamSynchronize(s));
        if (S->d_bmap) HIPCHECK(hipFree(S->d_bmap));
        for (int i = 0; i < 2; ++i)
            if (S->h_bmap_stage[i]) HIPCHECK(hipHostFree(S->h_bmap_stage[i]));
        S->bmap_cap = std::max<long long>((long long)bm.size(), 2 * S->bmap_cap);
        HIPCHECK(hipMalloc((void **)&S->d_bmap, S->bmap_cap * sizeof(int2)));
        for (int i = 0; i < 2; ++i)
            HIPCHECK(hipHostMalloc((void **)&S->h_bmap_stage[i], S->bmap_cap * sizeof(int2), hipHostMallocDefault));
    }
    // Tiles per round each solve's band is steered to: the resident solves share ROUND_TILES;
    // a busy server runs narrow bands (fewer redundant tile visits, the extra rounds are shared
    // by all solves), a lone solve runs wide ones (fewer rounds).
    // 2-D solves (8x8 tiles, fewer dependents) take four times the 3-D target, up to 1024: the
    // mixed stream's p99 fell from about 183 to 133 ms with the same throughput.
    const int tgt = std::max(16, std::min(256, ROUND_TILES / (int)S->running.size()));
    for (int i = 0; i < S->ndesc; ++i)
        S->h_desc[i].target = S->h_desc[i].model == MODEL_RIEMANN2 ? std::min(1024, 4 * tgt) : tgt;
    std::memcpy(S->h_desc_stage[par], S->h_desc, sizeof(SolveDesc) * S->ndesc);
    std::memcpy(S->h_bmap_stage[par], bm.data(), bm.size() * sizeof(int2));
    HIPCHECK(hipMemcpyAsync(S->d_desc, S->h_desc_stage[par], sizeof(SolveDesc) * S->ndesc, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(S->d_bmap, S->h_bmap_stage[par], bm.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    const unsigned grid = (unsigned)std::min<long long>((tiles + 3) / 4, 2048);
    for (int b = 0; b < batch; ++b) {
        const int parity = (int)(S->round & 1);
        hipLaunchKernelGGL(k_mschedule, dim3((unsigned)bm.size()), dim3(SCHED_BLOCK), 0, s, S->d_desc, S->d_bmap,
                           parity, S->d_list, &S->d_lc[parity]);
        HIPCHECK(hipGetLastError());
        hipLaunchKernelGGL(k_mfim, dim3(grid), dim3(256), 0, s, S->d_desc, S->ndesc, parity, S->d_list, S->list_cap,
                           &S->d_lc[parity], &S->d_lc[parity ^ 1]);
        HIPCHECK(hipGetLastError());
        S->round++;
    }
    for (Slot *sl : S->running) sl->rounds += batch;
    HIPCHECK(hipMemcpyAsync(S->h_count[par], S->d_count, sizeof(int) * 2 * S->ndesc, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(S->h_stats[par], S->d_stats, sizeof(unsigned long long) * S->ndesc, hipMemcpyDeviceToHost,
                            s));
    HIPCHECK(hipEventRecord(S->batch_done[par], s));
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.batches++;
    S->stats.rounds += batch;
}

// Read the counters of batch q (after its event): solves whose last round listed no tile have
// converged; export them.  Only solves still running and included in batch q are looked at.
// Their descriptors are switched off for the batches launched from now on; a batch already in
// flight may still list them and finds nothing pending.
void process_batch(hfm_server *S, long long q, int last_parity) {
    const int par = (int)(q & 1);
    HIPCHECK(hipEventSynchronize(S->batch_done[par]));
    std::vector<Slot *> done, still;
    for (Slot *sl : S->running) {
        const bool member = sl->first_batch <= q;
        const bool finished = member && S->h_count[par][2 * sl->idx + last_parity] == 0;
        const bool runaway = member && sl->rounds > 8LL * sl->rec->g.ntiles + 65536;
        if (!finished && !runaway) {
            still.push_back(sl);
            continue;
        }
        if (!finished) {
            sl->rec->res.status = 2;
            std::snprintf(sl->rec->res.error, sizeof(sl->rec->res.error),
                          "block fast-iterative method did not converge");
        }
        done.push_back(sl);
    }
    if (done.empty()) return;
    S->running.swap(still);
    for (Slot *sl : done) {
        S->h_desc[sl->idx].active = 0;
        finish_export(sl, S->h_stats[par]);
        HIPCHECK(launch_export_values<float>(sl->rec->g, sl->u, sl->values_dev, sl->rec->np, S->s_comp));
        if (sl->rec->values_out)
            HIPCHECK(hipMemcpyAsync(sl->rec->values_out, sl->values_dev, sl->rec->N * sizeof(double),
                                    hipMemcpyDeviceToHost, S->s_comp));
        // Recorded after every batch launched so far: once it completes, no kernel reads the
        // solve's memory any more.
        HIPCHECK(hipEventCreateWithFlags(&sl->export_done, hipEventDisableTiming));
        HIPCHECK(hipEventRecord(sl->export_done, S->s_comp));
        S->exporting.push_back(sl);
    }
}

void retire_exports(hfm_server *S) {
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
}

// The compute thread: joins set-up solves, keeps one batch of rounds in flight over the running
// solves while it reads the previous batch's counters, exports converged solves, frees them.
void compute_loop(hfm_server *S) {
    if (hipSetDevice(0) != hipSuccess) return;
    const int batch = 8;
    long long processed = 0;
    int last_parity[2] = {0, 0};
    while (true) {
        std::vector<Slot *> joining;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->running.empty() && S->exporting.empty() && processed == S->nlaunched)
                S->cv_ready.wait(lk, [&] { return S->stop || !S->ready.empty(); });
            if (S->stop && S->running.empty() && S->exporting.empty() && S->ready.empty() &&
                processed == S->nlaunched)
                return;
            while (!S->ready.empty()) {
                joining.push_back(S->ready.front());
                S->ready.pop_front();
            }
        }
        try {
            for (Slot *sl : joining) join_solve(S, sl);
            if (!S->running.empty()) {
                const long long q = S->nlaunched;
                launch_batch(S, q, batch);
                last_parity[q & 1] = (int)((S->round - 1) & 1);
                S->nlaunched = q + 1;
            }
            // keep the batch just launched in flight; drain everything when nothing was launched
            while (processed < S->nlaunched - (S->running.empty() ? 0 : 1)) {
                const long long q = processed++;
                process_batch(S, q, last_parity[q & 1]);
            }
            retire_exports(S);
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
            S->d_red[i] = dalloc<RedSlot>((size_t)RED_SLOTS * hfm_server::SETUP_GROUP);
            HIPCHECK(hipHostMalloc((void **)&S->h_red[i], sizeof(RedSlot) * RED_SLOTS * hfm_server::SETUP_GROUP,
                                   hipHostMallocDefault));
            HIPCHECK(hipHostMalloc((void **)&S->h_status[i], sizeof(int) * 2 * hfm_server::SETUP_GROUP,
                                   hipHostMallocDefault));
            HIPCHECK(hipHostMalloc((void **)&S->h_total[i], sizeof(int) * hfm_server::SETUP_GROUP,
                                   hipHostMallocDefault));
        }
        HIPCHECK(hipStreamCreateWithFlags(&S->s_comp, hipStreamNonBlocking));
        S->d_desc = dalloc<SolveDesc>(MAX_SLOTS);
        S->h_desc = new SolveDesc[MAX_SLOTS];
        std::memset(S->h_desc, 0, sizeof(SolveDesc) * MAX_SLOTS);
        for (int i = 0; i < 2; ++i) {
            HIPCHECK(hipHostMalloc((void **)&S->h_desc_stage[i], sizeof(SolveDesc) * MAX_SLOTS, hipHostMallocDefault));
            HIPCHECK(hipEventCreateWithFlags(&S->batch_done[i], hipEventDisableTiming));
        }
        S->d_lc = dalloc<int>(2);
        HIPCHECK(hipMemset(S->d_lc, 0, 2 * sizeof(int)));
        S->d_count = dalloc<int>(MAX_SLOTS * 2);
        S->d_pend = dalloc<unsigned>((size_t)MAX_SLOTS * 2 * PEND_WORDS);
        S->d_bound = dalloc<float>(MAX_SLOTS * 4);
        S->d_stats = dalloc<unsigned long long>(MAX_SLOTS);
        for (int i = 0; i < 2; ++i) {
            HIPCHECK(hipHostMalloc((void **)&S->h_count[i], sizeof(int) * MAX_SLOTS * 2, hipHostMallocDefault));
            HIPCHECK(hipHostMalloc((void **)&S->h_stats[i], sizeof(unsigned long long) * MAX_SLOTS,
                                   hipHostMallocDefault));
        }
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
    r->seed_blob = r->seed_tp;
    for (double v : r->seed_v) {
        long long bits;
        std::memcpy(&bits, &v, sizeof(bits));
        r->seed_blob.push_back(bits);
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
    for (int i = 0; i < hfm_server::NSETUP; ++i) {
        dfree(S->d_red[i]);
        if (S->h_red[i]) (void)hipHostFree(S->h_red[i]);
        if (S->h_status[i]) (void)hipHostFree(S->h_status[i]);
        if (S->h_total[i]) (void)hipHostFree(S->h_total[i]);
    }
    dfree(S->d_desc);
    dfree(S->d_lc);
    dfree(S->d_count);
    dfree(S->d_pend);
    dfree(S->d_bound);
    dfree(S->d_stats);
    delete[] S->h_desc;
    if (S->d_bmap) (void)hipFree(S->d_bmap);
    for (int i = 0; i < 2; ++i) {
        if (S->h_desc_stage[i]) (void)hipHostFree(S->h_desc_stage[i]);
        if (S->h_bmap_stage[i]) (void)hipHostFree(S->h_bmap_stage[i]);
        if (S->h_count[i]) (void)hipHostFree(S->h_count[i]);
        if (S->h_stats[i]) (void)hipHostFree(S->h_stats[i]);
        if (S->batch_done[i]) (void)hipEventDestroy(S->batch_done[i]);
    }
    if (S->s_comp) (void)hipStreamDestroy(S->s_comp);
    for (int i = 0; i < hfm_server::NSETUP; ++i)
        if (S->s_setup[i]) (void)hipStreamDestroy(S->s_setup[i]);
    delete S;
}

}  // extern "C"
