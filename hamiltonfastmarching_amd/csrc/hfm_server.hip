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
#include "hfm_setup_dev.h"
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
    const unsigned *seed_key;     // smallest key the seed marks left pending
    double *values_dev;           // row-major FP64 export of the values
    int joining;                  // 1 in the batch the solve joins: k_mjoin sets its scheduling state
};

// Jacobi sweeps per tile visit in the serving update kernel (an unconverged tile goes back to
// its solve's schedule).  Mixed stream, 128 in flight, three alternating pairs: 4 sweeps on
// Riemann3 tiles gave 5 % more sustained solves/s than 8 (2403 against 2281); on Riemann2
// tiles 4, 8 and 16 were within the run-to-run spread.
constexpr int SERVE_SWEEPS_2D = 8;
constexpr int SERVE_SWEEPS_3D = 4;

// Scheduling of one block-table entry: tiles [bm.y, bm.y + SCHED_BLOCK) of solve bm.x, by a
// SCHED_BLOCK-thread block; `b` (the entry's index) spreads the pending-minimum atomics.
__device__ inline void msched_entry(const SolveDesc *__restrict__ descs, int2 bm, int b, int parity,
                                    int *__restrict__ list, int *__restrict__ list_count) {
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
                               &d.pend[(parity ^ 1) * PEND_WORDS + (b & (PEND_SLOTS - 1)) * PEND_STRIDE], d.visits);
    if (take) list[pos] = (s << SOLVE_SHIFT) | t;
}

// Scheduling round over every running solve.  Block b covers tiles [bmap[b].y, bmap[b].y +
// SCHED_BLOCK) of solve bmap[b].x: the grid has one block per SCHED_BLOCK tiles of each running solve.
__global__ __launch_bounds__(SCHED_BLOCK) void k_mschedule(const SolveDesc *__restrict__ descs,
                                                           const int2 *__restrict__ bmap, int parity,
                                                           int *__restrict__ list, int *__restrict__ list_count) {
    msched_entry(descs, bmap[blockIdx.x], blockIdx.x, parity, list, list_count);
}

// Update round over the packed list of (solve, tile) entries of every resident solve, by blocks
// `blk` of `nblk` (each of blockDim.x / 64 waves, wave w with LDS row su[w]).  Block 0 also
// clears the next round's list counter and the counters of the round after.
__device__ inline void mfim_round(const SolveDesc *__restrict__ descs, int ndesc, int parity,
                                  const int *__restrict__ list, long long list_cap, const int *__restrict__ list_count,
                                  int *__restrict__ list_count_reset, int blk, int nblk, float (*su)[SU_ROW]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int n = *list_count;
    if (blk == 0) {
        if (threadIdx.x == 0) *list_count_reset = 0;
        for (int i = threadIdx.x; i < ndesc; i += blockDim.x)
            if (descs[i].active) descs[i].count[parity ^ 1] = 0;
        for (int i = threadIdx.x; i < ndesc * PEND_SLOTS; i += blockDim.x) {
            const SolveDesc &d = descs[i / PEND_SLOTS];
            if (d.active) d.pend[parity * PEND_WORDS + (i % PEND_SLOTS) * PEND_STRIDE] = KEY_NONE;
        }
    }
    const int gw = blk * nw + wv;
    const int slot = gw & (PEND_SLOTS - 1);
    const int stride = nblk * nw;
    // read this wave's first entry together with the count (entries past it are never used)
    int next = gw < list_cap ? list[gw] : 0;
    for (int i = gw; i < n; i += stride) {
        const int e = __builtin_amdgcn_readfirstlane(next);
        next = i + stride < n ? list[i + stride] : 0;
        const int s = e >> SOLVE_SHIFT;
        const int tile = e & ((1 << SOLVE_SHIFT) - 1);
        const SolveDesc &d = descs[s];
        unsigned wm = KEY_NONE;
        // a capped number of sweeps per visit; an unconverged tile goes back to its solve's schedule
        if (d.model == MODEL_RIEMANN3)
            (void)fim_tile<MODEL_RIEMANN3, float, SERVE_SWEEPS_3D>(d.ctx, tile, lane, su[wv], wm);
        else
            (void)fim_tile<MODEL_RIEMANN2, float, SERVE_SWEEPS_2D>(d.ctx, tile, lane, su[wv], wm);
        // each slot on its own 128-byte line: same-line atomics serialise
        if (lane == 0 && wm != KEY_NONE) atomicMin(&d.pend[(parity ^ 1) * PEND_WORDS + slot * PEND_STRIDE], wm);
    }
}

__global__ __launch_bounds__(256) void k_mfim(const SolveDesc *__restrict__ descs, int ndesc, int parity,
                                              const int *__restrict__ list, long long list_cap,
                                              const int *__restrict__ list_count,
                                              int *__restrict__ list_count_reset) {
    __shared__ float su[4][SU_ROW];
    mfim_round(descs, ndesc, parity, list, list_cap, list_count, list_count_reset, blockIdx.x, gridDim.x, su);
}

// ---------------------------------------------------------------------------------------------
// Persistent variant of a batch of rounds, kept for comparison with the per-round launches
// (off by default): one launch runs `nrounds` rounds, each a scheduling phase and an update
// phase over every running solve, separated by grid-wide barriers.  One SCHED_BLOCK-thread
// block per CU (the grid is the number of blocks that fit on the device at once, so every
// block is resident).  A barrier publishes the block's writes to the device (agent-scope
// release: the XCD's L2 written back), counts the block in, waits for the others, and
// acquires (the L2's stale lines invalidated).  The wait is bounded: past GRID_SPIN_LIMIT polls a
// block raises `abort`, and every block leaves at its next barrier, so the launch always ends.
// ---------------------------------------------------------------------------------------------
constexpr unsigned GRID_SPIN_LIMIT = 1u << 22;

__device__ inline bool grid_barrier(unsigned long long *bar, unsigned long long target, int *abort) {
    __shared__ int ok_s;
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(bar, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                ok = 0;
                break;
            }
            if (++spins > GRID_SPIN_LIMIT) {
                __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok_s = ok;
    }
    __syncthreads();
    return ok_s != 0;
}

__global__ __launch_bounds__(SCHED_BLOCK) void k_mbatch(const SolveDesc *__restrict__ descs,
                                                        const int2 *__restrict__ bmap, int nb, int ndesc,
                                                        int parity0, int nrounds, int *__restrict__ list,
                                                        long long list_cap, int *__restrict__ lc,
                                                        unsigned long long *bar, unsigned long long bar_base,
                                                        int *abort) {
    __shared__ float su[SCHED_BLOCK / 64][SU_ROW];
    unsigned long long target = bar_base;
    for (int r = 0; r < nrounds; ++r) {
        const int parity = (parity0 + r) & 1;
        for (int b = blockIdx.x; b < nb; b += gridDim.x) msched_entry(descs, bmap[b], b, parity, list, &lc[parity]);
        target += gridDim.x;
        if (!grid_barrier(bar, target, abort)) return;
        mfim_round(descs, ndesc, parity, list, list_cap, &lc[parity], &lc[parity ^ 1], blockIdx.x, gridDim.x, su);
        target += gridDim.x;
        if (!grid_barrier(bar, target, abort)) return;
    }
}

// ---------------------------------------------------------------------------------------------
// Setup of a group of admitted requests: each kernel below covers every request of the group in
// one launch.  The per-tile kernels run one 64-thread block per tile of each request (request j
// owns blocks [jobs[j].tile_base, jobs[j + 1].tile_base)), the seed kernel one block per seed
// (seed_base); the clear and scan kernels take the request from blockIdx.y / blockIdx.x.
// ---------------------------------------------------------------------------------------------
struct SetupReport {
    int status[2];   // Selling status (largest code) and smallest failing row-major index
    int total;       // dependency-list entries
    int npts;        // in-grid points with a nonzero stencil
    double tau_sum;  // sum of their one-step times 1/sqrt(sum_k w_k)
    double pad_;
};

// Tiles per block of the group scan (four per thread, 256 threads), and the sums the first scan
// kernel leaves per chunk.
constexpr int GSCAN_CHUNK = 1024;
struct ChunkSum {
    int total;   // dependency counts
    int npts;    // points with a nonzero stencil
    double tau;  // their one-step times
};

struct SetupJob {
    Geom g;
    int metric_is_field, dual, nseeds;
    long long np;
    InvHH ih;
    const double *metric;
    uint32_t *offs;
    float *w, *u;
    unsigned long long *fixedmask;
    unsigned *keys, *tmax;
    int *dep_cnt, *dep_start, *dep_fill, *deps;
    const long long *seed_p;
    const double *seed_v;
    unsigned *seed_key;
    SetupReport *rep;
    ChunkSum *chunks;  // ceil(ntiles / GSCAN_CHUNK)
    int tile_base, seed_base, chunk_base;
};

// Request of block b: the last one whose first block is <= b (requests without blocks in this
// launch have the same base as the next one and are passed over).
enum { BY_TILE, BY_SEED, BY_CHUNK };
__device__ inline int job_of(const SetupJob *jobs, int njobs, int b, int by) {
    auto base = [&](int j) { return by == BY_TILE ? jobs[j].tile_base : by == BY_SEED ? jobs[j].seed_base : jobs[j].chunk_base; };
    int j = 0;
    while (j + 1 < njobs && base(j + 1) <= b) ++j;
    return j;
}

// Sums over a 256-thread block, in a fixed order (every thread calls them; result on every thread).
__device__ inline int block_sum_i(int v, int *lds4) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    return lds4[0] + lds4[1] + lds4[2] + lds4[3];
}
__device__ inline double block_sum_d(double v, double *lds4) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    return (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
}

// Report words, seed key and the tile-graph counters of every request (blockIdx.y).
__global__ __launch_bounds__(256) void k_gsetup_clear(const SetupJob *__restrict__ jobs) {
    const SetupJob &J = jobs[blockIdx.y];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        J.rep->status[0] = 0;
        J.rep->status[1] = 0x7fffffff;
        *J.seed_key = KEY_NONE;
    }
    const int nt = J.g.ntiles;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += gridDim.x * blockDim.x) {
        J.dep_cnt[i] = 0;
        J.dep_fill[i] = 0;
    }
}

// Stencils of one tile's points (FP64 Selling, as k_riemann_stencil) and, from the same
// registers, the tile's source tiles (as the first pass of k_tile_deps): one count per source,
// and the list itself kept for the fill when it has at most 64 entries (64 ints a tile in u, its
// length in tmax).  The tile's sum and count of one-step times go to fixedmask and keys for the
// scan kernel.  k_gsetup_fill_init overwrites all four after they have been read.
template <int D>
__device__ inline void gsetup_stencil_tile(const SetupJob &J, int tile, int lane, int *set) {
    constexpr int K = D * (D + 1) / 2;
    constexpr int M = D == 2 ? MODEL_RIEMANN2 : MODEL_RIEMANN3;
    const Geom &g = J.g;
    const long long np = J.np, p = (long long)tile * 64 + lane;
    int c[3];
    tiled_coords(g, tile, lane, c);
    double tau = 0.;
    int cnt = 0;
    src_set_clear(set, lane);
    __syncthreads();
    if (!in_grid(g, c)) {
        for (int k = 0; k < K; ++k) {
            J.offs[k * np + p] = 0;
            J.w[k * np + p] = 0.f;
        }
    } else {
        long long lin;
        int8_t o8[K][D];
        double wd[K];
        const int st = riemann_point<D>(g, J.metric, J.metric_is_field, J.dual, J.ih, c, o8, wd, lin);
        double sw = 0.;
        for (int k = 0; k < K; ++k) {
            const float wk = (float)wd[k];
            J.offs[k * np + p] = packed_offset<D>(o8, k);
            J.w[k * np + p] = wk;
            sw += wd[k];
            if (wk != 0.f) src_set_add_sides<M>(g, c, tile, o8[k][0], o8[k][1], D > 2 ? o8[k][D > 2 ? 2 : 0] : 0, 2, set);
        }
        if (st != SELLING_OK) {
            atomicMax(&J.rep->status[0], st);
            atomicMin(&J.rep->status[1], (int)lin);
        }
        if (sw > 0.) {
            tau = 1. / sqrt(sw);
            cnt = 1;
        }
    }
    __syncthreads();
    const int nsrc = src_set_compact(set, lane);
    for (int i = lane; i < nsrc; i += 64) atomicAdd(&J.dep_cnt[set[i]], 1);
    if (nsrc <= 64 && lane < nsrc) reinterpret_cast<int *>(J.u)[p] = set[lane];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        tau += __shfl_xor(tau, o);
        cnt += __shfl_xor(cnt, o);
    }
    if (lane == 0) {
        J.tmax[tile] = (unsigned)nsrc;
        J.fixedmask[tile] = (unsigned long long)__double_as_longlong(tau);
        J.keys[tile] = (unsigned)cnt;
    }
}

__global__ __launch_bounds__(64) void k_gsetup_stencil(const SetupJob *__restrict__ jobs, int njobs) {
    __shared__ int set[HASH_SLOTS];
    const SetupJob &J = jobs[job_of(jobs, njobs, blockIdx.x, BY_TILE)];
    const int tile = blockIdx.x - J.tile_base;
    if (J.g.dim == 3) gsetup_stencil_tile<3>(J, tile, threadIdx.x, set);
    else gsetup_stencil_tile<2>(J, tile, threadIdx.x, set);
}

// Exclusive scan of every request's dependency counts into dep_start (dep_start[ntiles] =
// total) in two launches over chunks of GSCAN_CHUNK tiles, one 256-thread block per chunk of
// each request (chunk_base: the request's first block).  The counts are final when the first
// launch starts, so no block waits for another: the first launch sums each chunk (and its
// one-step times), the second takes its carry as the sum of the totals of the request's earlier
// chunks and scans its own chunk.  The last chunk's block writes the request's report.
__global__ __launch_bounds__(256) void k_gsetup_chunks(const SetupJob *__restrict__ jobs, int njobs) {
    __shared__ int li[4];
    __shared__ double ld[4];
    const SetupJob &J = jobs[job_of(jobs, njobs, blockIdx.x, BY_CHUNK)];
    const int c = blockIdx.x - J.chunk_base;
    const int n = J.g.ntiles;
    const int i0 = c * GSCAN_CHUNK + threadIdx.x * 4;
    int sum = 0, npts = 0;
    double tau = 0.;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (i0 + k < n) {
            sum += J.dep_cnt[i0 + k];
            npts += (int)J.keys[i0 + k];
            tau += __longlong_as_double((long long)J.fixedmask[i0 + k]);
        }
    }
    sum = block_sum_i(sum, li);
    npts = block_sum_i(npts, li);
    tau = block_sum_d(tau, ld);
    if (threadIdx.x == 0) J.chunks[c] = ChunkSum{sum, npts, tau};
}

__global__ __launch_bounds__(256) void k_gsetup_scan(const SetupJob *__restrict__ jobs, int njobs) {
    __shared__ int li[4];
    __shared__ double ld[4];
    const SetupJob &J = jobs[job_of(jobs, njobs, blockIdx.x, BY_CHUNK)];
    const int c = blockIdx.x - J.chunk_base;
    const int n = J.g.ntiles;
    const int nchunks = (n + GSCAN_CHUNK - 1) / GSCAN_CHUNK;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int carry = 0;
    for (int k = tid; k < c; k += blockDim.x) carry += J.chunks[k].total;
    carry = block_sum_i(carry, li);
    const int i0 = c * GSCAN_CHUNK + tid * 4;
    int v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = i0 + k < n ? J.dep_cnt[i0 + k] : 0;
    const int tsum = v[0] + v[1] + v[2] + v[3];
    int incl = tsum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    __syncthreads();
    if (lane == 63) li[wid] = incl;
    __syncthreads();
    int run = carry + incl - tsum;
    for (int w = 0; w < wid; ++w) run += li[w];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (i0 + k < n) J.dep_start[i0 + k] = run;
        run += v[k];
    }
    if (c != nchunks - 1) return;  // uniform over the block
    // last chunk: run on the last thread is the request's total
    int npts = 0;
    double tau = 0.;
    for (int k = tid; k < nchunks; k += blockDim.x) {
        npts += J.chunks[k].npts;
        tau += J.chunks[k].tau;
    }
    npts = block_sum_i(npts, li);
    tau = block_sum_d(tau, ld);
    if (tid == blockDim.x - 1) {
        J.dep_start[n] = run;
        J.rep->total = run;
        J.rep->npts = npts;
        J.rep->tau_sum = tau;
    }
}

// Dependency lists of one tile's sources (the second pass of k_tile_deps: from the kept list,
// or from the stored stencil when the tile has more than 64 sources), then the tile's initial
// state (as k_init): values +inf, padding fixed, no key.
template <int M>
__device__ inline void gsetup_fill_tile(const SetupJob &J, int tile, int lane, int *set) {
    const int ns = (int)J.tmax[tile];
    if (ns <= 64) {
        if (lane < ns) {
            const int src = reinterpret_cast<const int *>(J.u)[(long long)tile * 64 + lane];
            const int q = atomicAdd(&J.dep_fill[src], 1);
            J.deps[J.dep_start[src] + q] = tile;
        }
    } else {
        const StencilData<float> sd{J.offs, J.w, nullptr, nullptr, nullptr, J.np};
        const int nsrc = tile_sources<M, float>(J.g, sd, tile, lane, set);
        for (int i = lane; i < nsrc; i += 64) {
            const int src = set[i];
            const int q = atomicAdd(&J.dep_fill[src], 1);
            J.deps[J.dep_start[src] + q] = tile;
        }
    }
}

__global__ __launch_bounds__(64) void k_gsetup_fill_init(const SetupJob *__restrict__ jobs, int njobs) {
    __shared__ int set[HASH_SLOTS];
    const SetupJob &J = jobs[job_of(jobs, njobs, blockIdx.x, BY_TILE)];
    const int tile = blockIdx.x - J.tile_base;
    const int lane = threadIdx.x;
    if (J.g.dim == 3) gsetup_fill_tile<MODEL_RIEMANN3>(J, tile, lane, set);
    else gsetup_fill_tile<MODEL_RIEMANN2>(J, tile, lane, set);
    __syncthreads();  // the tile's kept list (in u and tmax) is read before it is overwritten
    const long long p = (long long)tile * 64 + lane;
    J.u[p] = dev_inf<float>();
    int c[3];
    tiled_coords(J.g, tile, lane, c);
    const unsigned long long pad = __ballot(!in_grid(J.g, c));
    if (lane == 0) {
        J.fixedmask[tile] = pad;
        J.keys[tile] = KEY_NONE;
        J.tmax[tile] = KEY_NONE;
    }
}

// One block per seed: its value and fixed bit (as k_set_seeds), and its tile and the tile's
// dependents marked pending with the seed's key (as k_seed_activate); the smallest key left
// pending goes to the request's seed key.
__global__ __launch_bounds__(64) void k_gsetup_seeds(const SetupJob *__restrict__ jobs, int njobs) {
    const SetupJob &J = jobs[job_of(jobs, njobs, blockIdx.x, BY_SEED)];
    const int i = blockIdx.x - J.seed_base;
    const int lane = threadIdx.x;
    const long long p = J.seed_p[i];
    const double v = J.seed_v[i];
    if (lane == 0) {
        J.u[p] = (float)v;
        atomicOr(&J.fixedmask[p >> 6], 1ull << (p & 63));
    }
    const int tile = (int)(p >> 6);
    const unsigned key = key_of((float)v);
    unsigned wmin = mark_pending(lane == 0 ? tile : -1, key, J.keys, J.tmax);
    const int b = J.dep_start[tile], e = J.dep_start[tile + 1];
    for (int j0 = b; j0 < e; j0 += 64) {
        const int j = j0 + lane;
        const unsigned m = mark_pending(j < e ? J.deps[j] : -1, key, J.keys, J.tmax);
        wmin = m < wmin ? m : wmin;
    }
    if (lane == 0 && wmin != KEY_NONE) atomicMin(J.seed_key, wmin);
}

// Scheduling state of the solves joining the running set at a round of the given parity (one
// block per descriptor; the others return): each seed key (smallest key the seed marks left
// pending) becomes the pending minimum of that round.
__global__ __launch_bounds__(64) void k_mjoin(const SolveDesc *__restrict__ descs, int parity) {
    const SolveDesc &d = descs[blockIdx.x];
    if (!d.joining) return;
    const int i = threadIdx.x;
    for (int j = i; j < 2 * PEND_SLOTS; j += 64)
        d.pend[(j / PEND_SLOTS) * PEND_WORDS + (j % PEND_SLOTS) * PEND_STRIDE] =
            (j == parity * PEND_SLOTS) ? *d.seed_key : KEY_NONE;
    if (i == 0) *d.visits = 0ull;
    if (i < 2) d.count[i] = 0;
    if (i < 4) d.bound[i] = i < 2 ? -__builtin_huge_valf() : d.delta;
}

// Descriptor indices of the converged solves exported by one launch.
constexpr int EXPORT_MAX = 32;
struct ExportList {
    int idx[EXPORT_MAX];
};

// Tiled FP32 values -> row-major FP64 of every listed solve (blockIdx.y: list entry).
__global__ __launch_bounds__(256) void k_mexport(const SolveDesc *__restrict__ descs, ExportList list) {
    const SolveDesc &d = descs[list.idx[blockIdx.y]];
    const Geom &g = d.ctx.g;
    const long long np = (long long)g.ntiles * 64;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long long)gridDim.x * blockDim.x) {
        int c[3];
        tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
        if (!in_grid(g, c)) continue;
        d.values_dev[((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2]] = (double)d.ctx.u[p];
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
    add(4);                 // seed key
    add(((size_t)(nt + GSCAN_CHUNK - 1) / GSCAN_CHUNK) * sizeof(ChunkSum));  // scan chunk sums
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
    SetupJob *d_jobs[NSETUP] = {};         // per thread: the group's request table
    SetupJob *h_jobs[NSETUP] = {};         // pinned source of its uploads
    SetupReport *d_rep[NSETUP] = {};       // per thread: SETUP_GROUP reports
    SetupReport *h_rep[NSETUP] = {};       // pinned copies
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
    // persistent batches (k_mbatch): on/off, grid (blocks resident at once), barrier counter
    // (monotonic: each launch waits for bar_count + 2 * rounds * grid arrivals), abort flag and
    // its per-batch host copies
    std::atomic<int> persistent{0};
    int grid_persist = 0;
    unsigned long long *d_bar = nullptr;
    unsigned long long bar_count = 0;
    int *d_abort = nullptr;
    int *h_abort[2] = {};
    // Device memory of the solves (arenas, dependency lists, staged metrics, the round list):
    // the server's own stream-ordered pool, which keeps freed memory for reuse (release
    // threshold: unlimited) until the server is destroyed.  With the default pool's threshold
    // of 0 a free returned memory to the system, and hipFreeAsync / hipMallocAsync calls of
    // 1 ms or more held up the compute and setup threads.
    hipMemPool_t pool = nullptr;
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

// Setup of a group of admitted requests on one setup stream, one launch per kernel for the whole
// group.  Phase 1: clears, stencils with the tile sources, scans; the reports (status, dependency
// total, one-step times) are copied back; one synchronisation.  Phase 2 allocates the dependency
// lists of the requests whose stencils succeeded, fills them and initialises values, then marks
// the seeds; one synchronisation.  Failed requests are reported in `err` (empty: ready).
struct SetupItem {
    std::shared_ptr<RequestRec> r;
    Slot *sl = nullptr;
    std::string err;
};

void setup_group(hfm_server *S, std::vector<SetupItem> &items, int which) {
    hipStream_t s = S->s_setup[which];
    const int n = (int)items.size();
    SetupJob *hj = S->h_jobs[which];
    int tiles = 0, max_nt = 0, nchunk = 0;
    for (int g = 0; g < n; ++g) {
        const std::shared_ptr<RequestRec> &r = items[g].r;
        Slot *sl = new Slot;
        items[g].sl = sl;
        sl->rec = r;
        HIPCHECK(hipMallocFromPoolAsync((void **)&sl->arena, arena_bytes(*r), S->pool, s));
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
        sl->seed_key = carve<unsigned>(p, 1);
        ChunkSum *chunks = carve<ChunkSum>(p, (r->g.ntiles + GSCAN_CHUNK - 1) / GSCAN_CHUNK);
        const double *metric = r->metric_dev;
        if (!metric) {
            const size_t mcount = r->metric_is_field ? (size_t)r->N * r->K : (size_t)r->K;
            HIPCHECK(hipMallocFromPoolAsync((void **)&sl->metric_stage, mcount * sizeof(double), S->pool, s));
            HIPCHECK(hipMemcpyAsync(sl->metric_stage, r->metric_host.data(), mcount * sizeof(double),
                                    hipMemcpyHostToDevice, s));
            metric = sl->metric_stage;
        }
        HIPCHECK(hipMemcpyAsync(sl->seed_p, r->seed_blob.data(), r->seed_blob.size() * 8, hipMemcpyHostToDevice, s));
        SetupJob &J = hj[g];
        J = SetupJob{};
        J.g = r->g;
        J.metric_is_field = r->metric_is_field;
        J.dual = r->dual;
        J.nseeds = (int)ns;
        J.np = r->np;
        const int D = r->g.dim;
        for (int i = 0; i < D; ++i)
            for (int j = 0; j <= i; ++j) J.ih.v[i * (i + 1) / 2 + j] = 1. / (r->h * r->h);
        J.metric = metric;
        J.offs = sl->offs;
        J.w = sl->w;
        J.u = sl->u;
        J.fixedmask = sl->fixedmask;
        J.keys = sl->keys;
        J.tmax = sl->tmax;
        J.dep_cnt = sl->dep_cnt;
        J.dep_start = sl->dep_start;
        J.dep_fill = sl->dep_fill;
        J.seed_p = sl->seed_p;
        J.seed_v = sl->seed_v;
        J.seed_key = sl->seed_key;
        J.rep = S->d_rep[which] + g;
        J.chunks = chunks;
        J.tile_base = tiles;
        J.chunk_base = nchunk;
        tiles += r->g.ntiles;
        nchunk += (r->g.ntiles + GSCAN_CHUNK - 1) / GSCAN_CHUNK;
        max_nt = std::max(max_nt, r->g.ntiles);
    }
    HIPCHECK(hipMemcpyAsync(S->d_jobs[which], hj, sizeof(SetupJob) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gsetup_clear, dim3((unsigned)std::min((max_nt + 255) / 256, 64), (unsigned)n), dim3(256), 0, s,
                       S->d_jobs[which]);
    hipLaunchKernelGGL(k_gsetup_stencil, dim3((unsigned)tiles), dim3(64), 0, s, S->d_jobs[which], n);
    hipLaunchKernelGGL(k_gsetup_chunks, dim3((unsigned)nchunk), dim3(256), 0, s, S->d_jobs[which], n);
    hipLaunchKernelGGL(k_gsetup_scan, dim3((unsigned)nchunk), dim3(256), 0, s, S->d_jobs[which], n);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(S->h_rep[which], S->d_rep[which], sizeof(SetupReport) * n, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    // Phase 2 over the requests whose stencils succeeded, compacted to the front of the table.
    int m = 0, seeds = 0;
    tiles = 0;
    for (int g = 0; g < n; ++g) {
        Slot *sl = items[g].sl;
        const SetupReport &rep = S->h_rep[which][g];
        if (rep.status[0] != SELLING_OK) {
            items[g].err = rep.status[0] == SELLING_NOT_SPD
                               ? "ObtuseSuperbase error : input must be positive definite"
                               : "ObtuseSuperbase error : did not terminate in 200 iterations";
            if (sl->metric_stage) HIPCHECK(hipFreeAsync(sl->metric_stage, s));
            HIPCHECK(hipFreeAsync(sl->arena, s));
            delete sl;
            items[g].sl = nullptr;
            continue;
        }
        HIPCHECK(hipMallocFromPoolAsync((void **)&sl->deps, sizeof(int) * std::max(rep.total, 1), S->pool, s));
        const double tau = rep.npts > 0 ? rep.tau_sum / rep.npts : 0.;
        sl->delta = (float)(8. * tau);
        if (m != g) hj[m] = hj[g];
        SetupJob &J = hj[m++];
        J.deps = sl->deps;
        J.tile_base = tiles;
        J.seed_base = seeds;
        tiles += J.g.ntiles;
        seeds += J.nseeds;
    }
    if (m > 0) {
        HIPCHECK(hipMemcpyAsync(S->d_jobs[which], hj, sizeof(SetupJob) * m, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gsetup_fill_init, dim3((unsigned)tiles), dim3(64), 0, s, S->d_jobs[which], m);
        if (seeds > 0) hipLaunchKernelGGL(k_gsetup_seeds, dim3((unsigned)seeds), dim3(64), 0, s, S->d_jobs[which], m);
        HIPCHECK(hipGetLastError());
    }
    for (int g = 0; g < n; ++g) {
        Slot *sl = items[g].sl;
        if (sl && sl->metric_stage) {
            HIPCHECK(hipFreeAsync(sl->metric_stage, s));
            sl->metric_stage = nullptr;
        }
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
    int idx;
    if (!S->free_idx.empty()) {
        idx = S->free_idx.back();
        S->free_idx.pop_back();
    } else {
        idx = S->ndesc++;
    }
    sl->idx = idx;
    sl->first_batch = S->nlaunched;
    unsigned *pend = S->d_pend + (size_t)idx * 2 * PEND_WORDS;
    float *bound = S->d_bound + (size_t)idx * 4;
    int *count = S->d_count + (size_t)idx * 2;
    unsigned long long *st = S->d_stats + idx;
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
    d.seed_key = sl->seed_key;
    d.values_dev = sl->values_dev;
    d.joining = 1;  // k_mjoin at the start of the next batch sets the scheduling state
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
        HIPCHECK(hipMallocFromPoolAsync((void **)&S->d_list, S->list_cap * sizeof(int), S->pool, s));
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
    bool joining = false;
    for (int i = 0; i < S->ndesc; ++i) {
        joining |= S->h_desc[i].joining != 0;
        S->h_desc[i].joining = 0;
    }
    std::memcpy(S->h_bmap_stage[par], bm.data(), bm.size() * sizeof(int2));
    HIPCHECK(hipMemcpyAsync(S->d_desc, S->h_desc_stage[par], sizeof(SolveDesc) * S->ndesc, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(S->d_bmap, S->h_bmap_stage[par], bm.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    const unsigned grid = (unsigned)std::min<long long>((tiles + 3) / 4, 2048);
    if (joining) {
        hipLaunchKernelGGL(k_mjoin, dim3((unsigned)S->ndesc), dim3(64), 0, s, S->d_desc, (int)(S->round & 1));
        HIPCHECK(hipGetLastError());
    }
    const bool persistent = S->persistent.load() && S->grid_persist > 0;
    if (persistent) {
        const int G = S->grid_persist;
        hipLaunchKernelGGL(k_mbatch, dim3((unsigned)G), dim3(SCHED_BLOCK), 0, s, S->d_desc, S->d_bmap,
                           (int)bm.size(), S->ndesc, (int)(S->round & 1), batch, S->d_list, S->list_cap, S->d_lc,
                           S->d_bar, S->bar_count, S->d_abort);
        HIPCHECK(hipGetLastError());
        S->bar_count += 2ull * (unsigned long long)batch * (unsigned long long)G;
        S->round += batch;
    }
    for (int b = 0; b < batch && !persistent; ++b) {
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
    HIPCHECK(hipMemcpyAsync(S->h_abort[par], S->d_abort, sizeof(int), hipMemcpyDeviceToHost, s));
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
    if (*S->h_abort[par])
        throw HipFail{"persistent round kernel: a grid-wide barrier did not complete (blocks not resident)"};
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
    // One export launch per EXPORT_MAX converged solves.  The descriptor table on the device is
    // the one uploaded with the latest batch, which still lists them.
    for (size_t b = 0; b < done.size(); b += EXPORT_MAX) {
        ExportList el{};
        long long np_max = 0;
        const int ne = (int)std::min<size_t>(EXPORT_MAX, done.size() - b);
        for (int i = 0; i < ne; ++i) {
            el.idx[i] = done[b + i]->idx;
            np_max = std::max(np_max, done[b + i]->rec->np);
        }
        hipLaunchKernelGGL(k_mexport, dim3((unsigned)std::min<long long>((np_max + 255) / 256, 1024), (unsigned)ne),
                           dim3(256), 0, S->s_comp, S->d_desc, el);
        HIPCHECK(hipGetLastError());
    }
    for (Slot *sl : done) {
        S->h_desc[sl->idx].active = 0;
        finish_export(sl, S->h_stats[par]);
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
        {
            int dev = 0;
            HIPCHECK(hipGetDevice(&dev));
            hipMemPoolProps props = {};
            props.allocType = hipMemAllocationTypePinned;
            props.handleTypes = hipMemHandleTypeNone;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = dev;
            HIPCHECK(hipMemPoolCreate(&S->pool, &props));
            uint64_t keep = UINT64_MAX;
            HIPCHECK(hipMemPoolSetAttribute(S->pool, hipMemPoolAttrReleaseThreshold, &keep));
        }
        S->budget = hbm_budget_bytes;
        S->max_resident = std::min(max_resident, MAX_SLOTS);
        S->stats.hbm_budget = hbm_budget_bytes;
        for (int i = 0; i < hfm_server::NSETUP; ++i) {
            HIPCHECK(hipStreamCreateWithFlags(&S->s_setup[i], hipStreamNonBlocking));
            S->d_jobs[i] = dalloc<SetupJob>(hfm_server::SETUP_GROUP);
            HIPCHECK(hipHostMalloc((void **)&S->h_jobs[i], sizeof(SetupJob) * hfm_server::SETUP_GROUP,
                                   hipHostMallocDefault));
            S->d_rep[i] = dalloc<SetupReport>(hfm_server::SETUP_GROUP);
            HIPCHECK(hipHostMalloc((void **)&S->h_rep[i], sizeof(SetupReport) * hfm_server::SETUP_GROUP,
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
        S->d_bar = dalloc<unsigned long long>(1);
        S->d_abort = dalloc<int>(1);
        HIPCHECK(hipMemset(S->d_bar, 0, sizeof(unsigned long long)));
        HIPCHECK(hipMemset(S->d_abort, 0, sizeof(int)));
        {
            // persistent grid: blocks of k_mbatch resident at once on the whole device
            int dev = 0, per_cu = 0;
            hipDeviceProp_t prop;
            HIPCHECK(hipGetDevice(&dev));
            HIPCHECK(hipGetDeviceProperties(&prop, dev));
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mbatch, SCHED_BLOCK, 0));
            S->grid_persist = per_cu * prop.multiProcessorCount;
        }
        for (int i = 0; i < 2; ++i) {
            HIPCHECK(hipHostMalloc((void **)&S->h_count[i], sizeof(int) * MAX_SLOTS * 2, hipHostMallocDefault));
            HIPCHECK(hipHostMalloc((void **)&S->h_abort[i], sizeof(int), hipHostMallocDefault));
            *S->h_abort[i] = 0;
            HIPCHECK(hipHostMalloc((void **)&S->h_stats[i], sizeof(unsigned long long) * MAX_SLOTS,
                                   hipHostMallocDefault));
        }
        for (int i = 0; i < hfm_server::NSETUP; ++i) S->th_setup[i] = std::thread(setup_loop, S, i);
        S->th_comp = std::thread(compute_loop, S);
    } catch (const HipFail &e) {
        set_error(e.msg);
        hfm_server_destroy(S);  // no worker thread started: frees what was allocated
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

int hfm_server_set_persistent(hfm_server *S, int on) {
    if (!S) {
        set_error("hfm_server_set_persistent: invalid argument");
        return 1;
    }
    if (on && S->grid_persist <= 0) {
        set_error("hfm_server_set_persistent: the device cannot hold the persistent round kernel");
        return 1;
    }
    S->persistent.store(on ? 1 : 0);
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
    if (S->d_list) {
        (void)hipFreeAsync(S->d_list, S->s_comp);
        (void)hipStreamSynchronize(S->s_comp);
    }
    for (int i = 0; i < hfm_server::NSETUP; ++i) {
        dfree(S->d_jobs[i]);
        if (S->h_jobs[i]) (void)hipHostFree(S->h_jobs[i]);
        dfree(S->d_rep[i]);
        if (S->h_rep[i]) (void)hipHostFree(S->h_rep[i]);
    }
    dfree(S->d_desc);
    dfree(S->d_lc);
    dfree(S->d_count);
    dfree(S->d_pend);
    dfree(S->d_bound);
    dfree(S->d_stats);
    dfree(S->d_bar);
    dfree(S->d_abort);
    delete[] S->h_desc;
    if (S->d_bmap) (void)hipFree(S->d_bmap);
    for (int i = 0; i < 2; ++i) {
        if (S->h_desc_stage[i]) (void)hipHostFree(S->h_desc_stage[i]);
        if (S->h_bmap_stage[i]) (void)hipHostFree(S->h_bmap_stage[i]);
        if (S->h_count[i]) (void)hipHostFree(S->h_count[i]);
        if (S->h_abort[i]) (void)hipHostFree(S->h_abort[i]);
        if (S->h_stats[i]) (void)hipHostFree(S->h_stats[i]);
        if (S->batch_done[i]) (void)hipEventDestroy(S->batch_done[i]);
    }
    if (S->s_comp) (void)hipStreamDestroy(S->s_comp);
    for (int i = 0; i < hfm_server::NSETUP; ++i)
        if (S->s_setup[i]) (void)hipStreamDestroy(S->s_setup[i]);
    // every allocation from the pool has been freed on a stream synchronised above
    if (S->pool) (void)hipMemPoolDestroy(S->pool);
    delete S;
}

}  // extern "C"
