// hfm_kernels.hip -- HIP kernels of the GPU Fast-Marching library (gfx950, wave64).
//
//   k_riemann_stencil   per-point FP64 Selling reduction -> int8 offsets + weights (tiled)
//   k_tile_deps         tile dependency graph: which tiles read a tile (one wave per tile)
//   k_fim               one round of the block fast-iterative method over the active-tile list
//   k_export_values     tiled values -> row-major FP64 'values'
//   k_flow              geodesic flow / active-neighbour bits from the converged values
//   k_adjoint_*         backward differentiation on the same tiles (reverse causal order)
//
// All kernels are launched through the host functions at the bottom of this file, which are the
// only symbols the plan code (hfm_plan.hip) uses.
#include "hfm_kernels.h"

#include <hip/hip_ext.h>

#include <algorithm>

#include "hfm_device.h"
#include "selling.h"
#include "hfm_tile.h"
#include "hfm_setup_dev.h"

namespace hfm {

// ---------------------------------------------------------------------------------------------
// Stencil construction (Riemann2 / Riemann3).
// ---------------------------------------------------------------------------------------------
template <int D, typename T>
__global__ __launch_bounds__(256) void k_riemann_stencil(Geom g, const double *__restrict__ metric, int metric_is_field,
                                                         int dual, InvHH ih, uint32_t *__restrict__ offs,
                                                         T *__restrict__ w, long long np, int *__restrict__ status,
                                                         RedSlot *__restrict__ red) {
    constexpr int K = D * (D + 1) / 2;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    double tau_v = 0., tau_c = 0.;
    int l1max = 0;
    if (!in_grid(g, c)) {
        for (int k = 0; k < K; ++k) {
            offs[k * np + p] = 0;
            w[k * np + p] = T(0);
        }
    } else {
        long long lin;
        int8_t o8[K][D];
        double wd[K];
        const int st = riemann_point<D>(g, metric, metric_is_field, dual, ih, c, o8, wd, lin);
        for (int k = 0; k < K; ++k) {
            offs[k * np + p] = packed_offset<D>(o8, k);
            w[k * np + p] = (T)wd[k];
        }
        if (st != SELLING_OK) {
            // status[0]: error code, status[1]: smallest failing row-major index
            atomicMax(&status[0], st);
            atomicMin(&status[1], (int)lin);
        }
        // MaxStencilWidth (HamiltonFastMarching.hxx:36-44): the largest L1 norm among the
        // reversed offsets, which hold a nonzero-weight side x +- e_k only where x +- e_k lies in
        // the domain (HFM_StencilRecompute.hxx:91-101; walls are attached later and do not count)
        for (int k = 0; k < K; ++k) {
            if (wd[k] == 0.) continue;
            const int o0 = o8[k][0], o1 = D > 1 ? o8[k][D > 1 ? 1 : 0] : 0, o2 = D > 2 ? o8[k][D > 2 ? 2 : 0] : 0;
            int r[3];
            if (!shift_coords(g, c, o0, o1, o2, r) && !shift_coords(g, c, -o0, -o1, -o2, r)) continue;
            int l1 = 0;
            for (int q = 0; q < D; ++q) l1 += o8[k][q] < 0 ? -o8[k][q] : o8[k][q];
            l1max = l1 > l1max ? l1 : l1max;
        }
        // Mean one-step time 1/sqrt(sum_k w_k): sets the width of the scheduling bands.
        double sw = 0.;
        for (int k = 0; k < K; ++k) sw += wd[k];
        if (sw > 0.) {
            tau_v = 1. / sqrt(sw);
            tau_c = 1.;
        }
    }
    wave_add_tau(tau_v, tau_c, red);
    wave_max_l1(l1max, red);
}

// ---------------------------------------------------------------------------------------------
// Tile dependency graph.  For tile T, its sources are the tiles holding x +- e_k for the
// points x of T (nonzero weights only).  deps(S) = tiles T != S having S as a source: the tiles
// to re-run when S changes.  Pass 0 counts, pass 1 fills (CSR offsets from an exclusive scan).
// ---------------------------------------------------------------------------------------------
template <int ModelT, typename T>
__global__ __launch_bounds__(64) void k_tile_deps(Geom g, StencilData<T> sd, int pass, int *__restrict__ dep_count,
                                                  const int *__restrict__ dep_start, int *__restrict__ dep_fill,
                                                  int *__restrict__ deps, int *__restrict__ srcbuf,
                                                  int *__restrict__ srcn) {
    __shared__ int set[HASH_SLOTS];
    const int tile = blockIdx.x;
    const int lane = threadIdx.x;
    // With srcbuf, pass 0 keeps each tile's source list when it has at most 64 entries (64 per
    // tile), and pass 1 fills from it instead of recomputing the stencil walk and the hash.
    if (pass == 1 && srcbuf) {
        const int ns = srcn[tile];
        if (ns <= 64) {
            if (lane < ns) {
                const int src = srcbuf[(long long)tile * 64 + lane];
                const int q = atomicAdd(&dep_fill[src], 1);
                deps[dep_start[src] + q] = tile;
            }
            return;
        }
    }
    const int nsrc = tile_sources<ModelT, T>(g, sd, tile, lane, set);
    for (int i = lane; i < nsrc; i += 64) {
        const int src = set[i];
        if (pass == 0) {
            atomicAdd(&dep_count[src], 1);
        } else {
            const int q = atomicAdd(&dep_fill[src], 1);
            deps[dep_start[src] + q] = tile;
        }
    }
    if (pass == 0 && srcbuf) {
        if (lane == 0) srcn[tile] = nsrc;
        if (nsrc <= 64 && lane < nsrc) srcbuf[(long long)tile * 64 + lane] = set[lane];
    }
}

// One block of 1024 sweeps the whole array chunk by chunk, carrying the running total; writes
// start[0..n].  Used for short arrays and for the chunk totals of the multi-block scan.
__global__ __launch_bounds__(1024) void k_exclusive_scan(const int *__restrict__ cnt, int *__restrict__ start, int n) {
    __shared__ int wsum[16];
    int carry = 0;
    for (int base = 0; base < n; base += SCAN_CHUNK) carry += scan_chunk(cnt, start, n, base, carry, wsum);
    if (threadIdx.x == 0) start[n] = carry;
}

// Multi-block scan, pass 1: the total of every chunk.
__global__ __launch_bounds__(1024) void k_scan_totals(const int *__restrict__ cnt, int n, int *__restrict__ totals) {
    __shared__ int wsum[16];
    const int t = scan_chunk(cnt, nullptr, n, blockIdx.x * SCAN_CHUNK, 0, wsum);
    if (threadIdx.x == 0) totals[blockIdx.x] = t;
}

// Pass 3: every chunk rescanned from its carry (the scanned chunk totals); start[n] = total.
__global__ __launch_bounds__(1024) void k_scan_apply(const int *__restrict__ cnt, const int *__restrict__ carries,
                                                     int *__restrict__ start, int n) {
    __shared__ int wsum[16];
    (void)scan_chunk(cnt, start, n, blockIdx.x * SCAN_CHUNK, carries[blockIdx.x], wsum);
    if (blockIdx.x == 0 && threadIdx.x == 0) start[n] = carries[gridDim.x];
}

// ---------------------------------------------------------------------------------------------
// Solve initialisation.
// ---------------------------------------------------------------------------------------------
// key_buffers: 1, or 2 for rounds that list their own tiles (keys[0..ntiles) and
// keys[ntiles..2 ntiles), taken from and marked into by alternate rounds).
template <typename T>
__global__ __launch_bounds__(256) void k_init(Geom g, T *__restrict__ u, unsigned long long *__restrict__ fixedmask,
                                              unsigned *__restrict__ keys, unsigned *__restrict__ tmax, long long np,
                                              int key_buffers) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    u[p] = dev_inf<T>();
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    const unsigned long long pad = __ballot(!in_grid(g, c));
    if ((threadIdx.x & 63) == 0) {
        const int tile = (int)(p >> 6);
        fixedmask[tile] = pad;
        keys[tile] = KEY_NONE;
        if (key_buffers > 1) keys[g.ntiles + tile] = KEY_NONE;
        tmax[tile] = KEY_NONE;
    }
}

// Solver state from given arrival times and active sets (row-major FP64): a reached point
// without active differences is a seed (fixed); padding stays +inf and fixed.
template <typename T>
__global__ __launch_bounds__(256) void k_load_state(Geom g, const double *__restrict__ values,
                                                    const double *__restrict__ active, T *__restrict__ u,
                                                    unsigned long long *__restrict__ fixedmask, long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;  // np is a multiple of 64: whole waves leave together
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    const bool in = in_grid(g, c);
    double v = __builtin_huge_val();
    bool fixed = !in;
    if (in) {
        const long long lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
        v = values[lin];
        fixed = active[lin] == 0. && v < __builtin_huge_val();
    }
    u[p] = (T)v;
    const unsigned long long m = __ballot(fixed);
    if ((threadIdx.x & 63) == 0) fixedmask[p >> 6] = m;
}

template <typename T>
hipError_t launch_load_state(const Geom &g, const double *values, const double *active, T *u,
                             unsigned long long *fixedmask, long long np, hipStream_t s) {
    hipLaunchKernelGGL((k_load_state<T>), dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, g, values, active, u,
                       fixedmask, np);
    return hipGetLastError();
}
template hipError_t launch_load_state<float>(const Geom &, const double *, const double *, float *,
                                             unsigned long long *, long long, hipStream_t);
template hipError_t launch_load_state<double>(const Geom &, const double *, const double *, double *,
                                              unsigned long long *, long long, hipStream_t);

template <typename T>
__global__ void k_set_seeds(const long long *__restrict__ seed_p, const double *__restrict__ seed_v, int nseeds,
                            T *__restrict__ u, unsigned long long *__restrict__ fixedmask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseeds) return;
    const long long p = seed_p[i];
    u[p] = (T)seed_v[i];
    atomicOr(&fixedmask[p >> 6], 1ull << (p & 63));
}

// Seed tiles and their dependents are the first pending tiles, keyed by the seed value: one
// wave per seed (block i takes seeds i, i + grid, ...).
__global__ __launch_bounds__(64) void k_seed_activate(const long long *__restrict__ seed_p,
                                                      const double *__restrict__ seed_v, int nseeds,
                                                      const int *__restrict__ dep_start, const int *__restrict__ deps,
                                                      unsigned *__restrict__ keys, const unsigned *__restrict__ tmax,
                                                      unsigned *__restrict__ pend_slots) {
    const int lane = threadIdx.x;
    unsigned wmin = KEY_NONE;
    for (int i = blockIdx.x; i < nseeds; i += gridDim.x) {
        const int tile = (int)(seed_p[i] >> 6);
        const unsigned key = key_of((float)seed_v[i]);
        unsigned m = mark_pending(lane == 0 ? tile : -1, key, keys, tmax);
        wmin = m < wmin ? m : wmin;
        const int b = dep_start[tile], e = dep_start[tile + 1];
        for (int j0 = b; j0 < e; j0 += 64) {
            const int j = j0 + lane;
            m = mark_pending(j < e ? deps[j] : -1, key, keys, tmax);
            wmin = m < wmin ? m : wmin;
        }
    }
    if (lane == 0 && wmin != KEY_NONE) atomicMin(&pend_slots[0], wmin);
}

// ---------------------------------------------------------------------------------------------
// Round scheduling, ordered by value bands.
//
// Every tile has a key: KEY_NONE, or a lower bound of the values its next visit can produce
// (the smallest new value among the changes that made it pending).  Round k lists the pending
// tiles whose key lies below the band bound B_k and consumes their keys; B_k stays at B_{k-1}
// while a pending key lies below it, and otherwise moves to (smallest pending key + delta).
// delta <= 0 lists every pending tile.  The ordering decides when work is done, never what the
// fixed point is.  The smallest pending key of round k+1 is accumulated in pend[(k+1)&1] by the
// keys this kernel leaves pending and by the marks of the update kernel of round k.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(SCHED_BLOCK) void k_schedule(int ntiles, unsigned *__restrict__ keys, SchedRound sr) {
    // the tile's key is loaded first, in parallel with the round parameters (their stores
    // below could otherwise alias it and order the load after them)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned k = KEY_NONE;
    if (t < ntiles) k = keys[t];
    const float d0 = *sr.delta0;
    float delta = d0;
    if (delta > 0.f) delta = adapt_delta(*sr.delta_in, *sr.prev_count, sr.target, 0.5f * d0, 64.f * d0);
    const float B = band_bound(key_value(read_pend(sr.pend_in, PEND_STRIDE)), *sr.bound_in, delta);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *sr.bound_out = B;
        *sr.delta_out = delta;
    }
    const bool take = band_take(k, B);
    const unsigned left = (!take && k != KEY_NONE) ? k : KEY_NONE;
    if (take) keys[t] = KEY_NONE;
    const int pos =
        block_take(take, left, sr.count, nullptr, &sr.pend_out[(blockIdx.x & (PEND_SLOTS - 1)) * PEND_STRIDE]);
    if (take) sr.list[pos] = t;
}

// ---------------------------------------------------------------------------------------------
// Update kernel of one round: every wave processes listed tiles with fim_tile (hfm_tile.h).
// ---------------------------------------------------------------------------------------------
template <int ModelT, typename T, int MAXSWEEP>
__global__ __launch_bounds__(256) void k_fim(Geom g, StencilData<T> sd, T *__restrict__ u,
                                             const unsigned long long *__restrict__ fixedmask,
                                             const int *__restrict__ dep_start, const int *__restrict__ deps,
                                             const int *__restrict__ list, const int *__restrict__ count,
                                             int *__restrict__ count_reset, unsigned *__restrict__ keys,
                                             unsigned *__restrict__ tmax, unsigned *__restrict__ pend_out,
                                             unsigned *__restrict__ pend_reset, unsigned long long *__restrict__ stats) {
    __shared__ T su[4][SU_ROW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int stride = gridDim.x * 4;
    int i = blockIdx.x * 4 + wv;
    // The list holds ntiles entries: read this wave's first entry together with the count (an
    // entry past the count is stale and never used) instead of after it.
    int next = i < g.ntiles ? list[i] : 0;
    const int n = *count;
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) *count_reset = 0;
        if (threadIdx.x < PEND_SLOTS) pend_reset[threadIdx.x * PEND_STRIDE] = KEY_NONE;
    }
    const TileCtx<T> ctx{g, sd, u, fixedmask, dep_start, deps, keys, tmax};
    unsigned long long my_sweeps = 0, my_visits = 0;
    unsigned wmin = KEY_NONE;
    for (; i < n; i += stride) {
        // the current entry goes to a scalar register before the next entry's load is issued,
        // so that nothing waits on the prefetch until the next iteration
        const int tile = __builtin_amdgcn_readfirstlane(next);
        next = list[i + stride < n ? i + stride : i];
        my_sweeps += fim_tile<ModelT, T, MAXSWEEP>(ctx, tile, lane, su[wv], wmin);
        my_visits += 1;
    }
    block_flush(wmin, my_visits, my_sweeps, pend_out, stats);
}

// ---------------------------------------------------------------------------------------------
// Update kernel that schedules its own round: block b owns tiles b, b + grid, b + 2 grid, ...;
// it lists those whose key lies below the round's band bound and its four waves visit them.
// Every block computes the same bound from the same inputs, and the separate scheduling kernel
// and its kernel boundary are gone from every round.
//
// Keys are double-buffered: round k takes from keys_in = keys[k % 2] and marks into
// keys_out = keys[(k + 1) % 2]; a pending tile it does not take moves its key to keys_out.  Only
// the owning block touches keys_in[t] during round k, so the take is a plain load and store.
// Every key a round takes was therefore written by an earlier launch, and every value that made
// it pending was stored before that launch ended: the kernel boundary (L2 write-back on every
// XCD at the end of a kernel, invalidation at the start of the next) makes those values visible
// to the visiting wave, with no cache maintenance inside the kernel.  A mark made during round k
// is consumed at the earliest by round k + 1.  Marking into a buffer nobody takes from in the
// same launch is what makes this hold: with a single buffer, a block could consume a mark made
// by a wave on another XCD in the same launch, then read that wave's values through a stale L2.
//
// The smallest pending key of round k + 1 is exact (leftover keys and marks both fold into
// pend_out), and the band bound exceeds it, so an empty round proves convergence.
// Round k reads pend[k % 4], accumulates pend[(k + 1) % 4] and counts[k % 4], and clears
// pend[(k + 2) % 4] and counts[(k + 1) % 4], which no block of round k reads or writes.
// ---------------------------------------------------------------------------------------------
template <int ModelT, typename T, int MAXSWEEP>
__global__ __launch_bounds__(256) void k_fim_sched(Geom g, StencilData<T> sd, T *__restrict__ u,
                                                   const unsigned long long *__restrict__ fixedmask,
                                                   const int *__restrict__ dep_start, const int *__restrict__ deps,
                                                   unsigned *__restrict__ keys_in, unsigned *__restrict__ keys_out,
                                                   unsigned *__restrict__ tmax, SchedRound sr,
                                                   int *__restrict__ count_reset,
                                                   unsigned *__restrict__ pend_reset,
                                                   unsigned long long *__restrict__ stats) {
    __shared__ T su[4][SU_ROW];
    __shared__ int blist[256];
    __shared__ int wcount[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int G = gridDim.x;
    const int owned = (g.ntiles - (int)blockIdx.x + G - 1) / G;
    // the first key is loaded before the round parameters it is compared with
    unsigned k0 = (int)threadIdx.x < owned ? keys_in[blockIdx.x + threadIdx.x * G] : KEY_NONE;
    const float d0 = *sr.delta0;
    float delta = d0;
    if (delta > 0.f) delta = adapt_delta(*sr.delta_in, *sr.prev_count, sr.target, 0.5f * d0, 64.f * d0);
    const float B = band_bound(key_value(read_pend(sr.pend_in, PEND_STRIDE)), *sr.bound_in, delta);
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            *sr.bound_out = B;
            *sr.delta_out = delta;
            *count_reset = 0;
        }
        if (threadIdx.x < PEND_SLOTS) pend_reset[threadIdx.x * PEND_STRIDE] = KEY_NONE;
    }
    const TileCtx<T> ctx{g, sd, u, fixedmask, dep_start, deps, keys_out, tmax};
    unsigned long long my_sweeps = 0, my_visits = 0;
    unsigned wmin = KEY_NONE;
    for (int c0 = 0; c0 < owned; c0 += 256) {
        const int j = c0 + (int)threadIdx.x;
        const int t = (int)blockIdx.x + j * G;
        const unsigned k = c0 == 0 ? k0 : (j < owned ? keys_in[t] : KEY_NONE);
        const bool take = band_take(k, B);
        const unsigned left = (!take && k != KEY_NONE) ? k : KEY_NONE;
        if (k != KEY_NONE) keys_in[t] = KEY_NONE;
        if (left != KEY_NONE) atomicMin(&keys_out[t], left);
        const unsigned long long bm = __ballot(take);
        const unsigned lmin = wave_min_u32(left);
        wmin = lmin < wmin ? lmin : wmin;
        if (lane == 0) wcount[wv] = __popcll(bm);
        __syncthreads();
        int base = 0, total = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            base += w < wv ? wcount[w] : 0;
            total += wcount[w];
        }
        const unsigned long long below = lane == 0 ? 0ull : (bm & (~0ull >> (64 - lane)));
        if (take) blist[base + __popcll(below)] = t;
        __syncthreads();
        if (threadIdx.x == 0 && total) atomicAdd(sr.count, total);
        for (int e = wv; e < total; e += 4) {
            const int tile = __builtin_amdgcn_readfirstlane(blist[e]);
            my_sweeps += fim_tile<ModelT, T, MAXSWEEP>(ctx, tile, lane, su[wv], wmin);
            my_visits += 1;
        }
        __syncthreads();  // blist and wcount are reused by the next chunk
    }
    block_flush(wmin, my_visits, my_sweeps, sr.pend_out, stats);
}

// ---------------------------------------------------------------------------------------------
// Exports.
// ---------------------------------------------------------------------------------------------
// Number of reached grid points (finite values): nAccepted of a solve without stopping criteria.
template <typename T>
__global__ __launch_bounds__(256) void k_count_reached(Geom g, const T *__restrict__ u, long long np,
                                                       RedSlot *__restrict__ red) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double c = 0.;
    if (p < np) {
        int c3[3];
        tiled_coords(g, (int)(p >> 6), (int)(p & 63), c3);
        c = (in_grid(g, c3) && u[p] < dev_inf<T>()) ? 1. : 0.;
    }
    wave_add_tau(0., c, red);
}

template <typename T> hipError_t launch_count_reached(const Geom &g, const T *u, long long np, RedSlot *red, hipStream_t s) {
    hipLaunchKernelGGL((k_count_reached<T>), dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, g, u, np, red);
    return hipGetLastError();
}
template hipError_t launch_count_reached<float>(const Geom &, const float *, long long, RedSlot *, hipStream_t);
template hipError_t launch_count_reached<double>(const Geom &, const double *, long long, RedSlot *, hipStream_t);

template <typename T>
__global__ __launch_bounds__(256) void k_export_values(Geom g, const T *__restrict__ u, double *__restrict__ out,
                                                       long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    if (!in_grid(g, c)) return;
    const long long lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
    out[lin] = (double)u[p];
}

// Geodesic flow and active-neighbour bits, recomputed from the converged values with the
// accepted (v < u) differences (Headers/Base/Implementation/HamiltonFastMarching.hxx:187-272,
// 363-378, EulerianStencil.hxx:155-233).  flow = -ReDim(sum_k w_k (u - v_k)^+ e_k).
// Whether a nonzero-weight difference of the lane's stencil reaches offset (o0, o1, o2): a
// forward one with that offset, or a symmetric one with that offset or its opposite.
template <int ModelT, typename T>
__device__ inline bool shares_offset(const LaneStencil<ModelT, T> &st, int k, int o0, int o1, int o2) {
    constexpr int NF = ModelShape<ModelT>::NFWD, NC = NF + ModelShape<ModelT>::NSYM;
    for (int j = 0; j < NC; ++j) {
        if (j == k || st.w[j] == T(0)) continue;
        int p0, p1, p2;
        unpack_offset(st.off[j], p0, p1, p2);
        if (p0 == o0 && p1 == o1 && p2 == o2) return true;
        if (j >= NF && p0 == -o0 && p1 == -o1 && p2 == -o2) return true;
    }
    return false;
}

template <int ModelT, typename T>
__global__ __launch_bounds__(256) void k_flow(Geom g, StencilData<T> sd, const T *__restrict__ u,
                                              const unsigned long long *__restrict__ fixedmask, FlowScale fs,
                                              double *__restrict__ flow, double *__restrict__ active_out,
                                              double *__restrict__ aoffs_out, long long np) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int tile = (int)(p >> 6), lane = (int)(p & 63);
    int c[3];
    tiled_coords(g, tile, lane, c);
    if (!in_grid(g, c)) return;
    const long long lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
    const T inf = dev_inf<T>();
    LaneStencil<ModelT, T> st;
    st.load(g, sd, p, c);
    T v[NC];
    int side[NC];
    for (int k = 0; k < NC; ++k) {
        int o0, o1, o2;
        unpack_offset(st.off[k], o0, o1, o2);
        v[k] = inf;
        side[k] = 0;
        // A neighbour enters the reversed stencil only through a nonzero-weight difference
        // (HFM_StencilRecompute.hxx:92, HFM_StencilShare.hxx:153), but its acceptance activates
        // every difference with that offset (EulerianStencil.hxx:114-137): a zero-weight
        // difference is active only if a nonzero-weight one shares its (signed) offset.
        // Walls: an invisible side is never updated from (VisibleOffset).
        const bool pa = st.visible(k, 0) && (st.w[k] != T(0) || shares_offset<ModelT>(st, k, o0, o1, o2));
        const bool pb = k >= NF && st.visible(k, 1) &&
                        (st.w[k] != T(0) || shares_offset<ModelT>(st, k, -o0, -o1, -o2));
        int r[3];
        T va = inf, vb = inf;
        if (pa && shift_coords_m<ModelT>(g, c, o0, o1, o2, r)) va = u[(long long)tile_of(g, r) * 64 + local_of(g, r)];
        if (pb && shift_coords_m<ModelT>(g, c, -o0, -o1, -o2, r)) vb = u[(long long)tile_of(g, r) * 64 + local_of(g, r)];
        v[k] = va <= vb ? va : vb;
        side[k] = va <= vb ? 1 : -1;
    }
    const T uu = u[p];
    const bool fixed = (fixedmask[tile] >> lane) & 1ull;
    uint32_t used = 0;
    if (!fixed && uu < inf) {
        // differences with v < u, in the local solver's order
        T vv[NC];
        for (int k = 0; k < NC; ++k) vv[k] = v[k] < uu ? v[k] : inf;
        used = active_slots<NC, T>(vv, st.w);
    }
    double f[3] = {0., 0., 0.};
    double bits = 0.;
    if (used) {
        // Recompute the value from the used differences (minVal = their smallest value).
        double m = __builtin_huge_val();
        for (int k = 0; k < NC; ++k)
            if (used >> k & 1u) m = (double)v[k] < m ? (double)v[k] : m;
        double a = 0, b = 0, cc = -1;
        for (int k = 0; k < NC; ++k) {
            if (!(used >> k & 1u)) continue;
            const double w = (double)st.w[k], vr = (double)v[k] - m;
            a += w;
            b += w * vr;
            cc += (w * vr) * vr;
        }
        const double d2 = b * b - a * cc;
        const double value = m + (b + sqrt(d2 > 0. ? d2 : 0.)) / a;
        unsigned long long ab = 0;
        for (int k = 0; k < NC; ++k) {
            if (!(used >> k & 1u)) continue;
            int o0, o1, o2;
            unpack_offset(st.off[k], o0, o1, o2);
            const double wpd = (double)st.w[k] * fmax(0., value - (double)v[k]);
            f[0] += wpd * (double)(side[k] * o0);
            f[1] += wpd * (double)(side[k] * o1);
            f[2] += wpd * (double)(side[k] * o2);
            const int bit = k < NF ? k : NF + 2 * (k - NF) + (side[k] > 0 ? 0 : 1);
            ab |= 1ull << bit;
        }
        bits = (double)ab;
    }
    if (flow) {
        for (int q = 0; q < g.dim; ++q) flow[lin * g.dim + q] = -(f[q] * fs.h[q]);
    }
    if (active_out) active_out[lin] = bits;
    if (aoffs_out) {
        // activeOffsets (HFMInterface.hxx:554-577): the active differences in slot order (a
        // symmetric one with the sign of its active side), each offset packed big-endian as
        // sum_q uint8(o_q) << 8 (DIM - 1 - q) into a double; unused entries 0.
        int j = 0;
        for (int k = 0; k < NC; ++k) {
            if (!(used >> k & 1u)) continue;
            int o[3];
            unpack_offset(st.off[k], o[0], o[1], o[2]);
            unsigned long long packed = 0;
            for (int q = 0; q < S::DIM; ++q)
                packed |= (unsigned long long)(uint8_t)(int8_t)(side[k] * o[q]) << (8 * (S::DIM - 1 - q));
            aoffs_out[lin * NC + j++] = (double)packed;
        }
        for (; j < NC; ++j) aoffs_out[lin * NC + j] = 0.;
    }
}

// ---------------------------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------------------------
static inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
hipError_t launch_riemann_stencil(const Geom &g, const double *metric, int is_field, int dual, const InvHH &ih,
                                  uint32_t *offs, T *w, long long np, int *status, RedSlot *red, hipStream_t s) {
    if (g.dim == 2)
        hipLaunchKernelGGL((k_riemann_stencil<2, T>), dim3(blocks_for(np, 256)), dim3(256), 0, s, g, metric, is_field,
                           dual, ih, offs, w, np, status, red);
    else
        hipLaunchKernelGGL((k_riemann_stencil<3, T>), dim3(blocks_for(np, 256)), dim3(256), 0, s, g, metric, is_field,
                           dual, ih, offs, w, np, status, red);
    return hipGetLastError();
}
template hipError_t launch_riemann_stencil<float>(const Geom &, const double *, int, int, const InvHH &, uint32_t *,
                                                  float *, long long, int *, RedSlot *, hipStream_t);
template hipError_t launch_riemann_stencil<double>(const Geom &, const double *, int, int, const InvHH &, uint32_t *,
                                                   double *, long long, int *, RedSlot *, hipStream_t);

template <int M, typename T>
static void deps_pass(const Geom &g, const StencilData<T> &sd, int pass, int *cnt, const int *start, int *fill,
                      int *deps, int *srcbuf, int *srcn, hipStream_t s) {
    hipLaunchKernelGGL((k_tile_deps<M, T>), dim3(g.ntiles), dim3(64), 0, s, g, sd, pass, cnt, start, fill, deps, srcbuf,
                       srcn);
}

template <typename T>
hipError_t launch_tile_deps(int model, const Geom &g, const StencilData<T> &sd, int pass, int *cnt, const int *start,
                            int *fill, int *deps, int *srcbuf, int *srcn, hipStream_t s) {
    switch (model) {
        case MODEL_RIEMANN2: deps_pass<MODEL_RIEMANN2, T>(g, sd, pass, cnt, start, fill, deps, srcbuf, srcn, s); break;
        case MODEL_RIEMANN3: deps_pass<MODEL_RIEMANN3, T>(g, sd, pass, cnt, start, fill, deps, srcbuf, srcn, s); break;
        case MODEL_RSF2: deps_pass<MODEL_RSF2, T>(g, sd, pass, cnt, start, fill, deps, srcbuf, srcn, s); break;
        case MODEL_ELASTICA2: deps_pass<MODEL_ELASTICA2, T>(g, sd, pass, cnt, start, fill, deps, srcbuf, srcn, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_tile_deps<float>(int, const Geom &, const StencilData<float> &, int, int *, const int *,
                                            int *, int *, int *, int *, hipStream_t);
template hipError_t launch_tile_deps<double>(int, const Geom &, const StencilData<double> &, int, int *, const int *,
                                             int *, int *, int *, int *, hipStream_t);

long long scan_work_ints(long long n) { return 2 * ((n + SCAN_CHUNK - 1) / SCAN_CHUNK) + 1; }

hipError_t launch_exclusive_scan(const int *cnt, int *start, int n, hipStream_t s, int *work) {
    const int nchunks = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (!work || nchunks <= 2) {
        hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(1024), 0, s, cnt, start, n);
        return hipGetLastError();
    }
    int *totals = work, *carries = work + nchunks;  // nchunks + (nchunks + 1) ints
    hipLaunchKernelGGL(k_scan_totals, dim3(nchunks), dim3(1024), 0, s, cnt, n, totals);
    hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(1024), 0, s, totals, carries, nchunks);
    hipLaunchKernelGGL(k_scan_apply, dim3(nchunks), dim3(1024), 0, s, cnt, carries, start, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_init(const Geom &g, T *u, unsigned long long *fixedmask, unsigned *keys, unsigned *tmax,
                       long long np, const long long *seed_p, const double *seed_v, int nseeds, hipStream_t s,
                       int key_buffers) {
    hipLaunchKernelGGL((k_init<T>), dim3(blocks_for(np, 256)), dim3(256), 0, s, g, u, fixedmask, keys, tmax, np,
                       key_buffers);
    if (nseeds > 0)
        hipLaunchKernelGGL((k_set_seeds<T>), dim3(blocks_for(nseeds, 64)), dim3(64), 0, s, seed_p, seed_v, nseeds, u,
                           fixedmask);
    return hipGetLastError();
}
template hipError_t launch_init<float>(const Geom &, float *, unsigned long long *, unsigned *, unsigned *,
                                       long long, const long long *, const double *, int, hipStream_t, int);
template hipError_t launch_init<double>(const Geom &, double *, unsigned long long *, unsigned *, unsigned *,
                                        long long, const long long *, const double *, int, hipStream_t, int);

hipError_t launch_seed_activate(const long long *seed_p, const double *seed_v, int nseeds, const int *dep_start,
                                const int *deps, unsigned *keys, const unsigned *tmax, unsigned *pend, hipStream_t s) {
    hipLaunchKernelGGL(k_seed_activate, dim3((unsigned)std::min(std::max(nseeds, 1), 1024)), dim3(64), 0, s, seed_p,
                       seed_v, nseeds, dep_start, deps, keys, tmax, pend);
    return hipGetLastError();
}

hipError_t launch_schedule(int ntiles, unsigned *keys, const SchedRound &sr, hipStream_t s) {
    hipLaunchKernelGGL(k_schedule, dim3(blocks_for(ntiles, SCHED_BLOCK)), dim3(SCHED_BLOCK), 0, s, ntiles, keys, sr);
    return hipGetLastError();
}

// With ev0/ev1 the launch goes through hipExtLaunchKernelGGL, whose events take the begin and
// end timestamps of the dispatch itself (the kernel's execution time, without the dispatch gap
// that an event recorded before the launch would include).
template <int M, typename T, int MS>
static void fim_sweeps(const FimArgs<T> &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    if (a.fused) {
        if (ev0)
            hipExtLaunchKernelGGL((k_fim_sched<M, T, MS>), dim3(a.grid), dim3(256), 0, s, ev0, ev1, 0, a.g, a.sd, a.u,
                                  a.fixedmask, a.dep_start, a.deps, a.keys, a.keys_out, a.tmax, a.sr,
                                  a.count_reset, a.pend_reset, a.stats);
        else
            hipLaunchKernelGGL((k_fim_sched<M, T, MS>), dim3(a.grid), dim3(256), 0, s, a.g, a.sd, a.u, a.fixedmask,
                               a.dep_start, a.deps, a.keys, a.keys_out, a.tmax, a.sr, a.count_reset, a.pend_reset,
                               a.stats);
        return;
    }
    if (ev0)
        hipExtLaunchKernelGGL((k_fim<M, T, MS>), dim3(a.grid), dim3(256), 0, s, ev0, ev1, 0, a.g, a.sd, a.u,
                              a.fixedmask, a.dep_start, a.deps, a.list, a.count, a.count_reset, a.keys, a.tmax,
                              a.pend_out, a.pend_reset, a.stats);
    else
        hipLaunchKernelGGL((k_fim<M, T, MS>), dim3(a.grid), dim3(256), 0, s, a.g, a.sd, a.u, a.fixedmask,
                           a.dep_start, a.deps, a.list, a.count, a.count_reset, a.keys, a.tmax, a.pend_out,
                           a.pend_reset, a.stats);
}

template <int M, typename T>
static void fim_one(const FimArgs<T> &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    if (a.max_sweeps <= 4) fim_sweeps<M, T, 4>(a, s, ev0, ev1);
    else if (a.max_sweeps <= 8) fim_sweeps<M, T, 8>(a, s, ev0, ev1);
    else if (a.max_sweeps <= 16) fim_sweeps<M, T, 16>(a, s, ev0, ev1);
    else fim_sweeps<M, T, 32>(a, s, ev0, ev1);
}

template <typename T>
hipError_t launch_fim(int model, const FimArgs<T> &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    switch (model) {
        case MODEL_RIEMANN2: fim_one<MODEL_RIEMANN2, T>(a, s, ev0, ev1); break;
        case MODEL_RIEMANN3: fim_one<MODEL_RIEMANN3, T>(a, s, ev0, ev1); break;
        case MODEL_RSF2: fim_one<MODEL_RSF2, T>(a, s, ev0, ev1); break;
        case MODEL_ELASTICA2: fim_one<MODEL_ELASTICA2, T>(a, s, ev0, ev1); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_fim<float>(int, const FimArgs<float> &, hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t launch_fim<double>(int, const FimArgs<double> &, hipStream_t, hipEvent_t, hipEvent_t);

template <typename T>
hipError_t launch_export_values(const Geom &g, const T *u, double *out, long long np, hipStream_t s) {
    hipLaunchKernelGGL((k_export_values<T>), dim3(blocks_for(np, 256)), dim3(256), 0, s, g, u, out, np);
    return hipGetLastError();
}
template hipError_t launch_export_values<float>(const Geom &, const float *, double *, long long, hipStream_t);
template hipError_t launch_export_values<double>(const Geom &, const double *, double *, long long, hipStream_t);

template <int M, typename T>
static void flow_one(const Geom &g, const StencilData<T> &sd, const T *u, const unsigned long long *fm,
                     const FlowScale &fs, double *flow, double *act, double *aoffs, long long np, hipStream_t s) {
    hipLaunchKernelGGL((k_flow<M, T>), dim3(blocks_for(np, 256)), dim3(256), 0, s, g, sd, u, fm, fs, flow, act, aoffs,
                       np);
}

template <typename T>
hipError_t launch_flow(int model, const Geom &g, const StencilData<T> &sd, const T *u, const unsigned long long *fm,
                       const FlowScale &fs, double *flow, double *act, double *aoffs, long long np, hipStream_t s) {
    switch (model) {
        case MODEL_RIEMANN2: flow_one<MODEL_RIEMANN2, T>(g, sd, u, fm, fs, flow, act, aoffs, np, s); break;
        case MODEL_RIEMANN3: flow_one<MODEL_RIEMANN3, T>(g, sd, u, fm, fs, flow, act, aoffs, np, s); break;
        case MODEL_RSF2: flow_one<MODEL_RSF2, T>(g, sd, u, fm, fs, flow, act, aoffs, np, s); break;
        case MODEL_ELASTICA2: flow_one<MODEL_ELASTICA2, T>(g, sd, u, fm, fs, flow, act, aoffs, np, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_flow<float>(int, const Geom &, const StencilData<float> &, const float *,
                                       const unsigned long long *, const FlowScale &, double *, double *, double *,
                                       long long, hipStream_t);
template hipError_t launch_flow<double>(int, const Geom &, const StencilData<double> &, const double *,
                                        const unsigned long long *, const FlowScale &, double *, double *, double *,
                                        long long, hipStream_t);

// ---------------------------------------------------------------------------------------------
// Final residual check of a converged solve, in FP64.  The discrete solution is the fixed point
// u = Lambda(u) of the local Hopf-Lax solve (EulerianStencil.hxx:30-72,104-153): at every reached
// point that is not a seed, u(x) is the root of sum_k w_k max(0, u - v_k)^2 = 1 over the smaller
// neighbour value v_k of each difference.  This recomputes that root in FP64 (the reference's
// operation order) from the stored values and weights and reports, over all such points, the
// largest |Lambda(u) - u| / |u| and the number of points where exactly one of the two is finite.
// A point the iteration left unconverged shows as a large deviation.
// ---------------------------------------------------------------------------------------------
template <int ModelT, typename T>
__global__ __launch_bounds__(256) void k_residual(Geom g, StencilData<T> sd, const T *__restrict__ u,
                                                  const unsigned long long *__restrict__ fixedmask, long long np,
                                                  ResSlot *__restrict__ out) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double dev = 0.;
    int bad = 0;
    if (p < np) {
        const int tile = (int)(p >> 6), lane = (int)(p & 63);
        int c[3];
        tiled_coords(g, tile, lane, c);
        const bool fixed = (fixedmask[tile] >> lane) & 1ull;
        if (in_grid(g, c) && !fixed) {
            const double inf = __builtin_huge_val();
            LaneStencil<ModelT, T> st;
            st.load(g, sd, p, c);
            double v[NC], w[NC];
            for (int k = 0; k < NC; ++k) {
                int o0, o1, o2;
                unpack_offset(st.off[k], o0, o1, o2);
                w[k] = (double)st.w[k];
                v[k] = inf;
                if (st.w[k] == T(0)) continue;
                int r[3];
                if (st.visible(k, 0) && shift_coords_m<ModelT>(g, c, o0, o1, o2, r))
                    v[k] = (double)u[(long long)tile_of(g, r) * 64 + local_of(g, r)];
                if (k >= NF && st.visible(k, 1) && shift_coords_m<ModelT>(g, c, -o0, -o1, -o2, r)) {
                    const double vb = (double)u[(long long)tile_of(g, r) * 64 + local_of(g, r)];
                    v[k] = vb < v[k] ? vb : v[k];
                }
            }
            uint32_t used = 0;
            const double lam = local_solve<NC, double>(v, w, used);
            const double uu = (double)u[p];
            const bool fl = lam < inf, fu = uu < inf;
            if (fl != fu) {
                bad = 1;
            } else if (fl) {
                const double den = fmax(fabs(uu), fabs(lam));
                dev = den > 0. ? fabs(lam - uu) / den : 0.;
            }
        }
    }
    // nonnegative doubles order like their bit patterns: wave maximum on the bits
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double d = __shfl_xor(dev, o);
        dev = d > dev ? d : dev;
        bad += __shfl_xor(bad, o);
    }
    if ((threadIdx.x & 63) == 0 && (dev > 0. || bad)) {
        const unsigned wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        ResSlot *r = out + (wave & (RED_SLOTS - 1));
        atomicMax(&r->maxbits, (unsigned long long)__double_as_longlong(dev));
        atomicAdd(&r->mismatch, (unsigned long long)bad);
    }
}

template <int M, typename T>
static void residual_one(const Geom &g, const StencilData<T> &sd, const T *u, const unsigned long long *fm,
                         long long np, ResSlot *out, hipStream_t s) {
    hipLaunchKernelGGL((k_residual<M, T>), dim3(blocks_for(np, 256)), dim3(256), 0, s, g, sd, u, fm, np, out);
}

template <typename T>
hipError_t launch_residual(int model, const Geom &g, const StencilData<T> &sd, const T *u,
                           const unsigned long long *fm, long long np, ResSlot *out, hipStream_t s) {
    switch (model) {
        case MODEL_RIEMANN2: residual_one<MODEL_RIEMANN2, T>(g, sd, u, fm, np, out, s); break;
        case MODEL_RIEMANN3: residual_one<MODEL_RIEMANN3, T>(g, sd, u, fm, np, out, s); break;
        case MODEL_RSF2: residual_one<MODEL_RSF2, T>(g, sd, u, fm, np, out, s); break;
        case MODEL_ELASTICA2: residual_one<MODEL_ELASTICA2, T>(g, sd, u, fm, np, out, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_residual<float>(int, const Geom &, const StencilData<float> &, const float *,
                                           const unsigned long long *, long long, ResSlot *, hipStream_t);
template hipError_t launch_residual<double>(int, const Geom &, const StencilData<double> &, const double *,
                                            const unsigned long long *, long long, ResSlot *, hipStream_t);

}  // namespace hfm
