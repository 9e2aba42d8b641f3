// hfm_adjoint.hip -- backward differentiation of the arrival times w.r.t. the cost field.
//
// The reference (FirstVariation::BackwardVariation, Headers/ExtraAlgorithms/FirstVariation.h:
// 341-385, with ValueVariation :198-251) sweeps the points in strictly decreasing value order
// and, at each point x with adjoint weight lambda(x):
//   w'_k = Weight_k * (u(x) - v_k)^+ over x's active differences, S = sum_k w'_k,
//   costSensitivity(x) = lambda(x) * var(x),   var = (mult * sum_k w'_k (u - v_k)) / S
//                                               (mult = speed = 1/cost; 1 for Riemann models),
//   lambda(neighbour_k) += lambda(x) * w'_k / S,
//   a point without active differences (a seed) is reported in seedSensitivity.
// The active differences are the ones the forward solve accepted (v_k < u(x)), recomputed here
// from the converged values with the forward kernels' local solver.
//
// The sweep order is a topological order of the DAG x -> neighbour_k (edges go to strictly
// smaller values).  On the GPU the same sums are accumulated with Kahn's algorithm: every point
// counts its predecessors (in-degree); a round processes the points whose predecessors have all
// been processed (their lambda is final), pushes lambda(x) w'_k / S with atomic adds, then
// decrements the neighbours' in-degrees and appends those reaching zero to the next round.
// The result equals the reference's up to the summation order of the atomic adds.
#include <algorithm>
#include <map>
#include <vector>

#include "hfm_internal.h"
#include "hfm_kernels.h"

namespace hfm {

// Active differences of tiled point p recomputed from converged values.  Returns the number of
// active entries; nb[] their tiled indices, wp[] the w'_k, and *wsum, *var.  After a serial solve
// (xact: that loop's activeNeighs) the active differences are the loop's, in stencil order, as
// ValueVariation reads them (FirstVariation.h:198-251: the stored flags and values, first-order
// weights w max(0, u - v), also after order 2/3 and factoring).
template <int ModelT, typename T>
__device__ inline int active_set(const Geom &g, const StencilData<T> &sd, const T *__restrict__ u, long long p,
                                 const int *c, bool fixed, const unsigned long long *xact, long long *nb, double *wp,
                                 double *wsum, double *var) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    const T inf = dev_inf<T>();
    const T uu = u[p];
    *wsum = 0.;
    *var = 0.;
    if (xact) {
        const unsigned long long act = xact[p];
        if (!act || !(uu < inf)) return 0;  // seeds and unreached points have no active difference
        LaneStencil<ModelT, T> st;
        st.load(g, sd, p, c);
        const double mult = S::SHARED ? (double)sd.speed[p] : 1.;
        int n = 0;
        double ws = 0., acc = 0.;
        for (int k = 0; k < NC; ++k) {
            int o0, o1, o2;
            unpack_offset(st.off[k], o0, o1, o2);
            for (int sd2 = 0; sd2 < (k < NF ? 1 : 2); ++sd2) {
                if (!(act >> side_bit<NF>(k, sd2) & 1ull)) continue;
                const int sg = sd2 == 0 ? 1 : -1;
                int r[3];
                shift_coords_m<ModelT>(g, c, sg * o0, sg * o1, sg * o2, r);  // active: in the grid
                const long long q = (long long)tile_of(g, r) * 64 + local_of(g, r);
                const double d = fmax(0., (double)uu - (double)u[q]);
                const double w = (double)st.w[k] * d;
                ws += w;
                acc += w * d;
                nb[n] = q;
                wp[n] = w;
                ++n;  // the loop activates one side of a symmetric difference: n <= NC
            }
        }
        *wsum = ws;
        *var = ws > 0. ? (1. / ws) * (mult * acc) : 0.;
        return n;
    }
    if (fixed || !(uu < inf)) return 0;
    LaneStencil<ModelT, T> st;
    st.load(g, sd, p, c);
    T v[NC];
    long long idx[NC];
    for (int k = 0; k < NC; ++k) {
        int o0, o1, o2;
        unpack_offset(st.off[k], o0, o1, o2);
        v[k] = inf;
        idx[k] = -1;
        if (st.w[k] == T(0)) continue;
        int r[3];
        T va = inf, vb = inf;
        long long ia = -1, ib = -1;
        if (st.visible(k, 0) && shift_coords_m<ModelT>(g, c, o0, o1, o2, r)) {
            ia = (long long)tile_of(g, r) * 64 + local_of(g, r);
            va = u[ia];
        }
        if (k >= NF && st.visible(k, 1) && shift_coords_m<ModelT>(g, c, -o0, -o1, -o2, r)) {
            ib = (long long)tile_of(g, r) * 64 + local_of(g, r);
            vb = u[ib];
        }
        const bool a_first = va <= vb;
        v[k] = a_first ? va : vb;
        idx[k] = a_first ? ia : ib;
        if (!(v[k] < uu)) v[k] = inf;
    }
    const uint32_t used = active_slots<NC, T>(v, st.w);
    const double mult = S::SHARED ? (double)sd.speed[p] : 1.;
    int n = 0;
    double ws = 0., acc = 0.;
    for (int k = 0; k < NC; ++k) {
        if (!(used >> k & 1u)) continue;
        const double d = fmax(0., (double)uu - (double)v[k]);
        const double w = (double)st.w[k] * d;
        ws += w;
        acc += w * d;
        nb[n] = idx[k];
        wp[n] = w;
        ++n;
    }
    *wsum = ws;
    // ValueVariation (FirstVariation.h:231-250): invWeight * (mult * sum), invWeight = 1 / S
    *var = ws > 0. ? (1. / ws) * (mult * acc) : 0.;
    return n;
}

// Lanes given to one point in a round of the reverse sweep: one per active edge it may have
// (the model's differences, rounded up to a power of two); a wave takes 64 / ADJ_SLOTS points.
template <int ModelT> constexpr int adj_slots() {
    constexpr int NC = ModelShape<ModelT>::NFWD + ModelShape<ModelT>::NSYM;
    return NC <= 4 ? 4 : NC <= 8 ? 8 : NC <= 16 ? 16 : 32;
}

template <int ModelT, typename T>
__global__ __launch_bounds__(256) void k_adj_prepare(Geom g, StencilData<T> sd, const T *__restrict__ u,
                                                     const unsigned long long *__restrict__ fixedmask,
                                                     const unsigned long long *__restrict__ xact,
                                                     int *__restrict__ indeg, double *__restrict__ lambda,
                                                     double *__restrict__ sens, int *__restrict__ ecnt,
                                                     unsigned *__restrict__ eidx, double *__restrict__ var_out,
                                                     double *__restrict__ ew, long long np) {
    constexpr int NC = ModelShape<ModelT>::NFWD + ModelShape<ModelT>::NSYM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    lambda[p] = 0.;
    sens[p] = 0.;
    ecnt[p] = 0;
    var_out[p] = 0.;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    if (!in_grid(g, c)) return;
    const bool fixed = (fixedmask[p >> 6] >> (p & 63)) & 1ull;
    long long nb[NC];
    double wp[NC], ws, var;
    const int n = active_set<ModelT, T>(g, sd, u, p, c, fixed, xact, nb, wp, &ws, &var);
    ecnt[p] = n;
    var_out[p] = var;
    const double inv = ws > 0. ? 1. / ws : 0.;
    for (int i = 0; i < n; ++i) {
        atomicAdd(&indeg[nb[i]], 1);
        eidx[(long long)i * np + p] = (unsigned)nb[i];  // tiled indices fit 32 bits (at most 2^26 tiles)
        ew[(long long)i * np + p] = wp[i] * inv;
    }
}

// Wave-aggregated append of tiled indices into a list.
__device__ inline void adj_push(long long q, int lane, long long *__restrict__ list, int *__restrict__ count) {
    const bool want = q >= 0;
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return;
    int base = 0;
    if (lane == 0) base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, 0);
    if (want) {
        const unsigned long long below = lane == 0 ? 0ull : (mask & (~0ull >> (64 - lane)));
        list[base + __popcll(below)] = q;
    }
}

// Wave-aggregated append of up to 8 tiled indices per lane (items[q] for the set bits q of
// mask): one counter atomic per wave, positions from a prefix sum over the lanes.
__device__ inline void wave_push8(const long long *items, unsigned mask, int lane, long long *__restrict__ list,
                                  int *__restrict__ count) {
    const int c = __popc(mask);
    int incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    const int total = __shfl(incl, 63);
    if (total == 0) return;  // wave-uniform
    int base = 0;
    if (lane == 0) base = atomicAdd(count, total);
    int pos = __shfl(base, 0) + incl - c;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (mask >> q & 1u) list[pos++] = items[q];
}

__global__ __launch_bounds__(256) void k_adj_frontier(Geom g, const int *__restrict__ indeg, long long np,
                                                      long long *__restrict__ list, int *__restrict__ count) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;  // np is a multiple of 64: whole waves leave together
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    const bool start = in_grid(g, c) && indeg[p] == 0;
    adj_push(start ? p : -1, threadIdx.x & 63, list, count);
}

__global__ void k_adj_targets(const long long *__restrict__ tp, const double *__restrict__ tw, int n,
                              double *__restrict__ lambda) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lambda[tp[i]] = tw[i];
}

// One level of the reverse sweep: every listed point (all its predecessors processed, so its
// lambda is final) writes its sensitivity, pushes lambda * w'/S along its active edges, and
// decrements its neighbours' in-degrees; those reaching zero form the next level's list.  A
// point's edges are spread over ADJ_SLOTS lanes, one edge each: every lane then waits for at
// most one returning atomic per pass (with one lane per point, a point's 27 decrements on
// Elastica2 were 27 round trips in sequence).
template <int ModelT, typename T>
__global__ __launch_bounds__(256) void k_adj_round(int *__restrict__ indeg, double *__restrict__ lambda,
                                                   double *__restrict__ sens, const int *__restrict__ ecnt,
                                                   const unsigned *__restrict__ eidx, const double *__restrict__ var_in,
                                                   const double *__restrict__ ew, long long np,
                                                   const long long *__restrict__ list_in,
                                                   const int *__restrict__ count_in, long long *__restrict__ list_out,
                                                   int *__restrict__ count_out, int *__restrict__ count_reset,
                                                   long long *__restrict__ seeds_idx, double *__restrict__ seeds_w,
                                                   int *__restrict__ nseeds, int max_seeds) {
    constexpr int SL = adj_slots<ModelT>();
    constexpr int GP = 64 / SL;  // points per wave pass
    const int n = *count_in;
    if (blockIdx.x == 0 && threadIdx.x == 0) *count_reset = 0;
    const int lane = threadIdx.x & 63;
    const int g = lane / SL, k = lane % SL;
    // Uniform trip count per wave so that the wave-level pushes see every lane.
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int base = wave * GP; base < n; base += nwaves * GP) {
        const int i = base + g;
        long long t = -1;
        if (i < n) {
            const long long p = list_in[i];
            const int na = ecnt[p];
            const double lam = lambda[p];
            if (k < na) {
                t = eidx[(long long)k * np + p];
                if (lam != 0.) atomicAdd(&lambda[t], lam * ew[(long long)k * np + p]);
            }
            if (k == 0 && lam != 0.) {
                sens[p] = lam * var_in[p];
                if (na == 0) {
                    const int s = atomicAdd(nseeds, 1);
                    if (s < max_seeds) {
                        seeds_idx[s] = p;
                        seeds_w[s] = lam;
                    }
                }
            }
        }
        // the next launch processes the neighbours that became ready, and sees this launch's
        // lambda updates
        const bool ready = t >= 0 && atomicSub(&indeg[t], 1) == 1;
        adj_push(ready ? t : -1, lane, list_out, count_out);
    }
}

__global__ void k_export_sens(Geom g, const double *__restrict__ s, double *__restrict__ out, long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    if (!in_grid(g, c)) return;
    out[((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2]] = s[p];
}

// ---------------------------------------------------------------------------------------------
// Forward differentiation (FirstVariation::ForwardVariation, FirstVariation.h:389-429).  The
// reference visits the points in increasing (value, index) order and sets, for k < nVar,
//   res(x, k) += var(x) * costVariation(x, k) + sum_i w_i / S * res(neighbour_i, k)
// over x's active differences in stencil order (ValueVariation, :198-251).  A point only reads
// its own active neighbours, all with strictly smaller values, so any topological order of the
// DAG x -> neighbour_i gives the same sums.  The GPU runs Kahn's algorithm in pull form: a
// point is processed once all its active neighbours are (pending count zero), gathers their
// final results in stencil order (the reference's operation order, no atomics on the sums),
// then decrements the pending counts of its dependants, found through a reverse adjacency
// (CSR of the edges neighbour -> x) built once per call.
// ---------------------------------------------------------------------------------------------

// Pending count of every point (its number of active neighbours) and in-degree of the reverse
// adjacency.
template <int ModelT, typename T>
__global__ __launch_bounds__(256) void k_fwd_prepare(Geom g, StencilData<T> sd, const T *__restrict__ u,
                                                     const unsigned long long *__restrict__ fixedmask,
                                                     const unsigned long long *__restrict__ xact,
                                                     int *__restrict__ pending, int *__restrict__ rcount,
                                                     long long np) {
    constexpr int NC = ModelShape<ModelT>::NFWD + ModelShape<ModelT>::NSYM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    if (!in_grid(g, c)) {
        pending[p] = 0;
        return;
    }
    const bool fixed = (fixedmask[p >> 6] >> (p & 63)) & 1ull;
    long long nb[NC];
    double wp[NC], ws, var;
    const int n = active_set<ModelT, T>(g, sd, u, p, c, fixed, xact, nb, wp, &ws, &var);
    pending[p] = n;
    for (int i = 0; i < n; ++i) atomicAdd(&rcount[nb[i]], 1);
}

// Reverse adjacency: dependants of point q in rev[rstart[q] .. rstart[q + 1]).  rcount (the
// in-degrees from k_fwd_prepare) is counted back down to zero as the slots are taken.
template <int ModelT, typename T>
__global__ __launch_bounds__(256) void k_fwd_fill(Geom g, StencilData<T> sd, const T *__restrict__ u,
                                                  const unsigned long long *__restrict__ fixedmask,
                                                  const unsigned long long *__restrict__ xact,
                                                  const int *__restrict__ rstart, int *__restrict__ rcount,
                                                  int *__restrict__ rev, long long np) {
    constexpr int NC = ModelShape<ModelT>::NFWD + ModelShape<ModelT>::NSYM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    if (!in_grid(g, c)) return;
    const bool fixed = (fixedmask[p >> 6] >> (p & 63)) & 1ull;
    long long nb[NC];
    double wp[NC], ws, var;
    const int n = active_set<ModelT, T>(g, sd, u, p, c, fixed, xact, nb, wp, &ws, &var);
    for (int i = 0; i < n; ++i) {
        const int slot = atomicSub(&rcount[nb[i]], 1) - 1;
        rev[rstart[nb[i]] + slot] = (int)p;
    }
}

// Initial results: zero, then the seeds' value variations (later seeds overwrite earlier ones
// at the same point, as the reference's loop over the seeds does, FirstVariation.h:116-119; the
// host passes one entry per point, the last one).
__global__ void k_fwd_seeds(const long long *__restrict__ sp, const double *__restrict__ sv, int nseeds, int nvar,
                            double *__restrict__ res) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseeds * nvar) return;
    const int s = i / nvar, k = i - s * nvar;
    res[sp[s] * nvar + k] = sv[(long long)s * nvar + k];
}

template <int ModelT, typename T>
__global__ __launch_bounds__(256) void k_fwd_round(Geom g, StencilData<T> sd, const T *__restrict__ u,
                                                   const unsigned long long *__restrict__ fixedmask,
                                                   const unsigned long long *__restrict__ xact,
                                                   int *__restrict__ pending, const int *__restrict__ rstart,
                                                   const int *__restrict__ rev, const double *__restrict__ cv,
                                                   int nvar, double *__restrict__ res,
                                                   const long long *__restrict__ list_in,
                                                   const int *__restrict__ count_in, long long *__restrict__ list_out,
                                                   int *__restrict__ count_out, int *__restrict__ count_reset) {
    constexpr int NC = ModelShape<ModelT>::NFWD + ModelShape<ModelT>::NSYM;
    const int n = *count_in;
    if (blockIdx.x == 0 && threadIdx.x == 0) *count_reset = 0;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int base = wave * 64; base < n; base += nwaves * 64) {
        const int i = base + lane;
        int beg = 0, cnt = 0;
        if (i < n) {
            const long long p = list_in[i];
            int c[3];
            tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
            const bool fixed = (fixedmask[p >> 6] >> (p & 63)) & 1ull;
            long long nb[NC];
            double wp[NC], ws = 0., var = 0.;
            const int na = active_set<ModelT, T>(g, sd, u, p, c, fixed, xact, nb, wp, &ws, &var);
            // neighbour weights w_i / S as the reference normalises them (weight *= 1 / S)
            const double inv = ws > 0. ? 1. / ws : 0.;
            const long long lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
            for (int k = 0; k < nvar; ++k) {
                double r = res[p * nvar + k];
                r += var * (cv ? cv[lin * nvar + k] : 0.);
                if (ws > 0.)
                    for (int j = 0; j < na; ++j) r += (wp[j] * inv) * res[nb[j] * nvar + k];
                res[p * nvar + k] = r;
            }
            beg = rstart[p];
            cnt = rstart[p + 1] - beg;
        }
        // Dependants' pending counts, eight per lane at a time (independent atomics in flight), then
        // one aggregated append of those that became ready.  They are processed by the next launch,
        // which sees this launch's results.
        for (int j0 = 0; __ballot(j0 < cnt) != 0ull; j0 += 8) {
            long long items[8];
            unsigned mask = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q) items[q] = j0 + q < cnt ? (long long)rev[beg + j0 + q] : -1;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (items[q] >= 0 && atomicSub(&pending[items[q]], 1) == 1) mask |= 1u << q;
            wave_push8(items, mask, lane, list_out, count_out);
        }
    }
}

__global__ void k_fwd_export(Geom g, const double *__restrict__ res, int nvar, double *__restrict__ out,
                             long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    if (!in_grid(g, c)) return;
    const long long lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
    for (int k = 0; k < nvar; ++k) out[lin * nvar + k] = res[p * nvar + k];
}

}  // namespace hfm

using namespace hfm;

namespace {

// Round lists shared by the backward and forward sweeps.
void ensure_lists(hfm_plan *p) {
    if (!p->adj_counts) {
        p->adj_list[0] = dalloc<long long>(p->np);
        p->adj_list[1] = dalloc<long long>(p->np);
        p->adj_counts = dalloc<int>(4);
    }
}

template <int M, typename T> void adjoint_run(hfm_plan *p, const std::vector<long long> &tp,
                                              const std::vector<double> &tw, double *d_out, hipStream_t s,
                                              std::vector<long long> &sidx, std::vector<double> &sw, int max_seeds,
                                              int *nseed_out) {
    const long long np = p->np;
    const unsigned nb = (unsigned)((np + 255) / 256);
    const T *u = static_cast<const T *>(p->u);
    const unsigned long long *xact = p->exact_solved ? p->ex_active : nullptr;  // a serial solve's flags
    HIPCHECK(hipMemsetAsync(p->adj_indeg, 0, sizeof(int) * np, s));
    hipLaunchKernelGGL((k_adj_prepare<M, T>), dim3(nb), dim3(256), 0, s, p->g, p->sd<T>(), u, p->fixedmask, xact,
                       p->adj_indeg, p->adj_lambda, p->adj_sens, p->adj_ecnt, p->adj_eidx, p->adj_var, p->adj_ew,
                       np);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(p->adj_tp, tp.data(), tp.size() * sizeof(long long), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(p->adj_tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice, s));
    if (!tp.empty()) {
        hipLaunchKernelGGL(k_adj_targets, dim3((unsigned)((tp.size() + 63) / 64)), dim3(64), 0, s, p->adj_tp,
                           p->adj_tw, (int)tp.size(), p->adj_lambda);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemsetAsync(p->adj_counts, 0, 4 * sizeof(int), s));
    hipLaunchKernelGGL(k_adj_frontier, dim3(nb), dim3(256), 0, s, p->g, p->adj_indeg, np, p->adj_list[0],
                       &p->adj_counts[0]);
    HIPCHECK(hipGetLastError());
    // a level rarely lists more than a few thousand points: 256 blocks (1024 waves) cover 65536
    // entries a pass, and a small grid starts and drains quickly
    const unsigned grid = (unsigned)std::min<long long>((np + 255) / 256, 256);
    long long k = 0;
    const long long max_rounds = np + 16;
    while (true) {
        for (int b = 0; b < 8; ++b, ++k) {
            hipLaunchKernelGGL((k_adj_round<M, T>), dim3(grid), dim3(256), 0, s, p->adj_indeg, p->adj_lambda,
                               p->adj_sens, p->adj_ecnt, p->adj_eidx, p->adj_var, p->adj_ew, np,
                               p->adj_list[k & 1], &p->adj_counts[k % 3], p->adj_list[(k + 1) & 1],
                               &p->adj_counts[(k + 1) % 3], &p->adj_counts[(k + 2) % 3], p->adj_seed_idx,
                               p->adj_seed_w, &p->adj_counts[3], max_seeds);
            HIPCHECK(hipGetLastError());
        }
        HIPCHECK(hipMemcpyAsync(p->h_count, &p->adj_counts[k % 3], sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        if (*p->h_count == 0) break;
        if (k > max_rounds) throw HipFail{"adjoint sweep did not terminate"};
    }
    p->st.adjoint_rounds = k;
    hipLaunchKernelGGL(k_export_sens, dim3(nb), dim3(256), 0, s, p->g, p->adj_sens, d_out, np);
    HIPCHECK(hipGetLastError());
    int ns = 0;
    HIPCHECK(hipMemcpyAsync(&ns, &p->adj_counts[3], sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    const int nget = std::min(ns, max_seeds);
    std::vector<long long> ti(nget);
    std::vector<double> tv(nget);
    if (nget > 0) {
        HIPCHECK(hipMemcpy(ti.data(), p->adj_seed_idx, nget * sizeof(long long), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(tv.data(), p->adj_seed_w, nget * sizeof(double), hipMemcpyDeviceToHost));
    }
    // Tiled -> row-major index; reported in the reference's processing order: decreasing
    // (value, index) (BackwardVariation pops the largest key of its (value, index) map).
    std::vector<std::pair<std::pair<double, long long>, double>> keyed;
    for (int i = 0; i < nget; ++i) {
        int c[3];
        tiled_coords(p->g, (int)(ti[i] >> 6), (int)(ti[i] & 63), c);
        T uv;
        HIPCHECK(hipMemcpy(&uv, u + ti[i], sizeof(T), hipMemcpyDeviceToHost));
        keyed.push_back({{(double)uv, ((long long)c[0] * p->g.n[1] + c[1]) * p->g.n[2] + c[2]}, tv[i]});
    }
    std::sort(keyed.begin(), keyed.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    sidx.clear();
    sw.clear();
    for (auto &pr : keyed) {
        sidx.push_back(pr.first.second);
        sw.push_back(pr.second);
    }
    *nseed_out = ns;
}

template <int M, typename T> void forward_run(hfm_plan *p, int nvar, const double *cv,
                                              const std::vector<long long> &sp, const std::vector<double> &sv,
                                              double *d_out, hipStream_t s) {
    const long long np = p->np;
    const unsigned nb = (unsigned)((np + 255) / 256);
    const T *u = static_cast<const T *>(p->u);
    const unsigned long long *xact = p->exact_solved ? p->ex_active : nullptr;  // a serial solve's flags
    HIPCHECK(hipMemsetAsync(p->fwd_rcount, 0, sizeof(int) * np, s));
    hipLaunchKernelGGL((k_fwd_prepare<M, T>), dim3(nb), dim3(256), 0, s, p->g, p->sd<T>(), u, p->fixedmask, xact,
                       p->fwd_pending, p->fwd_rcount, np);
    HIPCHECK(hipGetLastError());
    HIPCHECK(launch_exclusive_scan(p->fwd_rcount, p->fwd_rstart, (int)np, s, p->fwd_scan_work));
    int nedges = 0;
    HIPCHECK(hipMemcpyAsync(&nedges, p->fwd_rstart + np, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (nedges > p->fwd_rev_cap) {
        dfree(p->fwd_rev);
        p->fwd_rev = nullptr;
        p->fwd_rev_cap = std::max<long long>(nedges, np);
        p->fwd_rev = dalloc<int>(p->fwd_rev_cap);
    }
    hipLaunchKernelGGL((k_fwd_fill<M, T>), dim3(nb), dim3(256), 0, s, p->g, p->sd<T>(), u, p->fixedmask, xact,
                       p->fwd_rstart, p->fwd_rcount, p->fwd_rev, np);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemsetAsync(p->fwd_res, 0, sizeof(double) * np * nvar, s));
    if (!sp.empty()) {
        HIPCHECK(hipMemcpyAsync(p->fwd_sp, sp.data(), sp.size() * sizeof(long long), hipMemcpyHostToDevice, s));
        HIPCHECK(hipMemcpyAsync(p->fwd_sv, sv.data(), sv.size() * sizeof(double), hipMemcpyHostToDevice, s));
        const int n = (int)sp.size() * nvar;
        hipLaunchKernelGGL(k_fwd_seeds, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p->fwd_sp, p->fwd_sv,
                           (int)sp.size(), nvar, p->fwd_res);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemsetAsync(p->adj_counts, 0, 4 * sizeof(int), s));
    hipLaunchKernelGGL(k_adj_frontier, dim3(nb), dim3(256), 0, s, p->g, p->fwd_pending, np, p->adj_list[0],
                       &p->adj_counts[0]);
    HIPCHECK(hipGetLastError());
    const unsigned grid = (unsigned)std::min<long long>((np + 255) / 256, 1024);
    long long k = 0;
    const long long max_rounds = np + 16;
    while (true) {
        for (int b = 0; b < 8; ++b, ++k) {
            hipLaunchKernelGGL((k_fwd_round<M, T>), dim3(grid), dim3(256), 0, s, p->g, p->sd<T>(), u, p->fixedmask, xact,
                               p->fwd_pending, p->fwd_rstart, p->fwd_rev, cv, nvar, p->fwd_res, p->adj_list[k & 1],
                               &p->adj_counts[k % 3], p->adj_list[(k + 1) & 1], &p->adj_counts[(k + 1) % 3],
                               &p->adj_counts[(k + 2) % 3]);
            HIPCHECK(hipGetLastError());
        }
        HIPCHECK(hipMemcpyAsync(p->h_count, &p->adj_counts[k % 3], sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        if (*p->h_count == 0) break;
        if (k > max_rounds) throw HipFail{"forward variation sweep did not terminate"};
    }
    p->st.forward_rounds = k;
    hipLaunchKernelGGL(k_fwd_export, dim3(nb), dim3(256), 0, s, p->g, p->fwd_res, nvar, d_out, np);
    HIPCHECK(hipGetLastError());
}

}  // namespace

int hfm_plan_adjoint_impl(hfm_plan *p, int ntargets, const double *targets, const double *weights,
                          double *d_sensitivity, int64_t *seed_index, double *seed_weight, int max_seeds,
                          int *nseed_out, void *stream) {
    if (!p || !d_sensitivity || (ntargets > 0 && !targets) || !nseed_out || (max_seeds > 0 && (!seed_index || !seed_weight))) {
        set_error("hfm_plan_adjoint: null argument");
        return 1;
    }
    if (!p->solved) {
        set_error("hfm_plan_adjoint: no solve has run");
        return 1;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        // Targets: ADim + floor; out of range is an error; a repeated index keeps its first
        // weight (std::map::insert in BackwardVariation, FirstVariation.h:351-354).
        std::map<long long, std::pair<long long, double>> seen;
        std::vector<long long> tp;
        std::vector<double> tw;
        const int dim = p->g.dim;
        for (int t = 0; t < ntargets; ++t) {
            int c[3] = {0, 0, 0};
            for (int i = 0; i < dim; ++i) {
                double q;
                if (p->shared && i == 2) q = (targets[t * dim + i] - p->origin[i]) / p->dependScale;
                else q = (targets[t * dim + i] - p->origin[i]) / p->h[i];
                const double f = std::floor(q);
                if (!(f > -2e9 && f < 2e9)) throw HipFail{"Error : inspectSensitivity data points are out of range"};
                int ci = (int)f;
                if (p->g.periodic >> i & 1) {
                    ci %= p->g.n[i];
                    if (ci < 0) ci += p->g.n[i];
                } else if (ci < 0 || ci >= p->g.n[i]) {
                    throw HipFail{"Error : inspectSensitivity data points are out of range"};
                }
                c[i] = ci;
            }
            const long long lin = ((long long)c[0] * p->g.n[1] + c[1]) * p->g.n[2] + c[2];
            const long long tpi = (long long)tile_of(p->g, c) * 64 + local_of(p->g, c);
            if (seen.insert({lin, {tpi, weights ? weights[t] : 1.}}).second) {
                tp.push_back(tpi);
                tw.push_back(weights ? weights[t] : 1.);
            }
        }
        if (!p->adj_lambda) {
            p->adj_lambda = dalloc<double>(p->np);
            p->adj_sens = dalloc<double>(p->np);
            p->adj_indeg = dalloc<int>(p->np);
            p->adj_ecnt = dalloc<int>(p->np);
            p->adj_var = dalloc<double>(p->np);
            // every point's active edges, one slot per difference of the model (Elastica2: 27 x
            // 12 bytes a point, 0.77 GB at 192 x 192 x 64)
            p->adj_eidx = dalloc<unsigned>((size_t)p->ndiff * p->np);
            p->adj_ew = dalloc<double>((size_t)p->ndiff * p->np);
        }
        ensure_lists(p);
        const int cap = std::max(max_seeds, 1);
        if (cap > p->adj_seed_cap) {
            dfree(p->adj_seed_idx);
            dfree(p->adj_seed_w);
            p->adj_seed_idx = nullptr;
            p->adj_seed_w = nullptr;
            p->adj_seed_idx = dalloc<long long>(cap);
            p->adj_seed_w = dalloc<double>(cap);
            p->adj_seed_cap = cap;
        }
        if ((int)tp.size() > p->adj_t_cap) {
            dfree(p->adj_tp);
            dfree(p->adj_tw);
            p->adj_tp = nullptr;
            p->adj_tw = nullptr;
            p->adj_t_cap = std::max<int>((int)tp.size(), 64);
            p->adj_tp = dalloc<long long>(p->adj_t_cap);
            p->adj_tw = dalloc<double>(p->adj_t_cap);
        }
        std::vector<long long> sidx;
        std::vector<double> sw;
        int ns = 0;
        const bool f32 = p->precision == 32;
        switch (p->model) {
            case MODEL_RIEMANN2:
                if (f32) adjoint_run<MODEL_RIEMANN2, float>(p, tp, tw, d_sensitivity, s, sidx, sw, cap, &ns);
                else adjoint_run<MODEL_RIEMANN2, double>(p, tp, tw, d_sensitivity, s, sidx, sw, cap, &ns);
                break;
            case MODEL_RIEMANN3:
                if (f32) adjoint_run<MODEL_RIEMANN3, float>(p, tp, tw, d_sensitivity, s, sidx, sw, cap, &ns);
                else adjoint_run<MODEL_RIEMANN3, double>(p, tp, tw, d_sensitivity, s, sidx, sw, cap, &ns);
                break;
            case MODEL_RSF2:
                if (f32) adjoint_run<MODEL_RSF2, float>(p, tp, tw, d_sensitivity, s, sidx, sw, cap, &ns);
                else adjoint_run<MODEL_RSF2, double>(p, tp, tw, d_sensitivity, s, sidx, sw, cap, &ns);
                break;
            case MODEL_ELASTICA2:
                if (f32) adjoint_run<MODEL_ELASTICA2, float>(p, tp, tw, d_sensitivity, s, sidx, sw, cap, &ns);
                else adjoint_run<MODEL_ELASTICA2, double>(p, tp, tw, d_sensitivity, s, sidx, sw, cap, &ns);
                break;
        }
        for (size_t i = 0; i < sidx.size() && (int)i < max_seeds; ++i) {
            seed_index[i] = sidx[i];
            seed_weight[i] = sw[i];
        }
        *nseed_out = ns;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

int hfm_plan_forward_variation_impl(hfm_plan *p, int nvar, const double *d_cost_variation, int nseeds,
                                    const double *seed_value_variation, double *d_value_variation, void *stream) {
    if (!p || !d_value_variation || (seed_value_variation && nseeds < 0)) {
        set_error("hfm_plan_forward_variation: null argument");
        return 1;
    }
    if (!p->solved) {
        set_error("hfm_plan_forward_variation: no solve has run");
        return 1;
    }
    if (nvar <= 0) {
        set_error("Forward variation error : inconsistent input sizes");
        return 1;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        // seedValueVariation (nVar x nseeds, row-major) onto the seeds of the last solve, in the
        // order they were passed; a later seed overwrites an earlier one at the same point.
        std::vector<long long> sp;
        std::vector<double> sv;
        if (seed_value_variation) {
            if (nseeds != (int)p->seed_inputs.size()) throw HipFail{"Error : Inconsistent size of seedValueVariation"};
            std::map<long long, int> last;
            for (int i = 0; i < nseeds; ++i)
                if (p->seed_inputs[i] >= 0) last[p->seed_inputs[i]] = i;
            for (const auto &kv : last) {
                sp.push_back(kv.first);
                for (int k = 0; k < nvar; ++k) sv.push_back(seed_value_variation[(size_t)k * nseeds + kv.second]);
            }
        }
        ensure_lists(p);
        if (!p->fwd_pending) {
            p->fwd_pending = dalloc<int>(p->np);
            p->fwd_rcount = dalloc<int>(p->np);
            p->fwd_rstart = dalloc<int>(p->np + 1);
            p->fwd_scan_work = dalloc<int>(scan_work_ints(p->np));
        }
        if ((long long)nvar * p->np > p->fwd_res_cap) {
            dfree(p->fwd_res);
            p->fwd_res = nullptr;
            p->fwd_res_cap = (long long)nvar * p->np;
            p->fwd_res = dalloc<double>(p->fwd_res_cap);
        }
        if ((int)sp.size() > p->fwd_seed_cap) {
            dfree(p->fwd_sp);
            p->fwd_sp = nullptr;
            p->fwd_seed_cap = std::max<int>((int)sp.size(), 16);
            p->fwd_sp = dalloc<long long>(p->fwd_seed_cap);
        }
        if ((long long)sv.size() > p->fwd_sv_cap) {
            dfree(p->fwd_sv);
            p->fwd_sv = nullptr;
            p->fwd_sv_cap = std::max<long long>((long long)sv.size(), 64);
            p->fwd_sv = dalloc<double>(p->fwd_sv_cap);
        }
        const bool f32 = p->precision == 32;
        switch (p->model) {
            case MODEL_RIEMANN2:
                if (f32) forward_run<MODEL_RIEMANN2, float>(p, nvar, d_cost_variation, sp, sv, d_value_variation, s);
                else forward_run<MODEL_RIEMANN2, double>(p, nvar, d_cost_variation, sp, sv, d_value_variation, s);
                break;
            case MODEL_RIEMANN3:
                if (f32) forward_run<MODEL_RIEMANN3, float>(p, nvar, d_cost_variation, sp, sv, d_value_variation, s);
                else forward_run<MODEL_RIEMANN3, double>(p, nvar, d_cost_variation, sp, sv, d_value_variation, s);
                break;
            case MODEL_RSF2:
                if (f32) forward_run<MODEL_RSF2, float>(p, nvar, d_cost_variation, sp, sv, d_value_variation, s);
                else forward_run<MODEL_RSF2, double>(p, nvar, d_cost_variation, sp, sv, d_value_variation, s);
                break;
            case MODEL_ELASTICA2:
                if (f32) forward_run<MODEL_ELASTICA2, float>(p, nvar, d_cost_variation, sp, sv, d_value_variation, s);
                else forward_run<MODEL_ELASTICA2, double>(p, nvar, d_cost_variation, sp, sv, d_value_variation, s);
                break;
        }
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}
