// hfm_geodesic.hip -- geodesic backtracking from tips: the discrete solver of
// Headers/Base/GeodesicDiscreteSolver.h:14-24 / Implementation/GeodesicDiscreteSolver.hxx:17-231,
// one wavefront per tip.
//
// The unit mass at the tip is split over the 2^d surrounding grid points and transported down
// the discrete geodesic flow: the point of largest (value, index) is taken out of the live set
// and its mass is spread over its active neighbours in proportion to w_i max(0, r - u_i)
// (HopfLaxRecompute at order 1, r the recomputed value).  The running weighted mean of the
// live points' indices gives the geodesic: a point is emitted whenever it has moved by
// geodesicStep.  A pass restarts from its last point when the mass spreads beyond
// geodesicVolumeBound (bound x 1.3 from the 5th restart, at most 50 restarts).
//
// Layout.  The live set (at most GEO_CAP points: the reference's std::map keyed by
// (value, index)) lives in the wave's LDS, GEO_CAP / 64 slots per lane; taking the largest key
// and finding a key are wave reductions / ballots.  The flow at a point is evaluated with one
// lane per stencil difference.  All bookkeeping (weight sum, index sum, covariance, steps) is
// FP64 in the reference's order of operations -- a vector divided by a scalar is a
// multiplication by the reciprocal (JMM_CPPLibs LinearAlgebra/VectorType.h:37-42) -- and runs
// uniformly on every lane.  On an FP64 solve the emitted points are those the reference's
// solver computes from the same values and active sets.
#include "hfm_kernels.h"

#include "hfm_device.h"

namespace hfm {

constexpr int GEO_CAP = 256;    // live mass points per tip
constexpr int GEO_STEPS = 256;  // steps accumulated between two emitted points

namespace {

struct Key {
    double val;
    int i0, i1, i2;
    int slot;  // -1: none
};

// (value, index) order of std::map<std::pair<double, IndexType>, double> (std::array compares
// lexicographically).
__device__ inline bool key_less(const Key &a, const Key &b) {
    if (a.val != b.val) return a.val < b.val;
    if (a.i0 != b.i0) return a.i0 < b.i0;
    if (a.i1 != b.i1) return a.i1 < b.i1;
    return a.i2 < b.i2;
}

__device__ inline Key shfl_key(const Key &k, int lane) {
    Key r;
    r.val = __shfl(k.val, lane);
    r.i0 = __shfl(k.i0, lane);
    r.i1 = __shfl(k.i1, lane);
    r.i2 = __shfl(k.i2, lane);
    r.slot = __shfl(k.slot, lane);
    return r;
}

struct GeoLDS {
    double val[GEO_CAP];
    double w[GEO_CAP];
    int idx[GEO_CAP][3];
    int live[GEO_CAP];
    double st_val[GEO_STEPS];
    double st_vec[GEO_STEPS][3];
    // flow at the current point, one entry per stencil difference
    uint32_t s_off[MAX_DIFF];
    double s_w[MAX_DIFF];
    double s_v[MAX_DIFF];
    int s_side[MAX_DIFF];
    // emitted flow entries and the values of their target points
    int f_off[MAX_DIFF][3];
    double f_w[MAX_DIFF];
    double f_val[MAX_DIFF];
};

// FP64 running sums of GeodesicDiscrete's UpdateSums (GeodesicDiscreteSolver.hxx:80-85).
struct Sums {
    double weightSum = 0., indexSum[3] = {0., 0., 0.}, cov[6] = {0., 0., 0., 0., 0., 0.};
    template <int D> __device__ void update(double weight, const int *ind) {
        weightSum += weight;
        double v[3];
#pragma unroll
        for (int i = 0; i < D; ++i) v[i] = (double)ind[i];
#pragma unroll
        for (int i = 0; i < D; ++i) indexSum[i] += v[i] * weight;
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) cov[i * (i + 1) / 2 + j] += (v[i] * v[j]) * weight;
    }
    template <int D> __device__ void mean(double *m) const {
        const double b = 1. / weightSum;
#pragma unroll
        for (int i = 0; i < D; ++i) m[i] = indexSum[i] * b;
    }
};

template <int D> __device__ inline double det_sym(const double *c) {
    auto at = [&](int i, int j) { return i >= j ? c[i * (i + 1) / 2 + j] : c[j * (j + 1) / 2 + i]; };
    if (D == 2) return at(0, 0) * at(1, 1) - at(1, 0) * at(0, 1);
    return at(0, 0) * (at(1, 1) * at(2, 2) - at(1, 2) * at(2, 1)) +
           at(0, 1) * (2. * at(1, 2) * at(2, 0) - at(1, 0) * at(2, 2)) - at(0, 2) * (at(1, 1) * at(2, 0));
}

// Grid index -> periodized in-grid coordinates (closed axes: false when outside).
template <int ModelT> __device__ inline bool periodize(const Geom &g, const int *idx, int *c) {
    constexpr int DIM = ModelShape<ModelT>::DIM, PER = ModelShape<ModelT>::PERIODIC;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= DIM) {
            c[k] = 0;
            continue;
        }
        int v = idx[k];
        if (PER >> k & 1) {
            v %= g.n[k];
            if (v < 0) v += g.n[k];
        } else if (v < 0 || v >= g.n[k]) {
            return false;
        }
        c[k] = v;
    }
    return true;
}

template <typename T> __device__ inline double value_at(const Geom &g, const T *u, const int *c) {
    return (double)u[(long long)tile_of(g, c) * 64 + local_of(g, c)];
}

// Stencil difference k of the point at tiled index p (theta index th for shared stencils), with
// the weight expression of the update kernel.
template <int ModelT, typename T>
__device__ inline void load_slot(const StencilData<T> &sd, long long p, int th, int k, uint32_t &off, T &w) {
    using S = ModelShape<ModelT>;
    constexpr int NC = S::NFWD + S::NSYM;
    if (!S::SHARED) {
        off = sd.offs[k * sd.np + p];
        w = sd.w[k * sd.np + p];
    } else {
        const T s = sd.speed[p];
        const T s2 = s * s;
        off = sd.tab_offs[th * NC + k];
        w = (T)sd.tab_w[th * NC + k] * s2;
    }
}

// Discrete flow at in-grid coordinates c (HopfLaxRecompute at order 1,
// EulerianStencil.hxx:155-233), in the active set the flow export uses: entries (signed
// offset, w_i max(0, r - v_i)) in stencil order into L.f_off / L.f_w; returns their number.
template <int ModelT, typename T>
__device__ int discrete_flow(const Geom &g, const StencilData<T> &sd, const T *u,
                             const unsigned long long *fixedmask, const int *c, GeoLDS &L, int lane) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    const int tile = tile_of(g, c), loc = local_of(g, c);
    const long long p = (long long)tile * 64 + loc;
    // Everything that depends on the point alone is loaded at once (its value, fixed bit and
    // visible sides with the stencil), so that a step costs two dependent memory latencies: the
    // stencil, then the neighbour values.
    const T uu = u[p];
    const bool fixed = (fixedmask[tile] >> loc) & 1ull;
    const uint32_t vm = sd.vis ? sd.vis[p] : 0xffffffffu;
    if (lane < NC) {
        uint32_t off;
        T w;
        load_slot<ModelT, T>(sd, p, c[2], lane, off, w);
        L.s_off[lane] = off;
        L.s_w[lane] = (double)w;
    }
    __syncthreads();
    if (lane < NC) {
        const int k = lane;
        const T inf = dev_inf<T>();
        int o0, o1, o2;
        unpack_offset(L.s_off[k], o0, o1, o2);
        // zero-weight differences are active only through a nonzero-weight one with their offset
        auto shares = [&](int a0, int a1, int a2) {
            for (int j = 0; j < NC; ++j) {
                if (j == k || L.s_w[j] == 0.) continue;
                int p0, p1, p2;
                unpack_offset(L.s_off[j], p0, p1, p2);
                if (p0 == a0 && p1 == a1 && p2 == a2) return true;
                if (j >= NF && p0 == -a0 && p1 == -a1 && p2 == -a2) return true;
            }
            return false;
        };
        const bool nz = L.s_w[k] != 0.;
        // walls: sides that are not visible never enter the active set
        const bool pa = (vm >> side_bit<NF>(k, 0) & 1u) && (nz || (S::SHARED && shares(o0, o1, o2)));
        const bool pb = k >= NF && (vm >> side_bit<NF>(k, 1) & 1u) && (nz || (S::SHARED && shares(-o0, -o1, -o2)));
        T va = inf, vb = inf;
        int r[3];
        if (pa && shift_coords_m<ModelT>(g, c, o0, o1, o2, r)) va = u[(long long)tile_of(g, r) * 64 + local_of(g, r)];
        if (pb && shift_coords_m<ModelT>(g, c, -o0, -o1, -o2, r)) vb = u[(long long)tile_of(g, r) * 64 + local_of(g, r)];
        L.s_v[k] = (double)(va <= vb ? va : vb);
        L.s_side[k] = va <= vb ? 1 : -1;
    }
    __syncthreads();
    if (fixed || !(uu < dev_inf<T>())) return 0;
    // active set: the local solver's accepted differences among those below u (as the flow export)
    T vv[NC], ww[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const T v = (T)L.s_v[k];
        vv[k] = v < uu ? v : dev_inf<T>();
        ww[k] = (T)L.s_w[k];
    }
    uint32_t used = 0;
    (void)local_solve<NC, T>(vv, ww, used);
    if (!used) return 0;
    double m = __builtin_huge_val();
    for (int k = 0; k < NC; ++k)
        if (used >> k & 1u) m = L.s_v[k] < m ? L.s_v[k] : m;
    double a = 0, b = 0, cc = -1;
    for (int k = 0; k < NC; ++k) {
        if (!(used >> k & 1u)) continue;
        const double w = L.s_w[k], vr = L.s_v[k] - m;
        a += w;
        b += w * vr;
        cc += (w * vr) * vr;
    }
    const double d2 = b * b - a * cc;
    const double value = m + (b + sqrt(d2 > 0. ? d2 : 0.)) / a;
    int n = 0;
    for (int k = 0; k < NC; ++k) {
        if (!(used >> k & 1u)) continue;
        if (lane == 0) {
            int o[3];
            unpack_offset(L.s_off[k], o[0], o[1], o[2]);
            for (int q = 0; q < 3; ++q) L.f_off[n][q] = L.s_side[k] * o[q];
            L.f_w[n] = L.s_w[k] * fmax(0., value - L.s_v[k]);
            L.f_val[n] = L.s_v[k];  // the value at the flow's target point, as loaded above
        }
        ++n;
    }
    __syncthreads();
    return n;
}

struct GeoParams {
    double step, threshold, volumeBound;
    double scale[3], origin[3];
    long long max_points;  // per tip
    long long max_iter;    // per pass
};

// Status codes per tip.
enum GeoStatus { GEO_OK = 0, GEO_FAILED = 1, GEO_SET_OVERFLOW = 2, GEO_STEPS_OVERFLOW = 3, GEO_NO_END = 4 };

template <int ModelT, typename T>
__global__ __launch_bounds__(64) void k_geodesic(Geom g, StencilData<T> sd, const T *__restrict__ u,
                                                 const unsigned long long *__restrict__ fixedmask, GeoParams gp,
                                                 const double *__restrict__ tips, double *__restrict__ points,
                                                 long long *__restrict__ lengths, int *__restrict__ status) {
    constexpr int D = ModelShape<ModelT>::DIM;
    constexpr int PER_LANE = GEO_CAP / 64;
    __shared__ GeoLDS L;
    const int tip = blockIdx.x, lane = threadIdx.x;
    long long npts = 0;
    double last[3] = {0., 0., 0.};
    int st = GEO_OK;
    auto emit = [&](const double *q) {
        if (npts < gp.max_points && lane == 0)
            for (int c = 0; c < D; ++c) points[((long long)tip * gp.max_points + npts) * D + c] = q[c] * gp.scale[c] + gp.origin[c];
        ++npts;
        for (int c = 0; c < D; ++c) last[c] = q[c];
    };
    {
        double t[3] = {0., 0., 0.};
        for (int c = 0; c < D; ++c) t[c] = tips[tip * D + c];
        emit(t);
    }
    int nRestarts = 0;
    double vb = gp.volumeBound;
    while (nRestarts < 50 && st == GEO_OK) {
        // ---- one pass of GeodesicDiscrete from `last` (GeodesicDiscreteSolver.hxx:56-231)
        bool restart = false;
        for (int j = 0; j < PER_LANE; ++j) L.live[lane + 64 * j] = 0;
        __syncthreads();
        int nlive = 0;
        Sums sums;
        double indexAvg[3] = {0., 0., 0.}, indexBase[3] = {0., 0., 0.};
        double stepNorm = 0.;
        int nsteps = 0;
        {
            int indexRef[3] = {0, 0, 0};
            double pointRef[3] = {0., 0., 0.};
            for (int j = 0; j < D; ++j) {
                indexRef[j] = (int)floor(last[j]);
                pointRef[j] = indexRef[j] + 0.5;
            }
            for (int i = 0; i < (1 << D); ++i) {
                double weight = 1;
                int index[3] = {indexRef[0], indexRef[1], indexRef[2]};
                for (int j = 0; j < D; ++j) {
                    const double diff = last[j] - pointRef[j];
                    if (i & (1 << j)) {
                        index[j] += diff > 0 ? 1 : -1;
                        weight *= fabs(diff);
                    } else {
                        weight *= 1. - fabs(diff);
                    }
                }
                int c[3];
                if (!periodize<ModelT>(g, index, c)) continue;
                if (lane == 0) {
                    L.val[nlive] = value_at(g, u, c);
                    L.w[nlive] = weight;
                    for (int q = 0; q < 3; ++q) L.idx[nlive][q] = q < D ? index[q] : 0;
                    L.live[nlive] = 1;
                }
                ++nlive;
                sums.update<D>(weight, index);
            }
            __syncthreads();
            if (sums.weightSum == 0) break;
            sums.mean<D>(indexAvg);
            for (int q = 0; q < 3; ++q) indexBase[q] = indexAvg[q];
        }
        long long iter = 0;
        while (nlive > 0 && sums.weightSum > 1e-5) {
            if (++iter > gp.max_iter) {
                st = GEO_NO_END;
                break;
            }
            // take the live point of largest (value, index)
            Key best{0., 0, 0, 0, -1};
            for (int j = 0; j < PER_LANE; ++j) {
                const int sl = lane + 64 * j;
                if (!L.live[sl]) continue;
                const Key k{L.val[sl], L.idx[sl][0], L.idx[sl][1], L.idx[sl][2], sl};
                if (best.slot < 0 || key_less(best, k)) best = k;
            }
            for (int o = 32; o > 0; o >>= 1) {
                const Key other = shfl_key(best, lane ^ o);
                if (other.slot >= 0 && (best.slot < 0 || key_less(best, other))) best = other;
            }
            const int slot = best.slot;
            if (slot < 0) {  // the live count and the live flags disagree: never expected
                st = GEO_NO_END;
                break;
            }
            const double value = best.val;
            const int index[3] = {best.i0, best.i1, best.i2};
            const double weight = L.w[slot];
            __syncthreads();
            if (lane == 0) L.live[slot] = 0;
            --nlive;
            __syncthreads();
            if (weight == 0) continue;
            sums.update<D>(-weight, index);
            int c[3];
            if (!periodize<ModelT>(g, index, c)) continue;  // live points are in the domain
            const int nf = discrete_flow<ModelT, T>(g, sd, u, fixedmask, c, L, lane);
            double wSum = 0;
            for (int k = 0; k < nf; ++k) wSum += L.f_w[k];
            if (wSum == 0) continue;
            // the values at the flow's target points (L.f_val) came with the flow: a used
            // difference's value is its target's, loaded by discrete_flow
            for (int k = 0; k < nf; ++k) {
                const double w = weight * L.f_w[k] / wSum;
                int ind[3] = {0, 0, 0};
                for (int q = 0; q < D; ++q) ind[q] = index[q] + L.f_off[k][q];
                const double val = L.f_val[k];
                int found = -1, freeslot = -1;
                {
                    int myf = -1, myfree = -1;
                    for (int j = 0; j < PER_LANE; ++j) {
                        const int sl = lane + 64 * j;
                        if (L.live[sl]) {
                            if (L.val[sl] == val && L.idx[sl][0] == ind[0] && L.idx[sl][1] == ind[1] &&
                                L.idx[sl][2] == ind[2])
                                myf = sl;
                        } else if (myfree < 0) {
                            myfree = sl;
                        }
                    }
                    const unsigned long long bf = __ballot(myf >= 0), bfree = __ballot(myfree >= 0);
                    if (bf) found = __shfl(myf, __ffsll((long long)bf) - 1);
                    if (bfree) freeslot = __shfl(myfree, __ffsll((long long)bfree) - 1);
                }
                if (found < 0) {
                    if (w < sums.weightSum * gp.threshold) continue;
                    if (freeslot < 0) {
                        st = GEO_SET_OVERFLOW;
                        break;
                    }
                    if (lane == 0) {
                        L.val[freeslot] = val;
                        L.w[freeslot] = w;
                        for (int q = 0; q < 3; ++q) L.idx[freeslot][q] = ind[q];
                        L.live[freeslot] = 1;
                    }
                    ++nlive;
                } else if (lane == 0) {
                    L.w[found] += w;
                }
                __syncthreads();
                sums.update<D>(w, ind);
            }
            if (st != GEO_OK) break;
            if (sums.weightSum <= 0) {
                double q[3];
                for (int i = 0; i < D; ++i) q[i] = indexAvg[i] + 0.5;
                emit(q);
                break;
            }
            double oldAvg[3] = {indexAvg[0], indexAvg[1], indexAvg[2]};
            sums.mean<D>(indexAvg);
            double step[3] = {0., 0., 0.};
            for (int i = 0; i < D; ++i) step[i] = indexAvg[i] - oldAvg[i];
            if (nsteps >= GEO_STEPS) {
                st = GEO_STEPS_OVERFLOW;
                break;
            }
            if (lane == 0) {
                L.st_val[nsteps] = value;
                for (int i = 0; i < 3; ++i) L.st_vec[nsteps][i] = step[i];
            }
            ++nsteps;
            __syncthreads();
            double sq = 0;
            for (int i = 0; i < D; ++i) sq += step[i] * step[i];
            stepNorm += sqrt(sq);
            if (nlive == 1) {
                double q[3];
                for (int i = 0; i < D; ++i) q[i] = indexAvg[i] + 0.5;
                emit(q);
                for (int i = 0; i < 3; ++i) indexBase[i] = indexAvg[i];
                stepNorm = 0;
                nsteps = 0;
                continue;
            }
            if (stepNorm < gp.step && nlive > 1) continue;
            const double valInc = L.st_val[0] - value;
            if (valInc <= 0) continue;
            const double delta = valInc / D;
            double p[3] = {indexBase[0], indexBase[1], indexBase[2]};
            for (int s = 0; s < nsteps; ++s) {
                const double a = (L.st_val[s] - value) + delta;
                const double rb = 1. / (valInc + delta);
                for (int i = 0; i < D; ++i) p[i] += (L.st_vec[s][i] * a) * rb;
            }
            {
                double q[3];
                for (int i = 0; i < D; ++i) q[i] = p[i] + 0.5;
                emit(q);
            }
            for (int i = 0; i < 3; ++i) indexBase[i] = indexAvg[i];
            stepNorm = 0;
            nsteps = 0;
            // restart when the mass is too spread: det(cov / W - m m^T + I) > bound^2
            double var[6], mm[3];
            const double bw = 1. / sums.weightSum;
            for (int k = 0; k < D * (D + 1) / 2; ++k) var[k] = sums.cov[k] * bw;
            sums.mean<D>(mm);
            for (int i = 0; i < D; ++i)
                for (int j = 0; j <= i; ++j) {
                    var[i * (i + 1) / 2 + j] -= mm[i] * mm[j];
                    var[i * (i + 1) / 2 + j] += (i == j) ? 1. : 0.;
                }
            if (det_sym<D>(var) > vb * vb) {
                double q[3];
                for (int i = 0; i < D; ++i) q[i] = indexAvg[i] + 0.5;
                emit(q);
                restart = true;
                break;
            }
        }
        if (!restart || st != GEO_OK) break;
        ++nRestarts;
        if (nRestarts >= 5) vb *= 1.3;
    }
    if (st == GEO_OK && nRestarts == 50) st = GEO_FAILED;
    if (lane == 0) {
        lengths[tip] = npts;
        status[tip] = st;
    }
}

}  // namespace

template <int M, typename T>
static void geo_one(const Geom &g, const StencilData<T> &sd, const T *u, const unsigned long long *fm,
                    const GeoParamsHost &h, int ntips, const double *tips, double *points, long long *lengths,
                    int *status, hipStream_t s) {
    GeoParams gp;
    gp.step = h.step;
    gp.threshold = h.threshold;
    gp.volumeBound = h.volumeBound;
    for (int i = 0; i < 3; ++i) {
        gp.scale[i] = h.scale[i];
        gp.origin[i] = h.origin[i];
    }
    gp.max_points = h.max_points;
    gp.max_iter = h.max_iter;
    hipLaunchKernelGGL((k_geodesic<M, T>), dim3(ntips), dim3(64), 0, s, g, sd, u, fm, gp, tips, points, lengths,
                       status);
}

template <typename T>
hipError_t launch_geodesics(int model, const Geom &g, const StencilData<T> &sd, const T *u,
                            const unsigned long long *fm, const GeoParamsHost &h, int ntips, const double *tips,
                            double *points, long long *lengths, int *status, hipStream_t s) {
    if (ntips <= 0) return hipSuccess;
    switch (model) {
        case MODEL_RIEMANN2: geo_one<MODEL_RIEMANN2, T>(g, sd, u, fm, h, ntips, tips, points, lengths, status, s); break;
        case MODEL_RIEMANN3: geo_one<MODEL_RIEMANN3, T>(g, sd, u, fm, h, ntips, tips, points, lengths, status, s); break;
        case MODEL_RSF2: geo_one<MODEL_RSF2, T>(g, sd, u, fm, h, ntips, tips, points, lengths, status, s); break;
        case MODEL_ELASTICA2: geo_one<MODEL_ELASTICA2, T>(g, sd, u, fm, h, ntips, tips, points, lengths, status, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_geodesics<float>(int, const Geom &, const StencilData<float> &, const float *,
                                            const unsigned long long *, const GeoParamsHost &, int, const double *,
                                            double *, long long *, int *, hipStream_t);
template hipError_t launch_geodesics<double>(int, const Geom &, const StencilData<double> &, const double *,
                                             const unsigned long long *, const GeoParamsHost &, int, const double *,
                                             double *, long long *, int *, hipStream_t);

}  // namespace hfm
