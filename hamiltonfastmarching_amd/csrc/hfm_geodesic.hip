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
// (value, index)) lives in the wave's registers, GEO_CAP / 64 rows of one slot per lane.  A slot
// holds the point's value as an order-preserving 64-bit key, its grid index packed into 64 bits
// (lexicographic order = unsigned order), and its mass; an empty slot holds key 0 and index
// ~0, and which slots are taken is a 64-bit lane mask per row in scalar registers.  An index
// fixes the point and hence its value, so a point is found by its packed index alone (one
// 64-bit compare per row and a ballot).  Taking the largest key is a per-lane pass over the rows
// plus two DPP maxima; a slot is written under one scalar branch on its row.  The flow at a
// point comes from a cache in LDS that holds the flows of every point of the last tiles the mass
// visited, each tile computed one lane per point when it is first needed: a step then waits on
// no memory access unless the mass enters a new tile.  All bookkeeping (weight sum, index sum,
// covariance, steps) is FP64 in the reference's order of operations -- a vector divided by a
// scalar is a multiplication by the reciprocal (JMM_CPPLibs LinearAlgebra/VectorType.h:37-42) --
// and runs uniformly on every lane.  On an FP64 solve the emitted
// points are those the reference's solver computes from the same values and active sets.  After
// a serial solve (order 2/3, static factoring, stopping criteria) the flow at a point is that
// loop's recompute over its accepted set instead (ExactFlow), as the reference takes it.
#include "hfm_kernels.h"

#include <type_traits>

#include "hfm_device.h"
#include "hfm_recompute.h"
#include "hfm_tile.h"

namespace hfm {

constexpr int GEO_CAP = 256;    // live mass points per tip
constexpr int GEO_STEPS = 256;  // steps accumulated between two emitted points

namespace {

// Grid indices packed 21 bits per axis with a bias of 2^20: the unsigned order of the packed
// keys is the lexicographic order of the indices (std::array's, as the reference's map key).
constexpr int GEO_IBITS = 21;
constexpr int GEO_IBIAS = 1 << (GEO_IBITS - 1);
constexpr unsigned long long GEO_IMASK = (1ull << GEO_IBITS) - 1;
constexpr unsigned long long GEO_INONE = ~0ull;  // empty slot (above every packed index)
__device__ inline bool pack_index(const int *ind, unsigned long long &key) {
    unsigned long long k = 0;
    bool ok = true;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int v = ind[q] + GEO_IBIAS;
        ok &= (unsigned)v <= (unsigned)GEO_IMASK;
        k = (k << GEO_IBITS) | ((unsigned long long)(unsigned)v & GEO_IMASK);
    }
    key = k;
    return ok;
}
__device__ inline void unpack_index(unsigned long long key, int *ind) {
    ind[2] = (int)(key & GEO_IMASK) - GEO_IBIAS;
    ind[1] = (int)((key >> GEO_IBITS) & GEO_IMASK) - GEO_IBIAS;
    ind[0] = (int)((key >> (2 * GEO_IBITS)) & GEO_IMASK) - GEO_IBIAS;
}
// The packed key of index + o is key + packed_delta(o) (two's complement): no field carries as
// long as every index of the key lies within +-(2^20 - 128) and |o| <= 127, which
// index_has_margin checks before a point's targets are formed.
__device__ inline unsigned long long packed_delta(int o0, int o1, int o2) {
    return ((unsigned long long)(long long)o0 << (2 * GEO_IBITS)) + ((unsigned long long)(long long)o1 << GEO_IBITS) +
           (unsigned long long)(long long)o2;
}
__device__ inline bool index_has_margin(const int *ind) {
    constexpr int lim = GEO_IBIAS - 128;
    return (ind[0] > -lim) & (ind[0] < lim) & (ind[1] > -lim) & (ind[1] < lim) & (ind[2] > -lim) & (ind[2] < lim);
}

struct GeoLDS {
    // the steps since the last emitted point (value, index step); every lane writes the same
    // entry, so every lane reads back what it wrote itself
    double st_val[GEO_STEPS];
    double st_vec[GEO_STEPS][3];
    // flow at one point evaluated one lane per stencil difference (points whose entries do not
    // fit in their tile's cache slot)
    uint32_t s_off[MAX_DIFF];
    double s_w[MAX_DIFF];
    double s_v[MAX_DIFF];
    int s_side[MAX_DIFF];
    int f_off[MAX_DIFF][3];
    double f_w[MAX_DIFF];
    double f_val[MAX_DIFF];
};

// Flow cache: the discrete flow of every point of the last few tiles the mass visited, computed
// one lane per point when a tile is first needed.  Entries of a tile are packed by point
// (prefix sum of the per-point counts); a point whose entries would pass the slot's capacity
// is marked GEO_UNCACHED and evaluated on its own when it is taken out.
constexpr int GEO_UNCACHED = 0xff;
template <int ModelT, typename T> struct FlowCacheShape {
    static constexpr int NC = ModelShape<ModelT>::NFWD + ModelShape<ModelT>::NSYM;
    static constexpr int ECAP = 64 * (NC < 6 ? NC : 6);  // entries per tile
    // every point's entries at a fixed place (point * NC): its count and entries are read in
    // one LDS round trip, and no point is ever left uncached
    static constexpr bool FIXED = NC <= 6;
    static constexpr int SLOT_BYTES = ECAP * (4 + 8 + (int)sizeof(T)) + 64 * (8 + 4);
    static constexpr int BUDGET = 64 * 1024;  // with GeoLDS, two workgroups fit in a CU's 160 KB
    static constexpr int NSLOT = BUDGET / SLOT_BYTES < 12 ? BUDGET / SLOT_BYTES : 12;
    static_assert(NSLOT >= 4, "flow cache too small");
};
template <int ModelT, typename T> struct FlowCache {
    using CS = FlowCacheShape<ModelT, T>;
    double wsum[CS::NSLOT][64];         // sum of the point's flow weights, in entry order
    double e_w[CS::NSLOT][CS::ECAP];    // w_i max(0, r - u_i)
    T e_v[CS::NSLOT][CS::ECAP];         // value at the flow's target point
    uint32_t e_off[CS::NSLOT][CS::ECAP];  // signed offset to the target (pack_offset)
    uint32_t info[CS::NSLOT][64];         // first entry << 8 | number of entries (or GEO_UNCACHED)
    // (with the fixed layout the first entry is point * NC)
};

// Wave-uniform copies of lane l's value (l uniform).
__device__ inline int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
// A condition or value every lane holds, marked as such: branches on it are scalar branches and
// the values merged after them stay in scalar registers (a uniform condition is tested through
// its own lane mask).
__device__ inline bool uni(bool c) { return __ballot(c) != 0ull; }
__device__ inline double uni_d(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double rl_d(double v, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Unsigned 64-bit key in the order of the double values (-0 and +0 equal); 0 is below every value.
__device__ inline unsigned long long order_key(double v) {
    v = v + 0.;
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// FP64 running sums of GeodesicDiscrete's UpdateSums (GeodesicDiscreteSolver.hxx:80-85), held
// alike by every lane: ten independent accumulations that the wave issues back to back.
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
                             const unsigned long long *fixedmask, const ExactFlow &xf, const int *c, GeoLDS &L,
                             int lane) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    if constexpr (std::is_same<T, double>::value) {
        if (xf.on) {  // after a serial solve: its recompute, evaluated by lane 0
            int n = 0;
            if (lane == 0) {
                const long long p = (long long)tile_of(g, c) * 64 + local_of(g, c);
                const unsigned long long act = xf.active[p];
                RecomputeEntry re[2 * NC];
                double value = 0.;
                if (act) n = recompute<ModelT>(g, sd, u, xf.acc, p, c, act, xf.order, xf.fac, re, &value);
                for (int i = 0; i < n; ++i) {
                    int r[3];
                    shift_coords_m<ModelT>(g, c, re[i].o[0], re[i].o[1], re[i].o[2], r);
                    for (int q = 0; q < 3; ++q) L.f_off[i][q] = re[i].o[q];
                    L.f_w[i] = flow_weight(re[i], value);
                    L.f_val[i] = u[(long long)tile_of(g, r) * 64 + local_of(g, r)];
                }
            }
            n = __shfl(n, 0);
            __syncthreads();
            return n;
        }
    }
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
    const uint32_t used = active_slots<NC, T>(vv, ww);
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

// The discrete flow of a block solve at one point (tiled index p, coordinates c): the
// differences the local solver accepts among those below u (the flow export's active set), each
// with its stencil offset and weight, the value on its chosen side sv and that side, and the
// recomputed value (HopfLaxRecompute at order 1, EulerianStencil.hxx:155-233).  Returns the
// used differences as a bit mask (0: a seed, an unreached point or a point past the grid).
template <int ModelT, typename T>
__device__ inline uint32_t block_point_flow(const Geom &g, const StencilData<T> &sd, const T *u,
                                            const unsigned long long *fixedmask, long long p, const int *c,
                                            bool inside, uint32_t *off, T *w, double *sv, int *side, double &value) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    const T inf = dev_inf<T>();
    const T uu = u[p];
    const bool fixed = (fixedmask[p >> 6] >> (p & 63)) & 1ull;
    const uint32_t vm = sd.vis ? sd.vis[p] : 0xffffffffu;
    const int th = inside ? c[2] : 0;  // theta row of a shared stencil (points past the grid: any row)
#pragma unroll
    for (int k = 0; k < NC; ++k) load_slot<ModelT, T>(sd, p, th, k, off[k], w[k]);
    // zero-weight differences are active only through a nonzero-weight one with their offset
    // (shared stencils); rare, so the table is read again rather than kept in registers
    auto shares = [&](int k, int a0, int a1, int a2) {
        for (int j = 0; j < NC; ++j) {
            if (j == k) continue;
            uint32_t oj;
            T wj;
            load_slot<ModelT, T>(sd, p, th, j, oj, wj);
            if ((double)wj == 0.) continue;
            int p0, p1, p2;
            unpack_offset(oj, p0, p1, p2);
            if (p0 == a0 && p1 == a1 && p2 == a2) return true;
            if (j >= NF && p0 == -a0 && p1 == -a1 && p2 == -a2) return true;
        }
        return false;
    };
    bool pa[NC], pb[NC];
    T va[NC], vb[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        int o0, o1, o2, r[3];
        unpack_offset(off[k], o0, o1, o2);
        const bool nz = (double)w[k] != 0.;
        pa[k] = inside && (vm >> side_bit<NF>(k, 0) & 1u) && (nz || (S::SHARED && shares(k, o0, o1, o2)));
        pb[k] = inside && k >= NF && (vm >> side_bit<NF>(k, 1) & 1u) &&
                (nz || (S::SHARED && shares(k, -o0, -o1, -o2)));
        pa[k] = pa[k] && shift_coords_m<ModelT>(g, c, o0, o1, o2, r);
        va[k] = u[pa[k] ? (long long)tile_of(g, r) * 64 + local_of(g, r) : p];
        if (k >= NF) {
            pb[k] = pb[k] && shift_coords_m<ModelT>(g, c, -o0, -o1, -o2, r);
            vb[k] = u[pb[k] ? (long long)tile_of(g, r) * 64 + local_of(g, r) : p];
        } else {
            vb[k] = inf;
        }
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const T a = pa[k] ? va[k] : inf, b = pb[k] ? vb[k] : inf;
        sv[k] = (double)(a <= b ? a : b);
        side[k] = a <= b ? 1 : -1;
    }
    uint32_t used = 0;
    if (inside && !fixed && uu < inf) {
        T vv[NC], ww[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const T v = (T)sv[k];
            vv[k] = v < uu ? v : inf;
            ww[k] = w[k];
        }
        used = active_slots<NC, T>(vv, ww);
    }
    value = 0;
    if (used) {
        double m = __builtin_huge_val();
#pragma unroll
        for (int k = 0; k < NC; ++k)
            if (used >> k & 1u) m = sv[k] < m ? sv[k] : m;
        double a = 0, b = 0, cc = -1;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            if (!(used >> k & 1u)) continue;
            const double wk = (double)w[k], vr = sv[k] - m;
            a += wk;
            b += wk * vr;
            cc += (wk * vr) * vr;
        }
        const double d2 = b * b - a * cc;
        value = m + (b + sqrt(d2 > 0. ? d2 : 0.)) / a;
    }
    return used;
}

// The discrete flow of the 64 points of a tile, one lane per point, into cache slot `slot`: the
// same active set and weights as discrete_flow, in the same operations.  Every lane issues its
// stencil loads, then all its neighbour-value loads, before it uses any of them.
template <int ModelT, typename T>
__device__ void fill_flow(const Geom &g, const StencilData<T> &sd, const T *u, const unsigned long long *fixedmask,
                          const ExactFlow &xf, int tile, int slot, FlowCache<ModelT, T> &C, int lane) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    using CS = FlowCacheShape<ModelT, T>;
    if constexpr (std::is_same<T, double>::value) {
        if (xf.on) {  // after a serial solve: its recompute, one lane per point
            int c[3];
            tiled_coords(g, tile, lane, c);
            const long long p = (long long)tile * 64 + lane;
            RecomputeEntry re[2 * NC];
            double value = 0.;
            int n = 0;
            if (c[0] < g.n[0] && c[1] < g.n[1] && c[2] < g.n[2]) {
                const unsigned long long act = xf.active[p];
                if (act) n = recompute<ModelT>(g, sd, u, xf.acc, p, c, act, xf.order, xf.fac, re, &value);
            }
            // the loop activates at most one side of a symmetric difference: at most NC entries
            int st = lane * NC;
            bool fits = n <= NC;
            if (!CS::FIXED) {
                int incl = n;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const int t = __shfl_up(incl, d);
                    if (lane >= d) incl += t;
                }
                st = incl - n;
                fits = incl <= CS::ECAP;
            }
            double ws = 0;
            for (int i = 0; i < n; ++i) {
                const double fw = flow_weight(re[i], value);
                ws += fw;
                if (fits) {
                    int r[3];
                    shift_coords_m<ModelT>(g, c, re[i].o[0], re[i].o[1], re[i].o[2], r);
                    C.e_off[slot][st + i] = pack_offset(re[i].o[0], re[i].o[1], re[i].o[2]);
                    C.e_w[slot][st + i] = fw;
                    C.e_v[slot][st + i] = u[(long long)tile_of(g, r) * 64 + local_of(g, r)];
                }
            }
            C.wsum[slot][lane] = ws;
            C.info[slot][lane] = fits ? (uint32_t)st << 8 | (uint32_t)n : (uint32_t)GEO_UNCACHED;
            return;
        }
    }
    int c[3];
    tiled_coords(g, tile, lane, c);
    const bool inside = c[0] < g.n[0] && c[1] < g.n[1] && c[2] < g.n[2];
    const long long p = (long long)tile * 64 + lane;
    uint32_t off[NC];
    T w[NC];
    double sv[NC];
    int side[NC];
    double value = 0;
    const uint32_t used = block_point_flow<ModelT, T>(g, sd, u, fixedmask, p, c, inside, off, w, sv, side, value);
    const int n = __builtin_popcount(used);
    int st = lane * NC;
    bool fits = true;
    if (!CS::FIXED) {
        int incl = n;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d);
            if (lane >= d) incl += t;
        }
        st = incl - n;
        fits = incl <= CS::ECAP;
    }
    double ws = 0;
    int e = st;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        if (!(used >> k & 1u)) continue;
        int o0, o1, o2;
        unpack_offset(off[k], o0, o1, o2);
        const double fw = (double)w[k] * fmax(0., value - sv[k]);
        ws += fw;
        if (fits) {
            C.e_off[slot][e] = pack_offset(side[k] * o0, side[k] * o1, side[k] * o2);
            C.e_w[slot][e] = fw;
            C.e_v[slot][e] = (T)sv[k];
        }
        ++e;
    }
    C.wsum[slot][lane] = ws;
    C.info[slot][lane] = fits ? (uint32_t)st << 8 | (uint32_t)n : (uint32_t)GEO_UNCACHED;
}

struct GeoParams {
    double step, threshold, volumeBound;
    double scale[3], origin[3];
    long long max_points;  // per tip
    long long max_iter;    // per pass
};

// Status codes per tip.
enum GeoStatus { GEO_OK = 0, GEO_FAILED = 1, GEO_SET_OVERFLOW = 2, GEO_STEPS_OVERFLOW = 3, GEO_NO_END = 4,
                 GEO_INDEX_RANGE = 5 };

template <int ModelT, typename T>
__global__ __launch_bounds__(64) void k_geodesic(Geom g, StencilData<T> sd, const T *__restrict__ u,
                                                 const unsigned long long *__restrict__ fixedmask, ExactFlow xf,
                                                 GeoParams gp,
                                                 const double *__restrict__ tips, double *__restrict__ points,
                                                 long long *__restrict__ lengths, int *__restrict__ status) {
    constexpr int D = ModelShape<ModelT>::DIM;
    constexpr int ROWS = GEO_CAP / 64;
    using CS = FlowCacheShape<ModelT, T>;
    constexpr int NC = CS::NC;
    __shared__ GeoLDS L;
    __shared__ FlowCache<ModelT, T> C;
    const int tip = blockIdx.x, lane = threadIdx.x;
    // The live set: row j of every lane's registers; live[j] (wave-uniform) marks the taken slots.
    unsigned long long okey[ROWS], ikey[ROWS], live[ROWS];
    double wt[ROWS];
    int nlive = 0;
    // Write a slot (row r and lane l uniform) under one scalar branch on its row.
    auto slot_put = [&](int r, int l, unsigned long long ok, unsigned long long ik, double w) {
#pragma unroll
        for (int j = 0; j < ROWS; ++j)
            if (j == r) {
                if (lane == l) {
                    okey[j] = ok;
                    ikey[j] = ik;
                    wt[j] = w;
                }
                live[j] |= 1ull << l;
            }
    };
    auto slot_clear = [&](int r, int l) {
#pragma unroll
        for (int j = 0; j < ROWS; ++j)
            if (j == r) {
                if (lane == l) {
                    okey[j] = 0ull;
                    ikey[j] = GEO_INONE;
                }
                live[j] &= ~(1ull << l);
            }
    };
    auto slot_add = [&](int r, int l, double w) {
#pragma unroll
        for (int j = 0; j < ROWS; ++j)
            if (j == r && lane == l) wt[j] += w;
    };
    // Flow cache directory: lane s < NSLOT holds the tile in slot s (-1: empty) and its last use.
    int my_tag = -1;
    unsigned my_use = 0, clock = 0;
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
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            okey[j] = 0ull;
            ikey[j] = GEO_INONE;
            wt[j] = 0.;
            live[j] = 0ull;
        }
        nlive = 0;
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
                if (!uni(periodize<ModelT>(g, index, c))) continue;
                unsigned long long ik;
                if (!uni(pack_index(index, ik))) {
                    st = GEO_INDEX_RANGE;
                    break;
                }
                slot_put(0, nlive, order_key(value_at(g, u, c)), ik, weight);  // nlive < 8: row 0
                ++nlive;
                sums.update<D>(weight, index);
            }
            if (st != GEO_OK) break;
            if (uni(sums.weightSum == 0)) break;
            sums.mean<D>(indexAvg);
            for (int q = 0; q < 3; ++q) indexBase[q] = indexAvg[q];
        }
        long long iter = 0;
        while (uni(nlive > 0 && sums.weightSum > 1e-5)) {
            if (++iter > gp.max_iter) {
                st = GEO_NO_END;
                break;
            }
            // Take the live point of largest (value, index): each lane's largest row (empty slots
            // hold key 0 and index ~0, and lose to any taken slot), then the largest key over the
            // wave (two 32-bit DPP maxima); equal keys in several lanes are ordered by index.
            unsigned long long bo = okey[0], bi = ikey[0];
            double bw = wt[0];
            int br = 0;
#pragma unroll
            for (int j = 1; j < ROWS; ++j) {
                const bool gt = (okey[j] > bo) | ((okey[j] == bo) & (ikey[j] > bi));
                bo = gt ? okey[j] : bo;
                bi = gt ? ikey[j] : bi;
                bw = gt ? wt[j] : bw;
                br = gt ? j : br;
            }
            const unsigned khi = (unsigned)(bo >> 32), klo = (unsigned)bo;
            const unsigned mhi = wave_max_u32(khi);
            const unsigned mlo = wave_max_u32(khi == mhi ? klo : 0u);
            const unsigned long long tie = __ballot((khi == mhi) & (klo == mlo));
            if ((mhi | mlo) == 0u) {  // the live count and the live slots disagree: never expected
                st = GEO_NO_END;
                break;
            }
            int wl = __ffsll((long long)tie) - 1;
            unsigned long long wik = (unsigned long long)(unsigned)rl_i((int)(unsigned)(bi >> 32), wl) << 32 |
                                     (unsigned)rl_i((int)(unsigned)bi, wl);
            for (unsigned long long r = tie & (tie - 1); r; r &= r - 1) {
                const int l = __ffsll((long long)r) - 1;
                const unsigned long long cik = (unsigned long long)(unsigned)rl_i((int)(unsigned)(bi >> 32), l) << 32 |
                                               (unsigned)rl_i((int)(unsigned)bi, l);
                if (cik > wik) {
                    wl = l;
                    wik = cik;
                }
            }
            const int wr = rl_i(br, wl);
            const unsigned long long wkey = (unsigned long long)mhi << 32 | mlo;
            const double value =
                __longlong_as_double((long long)((wkey >> 63) ? (wkey & 0x7fffffffffffffffull) : ~wkey));
            const double weight = rl_d(bw, wl);
            int index[3];
            unpack_index(wik, index);
            slot_clear(wr, wl);
            --nlive;
            if (uni(weight == 0)) continue;
            if (!uni(index_has_margin(index))) {
                st = GEO_INDEX_RANGE;
                break;
            }
            sums.update<D>(-weight, index);
            int c[3];
            if (!uni(periodize<ModelT>(g, index, c))) continue;  // live points are in the domain
            // The point's flow: from its tile's cache slot, the least recently used slot being
            // filled first when the tile is not cached.
            const int tile = tile_of(g, c), loc = local_of(g, c);
            const unsigned long long hit = __ballot(lane < CS::NSLOT && my_tag == tile);
            int slot;
            if (hit) {
                slot = __ffsll((long long)hit) - 1;
            } else {
                const unsigned oldest = wave_min_u32(lane < CS::NSLOT ? my_use : 0xffffffffu);
                slot = __ffsll((long long)__ballot(lane < CS::NSLOT && my_use == oldest)) - 1;
                fill_flow<ModelT, T>(g, sd, u, fixedmask, xf, tile, slot, C, lane);
                if (lane == slot) my_tag = tile;
                __syncthreads();
            }
            ++clock;
            if (lane == slot) my_use = clock;
            // the point's flow entries, entry k in lane k
            int nf;
            double my_w = 0., my_v = 0., wSum;
            uint32_t my_off = 0;
            if (CS::FIXED) {
                // count, weight sum and entries in one round trip (entries past the count unused)
                const int e = loc * NC + (lane < NC ? lane : 0);
                const unsigned info = C.info[slot][loc];
                const double ws = C.wsum[slot][loc];
                const double ew = C.e_w[slot][e];
                const T ev = C.e_v[slot][e];
                const uint32_t eo = C.e_off[slot][e];
                nf = (int)(__builtin_amdgcn_readfirstlane((int)info) & 0xff);
                wSum = uni_d(ws);
                if (lane < nf) {
                    my_w = ew;
                    my_v = (double)ev;
                    my_off = eo;
                }
            } else {
                const unsigned info = (unsigned)__builtin_amdgcn_readfirstlane((int)C.info[slot][loc]);
                wSum = uni_d(C.wsum[slot][loc]);
                if (info != (unsigned)GEO_UNCACHED) {
                    nf = (int)(info & 0xffu);
                    const int e = (int)(info >> 8) + lane;
                    if (lane < nf) {
                        my_w = C.e_w[slot][e];
                        my_v = (double)C.e_v[slot][e];
                        my_off = C.e_off[slot][e];
                    }
                } else {
                    nf = discrete_flow<ModelT, T>(g, sd, u, fixedmask, xf, c, L, lane);
                    wSum = 0;
                    for (int k = 0; k < nf; ++k) wSum += L.f_w[k];
                    wSum = uni_d(wSum);
                    if (lane < nf) {
                        my_w = L.f_w[lane];
                        my_v = L.f_val[lane];
                        my_off = pack_offset(L.f_off[lane][0], L.f_off[lane][1], L.f_off[lane][2]);
                    }
                }
            }
            if (uni(wSum == 0)) continue;
            const double my_share = weight * my_w / wSum;  // entry k's mass, in lane k
            // entry k's target in lane k: its index and packed key offset
            int my_ind[3] = {index[0], index[1], index[2]};
            unsigned long long my_delta;
            {
                int o[3];
                unpack_offset(my_off, o[0], o[1], o[2]);
                for (int q = 0; q < D; ++q) my_ind[q] += o[q];
                my_delta = packed_delta(o[0], o[1], D > 2 ? o[2] : 0);
            }
            // spread the mass over the flow's targets; a flow entry carries its target's value
            for (int k = 0; k < nf; ++k) {
                const double w = rl_d(my_share, k);
                int ind[3] = {0, 0, 0};
                for (int q = 0; q < D; ++q) ind[q] = rl_i(my_ind[q], k);
                const unsigned long long tk =
                    wik + ((unsigned long long)(unsigned)rl_i((int)(unsigned)(my_delta >> 32), k) << 32 |
                           (unsigned)rl_i((int)(unsigned)my_delta, k));
                // the target's slot, if live: the first row holding its index
                int fr = -1;
                unsigned long long fm = 0ull;
#pragma unroll
                for (int j = 0; j < ROWS; ++j) {
                    const unsigned long long m = __ballot(ikey[j] == tk);
                    if (fr < 0 && m) {
                        fr = j;
                        fm = m;
                    }
                }
                if (fm) {
                    slot_add(fr, __ffsll((long long)fm) - 1, w);
                } else {
                    if (uni(w < sums.weightSum * gp.threshold)) continue;
                    // a free slot: the first free lane of the first row that has one
                    int r = -1;
                    unsigned long long fmask = 0ull;
#pragma unroll
                    for (int j = 0; j < ROWS; ++j)
                        if (r < 0 && ~live[j]) {
                            r = j;
                            fmask = ~live[j];
                        }
                    if (r < 0) {
                        st = GEO_SET_OVERFLOW;
                        break;
                    }
                    slot_put(r, __ffsll((long long)fmask) - 1, order_key(rl_d(my_v, k)), tk, w);
                    ++nlive;
                }
                sums.update<D>(w, ind);
            }
            if (st != GEO_OK) break;
            if (uni(sums.weightSum <= 0)) {
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
            L.st_val[nsteps] = value;
            for (int i = 0; i < 3; ++i) L.st_vec[nsteps][i] = step[i];
            ++nsteps;
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
            if (uni(stepNorm < gp.step && nlive > 1)) continue;
            const double valInc = L.st_val[0] - value;
            if (uni(valInc <= 0)) continue;
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
            const double bw2 = 1. / sums.weightSum;
            for (int k = 0; k < D * (D + 1) / 2; ++k) var[k] = sums.cov[k] * bw2;
            sums.mean<D>(mm);
            for (int i = 0; i < D; ++i)
                for (int j = 0; j <= i; ++j) {
                    var[i * (i + 1) / 2 + j] -= mm[i] * mm[j];
                    var[i * (i + 1) / 2 + j] += (i == j) ? 1. : 0.;
                }
            if (uni(det_sym<D>(var) > vb * vb)) {
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

// ---------------------------------------------------------------------------------------------
// The ODE geodesic solver (geodesicSolver = "ODE"; Headers/Base/GeodesicODESolver.h:17-40,
// Implementation/GeodesicODESolver.hxx:10-261), one wavefront per tip.  From the tip it takes
// midpoint steps of geodesicStep along the normalised flow.  The flow at a point averages the
// GeodesicFlow of the 2^d surrounding grid points (HamiltonFastMarching.hxx:363-378: the
// recompute's value, its width and sum_i weight_i offset_i) with the interpolation weights, over
// the corners whose value lies within causalityTolerance widths of the smallest; lane j handles
// corner j and lane order is the corners' order in every sum.  The walk stops at a seed (zero
// flow), when the L-infinity grid distance to the seeds (capped at 6) has grown against
// sqrt(d) / step steps before, or when a corner's flow has been read more than 6 sqrt(d) / step
// times (a stall: the tip fails).  The read counts live in a per-tip hash table in device memory.
// ---------------------------------------------------------------------------------------------
constexpr int ODE_HIST = 1024;        // distance history: sqrt(d) / geodesicStep entries at most
constexpr int ODE_DIST_MAX = 32767;   // the reference's unreached distance (largest short)
constexpr int ODE_DIST_BOUND = 6;     // the distance's upper bound
constexpr int ODE_LDS_SLOTS = 2048;   // read counts and flows kept in LDS before the device-memory table
constexpr int ODE_LDS_PROBES = 64;
static_assert(ODE_LDS_SLOTS == 1 << 11, "the LDS table's hash takes 11 bits");
enum { GEO_HASH_FULL = 6, GEO_HISTORY = 7 };

// LInfDistance (GeodesicODESolver.hxx:221-258; upper bound 6, the value the solver's
// constructor uses) is a breadth-first search over the 3^d - 1 neighbours that stops at the
// bound.  Without walls its level is the Chebyshev distance (periodic axes wrapped) to the
// nearest seed when that is at most 6, and unreached beyond: each thread here takes one offset
// of the (13)^d box around one seed and lowers the point's distance to max |offset_i|.
template <int ModelT>
__global__ __launch_bounds__(256) void k_ode_dist_box(Geom g, int *__restrict__ dist,
                                                      const long long *__restrict__ seeds, int nseeds) {
    constexpr int D = ModelShape<ModelT>::DIM;
    constexpr int W = 2 * ODE_DIST_BOUND + 1, BOX = D == 2 ? W * W : W * W * W;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nseeds * BOX) return;
    const int s = (int)(t / BOX), k = (int)(t % BOX);
    const int o0 = k % W - ODE_DIST_BOUND, o1 = (k / W) % W - ODE_DIST_BOUND,
              o2 = D == 3 ? k / (W * W) - ODE_DIST_BOUND : 0;
    const long long sp = seeds[s];
    int c[3], r[3];
    tiled_coords(g, (int)(sp >> 6), (int)(sp & 63), c);
    if (!shift_coords_m<ModelT>(g, c, o0, o1, o2, r)) return;
    const int a0 = o0 < 0 ? -o0 : o0, a1 = o1 < 0 ? -o1 : o1, a2 = o2 < 0 ? -o2 : o2;
    const int d = a0 > a1 ? (a0 > a2 ? a0 : a2) : (a1 > a2 ? a1 : a2);
    atomicMin(&dist[(long long)tile_of(g, r) * 64 + local_of(g, r)], d);
}

__global__ void k_ode_dist_init(int *__restrict__ dist, long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) dist[p] = ODE_DIST_MAX;
}


// GeodesicFlow at grid point q (coordinates c): value, width and flow vector.
template <int ModelT, typename T>
__device__ inline void ode_point_flow(const Geom &g, const StencilData<T> &sd, const T *u,
                                      const unsigned long long *fixedmask, const ExactFlow &xf, long long q,
                                      const int *c, double &value, double &width, double *f) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    f[0] = f[1] = f[2] = 0.;
    value = (double)u[q];
    width = 0.;
    if constexpr (std::is_same<T, double>::value) {
        if (xf.on) {  // after a serial solve: its recompute over its accepted set
            const unsigned long long act = xf.active[q];
            if (!act) return;
            RecomputeEntry re[2 * NC];
            double r = 0.;
            const int n = recompute<ModelT>(g, sd, u, xf.acc, q, c, act, xf.order, xf.fac, re, &r);
            constexpr double mult[4] = {0., 1., 3. / 2., 11. / 6.};
            double wid = 0., ws = 0.;
            for (int i = 0; i < n; ++i) {
                const double d = r - re[i].value;
                const double wpd = re[i].w * (d > 0. ? d : 0.);
                wid += wpd;
                const double fw = wpd * mult[re[i].ord];
                for (int k = 0; k < 3; ++k) f[k] += fw * (double)re[i].o[k];
            }
            for (int i = 0; i < n; ++i) ws += re[i].w;
            if (ws > 0) wid /= ws;
            value = r;
            width = wid;
            return;
        }
    }
    uint32_t off[NC];
    T w[NC];
    double sv[NC];
    int side[NC];
    double r = 0.;
    const uint32_t used = block_point_flow<ModelT, T>(g, sd, u, fixedmask, q, c, true, off, w, sv, side, r);
    if (!used) return;
    double wid = 0., ws = 0.;
    for (int k = 0; k < NC; ++k) {
        if (!(used >> k & 1u)) continue;
        const double d = r - sv[k];
        const double wpd = (double)w[k] * (d > 0. ? d : 0.);
        wid += wpd;
        int o[3];
        unpack_offset(off[k], o[0], o[1], o[2]);
        for (int j = 0; j < 3; ++j) f[j] += wpd * (double)(side[k] * o[j]);
    }
    for (int k = 0; k < NC; ++k)
        if (used >> k & 1u) ws += (double)w[k];
    if (ws > 0) wid /= ws;
    value = r;
    width = wid;
}

template <int ModelT, typename T>
__global__ __launch_bounds__(64) void k_geodesic_ode(Geom g, StencilData<T> sd, const T *__restrict__ u,
                                                     const unsigned long long *__restrict__ fixedmask, ExactFlow xf,
                                                     GeoParams gp, double causality, const int *__restrict__ dist,
                                                     unsigned long long *__restrict__ hkeys, int *__restrict__ hcnt,
                                                     int hlog2, const double *__restrict__ tips,
                                                     double *__restrict__ points, long long *__restrict__ lengths,
                                                     int *__restrict__ status) {
    constexpr int D = ModelShape<ModelT>::DIM;
    constexpr int NCORNER = 1 << D;
    __shared__ long long hist[ODE_HIST];
    __shared__ unsigned long long s_keys[ODE_LDS_SLOTS];
    __shared__ int s_cnt[ODE_LDS_SLOTS];
    __shared__ double s_flow[ODE_LDS_SLOTS][5];  // value, width, flow: computed at the first read
    __shared__ int s_dist[ODE_LDS_SLOTS];         // the point's distance to the seeds
    __shared__ double s_corner[NCORNER][6];       // one evaluation's corners: weight, value, width, flow
    __shared__ long long s_cint[NCORNER][3];      // read count, distance, table full
    const int tip = blockIdx.x, lane = threadIdx.x;
    const long long hcap = 1LL << hlog2;
    unsigned long long *keys = hkeys + (long long)tip * hcap;
    int *cnts = hcnt + (long long)tip * hcap;
    double p[3] = {0., 0., 0.};
    for (int c = 0; c < D; ++c) p[c] = tips[tip * D + c];
    long long npts = 0;
    int st = GEO_OK;
    auto emit = [&](const double *q) {
        if (lane == 0 && npts < gp.max_points)
            for (int c = 0; c < D; ++c) points[((long long)tip * gp.max_points + npts) * D + c] = q[c] * gp.scale[c] + gp.origin[c];
        ++npts;
    };
    emit(p);
    const long long nStatMax = (long long)((6. * sqrt((double)D)) / gp.step);
    int K = 0;
    while (K <= ODE_HIST && K < sqrt((double)D) / gp.step) ++K;
    if (K > ODE_HIST) {
        if (lane == 0) {
            lengths[tip] = npts;
            status[tip] = GEO_HISTORY;
        }
        return;
    }
    for (int i = lane; i < K; i += 64) hist[i] = 2147483647LL;
    for (int i = lane; i < ODE_LDS_SLOTS; i += 64) {
        s_keys[i] = 0ull;
        s_cnt[i] = 0;
    }
    __syncthreads();
    int head = 0;
    // the flow at x, in every lane; false if the count table is full
    auto flow_at = [&](const double *x, double *fo, long long &nStat, long long &tol) {
        __syncthreads();  // the previous evaluation's LDS writes before this one's reads
        long long pIndex[3] = {0, 0, 0};
        double w[3] = {0., 0., 0.}, sgn[3] = {0., 0., 0.};
        for (int i = 0; i < D; ++i) {
            pIndex[i] = (long long)floor(x[i]);
            w[i] = x[i] - (double)pIndex[i] - 0.5;
            sgn[i] = w[i] > 0 ? 1 : -1;
            w[i] = fabs(w[i]);
        }
        double cw = 0., cval = __builtin_huge_val(), cwid = 0., cf[3] = {0., 0., 0.};
        long long ctol = 2147483647LL, cstat = 0;
        bool full = false;
        if (lane < NCORNER) {
            int base[3] = {0, 0, 0}, off[3] = {0, 0, 0};
            bool inrange = true;
            for (int j = 0; j < D; ++j) {
                off[j] = (lane >> j & 1) ? (int)sgn[j] : 0;
                inrange = inrange && pIndex[j] > -(1LL << 30) && pIndex[j] < (1LL << 30);
                base[j] = inrange ? (int)pIndex[j] : 0;
            }
            int r[3];
            if (inrange && shift_coords_m<ModelT>(g, base, off[0], off[1], off[2], r)) {
                const long long q = (long long)tile_of(g, r) * 64 + local_of(g, r);
                // the corner's read count: linear probing on key q + 1 (0 marks an empty slot),
                // first over at most ODE_LDS_PROBES slots of the LDS table, then, when those are
                // all taken by other keys, in the device-memory table.  Slots are never freed,
                // so a key that went to device memory finds the same full LDS chain again.
                const unsigned long long key = (unsigned long long)q + 1ull;
                const unsigned long long mix = key * 0x9E3779B97F4A7C15ull;
                int ls = -1;
                {
                    int h = (int)(mix >> 53);  // 11 bits: ODE_LDS_SLOTS = 2048
                    for (int probe = 0; probe < ODE_LDS_PROBES; ++probe) {
                        const unsigned long long k = s_keys[h];
                        if (k == key) {
                            ls = h;
                            break;
                        }
                        if (k == 0ull) {
                            const unsigned long long prev = atomicCAS(&s_keys[h], 0ull, key);
                            if (prev == 0ull || prev == key) {
                                ls = h;
                                break;
                            }
                        }
                        h = (h + 1) & (ODE_LDS_SLOTS - 1);
                    }
                }
                long long h = (long long)(mix >> (64 - hlog2));
                long long slot = -1;
                for (long long probe = 0; ls < 0 && probe < hcap; ++probe) {
                    const unsigned long long k = keys[h];
                    if (k == key) {
                        slot = h;
                        break;
                    }
                    if (k == 0ull) {
                        const unsigned long long prev = atomicCAS(&keys[h], 0ull, key);
                        if (prev == 0ull || prev == key) {
                            slot = h;
                            break;
                        }
                    }
                    h = (h + 1) & (hcap - 1);
                }
                if (ls < 0 && slot < 0) {
                    full = true;
                    ctol = dist[q];
                } else {
                    cstat = ls >= 0 ? atomicAdd(&s_cnt[ls], 1) : atomicAdd(&cnts[slot], 1);
                    cw = 1.;
                    for (int j = 0; j < D; ++j) cw *= off[j] ? w[j] : (1. - w[j]);
                    // a point's flow is a function of the solve alone: computed at its first read
                    // and kept with its count (the corners of one evaluation are distinct points,
                    // so no other lane reads a slot while it is written)
                    if (ls >= 0 && cstat > 0) {
                        cval = s_flow[ls][0];
                        cwid = s_flow[ls][1];
                        for (int k = 0; k < 3; ++k) cf[k] = s_flow[ls][2 + k];
                        ctol = s_dist[ls];
                    } else {
                        ctol = dist[q];
                        ode_point_flow<ModelT, T>(g, sd, u, fixedmask, xf, q, r, cval, cwid, cf);
                        if (ls >= 0) {
                            s_flow[ls][0] = cval;
                            s_flow[ls][1] = cwid;
                            for (int k = 0; k < 3; ++k) s_flow[ls][2 + k] = cf[k];
                            s_dist[ls] = (int)ctol;
                        }
                    }
                }
            }
        }
        // the corners in order: smallest value (first on ties), threshold, weighted mean
        double minVal = __builtin_huge_val(), thrW = 0.;
        long long mstat = 0, mtol = 2147483647LL;
        bool anyfull = false;
        double wv[NCORNER], vv[NCORNER], fv[NCORNER][3];
        // every lane reads the corners back from LDS (broadcast reads); the next evaluation
        // starts with a barrier, so these slots are not rewritten before they are read
        if (lane < NCORNER) {
            s_corner[lane][0] = cw;
            s_corner[lane][1] = cval;
            s_corner[lane][2] = cwid;
            for (int k = 0; k < 3; ++k) s_corner[lane][3 + k] = cf[k];
            s_cint[lane][0] = cstat;
            s_cint[lane][1] = ctol;
            s_cint[lane][2] = full ? 1 : 0;
        }
        __syncthreads();
        for (int i = 0; i < NCORNER; ++i) {
            wv[i] = s_corner[i][0];
            vv[i] = s_corner[i][1];
            const double wd = s_corner[i][2];
            for (int k = 0; k < D; ++k) fv[i][k] = s_corner[i][3 + k];
            const long long si = s_cint[i][0], ti = s_cint[i][1];
            mstat = si > mstat ? si : mstat;
            mtol = ti < mtol ? ti : mtol;
            anyfull = anyfull || s_cint[i][2] != 0;
            if (wv[i] > 0 && vv[i] < minVal) {
                minVal = vv[i];
                thrW = wd;
            }
        }
        nStat = mstat;
        tol = mtol;
        fo[0] = fo[1] = fo[2] = 0.;
        if (minVal == __builtin_huge_val()) return !anyfull;
        const double thr = minVal + causality * thrW;
        double wsum = 0.;
        for (int i = 0; i < NCORNER; ++i) {
            if (wv[i] <= 0) continue;
            if (vv[i] <= thr) {
                wsum += wv[i];
                for (int k = 0; k < D; ++k) fo[k] += wv[i] * fv[i][k];
            }
        }
        if (wsum > 0) {
            const double b = 1. / wsum;
            for (int k = 0; k < D; ++k) fo[k] *= b;
        }
        return !anyfull;
    };
    auto norm = [&](const double *f) {
        double s2 = 0.;
        for (int k = 0; k < D; ++k) s2 += f[k] * f[k];
        return sqrt(s2);
    };
    while (true) {
        double f[3];
        long long nStat, tol;
        if (!flow_at(p, f, nStat, tol)) {
            st = GEO_HASH_FULL;
            break;
        }
        const long long old = hist[head];
        __syncthreads();
        if (lane == 0) hist[head] = tol;
        __syncthreads();
        head = head + 1 == K ? 0 : head + 1;
        if (old < tol) break;  // the walk moves away from the seeds
        if (nStat > nStatMax) {
            st = GEO_FAILED;  // stall
            break;
        }
        double fn = norm(f);
        if (fn == 0) break;  // at a seed
        double q[3] = {p[0], p[1], p[2]};
        {
            const double a = 0.5 * gp.step, b = 1. / fn;
            for (int k = 0; k < D; ++k) q[k] = p[k] + (f[k] * a) * b;
        }
        long long ns2, tol2;  // the midpoint's counts and distance are not used
        if (!flow_at(q, f, ns2, tol2)) {
            st = GEO_HASH_FULL;
            break;
        }
        fn = norm(f);
        if (fn == 0) break;
        {
            const double b = 1. / fn;
            for (int k = 0; k < D; ++k) p[k] = p[k] + (f[k] * gp.step) * b;
        }
        emit(p);
        if (npts > 64 * gp.max_points) {  // far beyond the host's buffer: give up
            st = GEO_STEPS_OVERFLOW;
            break;
        }
    }
    if (lane == 0) {
        lengths[tip] = npts;
        status[tip] = st;
    }
}

}  // namespace

template <int M, typename T>
static void geo_one(const Geom &g, const StencilData<T> &sd, const T *u, const unsigned long long *fm,
                    const ExactFlow &xf, const GeoParamsHost &h, int ntips, const double *tips, double *points,
                    long long *lengths, int *status, hipStream_t s) {
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
    hipLaunchKernelGGL((k_geodesic<M, T>), dim3(ntips), dim3(64), 0, s, g, sd, u, fm, xf, gp, tips, points,
                       lengths, status);
}

template <typename T>
hipError_t launch_geodesics(int model, const Geom &g, const StencilData<T> &sd, const T *u,
                            const unsigned long long *fm, const ExactFlow &xf, const GeoParamsHost &h, int ntips,
                            const double *tips, double *points, long long *lengths, int *status, hipStream_t s) {
    if (ntips <= 0) return hipSuccess;
    switch (model) {
        case MODEL_RIEMANN2: geo_one<MODEL_RIEMANN2, T>(g, sd, u, fm, xf, h, ntips, tips, points, lengths, status, s); break;
        case MODEL_RIEMANN3: geo_one<MODEL_RIEMANN3, T>(g, sd, u, fm, xf, h, ntips, tips, points, lengths, status, s); break;
        case MODEL_RSF2: geo_one<MODEL_RSF2, T>(g, sd, u, fm, xf, h, ntips, tips, points, lengths, status, s); break;
        case MODEL_ELASTICA2: geo_one<MODEL_ELASTICA2, T>(g, sd, u, fm, xf, h, ntips, tips, points, lengths, status, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template <int M, typename T>
static void geo_ode_one(const Geom &g, const StencilData<T> &sd, const T *u, const unsigned long long *fm,
                        const ExactFlow &xf, const GeoParamsHost &h, double causality, const int *dist,
                        unsigned long long *hkeys, int *hcnt, int hlog2, int ntips, const double *tips, double *points,
                        long long *lengths, int *status, hipStream_t s) {
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
    hipLaunchKernelGGL((k_geodesic_ode<M, T>), dim3(ntips), dim3(64), 0, s, g, sd, u, fm, xf, gp, causality, dist,
                       hkeys, hcnt, hlog2, tips, points, lengths, status);
}

template <typename T>
hipError_t launch_geodesics_ode(int model, const Geom &g, const StencilData<T> &sd, const T *u,
                                const unsigned long long *fm, const ExactFlow &xf, const GeoParamsHost &h,
                                double causality, const int *dist, unsigned long long *hkeys, int *hcnt, int hlog2,
                                int ntips, const double *tips, double *points, long long *lengths, int *status,
                                hipStream_t s) {
    if (ntips <= 0) return hipSuccess;
#define HFM_ODE_CASE(M)                                                                                        \
    case M:                                                                                                    \
        geo_ode_one<M, T>(g, sd, u, fm, xf, h, causality, dist, hkeys, hcnt, hlog2, ntips, tips, points,     \
                          lengths, status, s);                                                                 \
        break;
    switch (model) {
        HFM_ODE_CASE(MODEL_RIEMANN2)
        HFM_ODE_CASE(MODEL_RIEMANN3)
        HFM_ODE_CASE(MODEL_RSF2)
        HFM_ODE_CASE(MODEL_ELASTICA2)
        default: return hipErrorInvalidValue;
    }
#undef HFM_ODE_CASE
    return hipGetLastError();
}

hipError_t launch_ode_distance(int model, const Geom &g, int *dist, const long long *seeds, int nseeds, long long np,
                               hipStream_t s) {
    const unsigned nb = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_ode_dist_init, dim3(nb), dim3(256), 0, s, dist, np);
    if (nseeds <= 0) return hipGetLastError();
    constexpr long long W = 2 * ODE_DIST_BOUND + 1;
    const long long box = model == MODEL_RIEMANN2 ? W * W : W * W * W;
    const unsigned bb = (unsigned)(((long long)nseeds * box + 255) / 256);
    switch (model) {
        case MODEL_RIEMANN2: hipLaunchKernelGGL((k_ode_dist_box<MODEL_RIEMANN2>), dim3(bb), dim3(256), 0, s, g, dist, seeds, nseeds); break;
        case MODEL_RIEMANN3: hipLaunchKernelGGL((k_ode_dist_box<MODEL_RIEMANN3>), dim3(bb), dim3(256), 0, s, g, dist, seeds, nseeds); break;
        case MODEL_RSF2: hipLaunchKernelGGL((k_ode_dist_box<MODEL_RSF2>), dim3(bb), dim3(256), 0, s, g, dist, seeds, nseeds); break;
        case MODEL_ELASTICA2: hipLaunchKernelGGL((k_ode_dist_box<MODEL_ELASTICA2>), dim3(bb), dim3(256), 0, s, g, dist, seeds, nseeds); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_geodesics_ode<float>(int, const Geom &, const StencilData<float> &, const float *,
                                                const unsigned long long *, const ExactFlow &, const GeoParamsHost &,
                                                double, const int *, unsigned long long *, int *, int, int,
                                                const double *, double *, long long *, int *, hipStream_t);
template hipError_t launch_geodesics_ode<double>(int, const Geom &, const StencilData<double> &, const double *,
                                                 const unsigned long long *, const ExactFlow &, const GeoParamsHost &,
                                                 double, const int *, unsigned long long *, int *, int, int,
                                                 const double *, double *, long long *, int *, hipStream_t);

template hipError_t launch_geodesics<float>(int, const Geom &, const StencilData<float> &, const float *,
                                            const unsigned long long *, const ExactFlow &, const GeoParamsHost &, int,
                                            const double *, double *, long long *, int *, hipStream_t);
template hipError_t launch_geodesics<double>(int, const Geom &, const StencilData<double> &, const double *,
                                             const unsigned long long *, const ExactFlow &, const GeoParamsHost &,
                                             int, const double *, double *, long long *, int *, hipStream_t);

}  // namespace hfm
