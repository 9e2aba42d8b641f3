// hfm_recompute.h -- the recompute of a point from its active differences with the accepted
// set of a serial solve (Headers/Base/Implementation/HamiltonFastMarching.hxx:187-272 with
// EulerianStencil.hxx:155-233): second / third order combinations and the static factor's
// corrections (Headers/ExtraAlgorithms/Implementation/StaticFactoring.hxx:11-55).  Used by the
// serial loop at acceptance, by its flow exports and by geodesic backtracking after it.
#pragma once

#include "hfm_device.h"

namespace hfm {

// QuadraticMax<ScalarType,1> (EulerianStencil.hxx:30-72): a, b, c relative to minVal, c from -1.
struct Quad {
    double minVal, a, b, c;
    __device__ void add(double value, double w) {
        const double v = value - minVal;
        const double wv = w * v, wv2 = (w * v) * v;
        a += w;
        b += wv;
        c += wv2;
    }
    __device__ double solve() const {
        if (a == 0.) return minVal + __builtin_huge_val();
        const double delta2 = b * b - a * c;
        const double delta = sqrt(delta2 > 0. ? delta2 : 0.);
        return minVal + (b + delta) / a;
    }
};

// One active difference of x as it enters the recompute: its value (the accepted neighbour's,
// or the second / third order combination), the order used, and the signed offset.
struct RecomputeEntry {
    double value, w;
    int ord;
    int o[3];
};

// SetIndex (StaticFactoring.hxx:21-30): the box index of grid point c into cur, and whether the
// factor applies there (in the box and in the mask).  NeedsRecompute (:11-19) is the same test:
// these schemes do not factor their first pass (HamiltonFastMarching.h:47-50).
__device__ inline bool factor_at(const Geom &g, const FactorDev &F, const int *c, int *cur) {
    if (!F.on) return false;
    // loops over the three axes with a constant bound, so that cur[] stays in registers
    bool in = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        cur[i] = i < g.dim ? c[i] - F.shift[i] : 0;
        in = in && (!F.sub || i >= g.dim || (cur[i] >= 0 && cur[i] < F.bd[i]));
    }
    if (!in) return false;
    return F.mask[((long long)cur[0] * F.bd[1] + cur[1]) * F.bd[2] + cur[2]] != 0;
}

// Correction (StaticFactoring.hxx:33-55) of the difference along the signed offset o at box index
// cur for the order ord the difference uses: the factor's directional derivative grad T . o plus
// the factor's own one-sided difference of that order with the opposite sign.  On a box, a
// difference that leaves it gets none; on the whole grid the angle axis of the curvature models
// wraps.
template <int ModelT>
__device__ inline double factor_corr(const Geom &g, const FactorDev &F, const int *cur, const int *o, int ord) {
    int last[3], i1[3], i2[3];
    bool in = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        last[i] = cur[i] + ord * o[i];
        i1[i] = cur[i] + o[i];
        i2[i] = cur[i] + 2 * o[i];
        in = in && (i >= g.dim || (last[i] >= 0 && last[i] < F.bd[i]));
    }
    if (F.sub) {
        if (!in) return 0.;
    } else if (ModelShape<ModelT>::SHARED) {
        const int n = F.bd[2];
        last[2] = ((last[2] % n) + n) % n;
        i1[2] = ((i1[2] % n) + n) % n;
        i2[2] = ((i2[2] % n) + n) % n;
    }
    auto lin = [&](const int *i) { return ((long long)i[0] * F.bd[1] + i[1]) * F.bd[2] + i[2]; };
    const long long c = lin(cur);
    double deriv = 0.;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (i < g.dim) deriv += F.grads[c * g.dim + i] * (double)o[i];
    const double v0 = F.vals[c], vl = F.vals[lin(last)];
    if (ord == 1) return deriv + (v0 - vl);
    const double v1 = F.vals[lin(i1)];
    if (ord == 2) return deriv + (1.5 * v0 - 2. * v1 + 0.5 * vl);
    const double v2 = F.vals[lin(i2)];
    return deriv + ((11. / 6.) * v0 - 3. * v1 + 1.5 * v2 - (1. / 3.) * vl);
}

// GetValueCorr of Recompute (HamiltonFastMarching.hxx:199-262) for the difference x -> x + o
// (o signed): the neighbour's value, or (order >= 2) (2/3)(2 v1 - v2 / 2) when x + 2 o is accepted
// and v2 <= v1, or (order >= 3) (6/11)(3 v1 - 3 v2 / 2 + v3 / 3) when x + 3 o also is and v3 <= v2,
// each plus the factor's correction of that order inside the bracket when the factor applies at x
// (fact, with fcur x's box index).  The small-correction tests apply to strictly causal
// schemes only (HamiltonFastMarching.h:46), which these are not.
template <int ModelT>
__device__ inline double value_corr(const Geom &g, const double *u, const unsigned char *acc, const int *c, const int *o,
                                    int order, const FactorDev &F, bool fact, const int *fcur, int &ord) {
    auto corr = [&](int k) { return fact ? factor_corr<ModelT>(g, F, fcur, o, k) : 0.; };
    int c1[3], c2[3], c3[3];
    shift_coords_m<ModelT>(g, c, o[0], o[1], o[2], c1);  // active: in the domain and accepted
    const long long q1 = (long long)tile_of(g, c1) * 64 + local_of(g, c1);
    const bool h2 = order >= 2 && shift_coords_m<ModelT>(g, c1, o[0], o[1], o[2], c2);
    const bool h3 = h2 && order >= 3 && shift_coords_m<ModelT>(g, c2, o[0], o[1], o[2], c3);
    const long long q2 = h2 ? (long long)tile_of(g, c2) * 64 + local_of(g, c2) : q1;
    const long long q3 = h3 ? (long long)tile_of(g, c3) * 64 + local_of(g, c3) : q1;
    // every value the tests may read, loaded at once (the indices are in the grid)
    const double v1 = u[q1], v2 = u[q2], v3 = u[q3];
    const bool a2 = acc[q2] != 0, a3 = acc[q3] != 0;
    ord = 1;
    if (h2 && a2 && !(v2 > v1)) {
        if (h3 && a3 && !(v3 > v2)) {
            ord = 3;
            return (6. / 11.) * ((3. * v1 - 1.5 * v2 + (1. / 3.) * v3) + corr(3));
        }
        ord = 2;
        return (2. / 3.) * ((2. * v1 - 0.5 * v2) + corr(2));
    }
    return v1 + corr(1);
}

// Recompute (HamiltonFastMarching.hxx:187-272, EulerianStencil.hxx:155-233) of the point at c
// from its active differences `act` in stencil order: the entries into e[], their number
// returned, and the recomputed value in *value.  The values enter one quadratic relative to their
// minimum with the weights scaled by 1, 9/4, 121/36 by order.
template <int ModelT>
__device__ inline int recompute(const Geom &g, const StencilData<double> &sd, const double *u, const unsigned char *acc,
                                long long p, const int *c, unsigned long long act, int order, const FactorDev &F,
                                RecomputeEntry *e, double *value) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    LaneStencil<ModelT, double> st;
    st.load(g, sd, p, c);
    int fcur[3];
    const bool fact = factor_at(g, F, c, fcur);
    int n = 0;
    for (int k = 0; k < NC; ++k) {
        int o0, o1, o2;
        unpack_offset(st.off[k], o0, o1, o2);
        for (int s = 0; s < (k < NF ? 1 : 2); ++s) {
            if (!(act >> side_bit<NF>(k, s) & 1ull)) continue;
            const int sg = s == 0 ? 1 : -1;
            RecomputeEntry &r = e[n++];
            r.o[0] = sg * o0;
            r.o[1] = sg * o1;
            r.o[2] = sg * o2;
            r.w = st.w[k];
            r.value = value_corr<ModelT>(g, u, acc, c, r.o, order, F, fact, fcur, r.ord);
        }
    }
    Quad q{__builtin_huge_val(), 0., 0., -1.};
    for (int i = 0; i < n; ++i) q.minVal = e[i].value < q.minVal ? e[i].value : q.minVal;
    constexpr double mult2[4] = {0., 1., 9. / 4., 121. / 36.};
    for (int i = 0; i < n; ++i) q.add(e[i].value, mult2[e[i].ord] * e[i].w);
    *value = q.solve();
    return n;
}

// The discrete flow's weight of recompute entry e for the recomputed value r
// (EulerianStencil.hxx:218-231): w max(0, r - v) times 1, 3/2, 11/6 by order.
__device__ inline double flow_weight(const RecomputeEntry &e, double r) {
    constexpr double mult[4] = {0., 1., 3. / 2., 11. / 6.};
    const double d = r - e.value;
    return (e.w * (d > 0. ? d : 0.)) * mult[e.ord];
}

}  // namespace hfm
