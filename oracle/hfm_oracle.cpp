// hfm_oracle.cpp -- CPU restatement of HamiltonFastMarching's hot path (test infrastructure).
//
// This file is the checker, never the product: it is built into oracle/_build/ and is only
// loaded by tests/, by the smoke check and by the CPU-baseline leg of bench.py.
// Every function names the reference lines it restates.  All arithmetic is FP64 in the
// reference's operation order; the file is compiled with -ffp-contract=off like the reference.
#include "hfm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <limits>
#include <array>
#include <map>
#include <queue>
#include <string>
#include <deque>
#include <set>
#include <vector>

namespace {

thread_local std::string g_err;
// Walls for the next problem set up (oracle_set_walls): one byte per point, empty for none.
thread_local std::vector<uint8_t> g_walls;
thread_local bool g_refine_walls = false;
thread_local bool g_plant_no_walk = false;  // planted mistake for the tests: no line walk
const double kInf = std::numeric_limits<double>::infinity();
const double kPi = 3.141592653589793238462643383279502884L;  // CommonTraits.h:46 (long double -> double)
const double kEps = std::numeric_limits<double>::epsilon();

struct OracleError {
    std::string msg;
};

// ---------------------------------------------------------------------------------------------
// Small FP64 linear algebra on packed symmetric matrices.
// Packing (SymmetricMatrixType.hxx:11-16): index(i,j) = i(i+1)/2 + j for i >= j.
// ---------------------------------------------------------------------------------------------
template <int D> struct Sym {
    static const int K = D * (D + 1) / 2;
    double m[K];
    static int lin(int i, int j) { if (i < j) std::swap(i, j); return i * (i + 1) / 2 + j; }
    double operator()(int i, int j) const { return m[lin(i, j)]; }
    double &at(int i, int j) { return m[lin(i, j)]; }
    // Trace: sequential sum of diagonal (SymmetricMatrixType.hxx:186-192).
    double trace() const { double r = 0; for (int i = 0; i < D; ++i) r += (*this)(i, i); return r; }
    // Determinant, same expansion as SymmetricMatrixType.hxx:199-209.
    double det() const {
        const Sym &c = *this;
        if (D == 2) return c(0, 0) * c(1, 1) - c(1, 0) * c(0, 1);
        return c(0, 0) * (c(1, 1) * c(2, 2) - c(1, 2) * c(2, 1))
             + c(0, 1) * (2. * c(1, 2) * c(2, 0) - c(1, 0) * c(2, 2))
             - c(0, 2) * (c(1, 1) * c(2, 0));
    }
    // Comatrix (SymmetricMatrixType.hxx:220-237).
    Sym comatrix() const {
        Sym r; const Sym &c = *this;
        if (D == 2) { r.m[0] = c(1, 1); r.m[1] = -c(1, 0); r.m[2] = c(0, 0); return r; }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j <= i; ++j) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                r.at(i, j) = c(i1, j1) * c(i2, j2) - c(i1, j2) * c(i2, j1);
            }
        return r;
    }
    // Inverse = Comatrix / Determinant, the division done as a reciprocal multiply
    // (SymmetricMatrixType.hxx:239-243, FriendOperators.h:35, VectorType.h:37-42).
    Sym inverse() const {
        Sym r = comatrix();
        const double b = 1. / det();
        for (int k = 0; k < K; ++k) r.m[k] *= b;
        return r;
    }
    // <u, M v> with the reference's summation order (SymmetricMatrixType.hxx:18-30).
    double sp(const int *u, const int *v) const {
        double sum = 0.;
        for (int i = 0; i < D; ++i) {
            double sumi = 0.;
            for (int j = 0; j < D; ++j) sumi += (*this)(i, j) * v[j];
            sum += u[i] * sumi;
        }
        return sum;
    }
};

// ---------------------------------------------------------------------------------------------
// Selling obtuse superbase + tensor decomposition (BasisReduction.hpp:12-15,65-157).
// ---------------------------------------------------------------------------------------------
const int kMaxIt = 200;  // BasisReduction.h:72

template <int D> struct Decomp {
    static const int K = D * (D + 1) / 2;
    int off[K][D];
    double w[K];
};

template <int D> Decomp<D> tensor_decomposition(const Sym<D> &m) {
    int sb[D + 1][D];
    for (int i = 0; i <= D; ++i)
        for (int j = 0; j < D; ++j) sb[i][j] = i < D ? (i == j) : -1;
    if (m.det() <= 0 || m.trace() <= 0)
        throw OracleError{"ObtuseSuperbase error : input must be positive definite"};
    bool reduced = true;
    for (int it = 0; reduced && it < kMaxIt; ++it) {
        reduced = false;
        for (int i = 0; i < D; ++i)
            for (int j = i + 1; j <= D; ++j) {
                if (m.sp(sb[i], sb[j]) > 0) {
                    reduced = true;
                    if (D == 2) {
                        const int k = 3 - i - j;
                        for (int c = 0; c < D; ++c) sb[k][c] = sb[i][c] - sb[j][c];
                    } else {
                        int kl[2], n = 0;
                        for (int q = 0; q <= D; ++q) if (q != i && q != j) kl[n++] = q;
                        for (int c = 0; c < D; ++c) { sb[kl[0]][c] += sb[i][c]; sb[kl[1]][c] += sb[i][c]; }
                    }
                    for (int c = 0; c < D; ++c) sb[i][c] *= -1;
                }
            }
    }
    if (reduced) throw OracleError{"ObtuseSuperbase error : did not terminate"};
    Decomp<D> r;
    if (D == 2) {
        for (int i = 0; i < 3; ++i) {
            const int j = (i + 1) % 3, k = (i + 2) % 3;
            r.off[i][0] = -sb[i][1];
            r.off[i][1] = sb[i][0];
            r.w[i] = -m.sp(sb[j], sb[k]);
        }
    } else {
        int s = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) {
                int rem[2], n = 0;
                for (int q = 0; q < 4; ++q) if (q != i && q != j) rem[n++] = q;
                const int *u = sb[rem[0]], *v = sb[rem[1]];
                for (int c = 0; c < 3; ++c) {
                    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                    r.off[s][c] = u[c1] * v[c2] - u[c2] * v[c1];
                }
                r.w[s] = -m.sp(sb[i], sb[j]);
                ++s;
            }
    }
    return r;
}

// ---------------------------------------------------------------------------------------------
// Stencils.
// ---------------------------------------------------------------------------------------------
const int kMaxDiff = 27;

struct Diff {
    int8_t off[3];
    double w;
};

// One stencil in the reference layout: forward differences first, then symmetric ones
// (EulerianStencil.h:23-40, bit layout of ActiveNeighFlagType).
struct Stencil {
    int nfwd = 0, nsym = 0;
    Diff d[kMaxDiff];
};

// Voronoi1Mat (CommonTraits.h:82-105) applied to D = Inverse(M) o [1/(h_i h_j)]
// (Riemannian.h:41-44,50-52,68-71).
template <int D>
void riemann_stencil(const double *mcomp, bool dual, const double *h, Stencil &st) {
    Sym<D> m;
    for (int k = 0; k < Sym<D>::K; ++k) m.m[k] = mcomp[k];
    if (!dual) m = m.inverse();
    Sym<D> gs;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j <= i; ++j) gs.at(i, j) = h[i] * h[j];
    for (int k = 0; k < Sym<D>::K; ++k) m.m[k] = m.m[k] * (1. / gs.m[k]);
    const double atol = 1000 * kEps * m.trace();
    const Decomp<D> dec = tensor_decomposition<D>(m);
    st.nfwd = 0;
    st.nsym = Sym<D>::K;
    for (int k = 0; k < Sym<D>::K; ++k) {
        Diff &df = st.d[k];
        df.w = dec.w[k] < atol ? 0. : dec.w[k];
        df.off[0] = df.off[1] = df.off[2] = 0;
        for (int c = 0; c < D; ++c) df.off[c] = (int8_t)dec.off[k][c];
    }
}

// Voronoi1Vec (CommonTraits.h:118-161), writing Dv diffs into out[] with the reduction's
// dimension Dv (offset components beyond Dv stay 0).
template <int Dv>
void voronoi1vec(const double *v, double eps, Diff *out) {
    const double angleTol = 1. / std::sqrt(double(Dv - 1));
    const double weightTol = 1000 * kEps;
    Sym<Dv> relaxed;
    double sqn = 0.;
    for (int i = 0; i < Dv; ++i) sqn += v[i] * v[i];
    for (int i = 0; i < Dv; ++i)
        for (int j = 0; j <= i; ++j) {
            const double exact = v[i] * v[j];
            const double id = (i == j) ? 1. : 0.;
            relaxed.at(i, j) = exact * (1. - eps * eps) + (sqn * id) * (eps * eps);
        }
    const Decomp<Dv> dec = tensor_decomposition<Dv>(relaxed);
    const double vNorm2 = sqn;
    const double wTol = vNorm2 * weightTol;
    const double angTol = (1 + angleTol * angleTol) / vNorm2;
    for (int k = 0; k < Sym<Dv>::K; ++k) {
        double off[Dv];
        for (int c = 0; c < Dv; ++c) off[c] = double(dec.off[k][c]);
        double scal = 0., offn = 0.;
        for (int c = 0; c < Dv; ++c) scal += v[c] * off[c];
        for (int c = 0; c < Dv; ++c) offn += off[c] * off[c];
        const double weight = dec.w[k];
        out[k].w = (weight < wTol || (scal * scal) * angTol < offn) ? 0. : weight;
        out[k].off[0] = out[k].off[1] = out[k].off[2] = 0;
        for (int c = 0; c < Dv; ++c) out[k].off[c] = (int8_t)(scal <= 0 ? dec.off[k][c] : -dec.off[k][c]);
    }
}

// StencilReedsSheppForward2::SetStencil (Curvature2.h:88-102): 3 forward + 1 symmetric.
void rsf2_stencil(int itheta, int ntheta, double gridScale, double xi, double eps, Stencil &st) {
    const double dependScale = 2 * kPi / ntheta;
    const double theta = itheta * dependScale;
    double v[2] = {std::cos(theta), std::sin(theta)};
    const double b = 1. / gridScale;  // Vector /= scalar is a reciprocal multiply
    v[0] *= b; v[1] *= b;
    st.nfwd = 3; st.nsym = 1;
    voronoi1vec<2>(v, eps, st.d);
    Diff &ang = st.d[3];
    ang.off[0] = 0; ang.off[1] = 0; ang.off[2] = -1;
    const double t = xi * dependScale;
    ang.w = 1. / (t * t);
}

// StencilElastica2<5>::SetStencil (Curvature2.h:178-217) with the Fejer weights of :233-234.
void elastica2_stencil(int itheta, int ntheta, double gridScale, double xi, double eps, Stencil &st) {
    const int nFejer = 5;
    const double fejer[5] = {0.167781, 0.525552, 0.613333, 0.525552, 0.167781};
    const double dependScale = 2 * kPi / ntheta;
    const double theta = (2. * kPi * itheta) / ntheta;
    st.nfwd = 27; st.nsym = 0;
    for (int l = 0; l < nFejer / 2; ++l) {
        const double phi = kPi * (l + 0.5) / nFejer;
        const double v[3] = {std::sin(phi) * std::cos(theta) / gridScale,
                             std::sin(phi) * std::sin(theta) / gridScale,
                             std::cos(phi) / (xi * dependScale)};
        const int s0 = 2 * l, s1 = 2 * l + 1;
        voronoi1vec<3>(v, eps, &st.d[6 * s0]);
        for (int i = 0; i < 6; ++i) {
            st.d[6 * s0 + i].w *= fejer[l];
            st.d[6 * s1 + i] = st.d[6 * s0 + i];
            st.d[6 * s1 + i].off[2] = (int8_t)(-st.d[6 * s1 + i].off[2]);
        }
    }
    const double v2[2] = {std::cos(theta) / gridScale, std::sin(theta) / gridScale};
    const int s = nFejer - 1;
    voronoi1vec<2>(v2, eps, &st.d[6 * s]);
    for (int i = 0; i < 3; ++i) st.d[6 * s + i].w *= fejer[nFejer / 2];
}

// ---------------------------------------------------------------------------------------------
// Grid and problem description.
// ---------------------------------------------------------------------------------------------
struct Problem {
    int model = 0;
    int dim = 2;
    int dims[3] = {1, 1, 1};
    bool periodic[3] = {false, false, false};
    long N = 0;
    double h[3] = {1, 1, 1};
    double origin[3] = {0, 0, 0};
    double dependScale = 1;  // curvature models: 2 pi / ntheta
    // Riemann: per-point stencils. Curvature: per-theta stencils + per-point speed.
    std::vector<Stencil> stencils;
    std::vector<double> speed;
    bool shared = false;

    const Stencil &stencil(long x) const {
        if (!shared) return stencils[x];
        return stencils[x % dims[2]];
    }
    double mult(long x) const { return shared ? speed[x] : 1.; }
    // EulerianDifference<...,1>::Weight = baseWeight * mult^2 (EulerianStencil.h:84-97);
    // the Riemann difference ignores the multiplier.
    double weight(const Diff &d, double mlt) const { return shared ? d.w * (mlt * mlt) : d.w; }

    void unravel(long x, int *idx) const {
        for (int i = dim - 1; i > 0; --i) { idx[i] = int(x % dims[i]); x /= dims[i]; }
        idx[0] = int(x);
    }
    long ravel(const int *idx) const {
        long r = idx[0];
        for (int i = 1; i < dim; ++i) r = r * dims[i] + idx[i];
        return r;
    }
    // PeriodicGrid::PeriodizeNoBase for Closed / Periodic axes (PeriodicGrid.hxx:120-140).
    bool periodize(int *idx) const {
        for (int i = 0; i < dim; ++i) {
            if (periodic[i]) {
                int r = idx[i] % dims[i];
                if (r < 0) r += dims[i];
                idx[i] = r;
            } else if (idx[i] < 0 || idx[i] >= dims[i]) return false;
        }
        return true;
    }
    // x + s*offset, or -1 when it leaves a closed axis.
    long shift(long x, const int8_t *off, int s) const {
        int idx[3] = {0, 0, 0};
        unravel(x, idx);
        for (int i = 0; i < dim; ++i) idx[i] += s * off[i];
        if (!periodize(idx)) return -1;
        return ravel(idx);
    }
    // ADim (HFM_ParamDefault.hxx:77-80 for Riemann; :106-119 for the R2xS1 bundle).
    void adim(const double *p, double *q) const {
        for (int i = 0; i < dim; ++i) {
            if (shared && i == 2) q[i] = (p[i] - origin[i]) / dependScale;
            else q[i] = (p[i] - origin[i]) / h[i];
        }
    }
    double redim_vec(double v, int i) const { return (shared && i == 2) ? v * dependScale : v * h[i]; }
    // Obstacles (ExtraAlgorithms/WallObstructionTest.h): empty when there are none.
    std::vector<uint8_t> walls;
    bool refineWalls = false;  // refineStencilAtWallBoundary
    bool plantNoWalk = false;  // planted mistake (tests only): skip the line walk
};

void setup_problem(Problem &P, int model, const int *dims, const double *gridScales, const double *origin,
                   const double *metric, int metric_is_field, int dual, double xi, double eps,
                   const double *cost, int cost_is_field) {
    P.model = model;
    P.dim = (model == ORACLE_RIEMANN2) ? 2 : 3;
    P.N = 1;
    for (int i = 0; i < P.dim; ++i) {
        if (dims[i] <= 0) throw OracleError{"BaseGrid error : domain size must be positive"};
        P.dims[i] = dims[i];
        P.N *= dims[i];
    }
    if (model == ORACLE_RIEMANN2 || model == ORACLE_RIEMANN3) {
        for (int i = 0; i < P.dim; ++i) {
            P.h[i] = gridScales[i];
            if (!(P.h[i] > 0)) throw OracleError{"Parametrization error : gridscales should be positive"};
            P.origin[i] = origin ? origin[i] : 0.;
        }
        P.shared = false;
        P.stencils.resize(P.N);
        const int K = P.dim == 2 ? 3 : 6;
        for (long x = 0; x < P.N; ++x) {
            const double *mc = metric + (metric_is_field ? x * K : 0);
            if (P.dim == 2) riemann_stencil<2>(mc, dual != 0, P.h, P.stencils[x]);
            else riemann_stencil<3>(mc, dual != 0, P.h, P.stencils[x]);
        }
    } else {
        const double gs = gridScales[0];
        if (!(gs > 0)) throw OracleError{"Parametrization error : gridscale should be positive"};
        P.h[0] = P.h[1] = gs;
        P.periodic[2] = true;
        P.dependScale = 2 * kPi / dims[2];
        P.h[2] = P.dependScale;
        P.origin[0] = origin ? origin[0] : 0.;
        P.origin[1] = origin ? origin[1] : 0.;
        P.origin[2] = -P.dependScale / 2.;
        P.shared = true;
        P.stencils.resize(dims[2]);
        for (int t = 0; t < dims[2]; ++t) {
            if (model == ORACLE_RSF2) rsf2_stencil(t, dims[2], gs, xi, eps, P.stencils[t]);
            else elastica2_stencil(t, dims[2], gs, xi, eps, P.stencils[t]);
        }
        P.speed.resize(P.N);
        for (long x = 0; x < P.N; ++x) P.speed[x] = 1. / cost[cost_is_field ? x : 0];  // DataSource_Inverse
    }
    if (!g_walls.empty()) {
        if ((long)g_walls.size() != P.N) throw OracleError{"Error : Field walls has inconsistent dimensions.\n"};
        P.walls = g_walls;
        P.refineWalls = g_refine_walls;
        P.plantNoWalk = g_plant_no_walk;
    }
}

// ---------------------------------------------------------------------------------------------
// Walls (WallObstructionTest.h:38-150).  The distance to the walls is the grid L1 distance
// (unit steps along the axes, periodic axes wrapped) up to maxDist = MaxStencilWidth, stored
// as the reference's ShortType int8 (127 = farther); a difference from an accepted point y to
// the point x = y + v is visible unless the discrete segment from y to x meets a wall.
// ---------------------------------------------------------------------------------------------
constexpr int kWallFar = 127;  // std::numeric_limits<int_least8_t>::max()

// ImplementIn (:38-88): a Dijkstra from the walls with unit steps, i.e. a breadth-first search.
std::vector<int8_t> wall_distance(const Problem &P, int maxWidth) {
    std::vector<int8_t> dist(P.N, (int8_t)kWallFar);
    const int maxDist = (int8_t)maxWidth;  // narrowed to ShortType as in the reference
    std::vector<long> fifo;
    for (long x = 0; x < P.N; ++x)
        if (P.walls[x]) {
            dist[x] = 0;
            fifo.push_back(x);
        }
    for (size_t head = 0; head < fifo.size(); ++head) {
        const long x = fifo[head];
        const int nb = dist[x] + 1;
        if (nb > maxDist) continue;
        int idx[3] = {0, 0, 0};
        P.unravel(x, idx);
        for (int i = 0; i < P.dim; ++i)
            for (int e = -1; e <= 1; e += 2) {
                int n[3] = {idx[0], idx[1], idx[2]};
                n[i] += e;
                if (!P.periodize(n)) continue;
                const long y = P.ravel(n);
                if (dist[y] <= nb) continue;
                dist[y] = (int8_t)nb;
                fifo.push_back(y);
            }
    }
    return dist;
}

// Visible(base, v, sum) (:96-150): base = accepted point, sum = base + v (periodized).
bool wall_visible(const Problem &P, const std::vector<int8_t> &dist, const int *base, const int *v, long sum) {
    int wd = dist[sum];
    if (wd == kWallFar) return true;
    if (wd == 0) return false;
    if (P.refineWalls && dist[P.ravel(base)] == 1) return true;
    int sign[3] = {0, 0, 0};
    int vSum = 0, vProd = 1, vMax = 0;
    for (int i = 0; i < P.dim; ++i)
        if (v[i] != 0) {
            sign[i] = v[i] > 0 ? 1 : -1;
            const int a = std::abs(v[i]);
            vSum += a;
            vProd *= a;
            vMax = std::max(vMax, a);
        }
    if (vSum <= wd || P.plantNoWalk) return true;
    // Straight discrete path from base towards sum: axis j steps when its next crossing
    // (vProd / |v_j| * (2 m + 1)) is reached by the threshold 2 i vProd / vMax.
    int pos[3] = {base[0], base[1], base[2]};
    long long next[3], inc[3];
    for (int i = 0; i < P.dim; ++i) {
        next[i] = v[i] == 0 ? (long long)std::numeric_limits<int>::max() : vProd / std::abs(v[i]);
        inc[i] = 2 * next[i];
    }
    for (int i = 1; i < vMax; ++i) {
        const long long threshold = (2LL * i * vProd) / vMax;
        for (int j = 0; j < P.dim; ++j)
            if (next[j] <= threshold) {
                next[j] += inc[j];
                --vSum;
                pos[j] += sign[j];
            }
        int pp[3] = {pos[0], pos[1], pos[2]};
        if (!P.periodize(pp)) return true;  // not reached: the segment stays in the domain
        wd = dist[P.ravel(pp)];
        if (vSum <= wd) return true;
        if (wd == 0) return false;
    }
    return true;
}

// ---------------------------------------------------------------------------------------------
// Fast marching (HamiltonFastMarching.hxx:59-169, EulerianStencil.hxx:30-153).
// ---------------------------------------------------------------------------------------------
struct Quad {
    double minVal = -kInf;
    double a = 0, b = 0, c = -1;
    void add(double value, double weight) {  // EulerianStencil.hxx:63-72
        const double w = weight, v = value - minVal;
        const double wv = w * v, wv2 = (w * v) * v;
        a += w; b += wv; c += wv2;
    }
    double solve() const {  // EulerianStencil.hxx:30-50, n = 1
        if (a == 0) return minVal + kInf;
        const double delta2 = b * b - a * c;
        const double delta = std::sqrt(std::max(0., delta2));
        const double res = (b + delta) / a;
        return minVal + res;
    }
};

struct Off3 {
    int8_t o[3];
    bool operator<(const Off3 &r) const {
        for (int i = 0; i < 3; ++i) if (o[i] != r.o[i]) return o[i] < r.o[i];
        return false;
    }
    bool operator==(const Off3 &r) const { return o[0] == r.o[0] && o[1] == r.o[1] && o[2] == r.o[2]; }
};

struct FMState {
    std::vector<double> values;
    std::vector<uint64_t> active;
    std::vector<char> accepted;
    std::vector<int8_t> wdist;  // distance to the walls (empty without walls)
};

// Reversed offsets (HFM_StencilRecompute.hxx:81-124; HFM_StencilShare.hxx:132-181): for each
// accepted point y, the sorted unique offsets e such that some x = y - e has a nonzero
// difference reaching y.  Stored as CSR.
void reversed_offsets(const Problem &P, std::vector<long> &start, std::vector<Off3> &offs) {
    const long nKeys = P.shared ? P.dims[2] : P.N;
    std::vector<std::vector<Off3>> lists(nKeys);
    auto insert = [&](long x, const int8_t *off, int s) {
        const long y = P.shift(x, off, s);
        if (y < 0) return;
        Off3 e;
        for (int c = 0; c < 3; ++c) e.o[c] = (int8_t)(s * off[c]);
        lists[P.shared ? y % P.dims[2] : y].push_back(e);
    };
    for (long key = 0; key < nKeys; ++key) {
        long x = key;
        if (P.shared) {  // IndexFromShortIndex: mid-domain (x,y), theta = key
            int idx[3] = {P.dims[0] / 2, P.dims[1] / 2, int(key)};
            x = P.ravel(idx);
        }
        const Stencil &st = P.stencil(x);
        for (int k = 0; k < st.nfwd; ++k) if (st.d[k].w != 0.) insert(x, st.d[k].off, 1);
        for (int k = st.nfwd; k < st.nfwd + st.nsym; ++k)
            if (st.d[k].w != 0.) { insert(x, st.d[k].off, 1); insert(x, st.d[k].off, -1); }
    }
    start.assign(nKeys + 1, 0);
    offs.clear();
    for (long key = 0; key < nKeys; ++key) {
        auto &l = lists[key];
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        start[key] = long(offs.size());
        offs.insert(offs.end(), l.begin(), l.end());
        std::vector<Off3>().swap(l);
    }
    start[nKeys] = long(offs.size());
}

struct QueueElement {
    long linearIndex;
    double value;
    bool operator<(const QueueElement &o) const { return value > o.value; }
};

// High-order recompute settings of the problems solved next on this thread (order 1: none).
int g_order = 1;
// Pops whose key lies below the previous pop's key, counted by the last solve (diagnostics).
long g_dips = 0;
// Stopping criteria of the solves run next (ExtraAlgorithms/CommonStoppingCriteria.h:57-92): row-major indices,
// (x, y) indices for the unoriented sets; g_naccepted = nAccepted of the last solve.
long g_stop_max = -1;
double g_stop_dist = std::numeric_limits<double>::infinity();
std::set<long> g_stop_any, g_stop_all, g_stop_any_u, g_stop_all_u;
long g_naccepted = 0;

// Static factoring of the problems solved next on this thread (StaticFactoring.h:20-63,
// StaticFactoring.hxx:11-55): factor values, gradients (index units, D per point) and mask over a
// box of the grid whose first grid index is `shift` (sub), or over the whole grid.
struct Factoring {
    bool on = false, sub = false;
    int bd[3] = {1, 1, 1}, shift[3] = {0, 0, 0};
    std::vector<double> vals, grads;
    std::vector<unsigned char> mask;
    long lin(const int *i, int dim) const {
        long r = i[0];
        for (int k = 1; k < dim; ++k) r = r * bd[k] + i[k];
        return r;
    }
    bool in_box(const int *i, int dim) const {
        for (int k = 0; k < dim; ++k)
            if (i[k] < 0 || i[k] >= bd[k]) return false;
        return true;
    }
};
Factoring g_fact;

// SetIndex / NeedsRecompute (StaticFactoring.hxx:11-30): x's box index into cur, and whether the
// factor applies at x (the recompute at acceptance is then needed at every order: these schemes
// do not factor their first pass, HamiltonFastMarching.h:47-50).
bool factor_at(const Problem &P, long x, int *cur) {
    if (!g_fact.on) return false;
    P.unravel(x, cur);
    for (int i = 0; i < P.dim; ++i) cur[i] -= g_fact.shift[i];
    if (g_fact.sub && !g_fact.in_box(cur, P.dim)) return false;
    return g_fact.mask[g_fact.lin(cur, P.dim)] != 0;
}

// Correction (StaticFactoring.hxx:33-55) of the difference along the signed offset `off` at the
// box index cur, for the order `ord` the difference uses: the factor's directional derivative
// plus the factor's own finite difference with the opposite sign.  Box indices are periodized on
// the whole grid only; on a box, a difference leaving it gets no correction.
double factor_correction(const Problem &P, const int *cur, const int8_t *off, int ord) {
    const Factoring &F = g_fact;
    const int D = P.dim;
    int last[3] = {0, 0, 0}, i1[3] = {0, 0, 0}, i2[3] = {0, 0, 0};
    for (int i = 0; i < D; ++i) {
        last[i] = cur[i] + ord * off[i];
        i1[i] = cur[i] + off[i];
        i2[i] = cur[i] + 2 * off[i];
    }
    if (F.sub) {
        if (!F.in_box(last, D)) return 0.;
    } else {
        P.periodize(last);
        P.periodize(i1);
        P.periodize(i2);
    }
    const long c = F.lin(cur, D);
    double deriv = 0.;
    for (int i = 0; i < D; ++i) deriv += F.grads[c * D + i] * double(off[i]);
    const double v0 = F.vals[c], vl = F.vals[F.lin(last, D)];
    if (ord == 1) return deriv + (v0 - vl);
    const double v1 = F.vals[F.lin(i1, D)];
    if (ord == 2) return deriv + (1.5 * v0 - 2 * v1 + 0.5 * vl);
    const double v2 = F.vals[F.lin(i2, D)];
    return deriv + ((11. / 6.) * v0 - 3. * v1 + 1.5 * v2 - (1. / 3.) * vl);
}

// VisibleOffset (HamiltonFastMarching.hxx:160-170): base + v periodized, -1 when it leaves a
// closed axis or a wall hides it.
long visible_shift(const Problem &P, const std::vector<int8_t> &wdist, long base, const int8_t *v) {
    const long x = P.shift(base, v, 1);
    if (x < 0 || wdist.empty()) return x;
    int bi[3] = {0, 0, 0}, vi[3] = {v[0], v[1], v[2]};
    P.unravel(base, bi);
    return wall_visible(P, wdist, bi, vi, x) ? x : -1;
}

// Recompute at acceptance (HamiltonFastMarching.hxx:187-272 with EulerianStencil.hxx:155-233):
// every active difference of x takes its accepted neighbour's value, or, at order >= 2, the
// second (third) order combination when the next point(s) along the offset are accepted and no
// larger (the causality tests; the small-correction tests apply to strictly causal schemes
// only, which these are not, HamiltonFastMarching.h:46), with the weight scaled by 9/4
// (121/36).  The values enter one quadratic relative to their minimum, in stencil order.
struct RecEntry {
    double value, w;
    int ord;
    int8_t off[3];  // signed offset
};
int recompute_entries(const Problem &P, const FMState &S, long x, int order, RecEntry *out, double *value) {
    const Stencil &st = P.stencil(x);
    const double mlt = P.mult(x);
    const uint64_t act = S.active[x];
    int cur[3] = {0, 0, 0};
    const bool fact = factor_at(P, x, cur);
    auto corr = [&](const int8_t *off, int ord) { return fact ? factor_correction(P, cur, off, ord) : 0.; };
    int n = 0;
    auto push = [&](const int8_t *off, int s, double w) {
        RecEntry &r = out[n++];
        for (int c = 0; c < 3; ++c) r.off[c] = (int8_t)(s * off[c]);
        const long a1 = P.shift(x, r.off, 1);  // active: accepted and in the domain
        const double v1 = S.values[a1];
        double val = v1 + corr(r.off, 1);
        int used = 1;
        if (std::min(order, 3) >= 2) {
            const long a2 = visible_shift(P, S.wdist, a1, r.off);
            if (a2 >= 0 && S.accepted[a2] && !(S.values[a2] > v1)) {
                const double v2 = S.values[a2];
                used = 2;
                val = (2. / 3.) * ((2. * v1 - 0.5 * v2) + corr(r.off, 2));
                if (order >= 3) {
                    const long a3 = visible_shift(P, S.wdist, a2, r.off);
                    if (a3 >= 0 && S.accepted[a3] && !(S.values[a3] > v2)) {
                        const double v3 = S.values[a3];
                        used = 3;
                        val = (6. / 11.) * ((3. * v1 - 1.5 * v2 + (1. / 3.) * v3) + corr(r.off, 3));
                    }
                }
            }
        }
        r.value = val;
        r.w = w;
        r.ord = used;
    };
    int iNeigh = 0;
    for (int k = 0; k < st.nfwd; ++k, ++iNeigh)
        if (act >> iNeigh & 1) push(st.d[k].off, 1, P.weight(st.d[k], mlt));
    for (int k = st.nfwd; k < st.nfwd + st.nsym; ++k, iNeigh += 2) {
        if (act >> iNeigh & 1) push(st.d[k].off, 1, P.weight(st.d[k], mlt));
        if (act >> (iNeigh + 1) & 1) push(st.d[k].off, -1, P.weight(st.d[k], mlt));
    }
    static const double mult2[4] = {0., 1., 9. / 4., 121. / 36.};
    Quad q;
    q.minVal = kInf;
    for (int i = 0; i < n; ++i) q.minVal = std::min(q.minVal, out[i].value);
    for (int i = 0; i < n; ++i) q.add(out[i].value, mult2[out[i].ord] * out[i].w);
    *value = q.solve();
    return n;
}

void fast_marching(const Problem &P, const std::vector<std::pair<long, double>> &seeds, FMState &S,
                   double *tStencil, double *tFM) {
    const clock_t t0 = clock();
    std::vector<long> start;
    std::vector<Off3> offs;
    reversed_offsets(P, start, offs);
    std::vector<int8_t> &wdist = S.wdist;
    wdist.clear();
    if (!P.walls.empty()) {
        int maxw = 0;  // MaxStencilWidth (HamiltonFastMarching.hxx:37-45)
        for (const Off3 &e : offs) maxw = std::max(maxw, std::abs(e.o[0]) + std::abs(e.o[1]) + std::abs(e.o[2]));
        wdist = wall_distance(P, maxw);
    }
    const clock_t t1 = clock();
    S.values.assign(P.N, kInf);
    S.active.assign(P.N, 0);
    S.accepted.assign(P.N, 0);
    std::vector<Quad> quads(P.N);
    std::priority_queue<QueueElement> queue;
    for (const auto &sv : seeds) {
        S.values[sv.first] = sv.second;
        queue.push({sv.first, sv.second});
    }
    g_dips = 0;
    g_naccepted = 0;
    double lastKey = -kInf;
    std::set<long> all = g_stop_all, allU = g_stop_all_u;
    while (!queue.empty()) {
        QueueElement top = queue.top();
        queue.pop();
        if (S.accepted[top.linearIndex]) continue;
        const long y = top.linearIndex;
        if (top.value < lastKey) ++g_dips;
        lastKey = top.value;
        // PostProcess (HamiltonFastMarching.hxx:86-99): the stopping criteria, then the recompute
        // (a point without active differences -- a seed -- keeps its value), then the stop.
        ++g_naccepted;
        bool stop = g_stop_max >= 0 && g_naccepted >= g_stop_max;
        if (!stop && g_stop_any.count(y)) stop = true;
        if (!stop && !all.empty()) {
            all.erase(y);
            if (all.empty()) stop = true;
        }
        const long xy = P.shared ? y / P.dims[2] : y;
        if (!stop && P.shared && g_stop_any_u.count(xy)) stop = true;
        if (!stop && P.shared && !allU.empty()) {
            allU.erase(xy);
            if (allU.empty()) stop = true;
        }
        if (!stop && g_stop_dist < kInf && S.values[y] > g_stop_dist) stop = true;
        int cur[3];
        if ((g_order > 1 || factor_at(P, y, cur)) && S.active[y]) {
            RecEntry re[2 * kMaxDiff];
            (void)recompute_entries(P, S, y, g_order, re, &top.value);
            S.values[y] = top.value;
        }
        if (stop) break;
        S.accepted[y] = 1;
        const long key = P.shared ? y % P.dims[2] : y;
        for (long r = start[key]; r < start[key + 1]; ++r) {
            const Off3 &e = offs[r];
            // ConditionalUpdate (HamiltonFastMarching.hxx:136-154)
            const long x = P.shift(y, e.o, -1);
            if (x < 0) continue;
            if (!wdist.empty()) {  // VisibleOffset (HamiltonFastMarching.hxx:160-170)
                int yi[3] = {0, 0, 0}, v[3] = {0, 0, 0};
                P.unravel(y, yi);
                for (int c = 0; c < P.dim; ++c) v[c] = -e.o[c];
                if (!wall_visible(P, wdist, yi, v, x)) continue;
            }
            if (S.accepted[x]) continue;
            if (S.active[x] == 0 && S.values[x] != kInf) continue;  // seeds are never updated
            if (S.values[x] <= top.value) continue;
            // HopfLaxUpdate (EulerianStencil.hxx:104-153)
            Quad &q = quads[x];
            if (q.minVal == -kInf) q.minVal = top.value;
            const Stencil &st = P.stencil(x);
            const double mlt = P.mult(x);
            uint64_t &act = S.active[x];
            int iNeigh = 0;
            for (int k = 0; k < st.nfwd; ++k, ++iNeigh) {
                Off3 d; std::memcpy(d.o, st.d[k].off, 3);
                if (d == e) { act |= (1ull << iNeigh); q.add(top.value, P.weight(st.d[k], mlt)); }
            }
            for (int k = st.nfwd; k < st.nfwd + st.nsym; ++k, iNeigh += 2) {
                Off3 d, md;
                std::memcpy(d.o, st.d[k].off, 3);
                for (int c = 0; c < 3; ++c) md.o[c] = (int8_t)(-d.o[c]);
                if (d == e && !(act >> (iNeigh + 1) & 1)) {
                    act |= (1ull << iNeigh); q.add(top.value, P.weight(st.d[k], mlt));
                }
                if (md == e && !(act >> iNeigh & 1)) {
                    act |= (1ull << (iNeigh + 1)); q.add(top.value, P.weight(st.d[k], mlt));
                }
            }
            const double value = q.solve();
            S.values[x] = value;
            queue.push({x, value});
        }
    }
    const clock_t t2 = clock();
    if (tStencil) *tStencil = double(t1 - t0) / CLOCKS_PER_SEC;
    if (tFM) *tFM = double(t2 - t1) / CLOCKS_PER_SEC;
}

// Active differences of x as (neighbour, weight, offset, sign) in the reference's order.
struct ActiveDiff { long neigh; double w; const int8_t *off; int s; };
int active_diffs(const Problem &P, const FMState &S, long x, ActiveDiff *out) {
    const Stencil &st = P.stencil(x);
    const double mlt = P.mult(x);
    const uint64_t act = S.active[x];
    int n = 0, iNeigh = 0;
    for (int k = 0; k < st.nfwd; ++k, ++iNeigh)
        if (act >> iNeigh & 1) out[n++] = {P.shift(x, st.d[k].off, 1), P.weight(st.d[k], mlt), st.d[k].off, 1};
    for (int k = st.nfwd; k < st.nfwd + st.nsym; ++k, iNeigh += 2) {
        if (act >> iNeigh & 1) out[n++] = {P.shift(x, st.d[k].off, 1), P.weight(st.d[k], mlt), st.d[k].off, 1};
        if (act >> (iNeigh + 1) & 1) out[n++] = {P.shift(x, st.d[k].off, -1), P.weight(st.d[k], mlt), st.d[k].off, -1};
    }
    return n;
}

// GeodesicFlow (HamiltonFastMarching.hxx:363-378) via Recompute / HopfLaxRecompute at order 1
// (EulerianStencil.hxx:155-233), exported as -ReDim(flow) (HFMInterface.hxx:544-553).
void geodesic_flow(const Problem &P, const FMState &S, double *flow) {
    RecEntry re[2 * kMaxDiff];
    static const double mult[4] = {0., 1., 3. / 2., 11. / 6.};  // EulerianStencil.hxx:226-228
    for (long x = 0; x < P.N; ++x) {
        double f[3] = {0, 0, 0};
        double value = 0.;
        const int n = S.active[x] ? recompute_entries(P, S, x, g_order, re, &value) : 0;
        for (int i = 0; i < n; ++i) {
            const double wpd = (re[i].w * std::max(0., value - re[i].value)) * mult[re[i].ord];
            for (int c = 0; c < P.dim; ++c) f[c] += wpd * double(re[i].off[c]);
        }
        for (int c = 0; c < P.dim; ++c) flow[x * P.dim + c] = -P.redim_vec(f[c], c);
    }
}

// FirstVariation::ValueVariation (FirstVariation.h:198-251) and BackwardVariation (:341-385).
double value_variation(const Problem &P, const FMState &S, long x,
                       std::vector<std::pair<long, double>> &neighbors) {
    ActiveDiff ad[2 * kMaxDiff];
    const int n = active_diffs(P, S, x, ad);
    const double value = S.values[x];
    double var = 0, weightSum = 0;
    neighbors.clear();
    for (int i = 0; i < n; ++i) {
        const double nv = S.values[ad[i].neigh];
        const double valueDiff = std::max(0., value - nv);
        const double w = ad[i].w * valueDiff;
        weightSum += w;
        neighbors.push_back({ad[i].neigh, w});
        var += w * valueDiff;
    }
    if (weightSum == 0) return var;
    const double invWeight = 1. / weightSum;
    for (auto &nb : neighbors) nb.second *= invWeight;
    return P.shared ? invWeight * (P.mult(x) * var) : invWeight * var;
}

void backward_variation(const Problem &P, const FMState &S, const std::vector<std::pair<long, double>> &base,
                        double *result, std::vector<std::pair<long, double>> &seedSens) {
    std::map<std::pair<double, long>, double> geo;
    for (const auto &iw : base) geo.insert({{S.values[iw.first], iw.first}, iw.second});
    std::vector<std::pair<long, double>> neighbors;
    while (!geo.empty()) {
        auto it = std::prev(geo.end());
        const double value = it->first.first; (void)value;
        const long index = it->first.second;
        const double weight = it->second;
        geo.erase(it);
        if (weight == 0) continue;
        const double var = value_variation(P, S, index, neighbors);
        if (neighbors.empty()) seedSens.push_back({index, weight});
        result[index] = weight * var;
        for (const auto &nb : neighbors) {
            const std::pair<double, long> key{S.values[nb.first], nb.first};
            auto jt = geo.find(key);
            if (jt == geo.end()) geo[key] = nb.second * weight;
            else jt->second += nb.second * weight;
        }
    }
}

// FirstVariation::ForwardVariation (FirstVariation.h:389-429): points in increasing
// (value, index) order; result(x, k) += var(x) * costVariation(x, k) + sum_i w_i result(n_i, k)
// over the normalised active neighbours of ValueVariation.  result starts at the seeds' value
// variations (FirstVariation.h:103-124), zero elsewhere.  Arrays are (N, nVar) row-major.
void forward_variation(const Problem &P, const FMState &S, int nVar, const double *costVar, double *result) {
    std::vector<std::pair<double, long>> order(P.N);
    for (long i = 0; i < P.N; ++i) order[i] = {S.values[i], i};
    std::sort(order.begin(), order.end());
    std::vector<std::pair<long, double>> neighbors;
    for (const auto &vi : order) {
        const long x = vi.second;
        const double var = value_variation(P, S, x, neighbors);
        for (int k = 0; k < nVar; ++k) {
            double &res = result[x * nVar + k];
            res += var * (costVar ? costVar[x * nVar + k] : 0.);
            for (const auto &nb : neighbors) res += nb.second * result[nb.first * nVar + k];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Geodesic backtracking, discrete solver (GeodesicDiscreteSolver.h:14-24, .hxx:17-231): the unit
// mass at the tip is split among the 2^d surrounding grid points and transported down the
// discrete flow (Recompute, order 1) in decreasing (value, index) order; the running weighted
// mean of the mass positions gives the geodesic points.  Arithmetic follows the reference:
// vector / scalar is a multiplication by the reciprocal (VectorType.h:37-42), symmetric
// matrices are packed, rank-one tensors are u_i u_j.
// ---------------------------------------------------------------------------------------------
struct GeoParams {
    double step = 0.25, threshold = 0.001, volumeBound = -1;  // volumeBound <= 0: 5 * 1.3^d
};
using Idx3 = std::array<int, 3>;

// Discrete flow of Recompute (HamiltonFastMarching.hxx:187-272 with HopfLaxRecompute,
// EulerianStencil.hxx:155-233), as GeodesicDiscreteSolver.hxx:131-132 takes it: the active
// differences in stencil order, weight (w_i max(0, r - v_i)) times 1, 3/2, 11/6 by the order the
// difference uses, r the recomputed value; at order 1 without a factor, v_i is the neighbour's
// value.
int discrete_flow(const Problem &P, const FMState &S, long x, int (*off)[3], double *w) {
    RecEntry re[2 * kMaxDiff];
    double value = 0.;
    const int n = S.active[x] ? recompute_entries(P, S, x, g_order, re, &value) : 0;
    static const double mult[4] = {0., 1., 3. / 2., 11. / 6.};
    for (int i = 0; i < n; ++i) {
        w[i] = (re[i].w * std::max(0., value - re[i].value)) * mult[re[i].ord];
        for (int c = 0; c < 3; ++c) off[i][c] = c < P.dim ? re[i].off[c] : 0;
    }
    return n;
}

struct GeoSums {
    int d = 2;
    double weightSum = 0, indexSum[3] = {0, 0, 0}, cov[6] = {0, 0, 0, 0, 0, 0};
    void update(double weight, const Idx3 &ind) {
        weightSum += weight;
        double v[3];
        for (int i = 0; i < d; ++i) v[i] = double(ind[i]);
        for (int i = 0; i < d; ++i) indexSum[i] += v[i] * weight;
        for (int i = 0; i < d; ++i)
            for (int j = 0; j <= i; ++j) cov[i * (i + 1) / 2 + j] += (v[i] * v[j]) * weight;
    }
    void mean(double *m) const {
        const double b = 1. / weightSum;
        for (int i = 0; i < d; ++i) m[i] = indexSum[i] * b;
    }
};

double det_sym(int d, const double *c) {
    auto at = [&](int i, int j) { return i >= j ? c[i * (i + 1) / 2 + j] : c[j * (j + 1) / 2 + i]; };
    if (d == 2) return at(0, 0) * at(1, 1) - at(1, 0) * at(0, 1);
    return at(0, 0) * (at(1, 1) * at(2, 2) - at(1, 2) * at(2, 1)) +
           at(0, 1) * (2. * at(1, 2) * at(2, 0) - at(1, 0) * at(2, 2)) - at(0, 2) * (at(1, 1) * at(2, 0));
}

// One pass of GeodesicDiscrete from geodesic.back(); returns true when a restart is needed.
bool geodesic_pass(const Problem &P, const FMState &S, std::vector<std::array<double, 3>> &geodesic,
                   const GeoParams &gp, double volumeBound) {
    const int d = P.dim;
    std::map<std::pair<double, Idx3>, double> geo;
    GeoSums sums;
    sums.d = d;
    std::vector<std::pair<double, std::array<double, 3>>> steps;
    double stepNorm = 0;
    double indexAvg[3] = {0, 0, 0}, indexBase[3] = {0, 0, 0};
    auto value_at = [&](Idx3 ind) {
        P.periodize(ind.data());
        return S.values[P.ravel(ind.data())];
    };
    auto push_point = [&](const double *q) {
        std::array<double, 3> pt{0, 0, 0};
        for (int i = 0; i < d; ++i) pt[i] = q[i] + 0.5;
        geodesic.push_back(pt);
    };
    {
        const std::array<double, 3> tip = geodesic.back();
        Idx3 indexRef{0, 0, 0};
        double pointRef[3] = {0, 0, 0};
        for (int j = 0; j < d; ++j) {
            indexRef[j] = (int)std::floor(tip[j]);
            pointRef[j] = indexRef[j] + 0.5;
        }
        for (int i = 0; i < (1 << d); ++i) {
            double weight = 1;
            Idx3 index = indexRef;
            for (int j = 0; j < d; ++j) {
                const double diff = tip[j] - pointRef[j];
                if (i & (1 << j)) {
                    index[j] += diff > 0 ? 1 : -1;
                    weight *= std::abs(diff);
                } else {
                    weight *= 1. - std::abs(diff);
                }
            }
            Idx3 per = index;
            if (!P.periodize(per.data())) continue;
            geo[{S.values[P.ravel(per.data())], index}] = weight;
            sums.update(weight, index);
        }
        if (sums.weightSum == 0) return false;
        sums.mean(indexAvg);
        std::copy(indexAvg, indexAvg + 3, indexBase);
    }
    const double weightSumLowerBound = 1e-5;
    int off[2 * kMaxDiff][3];
    double fw[2 * kMaxDiff];
    while (!geo.empty() && sums.weightSum > weightSumLowerBound) {
        const auto it = std::prev(geo.end());
        const double value = it->first.first;
        const Idx3 index = it->first.second;
        const double weight = it->second;
        geo.erase(it);
        if (weight == 0) continue;
        sums.update(-weight, index);
        Idx3 per = index;
        P.periodize(per.data());
        const int nf = discrete_flow(P, S, P.ravel(per.data()), off, fw);
        double wSum = 0;
        for (int k = 0; k < nf; ++k) wSum += fw[k];
        if (wSum == 0) continue;
        for (int k = 0; k < nf; ++k) {
            const double w = weight * fw[k] / wSum;
            Idx3 ind = index;
            for (int c = 0; c < d; ++c) ind[c] += off[k][c];
            const double val = value_at(ind);
            const auto jt = geo.find({val, ind});
            if (jt == geo.end()) {
                if (w < sums.weightSum * gp.threshold) continue;
                geo[{val, ind}] = w;
            } else {
                jt->second += w;
            }
            sums.update(w, ind);
        }
        if (sums.weightSum <= 0) {
            push_point(indexAvg);
            return false;
        }
        double oldAvg[3];
        std::copy(indexAvg, indexAvg + 3, oldAvg);
        sums.mean(indexAvg);
        std::array<double, 3> step{0, 0, 0};
        for (int i = 0; i < d; ++i) step[i] = indexAvg[i] - oldAvg[i];
        steps.push_back({value, step});
        double sq = 0;
        for (int i = 0; i < d; ++i) sq += step[i] * step[i];
        stepNorm += std::sqrt(sq);
        if (geo.size() == 1) {
            if (nf == 0) break;
            push_point(indexAvg);
            std::copy(indexAvg, indexAvg + 3, indexBase);
            stepNorm = 0;
            steps.clear();
            continue;
        }
        if (stepNorm < gp.step && geo.size() > 1) continue;
        const double valInc = steps.front().first - value;
        if (valInc <= 0) continue;
        const double delta = valInc / d;
        double p[3];
        std::copy(indexBase, indexBase + 3, p);
        for (const auto &vs : steps) {
            const double a = (vs.first - value) + delta;
            const double rb = 1. / (valInc + delta);
            for (int i = 0; i < d; ++i) p[i] += (vs.second[i] * a) * rb;
        }
        push_point(p);
        std::copy(indexAvg, indexAvg + 3, indexBase);
        stepNorm = 0;
        steps.clear();
        double var[6], m[3];
        const double b = 1. / sums.weightSum;
        for (int k = 0; k < d * (d + 1) / 2; ++k) var[k] = sums.cov[k] * b;
        sums.mean(m);
        for (int i = 0; i < d; ++i)
            for (int j = 0; j <= i; ++j) {
                var[i * (i + 1) / 2 + j] -= m[i] * m[j];
                var[i * (i + 1) / 2 + j] += (i == j) ? 1. : 0.;
            }
        if (det_sym(d, var) > volumeBound * volumeBound) {
            push_point(indexAvg);
            return true;
        }
    }
    return false;
}

// GeodesicDiscreteSolver::Run (GeodesicDiscreteSolver.hxx:41-54): restarts from the last point
// while the mass spreads too much, the volume bound growing 1.3x from the 5th restart; gives up
// after 50.  Returns true on failure.
bool geodesic_run(const Problem &P, const FMState &S, std::vector<std::array<double, 3>> &geodesic,
                  const GeoParams &gp) {
    const double vb0 = gp.volumeBound > 0 ? gp.volumeBound : 5. * std::pow(1.3, P.dim);
    int nRestarts = 0;
    double vb = vb0;
    while (nRestarts < 50) {
        if (!geodesic_pass(P, S, geodesic, gp, vb)) break;
        ++nRestarts;
        if (nRestarts >= 5) vb *= 1.3;
    }
    return nRestarts == 50;
}

// ---------------------------------------------------------------------------------------------
// The ODE geodesic solver (geodesicSolver = "ODE"; Headers/Base/GeodesicODESolver.h:17-40,
// Implementation/GeodesicODESolver.hxx:10-261): a midpoint (second order Runge-Kutta) walk of step
// geodesicStep along the normalised flow, the flow at a point being the 2^d corners' GeodesicFlow
// (HamiltonFastMarching.hxx:363-378) averaged with the interpolation weights over the corners
// whose value is within causalityTolerance widths of the smallest.  It stops at a seed (zero
// flow), when the L-infinity grid distance to the seeds (capped at 6) grows again over the last
// sqrt(d) / step steps, or when some corner's flow has been read more than 6 sqrt(d) / step
// times (stall, a failure).  Walls are not restated here.
// ---------------------------------------------------------------------------------------------
int g_geo_ode = 0;
double g_geo_causality = 4.;
long g_geo_failed = 0;  // stalled tips of the last ODE run

struct OdeFlowData {
    double value, width;
    double flow[3];
};

// GeodesicFlow at a grid point: Recompute's value and width (sum of w_i max(0, r - v_i) over the
// weights' sum) and the flow sum_i weight_i offset_i with the discrete flow's weights; a point
// without active differences keeps its value, width 0, no flow.
OdeFlowData ode_point_flow(const Problem &P, const FMState &S, long x) {
    OdeFlowData f{S.values[x], 0., {0., 0., 0.}};
    if (!S.active[x]) return f;
    RecEntry re[2 * kMaxDiff];
    double value = 0.;
    const int n = recompute_entries(P, S, x, g_order, re, &value);
    static const double mult[4] = {0., 1., 3. / 2., 11. / 6.};
    double width = 0., wsum = 0.;
    for (int i = 0; i < n; ++i) {
        const double wpd = re[i].w * std::max(0., value - re[i].value);
        width += wpd;
        const double fw = wpd * mult[re[i].ord];
        for (int c = 0; c < P.dim; ++c) f.flow[c] += fw * double(re[i].off[c]);
    }
    for (int i = 0; i < n; ++i) wsum += re[i].w;
    if (wsum > 0) width /= wsum;
    f.value = value;
    f.width = width;
    return f;
}

// LInfDistance (GeodesicODESolver.hxx:221-258) with the upper bound 6 the solver's constructor
// uses: breadth-first over the 3^d - 1 neighbours, points at the bound not expanded, unreached
// points at the largest short.
std::vector<int> ode_linf_distance(const Problem &P, const std::vector<long> &seeds) {
    const int upper = 6, kMax = std::numeric_limits<short>::max();
    std::vector<int> d(P.N, kMax);
    std::vector<long> front, next;
    for (long s : seeds) {
        d[s] = 0;
        front.push_back(s);
    }
    const int nOff = P.dim == 2 ? 9 : 27;
    for (int level = 0; level < upper && !front.empty(); ++level) {
        next.clear();
        for (long a : front) {
            int ai[3] = {0, 0, 0};
            P.unravel(a, ai);
            for (int k = 0; k < nOff; ++k) {
                int idx[3] = {ai[0] + k % 3 - 1, ai[1] + (k / 3) % 3 - 1, P.dim == 3 ? ai[2] + k / 9 - 1 : 0};
                if (k == (nOff - 1) / 2) continue;  // the null offset
                if (!P.periodize(idx)) continue;
                const long y = P.ravel(idx);
                if (d[y] <= level + 1) continue;
                d[y] = level + 1;
                next.push_back(y);
            }
        }
        front.swap(next);
    }
    return d;
}

// GeodesicODESolver::Run for one tip: returns true on failure (a stall).
bool geodesic_ode(const Problem &P, const FMState &S, const std::vector<int> &target,
                  std::vector<std::array<double, 3>> &geodesic, double step) {
    const int D = P.dim;
    std::map<long, std::pair<OdeFlowData, int>> cache;
    const long nStatMax = long((6 * std::sqrt(double(D))) / step);
    std::deque<long> history;
    for (int i = 0; i < std::sqrt(double(D)) / step; ++i) history.push_back(std::numeric_limits<int>::max());
    struct Avg {
        double flow[3];
        long nStat, tol;
    };
    auto flowAt = [&](const double *p) {
        long pIndex[3] = {0, 0, 0};
        double w[3] = {0, 0, 0}, sign[3] = {0, 0, 0};
        for (int i = 0; i < D; ++i) {
            pIndex[i] = (long)std::floor(p[i]);
            w[i] = p[i] - pIndex[i] - 0.5;
            sign[i] = w[i] > 0 ? 1 : -1;
            w[i] = std::abs(w[i]);
        }
        long minTol = std::numeric_limits<int>::max(), maxStat = 0;
        const OdeFlowData *fl[8];
        double wt[8];
        for (int i = 0; i < (1 << D); ++i) {
            int idx[3] = {0, 0, 0};
            int off[3] = {0, 0, 0};
            for (int j = 0; j < D; ++j) {
                off[j] = (i & (1 << j)) ? (int)sign[j] : 0;
                idx[j] = (int)(pIndex[j] + off[j]);
            }
            fl[i] = nullptr;
            wt[i] = 0.;
            if (!P.periodize(idx)) continue;
            const long q = P.ravel(idx);
            minTol = std::min(minTol, (long)target[q]);
            auto it = cache.find(q);
            if (it == cache.end()) it = cache.insert({q, {ode_point_flow(P, S, q), 0}}).first;
            double weight = 1;
            for (int j = 0; j < D; ++j) weight *= off[j] ? w[j] : (1 - w[j]);
            fl[i] = &it->second.first;
            wt[i] = weight;
            maxStat = std::max(maxStat, (long)it->second.second);
            ++it->second.second;
        }
        Avg r{{0., 0., 0.}, maxStat, minTol};
        double minVal = kInf;
        int iMin = -1;
        for (int i = 0; i < (1 << D); ++i)
            if (wt[i] > 0 && fl[i]->value < minVal) {
                minVal = fl[i]->value;
                iMin = i;
            }
        if (minVal == kInf) return r;
        const double thr = minVal + g_geo_causality * fl[iMin]->width;
        double wsum = 0;
        for (int i = 0; i < (1 << D); ++i) {
            if (wt[i] <= 0) continue;
            if (fl[i]->value <= thr) {
                wsum += wt[i];
                for (int c = 0; c < D; ++c) r.flow[c] += wt[i] * fl[i]->flow[c];
            }
        }
        if (wsum > 0) {
            const double b = 1. / wsum;
            for (int c = 0; c < D; ++c) r.flow[c] *= b;
        }
        return r;
    };
    auto norm = [&](const double *f) {
        double s2 = 0;
        for (int c = 0; c < D; ++c) s2 += f[c] * f[c];
        return std::sqrt(s2);
    };
    while (true) {
        const std::array<double, 3> p = geodesic.back();
        Avg fd = flowAt(p.data());
        const long old = history.front();
        history.pop_front();
        history.push_back(fd.tol);
        if (old < fd.tol) return false;  // moving away from the seeds
        if (fd.nStat > nStatMax) return true;  // stall
        double fn = norm(fd.flow);
        if (fn == 0) return false;  // at a seed
        std::array<double, 3> q = p;
        {
            const double a = 0.5 * step, b = 1. / fn;
            for (int c = 0; c < D; ++c) q[c] = p[c] + (fd.flow[c] * a) * b;
        }
        fd = flowAt(q.data());
        fn = norm(fd.flow);
        if (fn == 0) return false;
        std::array<double, 3> r = p;
        {
            const double b = 1. / fn;
            for (int c = 0; c < D; ++c) r[c] = p[c] + (fd.flow[c] * step) * b;
        }
        geodesic.push_back(r);
    }
}

// Seeds: ADim then floor, out-of-range seeds dropped, first value kept for duplicates
// (StaticFactoring.hxx:242-393, BaseGrid.h:61-67).
std::vector<std::pair<long, double>> make_seeds(const Problem &P, int nseeds, const double *seeds,
                                                const double *seedValues) {
    std::map<long, double> m;
    for (int s = 0; s < nseeds; ++s) {
        double q[3];
        P.adim(seeds + s * P.dim, q);
        int idx[3] = {0, 0, 0};
        for (int c = 0; c < P.dim; ++c) idx[c] = (int)std::floor(q[c]);
        if (!P.periodize(idx)) continue;
        m.insert({P.ravel(idx), seedValues ? seedValues[s] : 0.});
    }
    if (m.empty() && nseeds > 0) throw OracleError{"Error : seeds incorrectly set"};
    return std::vector<std::pair<long, double>>(m.begin(), m.end());
}

}  // namespace

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

void oracle_set_order(int order) { g_order = order < 1 ? 1 : (order > 3 ? 3 : order); }
long oracle_last_dips(void) { return g_dips; }
void oracle_set_stopping(long max_accepted, double stop_distance, int nany, const long *any, int nall,
                         const long *all, int nanyu, const long *anyu, int nallu, const long *allu) {
    g_stop_max = max_accepted;
    g_stop_dist = stop_distance;
    g_stop_any = std::set<long>(any, any + nany);
    g_stop_all = std::set<long>(all, all + nall);
    g_stop_any_u = std::set<long>(anyu, anyu + nanyu);
    g_stop_all_u = std::set<long>(allu, allu + nallu);
}
long oracle_last_naccepted(void) { return g_naccepted; }

// Geodesic solver of the geodesics computed next on this thread: ode = 0 the discrete solver,
// 1 the ODE solver with the given causality tolerance (geodesicCausalityTolerance, default 4).
void oracle_set_geodesic_solver(int ode, double causality_tolerance) {
    g_geo_ode = ode;
    g_geo_causality = causality_tolerance;
}
long oracle_last_geodesic_failures(void) { return g_geo_failed; }

// Static factor of the problems solved next on this thread (vals == NULL: none): box extents bd
// (the grid's when sub = 0), the box's first grid index shift, values and gradients (index units)
// row-major over the box, mask (NULL: everywhere).
void oracle_set_factoring(int dim, int sub, const int *bd, const int *shift, const double *vals,
                          const double *grads, const unsigned char *mask) {
    g_fact = Factoring();
    if (!vals) return;
    g_fact.on = true;
    g_fact.sub = sub != 0;
    long n = 1;
    for (int i = 0; i < dim; ++i) {
        g_fact.bd[i] = bd[i];
        g_fact.shift[i] = sub ? shift[i] : 0;
        n *= bd[i];
    }
    g_fact.vals.assign(vals, vals + n);
    g_fact.grads.assign(grads, grads + n * dim);
    if (mask) g_fact.mask.assign(mask, mask + n);
    else g_fact.mask.assign(n, 1);
}

// Walls of the problems set up next on this thread: n bytes (nonzero = wall), n = 0 for none.
void oracle_set_walls(long n, const unsigned char *walls, int refine) {
    g_walls.assign(walls, walls + (walls ? n : 0));
    g_refine_walls = (refine & 1) != 0;
    g_plant_no_walk = (refine & 2) != 0;
}

int oracle_riemann_stencils(int dim, const int *dims, const double *gridScales, const double *metric,
                            int metric_is_field, int dual, int8_t *offsets, double *weights) {
    try {
        g_err.clear();
        long N = 1;
        for (int i = 0; i < dim; ++i) N *= dims[i];
        const int K = dim == 2 ? 3 : 6;
        Stencil st;
        for (long x = 0; x < N; ++x) {
            const double *mc = metric + (metric_is_field ? x * K : 0);
            if (dim == 2) riemann_stencil<2>(mc, dual != 0, gridScales, st);
            else riemann_stencil<3>(mc, dual != 0, gridScales, st);
            for (int k = 0; k < K; ++k) {
                weights[x * K + k] = st.d[k].w;
                for (int c = 0; c < dim; ++c) offsets[(x * K + k) * dim + c] = st.d[k].off[c];
            }
        }
        return 0;
    } catch (const OracleError &e) {
        g_err = e.msg;
        return -1;
    }
}

int oracle_curvature_stencils(int model, int ntheta, double gridScale, double xi, double eps,
                              int8_t *offsets, double *weights, int *symmetric) {
    try {
        g_err.clear();
        Stencil st;
        for (int t = 0; t < ntheta; ++t) {
            if (model == ORACLE_RSF2) rsf2_stencil(t, ntheta, gridScale, xi, eps, st);
            else if (model == ORACLE_ELASTICA2) elastica2_stencil(t, ntheta, gridScale, xi, eps, st);
            else throw OracleError{"not a curvature model"};
            const int n = st.nfwd + st.nsym;
            for (int k = 0; k < n; ++k) {
                weights[t * n + k] = st.d[k].w;
                for (int c = 0; c < 3; ++c) offsets[(t * n + k) * 3 + c] = st.d[k].off[c];
                if (t == 0) symmetric[k] = k >= st.nfwd;
            }
        }
        return 0;
    } catch (const OracleError &e) {
        g_err = e.msg;
        return -1;
    }
}

int oracle_solve(int model, const int *dims, const double *gridScales, const double *origin,
                 const double *metric, int metric_is_field, int dual, double xi, double eps,
                 const double *cost, int cost_is_field, int nseeds, const double *seeds,
                 const double *seedValues, double *values, uint64_t *activeNeighs, double *flow,
                 double *timings) {
    try {
        g_err.clear();
        Problem P;
        const clock_t t0 = clock();
        setup_problem(P, model, dims, gridScales, origin, metric, metric_is_field, dual, xi, eps, cost,
                      cost_is_field);
        const clock_t t1 = clock();
        const auto sd = make_seeds(P, nseeds, seeds, seedValues);
        FMState S;
        double tS = 0, tF = 0;
        fast_marching(P, sd, S, &tS, &tF);
        if (values) std::copy(S.values.begin(), S.values.end(), values);
        if (activeNeighs) std::copy(S.active.begin(), S.active.end(), activeNeighs);
        if (flow) geodesic_flow(P, S, flow);
        if (timings) { timings[0] = double(t1 - t0) / CLOCKS_PER_SEC + tS; timings[1] = tF; }
        return 0;
    } catch (const OracleError &e) {
        g_err = e.msg;
        return -1;
    }
}

int oracle_solve_backward(int model, const int *dims, const double *gridScales, const double *origin,
                          const double *metric, int metric_is_field, int dual, double xi, double eps,
                          const double *cost, int cost_is_field, int nseeds, const double *seeds,
                          const double *seedValues, int ntargets, const double *targets,
                          const double *tweights, double *values, double *costSensitivity,
                          int64_t *seedSensIndex, double *seedSensWeight, int nseeds_max, int *nseedSens) {
    try {
        g_err.clear();
        Problem P;
        setup_problem(P, model, dims, gridScales, origin, metric, metric_is_field, dual, xi, eps, cost,
                      cost_is_field);
        const auto sd = make_seeds(P, nseeds, seeds, seedValues);
        FMState S;
        fast_marching(P, sd, S, nullptr, nullptr);
        if (values) std::copy(S.values.begin(), S.values.end(), values);
        std::vector<std::pair<long, double>> base;
        for (int t = 0; t < ntargets; ++t) {
            double q[3];
            P.adim(targets + t * P.dim, q);
            int idx[3] = {0, 0, 0};
            for (int c = 0; c < P.dim; ++c) idx[c] = (int)std::floor(q[c]);
            if (!P.periodize(idx)) throw OracleError{"Error : inspectSensitivity data points are out of range"};
            base.push_back({P.ravel(idx), tweights ? tweights[t] : 1.});
        }
        std::fill(costSensitivity, costSensitivity + P.N, 0.);
        std::vector<std::pair<long, double>> ss;
        backward_variation(P, S, base, costSensitivity, ss);
        int n = 0;
        for (const auto &p : ss) {
            if (n < nseeds_max) { seedSensIndex[n] = p.first; seedSensWeight[n] = p.second; }
            ++n;
        }
        *nseedSens = n;
        return 0;
    } catch (const OracleError &e) {
        g_err = e.msg;
        return -1;
    }
}

static int geodesics_impl(Problem &P, FMState &S, int ntips, const double *tips, double step, double threshold,
                          double volumeBound, double *points, long max_points, int64_t *lengths, long *npoints,
                          const std::vector<long> &seeds = std::vector<long>()) {
    GeoParams gp;
    gp.step = step;
    gp.threshold = threshold;
    gp.volumeBound = volumeBound;
    long np = 0;
    std::vector<int> target;
    if (g_geo_ode) target = ode_linf_distance(P, seeds);
    g_geo_failed = 0;
    for (int t = 0; t < ntips; ++t) {
        std::vector<std::array<double, 3>> geodesic;
        std::array<double, 3> tip{0, 0, 0};
        P.adim(tips + t * P.dim, tip.data());
        geodesic.push_back(tip);
        if (g_geo_ode) {
            if (geodesic_ode(P, S, target, geodesic, step)) ++g_geo_failed;
        } else {
            geodesic_run(P, S, geodesic, gp);
        }
        lengths[t] = (int64_t)geodesic.size();
        for (const auto &q : geodesic) {
            if (np < max_points)
                for (int c = 0; c < P.dim; ++c) {
                    // ReDim (HFM_ParamDefault.hxx:86-89 and :137-148)
                    const double sc = (P.shared && c == 2) ? P.dependScale : P.h[c];
                    points[np * P.dim + c] = q[c] * sc + P.origin[c];
                }
            ++np;
        }
    }
    *npoints = np;
    return 0;
}

int oracle_geodesics(int model, const int *dims, const double *gridScales, const double *origin,
                     const double *metric, int metric_is_field, int dual, double xi, double eps,
                     const double *cost, int cost_is_field, int nseeds, const double *seeds,
                     const double *seedValues, int ntips, const double *tips, double step, double threshold,
                     double volumeBound, double *points, long max_points, int64_t *lengths, long *npoints) {
    try {
        g_err.clear();
        Problem P;
        setup_problem(P, model, dims, gridScales, origin, metric, metric_is_field, dual, xi, eps, cost,
                      cost_is_field);
        const auto sd = make_seeds(P, nseeds, seeds, seedValues);
        FMState S;
        fast_marching(P, sd, S, nullptr, nullptr);
        std::vector<long> seedIdx;
        for (const auto &sv : sd) seedIdx.push_back(sv.first);
        return geodesics_impl(P, S, ntips, tips, step, threshold, volumeBound, points, max_points, lengths,
                              npoints, seedIdx);
    } catch (const OracleError &e) {
        g_err = e.msg;
        return -1;
    }
}

int oracle_geodesics_given(int model, const int *dims, const double *gridScales, const double *origin,
                           const double *metric, int metric_is_field, int dual, double xi, double eps,
                           const double *cost, int cost_is_field, const double *values,
                           const uint64_t *activeNeighs, int ntips, const double *tips, double step,
                           double threshold, double volumeBound, double *points, long max_points,
                           int64_t *lengths, long *npoints, int nseeds, const double *seeds) {
    try {
        g_err.clear();
        Problem P;
        setup_problem(P, model, dims, gridScales, origin, metric, metric_is_field, dual, xi, eps, cost,
                      cost_is_field);
        FMState S;
        S.values.assign(values, values + P.N);
        S.active.assign(activeNeighs, activeNeighs + P.N);
        std::vector<long> seedIdx;
        if (nseeds > 0)
            for (const auto &sv : make_seeds(P, nseeds, seeds, nullptr)) seedIdx.push_back(sv.first);
        return geodesics_impl(P, S, ntips, tips, step, threshold, volumeBound, points, max_points, lengths,
                              npoints, seedIdx);
    } catch (const OracleError &e) {
        g_err = e.msg;
        return -1;
    }
}

int oracle_forward_variation(int model, const int *dims, const double *gridScales, const double *origin,
                             const double *metric, int metric_is_field, int dual, double xi, double eps,
                             const double *cost, int cost_is_field, int nseeds, const double *seeds,
                             const double *seedValues, int nVar, const double *costVariation,
                             const double *seedValueVariation, double *values, double *valueVariation) {
    try {
        g_err.clear();
        Problem P;
        setup_problem(P, model, dims, gridScales, origin, metric, metric_is_field, dual, xi, eps, cost,
                      cost_is_field);
        const auto sd = make_seeds(P, nseeds, seeds, seedValues);
        FMState S;
        fast_marching(P, sd, S, nullptr, nullptr);
        if (values) std::copy(S.values.begin(), S.values.end(), values);
        std::fill(valueVariation, valueVariation + P.N * nVar, 0.);
        if (seedValueVariation) {
            // (nVar, nseeds) row-major; each seed's index from ADim + floor, later seeds overwrite
            for (int i = 0; i < nseeds; ++i) {
                double q[3];
                P.adim(seeds + i * P.dim, q);
                int idx[3] = {0, 0, 0};
                for (int c = 0; c < P.dim; ++c) idx[c] = (int)std::floor(q[c]);
                if (!P.periodize(idx)) continue;
                const long x = P.ravel(idx);
                for (int k = 0; k < nVar; ++k) valueVariation[x * nVar + k] = seedValueVariation[k * nseeds + i];
            }
        }
        forward_variation(P, S, nVar, costVariation, valueVariation);
        return 0;
    } catch (const OracleError &e) {
        g_err = e.msg;
        return -1;
    }
}

}  // extern "C"
