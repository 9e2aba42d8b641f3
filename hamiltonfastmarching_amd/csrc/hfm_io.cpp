// hfm_io.cpp -- the key/value IO object and the hfm_run() driver (C ABI hfm_io_*, hfm_run).
//
// hfm_run reproduces the phases and keys of HFMInterface<Traits>::Run
// (Headers/Base/Implementation/HFMInterface.hxx:296-579) for the four models of this library,
// on top of the device plan (hfm_plan_*):
//   SetupIO       verbosity, arrayOrdering                                  (:311-341)
//   SetupStencil  dims, gridScale(s), origin, metric | dualMetric, cost | speed, xi, eps
//                 (Riemannian.h:46-61, Curvature2.h:104-108,219-224, HFM_StencilShare.hxx:61-71)
//   SetupSolver   MaxStencilWidth, seeds / seedValues                      (:354-410)
//   RunSolver     FMCPUTime (wall-clock seconds of the GPU solve here)     (:422-450)
//   ExtractGeodesics  tips -> geodesicPoints, geodesicLengths, GeodesicCPUTime (:452-523)
//   ExportData    values, activeNeighs, geodesicFlow, activeOffsets         (:526-579)
//   Extra algorithms  walls, refineStencilAtWallBoundary (ExtraAlgorithms/WallObstructionTest.h:30-36)
//   Finally       inspectSensitivity -> costSensitivity_<i>, seedSensitivity_<i>;
//                 costVariation, seedValueVariation -> valueVariation
//                 (Headers/ExtraAlgorithms/FirstVariation.h:60-124)
// Arrays are FP64, stored row-major with element components last; arrayOrdering permutes the
// grid axes exactly like IO_::GetArray / SetArray (JMM_CPPLibs/JMM_CPPLibs/Output/IO.hxx:68-110,
// 153-240): n = P(s) with P = identity, T (swap axes 0,1), R (reverse) or T o R.
// Unused / defaulted keys are reported like BaseIO::UsageReport (BaseIO.hxx:64-104).
#include <sys/mman.h>

#include <hip/hip_runtime.h>

#include <array>
#include <chrono>
#include <cmath>
#include <limits>
#include <fstream>
#include <cstring>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hfm_amd.h"
#include "selling.h"

// Host copy of a curvature plan's per-theta stencil tables (hfm_plan.hip).
int hfm_plan_curvature_tables(const hfm_plan *p, std::vector<uint32_t> &offs, std::vector<double> &w);

// Reached (finite-valued) grid points of the last solve, counted on the device (hfm_plan.hip).
namespace hfm {
int plan_count_reached(hfm_plan *p, int64_t *n, void *stream);
}

// Geodesics from tips already in grid units (hfm_plan.hip).
int hfm_plan_geodesics_impl(hfm_plan *p, int ntips, const double *tips, bool tips_adim, double step,
                            double threshold, double volumeBound, double *points, int64_t max_points,
                            int64_t *lengths, int64_t *npoints, int *nfailed, void *stream, int solver = 0,
                            double causality = 4., int *tip_failed = nullptr);

namespace {

struct IOError {
    std::string msg;
};

struct Field {
    bool is_string = false;
    std::string str;
    std::vector<int64_t> dims;  // empty: scalar
    std::vector<double> data;
    bool computed = false;
};

// Host arrays of the IO store.  Arrays of 4 MB or more ask the kernel to back them with 2 MB pages
// (transparent huge pages on request, MADV_HUGEPAGE), which halves the cost of first touching a
// fresh 100 MB array; the request is advice only, and the contents are those of a plain vector.
void advise_huge(void *p, size_t bytes) {
    if (bytes < (4u << 20)) return;
    const uintptr_t huge = 2u << 20;
    const uintptr_t b = (reinterpret_cast<uintptr_t>(p) + huge - 1) & ~(huge - 1);
    const uintptr_t e = (reinterpret_cast<uintptr_t>(p) + bytes) & ~(huge - 1);
    if (e > b) (void)madvise(reinterpret_cast<void *>(b), e - b, MADV_HUGEPAGE);
}
std::vector<double> host_array(size_t n) {
    std::vector<double> v;
    v.reserve(n);
    advise_huge(v.data(), n * sizeof(double));
    v.resize(n);
    return v;
}

enum Ordering { ROW_MAJOR = 0, YXZ_ROW_MAJOR = 1, COLUMN_MAJOR = 2, YXZ_COLUMN_MAJOR = 3 };

}  // namespace

struct hfm_io {
    std::map<std::string, Field> fields;
    std::set<std::string> used, defaulted, visitedUnset;
    std::string err;
    std::vector<std::string> computed_order;
    int verbosity = 1;
    Ordering ordering = ROW_MAJOR;
    std::string log;

    bool has(const std::string &k) const { return fields.count(k) > 0; }
    const Field &raw(const std::string &k) {
        auto it = fields.find(k);
        if (it == fields.end()) {
            visitedUnset.insert(k);
            throw IOError{"Field " + k + " not found"};
        }
        used.insert(k);
        return it->second;
    }
    double scalar(const std::string &k) {
        const Field &f = raw(k);
        if (f.is_string) throw IOError{"IO error : key " + k + " is a string, not a numerical field."};
        if (f.data.size() != 1) throw IOError{"IO error : field " + k + " is not a scalar"};
        return f.data[0];
    }
    double scalar(const std::string &k, double def) {
        if (!has(k)) {
            defaulted.insert(k);
            return def;
        }
        return scalar(k);
    }
    std::string str(const std::string &k) {
        const Field &f = raw(k);
        if (!f.is_string) throw IOError{"IO error : key " + k + " is a numerical field, not a string"};
        return f.str;
    }
    std::string str(const std::string &k, const std::string &def) {
        if (!has(k)) {
            defaulted.insert(k);
            return def;
        }
        return str(k);
    }
    std::vector<double> vec(const std::string &k) {
        const Field &f = raw(k);
        if (f.is_string) throw IOError{"IO error : key " + k + " is a string, not a numerical field."};
        return f.data;
    }
    void set_computed(const std::string &k, Field f) {
        f.computed = true;
        if (!fields.count(k) || !fields[k].computed) computed_order.push_back(k);
        fields[k] = std::move(f);
    }
    void set_scalar(const std::string &k, double v) {
        Field f;
        f.data = {v};
        set_computed(k, std::move(f));
    }
    void set_string(const std::string &k, const std::string &v) {
        Field f;
        f.is_string = true;
        f.str = v;
        set_computed(k, std::move(f));
    }
};

namespace {

// The angle step of the curvature models' bundle axis, 2 pi / n_theta in FP64 as the reference
// forms it (mathPi a double, CommonTraits.h:30; Curvature2.h:108).
double theta_step(int64_t ntheta) {
    constexpr double pi = 3.141592653589793238462643383279502884;
    return 2 * pi / (double)ntheta;
}

// Index map between a stored array (as the user wrote it) and the internal row-major grid
// array: internal multi-index n = P(s) of stored multi-index s (only the first g axes move).
std::vector<int64_t> permute_dims(const std::vector<int64_t> &d, Ordering o, bool inverse = false) {
    std::vector<int64_t> r = d;
    auto T = [](std::vector<int64_t> &v) {
        if (v.size() >= 2) std::swap(v[0], v[1]);
    };
    auto R = [](std::vector<int64_t> &v) {
        const size_t n = v.size();
        for (size_t i = 0; i < n / 2; ++i) std::swap(v[i], v[n - 1 - i]);
    };
    switch (o) {
        case ROW_MAJOR: break;
        case YXZ_ROW_MAJOR: T(r); break;
        case COLUMN_MAJOR: R(r); break;
        case YXZ_COLUMN_MAJOR:  // P = T o R, P^-1 = R o T
            if (inverse) { T(r); R(r); } else { R(r); T(r); }
            break;
    }
    return r;
}

// Convert a field whose first axes are grid axes (ncomp trailing component values per point)
// between the stored layout (as the user wrote it) and the internal row-major grid layout.
// Internal multi-index n = P(s) for stored multi-index s.  dims_in: grid dims of src;
// dims_out receives the grid dims of the result.
std::vector<double> reorder(const std::vector<double> &src, const std::vector<int64_t> &dims_in, int64_t ncomp,
                            Ordering o, bool to_internal, std::vector<int64_t> &dims_out) {
    const std::vector<int64_t> stored = to_internal ? dims_in : permute_dims(dims_in, o, true);
    const std::vector<int64_t> internal = to_internal ? permute_dims(dims_in, o) : dims_in;
    dims_out = to_internal ? internal : stored;
    if (o == ROW_MAJOR || stored.size() < 2) return src;
    const size_t g = stored.size();
    int64_t total = 1;
    for (auto v : stored) total *= v;
    std::vector<double> dst = host_array(src.size());
    std::vector<int64_t> s(g);
    for (int64_t si = 0; si < total; ++si) {
        int64_t r = si;
        for (size_t k = g; k-- > 0;) {
            s[k] = r % stored[k];
            r /= stored[k];
        }
        const std::vector<int64_t> n = permute_dims(s, o);
        int64_t ni = 0;
        for (size_t k = 0; k < g; ++k) ni = ni * internal[k] + n[k];
        const int64_t a = to_internal ? si : ni;
        const int64_t b = to_internal ? ni : si;
        for (int64_t c = 0; c < ncomp; ++c) dst[b * ncomp + c] = src[a * ncomp + c];
    }
    return dst;
}

Ordering parse_ordering(const std::string &s) {
    if (s == "RowMajor") return ROW_MAJOR;
    if (s == "YXZ_RowMajor") return YXZ_ROW_MAJOR;
    if (s == "ColumnMajor") return COLUMN_MAJOR;
    if (s == "YXZ_ColumnMajor") return YXZ_COLUMN_MAJOR;
    throw IOError{"Error : unrecognized array ordering"};
}

struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) throw IOError{"hipMalloc failed"};
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    double *d() const { return static_cast<double *>(p); }
};

struct PlanGuard {
    hfm_plan *p = nullptr;
    ~PlanGuard() { hfm_plan_destroy(p); }
};

void check(int rc) {
    if (rc) throw IOError{hfm_last_error()};
}

// A grid field of element size ncomp: constant (0-D element), full grid, or for the curvature
// models a field over (x, y) only or over theta only (HFMInterface.hxx:23-32).  Returns the
// internal row-major full-grid array (or a single element if constant).
std::vector<double> grid_field(hfm_io *io, const std::string &key, const std::vector<int64_t> &dims, int64_t ncomp,
                               bool bundle, bool &is_field) {
    const Field &f = io->raw(key);
    if (f.is_string) throw IOError{"IO error : key " + key + " is a string, not a numerical field."};
    std::vector<int64_t> gd = f.dims;
    if (ncomp > 1) {
        if (gd.empty() || gd.back() != ncomp)
            throw IOError{"BaseIO input error first dimension " + std::to_string(gd.empty() ? 0 : gd.back()) +
                          " of field " + key + " does not match expected value " + std::to_string(ncomp) + "."};
        gd.pop_back();
    }
    if (gd.empty()) {
        is_field = false;
        return f.data;
    }
    std::vector<int64_t> internal;
    std::vector<double> v = reorder(f.data, gd, ncomp, io->ordering, true, internal);
    is_field = true;
    const size_t D = dims.size();
    if (internal.size() == D) {
        if (internal != dims) throw IOError{"Error : Field " + key + " has inconsistent dimensions."};
        return v;
    }
    if (bundle && (internal.size() == D - 1 || internal.size() == 1)) {
        const bool indep = internal.size() == D - 1;  // (x, y) field broadcast over theta
        if (indep && (internal[0] != dims[0] || internal[1] != dims[1]))
            throw IOError{"Error : Field " + key + " has inconsistent dimensions."};
        if (!indep && internal[0] != dims[2]) throw IOError{"Error : Field " + key + " has inconsistent dimensions."};
        std::vector<double> full = host_array((size_t)(dims[0] * dims[1] * dims[2]) * ncomp);
        for (int64_t i = 0; i < dims[0]; ++i)
            for (int64_t j = 0; j < dims[1]; ++j)
                for (int64_t t = 0; t < dims[2]; ++t) {
                    const int64_t dst = (i * dims[1] + j) * dims[2] + t;
                    const int64_t src = indep ? i * dims[1] + j : t;
                    for (int64_t c = 0; c < ncomp; ++c) full[dst * ncomp + c] = v[src * ncomp + c];
                }
        return full;
    }
    throw IOError{"Field " + key + " has incorrect depth.\n"};
}

// Store an internal row-major grid array as a computed key, in the IO's array ordering (moved,
// not copied, when no reordering is needed).
void set_grid_array(hfm_io *io, const std::string &key, std::vector<double> internal,
                    const std::vector<int64_t> &dims, int64_t ncomp) {
    Field f;
    std::vector<int64_t> stored;
    if (io->ordering == ROW_MAJOR || dims.size() < 2) {
        stored = dims;
        f.data = std::move(internal);
    } else {
        f.data = reorder(internal, dims, ncomp, io->ordering, false, stored);
    }
    f.dims = stored;
    if (ncomp > 1) f.dims.push_back(ncomp);
    io->set_computed(key, std::move(f));
}

// The input grid field read in place when it is already a full row-major grid of the expected
// dimensions (no copy of e.g. the 96 MB metric of a 128^3 grid), else converted by grid_field.
struct FieldView {
    const double *p = nullptr;
    size_t n = 0;
    std::vector<double> own;
    const double *data() const { return p; }
    size_t size() const { return n; }
};

FieldView grid_field_view(hfm_io *io, const std::string &key, const std::vector<int64_t> &dims, int64_t ncomp,
                          bool bundle, bool &is_field) {
    FieldView v;
    const Field &f = io->raw(key);
    if (!f.is_string && io->ordering == ROW_MAJOR) {
        std::vector<int64_t> want = dims;
        if (ncomp > 1) want.push_back(ncomp);
        if (f.dims == want) {
            is_field = true;
            v.p = f.data.data();
            v.n = f.data.size();
            return v;
        }
    }
    v.own = grid_field(io, key, dims, ncomp, bundle, is_field);
    v.p = v.own.data();
    v.n = v.own.size();
    return v;
}

// Distance guesses of the Riemannian stencil (Riemannian.h:62-71): the metric as given (primal,
// or dual and inverted) scaled by h_i h_j, at a grid index, or at a point (grid-index units)
// interpolated over the 2^d cells around it (BaseGrid.h:80-104): out-of-domain corners weighted 0
// without renormalisation (PeriodicGrid.hxx:28-40), the first corner always evaluated, the
// others only with a nonzero weight (MapWeightedSum.h:21-36).
template <int D> struct MetricGuess {
    using Sym = hfm::SymPacked<D>;
    static constexpr int K = Sym::K;
    const std::vector<int64_t> &dims;
    const double *h;
    const double *metric;
    bool is_field, dual;
    bool in_domain(const int64_t *idx) const {
        for (int i = 0; i < D; ++i)
            if (idx[i] < 0 || idx[i] >= dims[i]) return false;
        return true;
    }
    Sym raw(const int64_t *idx) const {
        Sym m;
        int64_t lin = 0;
        for (int i = 0; i < D; ++i) lin = lin * dims[i] + idx[i];
        for (int k = 0; k < K; ++k) m.m[k] = metric[(is_field ? lin : 0) * K + k];
        return m;
    }
    // Rescale<false>: primal metric times the rank-one tensor h h^T, component by component
    Sym rescale(Sym m) const {
        if (dual) m = hfm::sym_inverse<D>(m);
        for (int i = 0; i < D; ++i)
            for (int j = 0; j <= i; ++j) m.m[Sym::lin(i, j)] *= h[i] * h[j];
        return m;
    }
    Sym at_index(const int64_t *idx) const { return rescale(raw(idx)); }
    Sym at_point(const double *p) const {
        int64_t nearest[D];
        int dir[D];
        double wgt[D];
        for (int i = 0; i < D; ++i) {
            nearest[i] = (int64_t)std::floor(p[i]);
            const double diff = p[i] - (double)nearest[i] - 0.5;
            dir[i] = diff > 0 ? 1 : -1;
            wgt[i] = std::abs(diff);
        }
        Sym mp{};
        for (int c = 0; c < (1 << D); ++c) {
            int64_t q[D];
            double w = 1.;
            for (int j = 0; j < D; ++j) {
                q[j] = nearest[j];
                if (c >> j & 1) {
                    w *= wgt[j];
                    q[j] += dir[j];
                } else {
                    w *= 1. - wgt[j];
                }
            }
            if (!in_domain(q)) w = 0.;
            if (c == 0) {
                // the reference evaluates the first corner unconditionally; outside the domain there is
                // no metric, and no corner at all is valid (PeriodicGrid.hxx:37-38)
                if (!in_domain(q)) throw IOError{"Error : point has no valid neighbors in the grid domain"};
                const Sym m = raw(q);
                for (int k = 0; k < K; ++k) mp.m[k] = w * m.m[k];
            } else if (w != 0.) {
                const Sym m = raw(q);
                for (int k = 0; k < K; ++k) mp.m[k] += w * m.m[k];
            }
        }
        return rescale(mp);
    }
    // sqrt(<v, M v>) and M v / Norm(v) (a reciprocal multiply) in the reference's order
    // (SymmetricMatrixType.h:42-54, SymmetricMatrixType.hxx:18-44, VectorType.h:37-42)
    static double norm(const Sym &m, const double *v) {
        double sum = 0.;
        for (int i = 0; i < D; ++i) {
            double sumi = 0.;
            for (int j = 0; j < D; ++j) sumi += m(i, j) * v[j];
            sum += v[i] * sumi;
        }
        return std::sqrt(sum);
    }
    static void gradient(const Sym &m, const double *v, double *g) {
        for (int i = 0; i < D; ++i) {
            g[i] = 0.;
            for (int j = 0; j < D; ++j) g[i] += m(i, j) * v[j];
        }
        const double b = 1. / norm(m, v);
        for (int i = 0; i < D; ++i) g[i] *= b;
    }
};

// Seed spreading (StaticFactoring.hxx:284-345), Riemann models: every grid point within
// `radius` (grid units) of a seed p, and p's own cell, becomes a seed with value
// v + (|p - q|_{M(p)} + |p - q|_{M(q)}) / 2, where M(p) is the metric interpolated at p from the
// 2^d surrounding grid points (BaseGrid.h:80-100, out-of-domain corners weighted 0 without
// renormalisation, PeriodicGrid.hxx:28-40; MapWeightedSum.h) and M(q) the metric at q, both primal
// and scaled by h_i h_j (Riemannian.h:62-71).  Candidates come from the box of half-width
// 1 + ceil(radius) in row-major order; a point is taken once per seed.  `seeds` (physical) and
// `values` are replaced; spreadedSeeds / spreadedSeedValues are exported.
template <int D>
void spread_seeds_d(hfm_io *io, const std::vector<int64_t> &dims, const double *h, const std::vector<double> &origin,
                    const double *metric, bool is_field, bool dual, double radius, int step_width,
                    std::vector<double> &seeds, std::vector<double> &values) {
    using Sym = hfm::SymPacked<D>;
    constexpr int K = Sym::K;
    // Reversed offsets of a grid point y (HFM_StencilRecompute.hxx:81-124, read by
    // StaticFactoring.hxx:333-340): the offsets e, sorted and unique, such that the point y - e has
    // a nonzero-weight difference +-e_k equal to e (so that y = (y - e) + e), both points in the
    // grid.  The stencils are the FP64 Selling ones the device computes, evaluated here for the
    // points within step_width (MaxStencilWidth, a bound on every |e_i|) of y.
    std::map<int64_t, std::pair<std::array<int8_t, 3 * K>, std::array<double, K>>> stencil_cache;
    double invhh[6];
    for (int i = 0; i < D; ++i)
        for (int j = 0; j <= i; ++j) invhh[i * (i + 1) / 2 + j] = 1. / (h[i] * h[j]);
    auto stencil_of = [&](const int64_t *x) -> const std::pair<std::array<int8_t, 3 * K>, std::array<double, K>> & {
        int64_t lin = 0;
        for (int i = 0; i < D; ++i) lin = lin * dims[i] + x[i];
        auto it = stencil_cache.find(lin);
        if (it != stencil_cache.end()) return it->second;
        std::pair<std::array<int8_t, 3 * K>, std::array<double, K>> st{};
        int8_t o8[K][D];
        int mx = 0;
        hfm::riemann_point_stencil<D>(metric + (is_field ? lin * K : 0), dual, invhh, o8, st.second.data(), &mx);
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < D; ++c) st.first[3 * k + c] = o8[k][c];
        return stencil_cache.emplace(lin, st).first->second;
    };
    auto reversed_offsets = [&](const int64_t *y) {
        std::set<std::array<int, D>> offs;
        int64_t x[D];
        const int64_t side = 2 * (int64_t)step_width + 1;
        int64_t box = 1;
        for (int i = 0; i < D; ++i) box *= side;
        for (int64_t b = 0; b < box; ++b) {
            int64_t r = b;
            bool ok = true;
            for (int i = D - 1; i >= 0; --i) {
                x[i] = y[i] + (r % side) - step_width;
                r /= side;
                ok = ok && x[i] >= 0 && x[i] < dims[i];
            }
            if (!ok) continue;
            const auto &st = stencil_of(x);
            for (int k = 0; k < K; ++k) {
                if (st.second[k] == 0.) continue;
                for (int sg = 1; sg >= -1; sg -= 2) {
                    std::array<int, D> e;
                    bool hit = true;
                    for (int i = 0; i < D; ++i) {
                        e[i] = (int)(int8_t)(sg * st.first[3 * k + i]);
                        hit = hit && x[i] + e[i] == y[i];
                    }
                    if (hit) offs.insert(e);
                }
            }
        }
        return offs;
    };
    auto org = [&](int i) { return origin.empty() ? 0. : origin[i]; };
    const MetricGuess<D> guess{dims, h, metric, is_field, dual};
    auto in_domain = [&](const int64_t *idx) { return guess.in_domain(idx); };
    const int n_in = (int)(seeds.size() / D);
    std::vector<double> newPts, newVals;
    const int64_t bd = 1 + (int64_t)std::ceil(radius);
    const int64_t side = 1 + 2 * bd;
    int64_t boxn = 1;
    for (int i = 0; i < D; ++i) boxn *= side;
    for (int s = 0; s < n_in; ++s) {
        double p[D];
        for (int i = 0; i < D; ++i) p[i] = (seeds[(size_t)s * D + i] - org(i)) / h[i];
        const double value = values[s];
        int64_t nearest[D];
        for (int i = 0; i < D; ++i) nearest[i] = (int64_t)std::floor(p[i]);
        const Sym distp = guess.at_point(p);
        std::set<std::vector<int64_t>> seen;
        auto insert = [&](const int64_t *index, double rad) {
            double v[D], q[D];
            for (int i = 0; i < D; ++i) {
                q[i] = (double)index[i] + 0.5;
                v[i] = p[i] - q[i];
            }
            double e2 = 0.;
            for (int i = 0; i < D; ++i) e2 += v[i] * v[i];
            if (std::sqrt(e2) >= rad) return;
            if (!seen.insert(std::vector<int64_t>(index, index + D)).second) return;
            if (!in_domain(index)) return;
            const Sym distq = guess.at_index(index);
            for (int i = 0; i < D; ++i) newPts.push_back(q[i]);
            newVals.push_back(value + 0.5 * (MetricGuess<D>::norm(distp, v) + MetricGuess<D>::norm(distq, v)));
        };
        insert(nearest, std::numeric_limits<double>::infinity());
        for (int64_t b = 0; b < boxn; ++b) {
            int64_t idx[D];
            int64_t r = b;
            for (int i = D - 1; i >= 0; --i) {
                idx[i] = nearest[i] + (r % side) - bd;
                r /= side;
            }
            insert(idx, radius);
        }
        if (step_width <= 0) continue;
        // negative seedRadius: also every point one step of the reversed stencil away from a point
        // inserted above, visited in sorted index order and sorted offset order.  (An inserted
        // index outside the grid has no stencil: the reference would read its reversed offsets at
        // an index computed from outside the grid; here it is skipped.)
        const std::set<std::vector<int64_t>> seen2 = seen;
        for (const auto &y : seen2) {
            if (!in_domain(y.data())) continue;
            for (const auto &e : reversed_offsets(y.data())) {
                int64_t idx[D];
                for (int i = 0; i < D; ++i) idx[i] = y[i] + e[i];
                insert(idx, std::numeric_limits<double>::infinity());
            }
        }
    }
    const size_t n = newVals.size();
    seeds.resize(n * D);
    Field fp, fv;
    fp.dims = {(int64_t)n, (int64_t)D};
    for (size_t k = 0; k < n; ++k)
        for (int i = 0; i < D; ++i) {
            const double x = newPts[k * D + i] * h[i] + org(i);  // ReDim
            seeds[k * D + i] = x;
            fp.data.push_back(x);
        }
    values = newVals;
    fv.dims = {(int64_t)n};
    fv.data = newVals;
    io->set_computed("spreadedSeeds", std::move(fp));
    io->set_computed("spreadedSeedValues", std::move(fv));
}

// A static factor on the host, as hfm_plan_set_factoring takes it: extents and first grid index
// of its box, values, gradients (grid-index units, dim per point) and mask (empty: everywhere),
// row-major over the box.
struct FactorHost {
    bool on = false, sub = false;
    std::vector<int64_t> bd, shift;
    std::vector<double> vals, grads;
    std::vector<uint8_t> mask;
};

// ComputeFactor (StaticFactoring.hxx:116-239), Riemann models, from the seed p (grid-index
// units): over the box of grid points from max(0, floor(p - R)) to min(n - 1, ceil(p + R))
// (upper end excluded, as the reference sizes it), or the whole grid for R < 0.
//   "Seed" / "Key": T(x) = grad . v, grad = M v / |v|_M with v = x - p and M the metric at p
//     (so T = |v|_M); both 0 at v = 0.
//   "Both": T = g . v with g the mean of that gradient and the one of the metric at x -- the
//     metric read at the point's index within the box, as the reference reads it -- and the
//     gradient of T completed by centred differences of |v|_{M(x)} / 2 along each axis (one-sided
//     at the box's faces), accumulated point by point in the reference's order.
template <int D>
FactorHost compute_factor_d(const std::vector<int64_t> &dims, const double *h, const double *metric, bool is_field,
                            bool dual, const double *seed, int64_t R, bool both) {
    const MetricGuess<D> guess{dims, h, metric, is_field, dual};
    using Sym = typename MetricGuess<D>::Sym;
    FactorHost F;
    F.sub = R >= 0;
    F.bd.assign(D, 0);
    F.shift.assign(D, 0);
    int64_t size = 1;
    for (int i = 0; i < D; ++i) {
        if (R < 0) {
            F.bd[i] = dims[i];
        } else {
            const int64_t bottom = std::max<int64_t>(0, (int64_t)std::floor(seed[i] - (double)R));
            const int64_t top = std::min<int64_t>(dims[i] - 1, (int64_t)std::ceil(seed[i] + (double)R));
            F.shift[i] = bottom;
            F.bd[i] = top - bottom;
        }
        if (F.bd[i] <= 0) return FactorHost();  // an empty box: the factor is disabled
        size *= F.bd[i];
    }
    F.on = true;
    F.vals.assign((size_t)size, std::numeric_limits<double>::quiet_NaN());
    F.grads.assign((size_t)size * D, 0.);
    const Sym seedDist = guess.at_point(seed);
    auto unravel = [&](int64_t i, int64_t *idx) {
        for (int k = D - 1; k >= 0; --k) {
            idx[k] = i % F.bd[k];
            i /= F.bd[k];
        }
    };
    auto in_box = [&](const int64_t *idx) {
        for (int k = 0; k < D; ++k)
            if (idx[k] < 0 || idx[k] >= F.bd[k]) return false;
        return true;
    };
    auto conv = [&](const int64_t *idx) {
        int64_t r = idx[0];
        for (int k = 1; k < D; ++k) r = r * F.bd[k] + idx[k];
        return r;
    };
    // -Gradient(-v) of a guess (the sign convention of asymmetric norms), 0 at v = 0
    auto grad_of = [&](const Sym &m, const double *v, bool null, double *g) {
        if (null) {
            for (int k = 0; k < D; ++k) g[k] = 0.;
            return;
        }
        double mv[D], gg[D];
        for (int k = 0; k < D; ++k) mv[k] = -v[k];
        MetricGuess<D>::gradient(m, mv, gg);
        for (int k = 0; k < D; ++k) g[k] = -gg[k];
    };
    for (int64_t pi = 0; pi < size; ++pi) {
        int64_t loc[D];
        unravel(pi, loc);
        double v[D];
        bool null = true;
        for (int k = 0; k < D; ++k) {
            v[k] = ((double)(loc[k] + F.shift[k]) + 0.5) - seed[k];
            null = null && v[k] == 0.;
        }
        double g[D];
        if (!both) {
            grad_of(seedDist, v, null, g);
            double sp = 0.;
            for (int k = 0; k < D; ++k) {
                F.grads[(size_t)pi * D + k] = g[k];
                sp += g[k] * v[k];
            }
            F.vals[(size_t)pi] = sp;
            continue;
        }
        const Sym pDist = guess.at_index(loc);
        double seedG[D], pG[D];
        grad_of(seedDist, v, null, seedG);
        grad_of(pDist, v, null, pG);
        double sp = 0.;
        for (int k = 0; k < D; ++k) g[k] = 0.5 * (pG[k] + seedG[k]);
        for (int k = 0; k < D; ++k) sp += g[k] * v[k];
        F.vals[(size_t)pi] = sp;
        for (int k = 0; k < D; ++k) F.grads[(size_t)pi * D + k] += g[k];
        for (int k = 0; k < D; ++k)
            for (int eps = -1; eps <= 1; eps += 2) {
                int64_t q[D];
                for (int j = 0; j < D; ++j) q[j] = loc[j];
                q[k] += eps;
                if (in_box(q)) {
                    double w[D];
                    for (int j = 0; j < D; ++j) w[j] = -((((double)(q[j] + F.shift[j])) + 0.5) - seed[j]);
                    double qVal = MetricGuess<D>::norm(pDist, w) / 2.;
                    const int64_t qi = conv(q);
                    q[k] += eps;
                    if (in_box(q)) qVal /= 2;
                    F.grads[(size_t)qi * D + k] -= eps * qVal;
                } else {
                    double pgv = 0.;
                    for (int j = 0; j < D; ++j) pgv += pG[j] * v[j];
                    F.grads[(size_t)pi * D + k] += eps * pgv / 2.;
                }
            }
    }
    return F;
}

// ImportFactor (StaticFactoring.hxx:77-113): factoringValues, factoringGradients (physical,
// re-dimensioned into grid-index units by the scale of each axis), factoringMask (nonzero = on;
// default everywhere) and factoringIndexShift (the box's first grid index, truncated to integers;
// without it the arrays cover the whole grid), in the IO's array ordering.
FactorHost import_factor(hfm_io *io, const std::vector<int64_t> &dims, const double *scales, int dim) {
    auto box_array = [&](const std::string &key, int64_t ncomp, std::vector<int64_t> &bdims) {
        const Field &f = io->raw(key);
        if (f.is_string) throw IOError{"IO error : key " + key + " is a string, not a numerical field."};
        std::vector<int64_t> gd = f.dims;
        if (ncomp > 1) {
            if (gd.empty() || gd.back() != ncomp)
                throw IOError{"BaseIO input error first dimension " + std::to_string(gd.empty() ? 0 : gd.back()) +
                              " of field " + key + " does not match expected value " + std::to_string(ncomp) + "."};
            gd.pop_back();
        }
        if ((int)gd.size() != dim) throw IOError{"Field " + key + " has incorrect depth.\n"};
        return reorder(f.data, gd, ncomp, io->ordering, true, bdims);
    };
    FactorHost F;
    std::vector<int64_t> vd, gdims, md;
    F.vals = box_array("factoringValues", 1, vd);
    if (!io->has("factoringGradients")) throw IOError{"StaticFactoring error : factoringGradients are not provided"};
    F.grads = box_array("factoringGradients", dim, gdims);
    if (io->has("factoringMask")) {
        const std::vector<double> m = box_array("factoringMask", 1, md);
        F.mask.resize(m.size());
        for (size_t i = 0; i < m.size(); ++i) F.mask[i] = m[i] != 0. ? 1 : 0;
    } else {
        md = vd;
    }
    F.sub = io->has("factoringIndexShift");
    F.shift.assign(dim, 0);
    if (F.sub) {
        const std::vector<double> sh = io->vec("factoringIndexShift");
        if ((int)sh.size() != dim) throw IOError{"IO error : field factoringIndexShift has the wrong size"};
        for (int i = 0; i < dim; ++i) F.shift[i] = (int64_t)sh[i];
    }
    const std::vector<int64_t> &want = F.sub ? vd : dims;
    if (vd != want || gdims != want || md != want)
        throw IOError{"Static factoring error : inconsistent data dimension."};
    for (size_t k = 0; k < F.grads.size(); ++k) F.grads[k] *= scales[k % dim];
    F.bd = vd;
    F.on = !F.vals.empty();
    return F;
}

// seedValueVariation is given for the `seeds` as passed; the solver may run on more seeds
// (spread or unoriented).  Each solver seed takes the variation of the last given seed at its
// grid point (FirstVariation.h:110-119 writes them in order), zero otherwise.
std::vector<double> seed_variation_for_solver(const std::vector<double> &svv, int64_t nvar, int nin,
                                              const std::vector<double> &given, const std::vector<double> &solver,
                                              int dim, const std::vector<int64_t> &dims, const double *h,
                                              const std::vector<double> &origin, bool bundle, double thetaStep) {
    auto index_of = [&](const double *pt, std::vector<int64_t> &idx) {
        idx.assign(dim, 0);
        for (int i = 0; i < dim; ++i) {
            const bool th = bundle && i == 2;
            const double o = th ? -thetaStep / 2. : (origin.empty() ? 0. : origin[i]);
            const double q = std::floor((pt[i] - o) / (th ? thetaStep : h[i]));
            if (!(q > -2e9 && q < 2e9)) return false;
            int64_t c = (int64_t)q;
            if (th) {
                c %= dims[2];
                if (c < 0) c += dims[2];
            } else if (c < 0 || c >= dims[i]) {
                return false;
            }
            idx[i] = c;
        }
        return true;
    };
    std::map<std::vector<int64_t>, int> last;
    std::vector<int64_t> idx;
    for (int i = 0; i < nin; ++i)
        if (index_of(&given[(size_t)i * dim], idx)) last[idx] = i;
    const int nsol = (int)(solver.size() / dim);
    std::vector<double> out((size_t)nvar * nsol, 0.);
    for (int j = 0; j < nsol; ++j) {
        if (!index_of(&solver[(size_t)j * dim], idx)) continue;
        const auto it = last.find(idx);
        if (it == last.end()) continue;
        for (int64_t k = 0; k < nvar; ++k) out[(size_t)k * nsol + j] = svv[(size_t)k * nin + it->second];
    }
    return out;
}

double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void run_impl(hfm_io *io) {
    io->computed_order.clear();
    io->used.clear();
    io->defaulted.clear();
    io->visitedUnset.clear();
    // Run_SetupIO
    io->verbosity = (int)io->scalar("verbosity", 1);
    io->ordering = parse_ordering(io->str("arrayOrdering", "RowMajor"));
    // Dispatch on the model key (DispatchAndRun.h:224-353)
    const std::string model = io->str("model");
    const bool riemann = (model == "Riemann2" || model == "Riemann3");
    const bool bundle = (model == "ReedsSheppForward2" || model == "Elastica2" || model == "Elastica2_5");
    if (!riemann && !bundle) throw IOError{"Unrecognized model " + model};
    const int dim = model == "Riemann2" ? 2 : 3;
    // Run_SetupStencil
    const std::vector<double> dv = io->vec("dims");
    if ((int)dv.size() != dim) throw IOError{"IO error : field dims has the wrong size"};
    std::vector<int64_t> dims(dim);
    for (int i = 0; i < dim; ++i) dims[i] = (int64_t)dv[i];
    double h[3] = {0, 0, 0};
    if (riemann && io->has("gridScales")) {
        const std::vector<double> g = io->vec("gridScales");
        if ((int)g.size() != dim) throw IOError{"IO error : field gridScales has the wrong size"};
        for (int i = 0; i < dim; ++i) h[i] = g[i];
    } else {
        const double gs = io->scalar("gridScale");
        for (int i = 0; i < dim; ++i) h[i] = gs;
    }
    std::vector<double> origin;
    if (io->has("origin")) {
        origin = io->vec("origin");
        const size_t want = riemann ? dim : 2;
        if (origin.size() != want) throw IOError{"IO error : field origin has the wrong size"};
    } else {
        io->defaulted.insert("origin");
    }
    const int precision = (int)io->scalar("precision", 64.);
    const double xi = bundle ? io->scalar("xi", 1.) : 1.;
    const double eps = bundle ? io->scalar("eps", 0.1) : 0.1;
    // Read and validate every input field on the host before any device work.
    const int64_t K = dim * (dim + 1) / 2;
    bool is_field = false, hasDual = false;
    FieldView field;
    if (riemann) {
        hasDual = io->has("dualMetric");
        if (hasDual && io->has("metric")) throw IOError{"Error: both primal and dual metric provided"};
        field = grid_field_view(io, hasDual ? "dualMetric" : "metric", dims, K, false, is_field);
    } else if (io->has("speed")) {
        field.own = grid_field(io, "speed", dims, 1, true, is_field);
        for (auto &v : field.own) v = 1. / v;
        field.p = field.own.data();
        field.n = field.own.size();
    } else {
        field = grid_field_view(io, "cost", dims, 1, true, is_field);
    }
    if (precision != 32 && precision != 64) throw IOError{"precision must be 32 or 64"};
    // Walls (WallObstructionTest.h:30-36): an integral field read as bool (nonzero = wall),
    // constant, full, or for the curvature models over (x, y) or theta only.
    std::vector<uint8_t> walls;
    if (io->has("walls")) {
        bool wfield = false;
        const std::vector<double> w = grid_field(io, "walls", dims, 1, bundle, wfield);
        int64_t n = 1;
        for (auto d : dims) n *= d;
        walls.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) walls[i] = w[wfield ? i : 0] != 0. ? 1 : 0;
    }
    const bool refineWalls = io->scalar("refineStencilAtWallBoundary", 0.) != 0.;

    PlanGuard pg;
    check(hfm_plan_create(&pg.p, model.c_str(), dims.data(), h, origin.empty() ? nullptr : origin.data(), precision,
                          xi, eps));
    hfm_plan *plan = pg.p;
    int64_t N = 1;
    for (auto d : dims) N *= d;

    const double t0 = wall();
    {
        DevBuf dm(field.size() * sizeof(double));
        if (hipMemcpy(dm.p, field.data(), field.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            throw IOError{"hipMemcpy failed"};
        if (riemann) check(hfm_plan_set_metric(plan, dm.d(), is_field ? 1 : 0, hasDual ? 1 : 0, nullptr));
        else check(hfm_plan_set_cost(plan, dm.d(), is_field ? 1 : 0, nullptr));
    }
    if (!walls.empty()) {
        DevBuf dw(walls.size());
        if (hipMemcpy(dw.p, walls.data(), walls.size(), hipMemcpyHostToDevice) != hipSuccess)
            throw IOError{"hipMemcpy failed"};
        check(hfm_plan_set_walls(plan, static_cast<const uint8_t *>(dw.p), refineWalls ? 1 : 0, nullptr));
        if (hipDeviceSynchronize() != hipSuccess) throw IOError{"hipDeviceSynchronize failed"};
    }
    const double t1 = wall();
    io->set_scalar("StencilCPUTime", t1 - t0);
    hfm_plan_stats st;
    check(hfm_plan_get_stats(plan, &st));
    // Run_SetupSolver
    io->set_scalar("MaxStencilWidth", (double)st.max_stencil_width);
    // getStencils (HFMInterface.hxx:360-375): the stencils at the given points, printed as the
    // reference prints them (EulerianStencil.hxx:76-87, Macros/ExportArrow.h): for each point
    // {index,{"forward" -> {{diffs}},"symmetric" -> {{diffs}},}}, a difference being
    // {baseWeight,{offset}}; points out of the domain are skipped with a warning.
    if (io->has("getStencils")) {
        const std::vector<double> pts = io->vec("getStencils");
        if (pts.size() % dim) throw IOError{"IO error : getStencils has the wrong size"};
        std::vector<uint32_t> toffs;
        std::vector<double> tw;
        if (bundle) check(hfm_plan_curvature_tables(plan, toffs, tw));
        const int nfwd = model == "ReedsSheppForward2" ? 3 : (bundle ? 27 : 0);
        const int ndiff = bundle ? (int)(tw.size() / dims[2]) : dim * (dim + 1) / 2;
        std::ostringstream oss;
        oss << "{";
        for (size_t q = 0; q + dim <= pts.size(); q += dim) {
            int64_t idx[3] = {0, 0, 0};
            bool ok = true;
            for (int i = 0; i < dim; ++i) {
                const bool th = bundle && i == 2;
                const double step = th ? theta_step(dims[2]) : h[i];
                const double o = th ? -step / 2. : (origin.empty() ? 0. : origin[i]);
                const double a = std::floor((pts[q + i] - o) / step);
                if (!(a > -2e9 && a < 2e9)) { ok = false; break; }
                int64_t c = (int64_t)a;
                if (th) {
                    c %= dims[2];
                    if (c < 0) c += dims[2];
                } else if (c < 0 || c >= dims[i]) {
                    ok = false;
                    break;
                }
                idx[i] = c;
            }
            if (!ok) {
                std::ostringstream w;
                w << "GetStencil error : point {";
                for (int i = 0; i < dim; ++i) w << (i ? "," : "") << pts[q + i];
                w << "}is out of range.\n";
                io->log += w.str();
                continue;
            }
            std::vector<int> off(ndiff * 3, 0);
            std::vector<double> wt(ndiff, 0.);
            if (bundle) {
                for (int k = 0; k < ndiff; ++k) {
                    const uint32_t pk = toffs[(size_t)idx[2] * ndiff + k];
                    off[3 * k] = (int8_t)(pk & 0xff);
                    off[3 * k + 1] = (int8_t)((pk >> 8) & 0xff);
                    off[3 * k + 2] = (int8_t)((pk >> 16) & 0xff);
                    wt[k] = tw[(size_t)idx[2] * ndiff + k];
                }
            } else {
                int64_t lin = 0;
                for (int i = 0; i < dim; ++i) lin = lin * dims[i] + idx[i];
                const int K = ndiff;
                const double *mc = field.data() + (is_field ? lin * K : 0);
                double invhh[6];
                for (int i = 0; i < dim; ++i)
                    for (int j = 0; j <= i; ++j) invhh[i * (i + 1) / 2 + j] = 1. / (h[i] * h[j]);
                int mx = 0;
                if (dim == 2) {
                    int8_t o8[3][2];
                    hfm::riemann_point_stencil<2>(mc, hasDual, invhh, o8, wt.data(), &mx);
                    for (int k = 0; k < 3; ++k) off[3 * k] = o8[k][0], off[3 * k + 1] = o8[k][1];
                } else {
                    int8_t o8[6][3];
                    hfm::riemann_point_stencil<3>(mc, hasDual, invhh, o8, wt.data(), &mx);
                    for (int k = 0; k < 6; ++k)
                        for (int c = 0; c < 3; ++c) off[3 * k + c] = o8[k][c];
                }
            }
            auto diffs = [&](int k0, int k1) {
                oss << "{{";
                for (int k = k0; k < k1; ++k) {
                    oss << (k > k0 ? "," : "") << "{" << wt[k] << ",{";
                    for (int c = 0; c < dim; ++c) oss << (c ? "," : "") << off[3 * k + c];
                    oss << "}}";
                }
                oss << "}}";
            };
            oss << "{{";
            for (int i = 0; i < dim; ++i) oss << (i ? "," : "") << idx[i];
            oss << "},{\"forward\" -> ";
            diffs(0, nfwd);
            oss << ",\"symmetric\" -> ";
            diffs(nfwd, ndiff);
            oss << ",}},";
        }
        oss << "}";
        io->set_string("stencils", oss.str());
    }
    // pointToIndex -> indexFromPoint (IndexFromPoint(ADim(p)), unperiodized) and indexToPoint ->
    // pointFromIndex (ReDim(index + 1/2), the index cast by truncation) (HFMInterface.hxx:377-396)
    auto axis_scale = [&](int i) { return bundle && i == 2 ? theta_step(dims[2]) : h[i]; };
    auto axis_origin = [&](int i) {
        return bundle && i == 2 ? -(double)axis_scale(2) / 2. : (origin.empty() ? 0. : origin[i]);
    };
    if (io->has("pointToIndex")) {
        const std::vector<double> pts = io->vec("pointToIndex");
        if (pts.size() % dim) throw IOError{"IO error : pointToIndex has the wrong size"};
        Field f;
        f.dims = {(int64_t)(pts.size() / dim), (int64_t)dim};
        for (size_t q = 0; q < pts.size(); ++q) {
            const int i = (int)(q % dim);
            f.data.push_back(std::floor((pts[q] - axis_origin(i)) / (double)axis_scale(i)));
        }
        io->set_computed("indexFromPoint", std::move(f));
    }
    if (io->has("indexToPoint")) {
        const std::vector<double> idx = io->vec("indexToPoint");
        if (idx.size() % dim) throw IOError{"IO error : indexToPoint has the wrong size"};
        Field f;
        f.dims = {(int64_t)(idx.size() / dim), (int64_t)dim};
        for (size_t q = 0; q < idx.size(); ++q) {
            const int i = (int)(q % dim);
            const double c = (double)(int)idx[q] + 0.5;
            f.data.push_back(c * (double)axis_scale(i) + axis_origin(i));
        }
        io->set_computed("pointFromIndex", std::move(f));
    }
    // order (HFMInterface.hxx:358, read as an int): 2 / 3 recompute every point at its acceptance
    // with second / third order differences, solved by the reference's serial acceptance loop
    // on the device (hfm_plan_set_order): FP64 plans without walls.
    const int order = (int)io->scalar("order", 1.);
    if (order > 1) {
        if (precision != 64)
            throw IOError{"Error : order " + std::to_string(order) + " needs precision 64 (the serial acceptance "
                          "loop reproduces the reference's FP64 values)"};
        if (!walls.empty())
            throw IOError{"Error : order " + std::to_string(order) + " with walls is not supported by this library"};
        check(hfm_plan_set_order(plan, order));
    }
    // Static factoring (StaticFactoring.hxx:58-113): an imported factor now, a computed one
    // (factoringRadius != 0) once the seeds are known; both at once is an error.  A factored solve
    // runs the serial acceptance loop.
    FactorHost factor;
    const bool facRadiusKey = io->has("factoringRadius");
    const bool facCompute = facRadiusKey && io->scalar("factoringRadius", 10.) != 0.;
    if (io->has("factoringValues") && facCompute)
        throw IOError{"StaticFactoring error : factoringValues (import) and factoringRadius (compute) both "
                      "specified."};
    if (io->has("factoringValues")) {
        double scales[3] = {h[0], h[1], h[2]};
        if (bundle) scales[2] = theta_step(dims[2]);
        factor = import_factor(io, dims, scales, dim);
    }
    // Stopping criteria (ExtraAlgorithms/CommonStoppingCriteria.h:115-165): the targets' indices are
    // IndexFromPoint(ADim(p)), periodized, an out-of-range one dropped with a warning; an
    // unoriented target is its (x, y) at any theta.  They run the serial acceptance loop.
    bool stopping = false;
    {
        const bool hasAny = io->has("stopWhenAnyAccepted"), hasAll = io->has("stopWhenAllAccepted");
        const bool hasAnyU = bundle && io->has("stopWhenAnyAccepted_Unoriented");
        const bool hasAllU = bundle && io->has("stopWhenAllAccepted_Unoriented");
        const bool hasMax = io->has("nMaxAccepted"), hasDist = io->has("stopAtDistance");
        if (hasAny || hasAll || hasAnyU || hasAllU || hasMax || hasDist) {
            if (precision != 64)
                throw IOError{"Error : stopping criteria need precision 64 (the serial acceptance loop reproduces "
                              "the reference's FP64 values)"};
            if (!walls.empty()) throw IOError{"Error : stopping criteria with walls are not supported by this library"};
            auto targets = [&](const char *key, bool unoriented) {
                std::vector<int64_t> out;
                const int pd = unoriented ? 2 : dim;
                const std::vector<double> pts = io->vec(key);
                if (pts.size() % pd) throw IOError{std::string("IO error : ") + key + " has the wrong size"};
                for (size_t q = 0; q < pts.size(); q += pd) {
                    int64_t idx[3] = {0, 0, 0};
                    bool ok = true;
                    for (int i = 0; i < pd; ++i) {
                        const double f = std::floor((pts[q + i] - axis_origin(i)) / (double)axis_scale(i));
                        if (!(f > -2e9 && f < 2e9)) { ok = false; break; }
                        int64_t v = (int64_t)f;
                        if (bundle && i == 2) v = ((v % dims[2]) + dims[2]) % dims[2];
                        else if (v < 0 || v >= dims[i]) ok = false;
                        idx[i] = v;
                    }
                    if (!ok) {
                        // ExtraAlgorithms/CommonStoppingCriteria.h:110,130: the point as the reference prints it, no newline
                        std::ostringstream os;
                        os << (unoriented ? "Warning : unoriented target \"p\" -> {" : "Warning : target {");
                        for (int i = 0; i < pd; ++i) os << (i ? "," : "") << pts[q + i];
                        os << (unoriented ? "}, is out of range." : "} is out of range.");
                        io->log += os.str();
                        continue;
                    }
                    out.push_back(unoriented ? idx[0] * dims[1] + idx[1]
                                             : (dim == 2 ? idx[0] * dims[1] + idx[1]
                                                         : (idx[0] * dims[1] + idx[1]) * dims[2] + idx[2]));
                }
                return out;
            };
            const std::vector<int64_t> any = hasAny ? targets("stopWhenAnyAccepted", false) : std::vector<int64_t>();
            const std::vector<int64_t> all = hasAll ? targets("stopWhenAllAccepted", false) : std::vector<int64_t>();
            const std::vector<int64_t> anyU =
                hasAnyU ? targets("stopWhenAnyAccepted_Unoriented", true) : std::vector<int64_t>();
            const std::vector<int64_t> allU =
                hasAllU ? targets("stopWhenAllAccepted_Unoriented", true) : std::vector<int64_t>();
            const int64_t nmax = hasMax ? (int64_t)(size_t)io->scalar("nMaxAccepted") : -1;
            const double dist = hasDist ? io->scalar("stopAtDistance") : std::numeric_limits<double>::infinity();
            check(hfm_plan_set_stopping(plan, nmax, dist, (int)any.size(), any.data(), (int)all.size(), all.data(),
                                        (int)anyU.size(), anyU.data(), (int)allU.size(), allU.data()));
            stopping = true;
        }
    }
    // Seeds (StaticFactoring::SetSeeds, StaticFactoring.hxx:242-393): `seeds` / `seedValues`,
    // optionally spread over the grid points within `seedRadius` (Riemann models), then for the
    // curvature models `seeds_Unoriented` / `seedValues_Unoriented` at every theta.
    const double thetaStep = bundle ? theta_step(dims[2]) : 0.;
    auto redim = [&](double q, int i) {
        const double o = bundle && i == 2 ? -thetaStep / 2. : (origin.empty() ? 0. : origin[i]);
        return q * (bundle && i == 2 ? thetaStep : h[i]) + o;
    };
    std::vector<double> seeds, sv;   // physical points and values handed to the solver
    int nseeds_in = 0;               // number of `seeds` as given
    std::vector<double> seeds_given; // `seeds` as passed
    int nseeds = 0;
    // Run_RunSolver (HFMInterface.hxx:422-450): given both values and activeNeighs, the solver is
    // bypassed and everything downstream (geodesics, exports, differentiation) uses them.
    const bool bypass = io->has("values") && io->has("activeNeighs");
    if (bypass && (factor.on || facCompute))
        throw IOError{"Error : static factoring with cached values and activeNeighs is not supported by this library"};
    if (bypass) {
        bool vf = false, af = false;
        const std::vector<double> vals = grid_field(io, "values", dims, 1, false, vf);
        const std::vector<double> act = grid_field(io, "activeNeighs", dims, 1, false, af);
        if (!vf || !af) throw IOError{"Error : values and activeNeighs must be fields over the grid"};
        DevBuf dv(vals.size() * sizeof(double)), da(act.size() * sizeof(double));
        if (hipMemcpy(dv.p, vals.data(), vals.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(da.p, act.data(), act.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            throw IOError{"hipMemcpy failed"};
        check(hfm_plan_load_values(plan, dv.d(), da.d(), nullptr));
        if (io->verbosity >= 1) io->log += "Bypassing fast marching solver based on cached data.\n";
        io->set_scalar("nAccepted", 0.);  // nothing accepted: the solver did not run
    } else {
        if (io->has("seeds")) {
            seeds = io->vec("seeds");
            if (seeds.size() % dim) throw IOError{"IO error : seeds has the wrong size"};
            nseeds_in = (int)(seeds.size() / dim);
            seeds_given = seeds;
            if (io->has("seedValues")) {
                sv = io->vec("seedValues");
                if ((int)sv.size() != nseeds_in) throw IOError{"Error : Inconsistent size of seedValues."};
            } else {
                sv.assign(nseeds_in, 0.);
            }
            // SetSeeds (StaticFactoring.hxx:247-262,285): a computed factor needs exactly one seed;
            // seeds spread over seedRadius = 2 by default once factoring is on
            if (facRadiusKey && nseeds_in != 1)
                throw IOError{"Exactly one seed point must be specified for computing the factor, found " +
                              std::to_string(nseeds_in) + "."};
            const bool facDisabled = !factor.on && !facRadiusKey;
            const double seedRadius = io->scalar("seedRadius", facDisabled ? 0. : 2.);
            if (seedRadius != 0.) {
                if (!riemann) throw IOError{"Equation factoring error : no guess"};
                // negative: also one step of the reversed stencil (StaticFactoring.hxx:285-287)
                const int step = seedRadius < 0. ? std::max(1, (int)st.max_stencil_width) : 0;
                const double r = std::abs(seedRadius);
                if (dim == 2)
                    spread_seeds_d<2>(io, dims, h, origin, field.data(), is_field, hasDual, r, step, seeds, sv);
                else
                    spread_seeds_d<3>(io, dims, h, origin, field.data(), is_field, hasDual, r, step, seeds, sv);
            }
        }
        if (bundle && io->has("seeds_Unoriented")) {
            const std::vector<double> up = io->vec("seeds_Unoriented");
            if (up.size() % 2) throw IOError{"IO error : seeds_Unoriented has the wrong size"};
            const size_t nu = up.size() / 2;
            if (facRadiusKey && nu > 0)
                throw IOError{"Factor computation incompatible with unoriented seed points"};
            std::vector<double> uv(nu, 0.);
            if (io->has("seedValues_Unoriented")) uv = io->vec("seedValues_Unoriented");
            if (uv.size() != nu) throw IOError{"Error : Inconsistent size of seedValues_Unoriented."};
            // EquivalentPoints (HFMInterface.hxx:67-84): the base point at every theta index
            for (size_t i = 0; i < nu; ++i)
                for (int64_t t = 0; t < dims[2]; ++t) {
                    seeds.push_back(up[2 * i]);
                    seeds.push_back(up[2 * i + 1]);
                    seeds.push_back(redim((double)t + 0.5, 2));
                    sv.push_back(uv[i]);
                }
        }
        if (seeds.empty()) {
            io->log += "No seeds to run fast marching solver, and missing values or activeNeighs to compute geodesics.\n";
            return;
        }
        nseeds = (int)(seeds.size() / dim);
        // ComputeFactor (StaticFactoring.hxx:116-239) from the seed as given, then the factor on the plan
        if (facCompute) {
            if (!riemann) throw IOError{"Equation factoring error : no guess"};
            const int64_t R = (int64_t)io->scalar("factoringRadius", 10.);
            const std::string choice = io->str("factoringPointChoice", order < 3 ? "Seed" : "Both");
            const bool both = choice == "Both";
            if (!both && choice != "Seed" && choice != "Key")
                throw IOError{"StaticFactoring error : Unsupported factoringPointChoice " + choice + "."};
            if (both)
                io->log += "***** Warning ! *****\nfactoringPointChoice = 'Both' yields a factor that is best computed "
                           "in Python, via dictIn.SetFactor, which uses higher order interpolation methods."
                           "********************\n";
            double p0[3];
            for (int i = 0; i < dim; ++i) p0[i] = (seeds_given[i] - (origin.empty() ? 0. : origin[i])) / h[i];
            factor = dim == 2 ? compute_factor_d<2>(dims, h, field.p, is_field, hasDual, p0, R, both)
                              : compute_factor_d<3>(dims, h, field.p, is_field, hasDual, p0, R, both);
            if (io->scalar("exportFactoring", 0.) != 0. && factor.on) {
                set_grid_array(io, "factoringValues", factor.vals, factor.bd, 1);
                set_grid_array(io, "factoringGradients", factor.grads, factor.bd, dim);
                if (factor.sub) {
                    Field f;
                    f.dims = {(int64_t)dim};
                    for (int i = 0; i < dim; ++i) f.data.push_back((double)factor.shift[i]);
                    io->set_computed("factoringIndexShift", std::move(f));
                }
            }
        }
        if (factor.on) {
            if (precision != 64)
                throw IOError{"Error : static factoring needs precision 64 (the serial acceptance loop reproduces "
                              "the reference's FP64 values)"};
            if (!walls.empty()) throw IOError{"Error : static factoring with walls is not supported by this library"};
            check(hfm_plan_set_factoring(plan, factor.bd.data(), factor.sub ? factor.shift.data() : nullptr,
                                         factor.vals.data(), factor.grads.data(),
                                         factor.mask.empty() ? nullptr : factor.mask.data()));
        }
        // Run_RunSolver
        const double t2 = wall();
        check(hfm_plan_solve(plan, nseeds, seeds.data(), sv.data(), nullptr));
        const double t3 = wall();
        check(hfm_plan_get_stats(plan, &st));
        io->set_scalar("FMCPUTime", t3 - t2);
        io->set_scalar("FMGPUTime", st.solve_ms * 1e-3);
        io->set_scalar("nIterations", (double)st.iterations);
        io->set_scalar("nTileVisits", (double)st.tile_visits);
        if (io->verbosity >= 1) {
            std::ostringstream os;
            os << "Fast marching solver completed in " << (t3 - t2) << " s.\n";
            io->log += os.str();
        }
        // nAccepted (ExtraAlgorithms/CommonStoppingCriteria.h:162-165): the serial loop's count when it ran (the
        // pop that triggered a stop included); otherwise every point the heap accepts is a reached
        // one, i.e. a finite value
        if (order > 1 || stopping || factor.on) {
            io->set_scalar("nAccepted", (double)st.n_accepted);
        } else {
            int64_t reached = 0;
            check(hfm::plan_count_reached(plan, &reached, nullptr));
            io->set_scalar("nAccepted", (double)reached);
        }
    }
    // Run_ExtractGeodesics (HFMInterface.hxx:452-523): tips -> geodesicPoints, geodesicLengths;
    // for the curvature models tips_Unoriented -> geodesicPoints_Unoriented, geodesicLengths_Unoriented
    const double tg = wall();
    bool anyGeodesic = false;
    auto export_geodesics = [&](const std::string &suffix, const std::vector<double> &tips, bool tips_adim) {
        // HFMInterface.hxx:459-464: 'Discrete' (the default) or 'ODE'
        const std::string solver = io->str("geodesicSolver", "Discrete");
        if (solver != "Discrete" && solver != "ODE")
            throw IOError{"Error : Unrecognized geodesic solver " + solver + ".\n"};
        const bool ode = solver == "ODE";
        const int ntips = (int)(tips.size() / dim);
        const double step = io->scalar("geodesicStep", 0.25);
        double thr = 0., vb = 0., causality = 4.;
        if (ode) {
            // GeodesicODESolver.hxx:17-23; the target tolerance is read but the solver keeps the
            // bound 6 its constructor gave the seeds' distance
            causality = io->scalar("geodesicCausalityTolerance", 4.);
            (void)io->scalar("geodesicTargetTolerance", 6.);
        } else {
            thr = io->scalar("geodesicWeightThreshold", 0.001);
            vb = io->scalar("geodesicVolumeBound", 5. * std::pow(1.3, dim));
        }
        std::vector<int64_t> lens(ntips);
        std::vector<double> pts((size_t)4096 * dim);
        std::vector<int> tipFailed(ntips, 0);
        int64_t npts = 0;
        int nfail = 0;
        check(hfm_plan_geodesics_impl(plan, ntips, tips.data(), tips_adim, step, thr, vb, pts.data(), 4096,
                                      lens.data(), &npts, &nfail, nullptr, ode ? 1 : 0, causality,
                                      tipFailed.data()));
        if (npts > 4096) {
            pts.resize((size_t)npts * dim);
            check(hfm_plan_geodesics_impl(plan, ntips, tips.data(), tips_adim, step, thr, vb, pts.data(), npts,
                                          lens.data(), &npts, &nfail, nullptr, ode ? 1 : 0, causality,
                                          tipFailed.data()));
        }
        if (ode) {
            // GeodesicODESolver.hxx:32-36: one message per failed tip, the tip in physical units
            int64_t first = 0;
            for (int t = 0; t < ntips; ++t) {
                if (tipFailed[t]) {
                    std::ostringstream os;
                    os << "Geodesic ODE solver seems to have failed for tip {";
                    for (int c = 0; c < dim; ++c) os << (c ? "," : "") << pts[(size_t)first * dim + c];
                    os << "}.\n";
                    io->log += os.str();
                }
                first += lens[t];
            }
        }
        pts.resize((size_t)npts * dim);
        if (nfail && !ode && io->verbosity >= 1)
            io->log += "Geodesic extraction failed for " + std::to_string(nfail) + " tip(s)\n";
        const std::string sep = suffix.empty() ? "" : "_";
        Field fp;
        fp.dims = {npts, (int64_t)dim};
        fp.data = std::move(pts);
        io->set_computed("geodesicPoints" + sep + suffix, std::move(fp));
        Field fl;
        fl.dims = {(int64_t)ntips};
        for (int64_t l : lens) fl.data.push_back((double)l);
        io->set_computed("geodesicLengths" + sep + suffix, std::move(fl));
        anyGeodesic = true;
    };
    if (io->has("tips")) {
        const std::vector<double> tips = io->vec("tips");
        if (tips.size() % dim) throw IOError{"IO error : tips has the wrong size"};
        export_geodesics("", tips, false);
    }
    if (bundle && io->has("tips_Unoriented")) {
        // each base point starts at the theta of smallest arrival time (the last one among equal
        // values), in grid units (HFMInterface.hxx:502-521)
        const std::vector<double> up = io->vec("tips_Unoriented");
        if (up.size() % 2) throw IOError{"IO error : tips_Unoriented has the wrong size"};
        DevBuf dvals(N * sizeof(double));
        check(hfm_plan_values(plan, dvals.d(), nullptr));
        std::vector<double> vals = host_array(N);
        if (hipMemcpy(vals.data(), dvals.p, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            throw IOError{"hipMemcpy failed"};
        std::vector<double> tips;
        for (size_t i = 0; i + 1 < up.size(); i += 2) {
            const double ax = (up[i] - (origin.empty() ? 0. : origin[0])) / h[0];
            const double ay = (up[i + 1] - (origin.empty() ? 0. : origin[1])) / h[1];
            double best[3] = {ax, ay, 0.5};
            double vmin = std::numeric_limits<double>::infinity();
            const double fx = std::floor(ax), fy = std::floor(ay);
            const bool inside = fx >= 0 && fx < (double)dims[0] && fy >= 0 && fy < (double)dims[1];
            for (int64_t t = 0; inside && t < dims[2]; ++t) {
                const double v = vals[((int64_t)fx * dims[1] + (int64_t)fy) * dims[2] + t];
                if (vmin < v) continue;
                best[2] = (double)t + 0.5;
                vmin = v;
            }
            tips.insert(tips.end(), best, best + 3);
        }
        export_geodesics("Unoriented", tips, true);
    }
    if (anyGeodesic) io->set_scalar("GeodesicCPUTime", wall() - tg);
    // Run_ExportData
    const bool exportValues = io->scalar("exportValues", 0.) != 0.;
    const bool exportActive = io->scalar("exportActiveNeighs", 0.) != 0.;
    const bool exportFlow = io->scalar("exportGeodesicFlow", 0.) != 0.;
    if (exportValues) {
        DevBuf dvv(N * sizeof(double));
        check(hfm_plan_values(plan, dvv.d(), nullptr));
        std::vector<double> v = host_array(N);
        if (hipMemcpy(v.data(), dvv.p, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            throw IOError{"hipMemcpy failed"};
        set_grid_array(io, "values", std::move(v), dims, 1);
    }
    if (exportActive || exportFlow) {
        DevBuf dfl(exportFlow ? N * dim * sizeof(double) : 8);
        DevBuf dac(exportActive ? N * sizeof(double) : 8);
        check(hfm_plan_flow(plan, exportFlow ? dfl.d() : nullptr, exportActive ? dac.d() : nullptr, nullptr));
        if (exportActive) {
            std::vector<double> a = host_array(N);
            if (hipMemcpy(a.data(), dac.p, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                throw IOError{"hipMemcpy failed"};
            set_grid_array(io, "activeNeighs", std::move(a), dims, 1);
        }
        if (exportFlow) {
            std::vector<double> f = host_array(N * dim);
            if (hipMemcpy(f.data(), dfl.p, N * dim * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                throw IOError{"hipMemcpy failed"};
            set_grid_array(io, "geodesicFlow", std::move(f), dims, dim);
        }
    }
    if (io->scalar("exportActiveOffsets", 0.) != 0.) {
        const int ns = hfm_plan_n_active_slots(plan);
        DevBuf dao(N * ns * sizeof(double));
        check(hfm_plan_active_offsets(plan, dao.d(), nullptr));
        std::vector<double> a = host_array(N * ns);
        if (hipMemcpy(a.data(), dao.p, N * ns * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            throw IOError{"hipMemcpy failed"};
        set_grid_array(io, "activeOffsets", std::move(a), dims, ns);
    }
    // FirstVariation::Finally (backward part)
    if (io->has("inspectSensitivity")) {
        const std::vector<double> pts = io->vec("inspectSensitivity");
        if (pts.size() % dim) throw IOError{"IO error : inspectSensitivity has the wrong size"};
        const size_t npts = pts.size() / dim;
        std::vector<double> lengths;
        if (io->has("inspectSensitivityLengths")) {
            lengths = io->vec("inspectSensitivityLengths");
            double tot = 0;
            for (double l : lengths) tot += l;
            if (tot != (double)npts) throw IOError{"Error : inconsistent inspectSensitivityLengths"};
        } else {
            lengths.assign(npts, 1.);
        }
        std::vector<double> weights;
        if (io->has("inspectSensitivityWeights")) {
            weights = io->vec("inspectSensitivityWeights");
            if (weights.size() != npts) throw IOError{"Error: Inconsistent inspectSensitivityWeights size"};
        } else {
            weights.assign(npts, 1.);
        }
        DevBuf dsens(N * sizeof(double));
        size_t pos = 0;
        for (size_t i = 0; i < lengths.size(); ++i) {
            const int cnt = (int)lengths[i];
            const int maxs = 4096;
            std::vector<int64_t> si(maxs);
            std::vector<double> sw(maxs);
            int ns = 0;
            check(hfm_plan_adjoint(plan, cnt, pts.data() + pos * dim, weights.data() + pos, dsens.d(), si.data(),
                                   sw.data(), maxs, &ns, nullptr));
            pos += cnt;
            std::vector<double> s = host_array(N);
            if (hipMemcpy(s.data(), dsens.p, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                throw IOError{"hipMemcpy failed"};
            set_grid_array(io, "costSensitivity_" + std::to_string(i), std::move(s), dims, 1);
            // seedSensitivity: (ReDim(PointFromIndex(index)), weight) pairs (FirstVariation.h:90-98)
            Field f;
            const int n = std::min(ns, maxs);
            f.dims = {(int64_t)n, (int64_t)(dim + 1)};
            for (int k = 0; k < n; ++k) {
                int64_t r = si[k];
                int64_t idx[3] = {0, 0, 0};
                for (int q = dim - 1; q >= 0; --q) {
                    idx[q] = r % dims[q];
                    r /= dims[q];
                }
                for (int q = 0; q < dim; ++q) {
                    const double pt = idx[q] + 0.5;
                    double o;
                    if (bundle && q == 2) o = pt * (theta_step(dims[2])) +
                                              (-(theta_step(dims[2])) / 2.);
                    else o = pt * h[q] + (origin.empty() ? 0. : origin[q]);
                    f.data.push_back(o);
                }
                f.data.push_back(sw[k]);
            }
            io->set_computed("seedSensitivity_" + std::to_string(i), std::move(f));
        }
    }
    // FirstVariation::Finally (forward part, FirstVariation.h:102-121): costVariation (dims...,
    // nVar; every axis follows arrayOrdering) and seedValueVariation (nVar x nseeds, read
    // RowMajor) -> valueVariation (dims..., nVar)
    if (io->has("costVariation") || io->has("seedValueVariation")) {
        const size_t D = dims.size();
        int64_t nvc = -1, nvs = -1;
        std::vector<double> cv;
        if (io->has("costVariation")) {
            const Field &f = io->raw("costVariation");
            if (f.is_string) throw IOError{"IO error : key costVariation is a string, not a numerical field."};
            if (f.dims.size() != D + 1)
                throw IOError{"IO::GetArray error : Field costVariation has depth " + std::to_string(f.dims.size()) +
                              " instead of expected " + std::to_string(D + 1) + "."};
            std::vector<int64_t> internal;
            cv = reorder(f.data, f.dims, 1, io->ordering, true, internal);
            if (!std::equal(dims.begin(), dims.end(), internal.begin()))
                throw IOError{"Forward variation error : inconsistent input sizes"};
            nvc = internal[D];
        }
        std::vector<double> svv;
        if (io->has("seedValueVariation")) {
            const Field &f = io->raw("seedValueVariation");
            if (f.is_string) throw IOError{"IO error : key seedValueVariation is a string, not a numerical field."};
            if (f.dims.size() != 2)
                throw IOError{"IO::GetArray error : Field seedValueVariation has depth " +
                              std::to_string(f.dims.size()) + " instead of expected 2."};
            if (f.dims[1] != nseeds_in) throw IOError{"Error : Inconsistent size of seedValueVariation"};
            nvs = f.dims[0];
            svv = f.data;
            if (nseeds != nseeds_in)
                svv = seed_variation_for_solver(svv, nvs, nseeds_in, seeds_given, seeds, dim, dims, h, origin, bundle,
                                                thetaStep);
        }
        if (nvc >= 0 && nvs >= 0 && nvc != nvs) throw IOError{"Forward variation error : inconsistent input sizes"};
        const int64_t nvar = std::max(nvc, nvs);
        std::vector<double> vv = host_array((size_t)N * nvar);
        if (nvar > 0) {
            DevBuf dcv(cv.empty() ? 8 : cv.size() * sizeof(double));
            if (!cv.empty() &&
                hipMemcpy(dcv.p, cv.data(), cv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                throw IOError{"hipMemcpy failed"};
            DevBuf dvv(vv.size() * sizeof(double));
            check(hfm_plan_forward_variation(plan, (int)nvar, cv.empty() ? nullptr : dcv.d(), nseeds,
                                             svv.empty() ? nullptr : svv.data(), dvv.d(), nullptr));
            if (hipMemcpy(vv.data(), dvv.p, vv.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                throw IOError{"hipMemcpy failed"};
        }
        std::vector<int64_t> vdims(dims.begin(), dims.end());
        vdims.push_back(nvar);
        set_grid_array(io, "valueVariation", std::move(vv), vdims, 1);
    }
}

void usage_report(hfm_io *io) {
    std::ostringstream unusedUser, defaulted, visited;
    for (const auto &kv : io->fields)
        if (!kv.second.computed && !io->used.count(kv.first)) unusedUser << kv.first << " ";
    for (const auto &k : io->defaulted) defaulted << k << " ";
    for (const auto &k : io->visitedUnset) visited << k << " ";
    if (!unusedUser.str().empty()) io->set_string("unusedFromUser", unusedUser.str());
    if (!defaulted.str().empty()) io->set_string("defaulted", defaulted.str());
    if (!visited.str().empty()) io->set_string("visitedUnset", visited.str());
}

long long copy_out(const std::string &s, char *buf, size_t bufsize) {
    if (buf && bufsize) {
        const size_t n = std::min(bufsize - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return (long long)s.size();
}

}  // namespace

extern "C" {

hfm_io *hfm_io_create(void) { return new hfm_io; }
void hfm_io_destroy(hfm_io *io) { delete io; }

int hfm_io_set_string(hfm_io *io, const char *key, const char *value) {
    if (!io || !key || !value) return 1;
    Field f;
    f.is_string = true;
    f.str = value;
    io->fields[key] = std::move(f);
    return 0;
}

int hfm_io_set_scalar(hfm_io *io, const char *key, double value) {
    if (!io || !key) return 1;
    Field f;
    f.data = {value};
    io->fields[key] = std::move(f);
    return 0;
}

int hfm_io_set_array(hfm_io *io, const char *key, const double *data, int ndim, const int64_t *shape) {
    if (!io || !key || ndim < 0 || (ndim > 0 && !shape)) return 1;
    Field f;
    int64_t n = 1;
    for (int i = 0; i < ndim; ++i) {
        if (shape[i] < 0) {
            io->err = "negative dimension";
            return 1;
        }
        f.dims.push_back(shape[i]);
        n *= shape[i];
    }
    if (n > 0 && !data) return 1;
    f.data.reserve((size_t)n);
    advise_huge(f.data.data(), (size_t)n * sizeof(double));
    f.data.assign(data, data + n);
    io->fields[key] = std::move(f);
    return 0;
}

int hfm_io_set_array_f32(hfm_io *io, const char *key, const float *data, int ndim, const int64_t *shape) {
    if (!io || !key || ndim < 0 || (ndim > 0 && !shape)) return 1;
    Field f;
    int64_t n = 1;
    for (int i = 0; i < ndim; ++i) {
        if (shape[i] < 0) {
            io->err = "negative dimension";
            return 1;
        }
        f.dims.push_back(shape[i]);
        n *= shape[i];
    }
    if (n > 0 && !data) return 1;
    f.data.reserve((size_t)n);
    advise_huge(f.data.data(), (size_t)n * sizeof(double));
    f.data.assign(data, data + n);  // widened to FP64
    io->fields[key] = std::move(f);
    return 0;
}

int hfm_io_has_field(const hfm_io *io, const char *key) { return (io && key && io->has(key)) ? 1 : 0; }

int hfm_io_erase_field(hfm_io *io, const char *key) {
    if (!io || !key) return 1;
    io->fields.erase(key);
    return 0;
}

int hfm_io_field_kind(const hfm_io *io, const char *key) {
    if (!io || !key) return 0;
    auto it = io->fields.find(key);
    if (it == io->fields.end()) return 0;
    return it->second.is_string ? 1 : 2;
}

long long hfm_io_get_string(hfm_io *io, const char *key, char *buf, size_t bufsize) {
    if (!io || !key) return -1;
    auto it = io->fields.find(key);
    if (it == io->fields.end() || !it->second.is_string) {
        io->err = std::string("Field ") + key + " not found or not a string";
        return -1;
    }
    return copy_out(it->second.str, buf, bufsize);
}

int hfm_io_get_scalar(hfm_io *io, const char *key, double *value) {
    if (!io || !key || !value) return 1;
    auto it = io->fields.find(key);
    if (it == io->fields.end() || it->second.is_string || it->second.data.size() != 1) {
        io->err = std::string("Field ") + key + " not found or not a scalar";
        return 1;
    }
    *value = it->second.data[0];
    return 0;
}

int hfm_io_get_array_shape(hfm_io *io, const char *key, int *ndim, int64_t *shape, int max_ndim) {
    if (!io || !key || !ndim) return 1;
    auto it = io->fields.find(key);
    if (it == io->fields.end() || it->second.is_string) {
        io->err = std::string("Field ") + key + " not found or not numerical";
        return 1;
    }
    const auto &d = it->second.dims;
    *ndim = (int)d.size();
    if ((int)d.size() > max_ndim) return 1;
    for (size_t i = 0; i < d.size(); ++i) shape[i] = d[i];
    return 0;
}

int hfm_io_get_array(hfm_io *io, const char *key, double *out, size_t count) {
    if (!io || !key || (!out && count)) return 1;
    auto it = io->fields.find(key);
    if (it == io->fields.end() || it->second.is_string) {
        io->err = std::string("Field ") + key + " not found or not numerical";
        return 1;
    }
    const auto &v = it->second.data;
    if (count != v.size()) {
        io->err = "hfm_io_get_array: size mismatch";
        return 1;
    }
    std::memcpy(out, v.data(), count * sizeof(double));
    return 0;
}

long long hfm_io_computed_keys(hfm_io *io, char *buf, size_t bufsize) {
    if (!io) return -1;
    std::string s;
    for (const auto &k : io->computed_order) {
        if (!s.empty()) s += " ";
        s += k;
    }
    return copy_out(s, buf, bufsize);
}

const char *hfm_io_last_error(const hfm_io *io) { return io ? io->err.c_str() : "null io"; }

const char *hfm_io_messages(const hfm_io *io) { return io ? io->log.c_str() : ""; }

// FileHFM wire format (JMM_CPPLibs/JMM_CPPLibs/Output/FileIO.hxx:7-93): <prefix>_Format.txt lists,
// per key, the key, the number of dims (-1: a string on the next line, 0: a scalar), one dim per
// line and a blank line; <prefix>_Data.dat holds every numerical field's FP64 values in order.
int hfm_io_read_files(hfm_io *io, const char *prefix) {
    if (!io || !prefix) return 1;
    io->err.clear();
    try {
        const std::string pre(prefix);
        std::ifstream ifs(pre + "_Format.txt");
        if (!ifs.is_open())
            throw IOError{"FileIO error : Input format file could not be opened " + pre + "_Format.txt."};
        auto next_line = [&] { ifs.ignore(std::numeric_limits<std::streamsize>::max(), ifs.widen('\n')); };
        struct Pending {
            std::string key;
            Field f;
            size_t begin = 0, size = 0;
        };
        std::vector<Pending> items;
        size_t pos = 0;
        while (!ifs.eof()) {
            std::string key;
            ifs >> key;
            next_line();
            if (key.empty()) {
                if (ifs.eof()) break;
                throw IOError{"FileIO Error empty field name"};
            }
            if (ifs.eof() || ifs.fail()) throw IOError{"FileIO error : invalid input format file."};
            Pending it;
            it.key = key;
            long long ndims = 0;
            ifs >> ndims;
            next_line();
            if (ifs.eof() || ifs.fail() || ndims < -1)
                throw IOError{"FileIO error : invalid dimension value (" + std::to_string(ndims) + ") for key " + key +
                              "in input format file."};
            if (ndims == -1) {
                it.f.is_string = true;
                ifs >> it.f.str;
                next_line();
                if (ifs.eof() || ifs.fail())
                    throw IOError{"FileIO error : invalid string value for key " + key + "in input format file."};
            } else {
                size_t size = 1;
                for (long long i = 0; i < ndims; ++i) {
                    long long d = 0;
                    ifs >> d;
                    next_line();
                    if (ifs.eof() || ifs.fail() || d < 0)
                        throw IOError{"FileIO error : invalid dimension for key " + key + " in input format file."};
                    it.f.dims.push_back(d);
                    size *= (size_t)d;
                }
                it.begin = pos;
                it.size = size;
                pos += size;
            }
            next_line();
            items.push_back(std::move(it));
        }
        std::vector<double> data(pos);
        if (pos > 0) {
            std::ifstream dfs(pre + "_Data.dat", std::ios::binary);
            if (!dfs.is_open()) throw IOError{"FileIO error : input data file could not be opened " + pre + "_Data.dat"};
            dfs.seekg(0, dfs.end);
            if ((size_t)dfs.tellg() != pos * sizeof(double))
                throw IOError{"FileIO error : input data file has incorrect size."};
            dfs.seekg(0, dfs.beg);
            dfs.read(reinterpret_cast<char *>(data.data()), (std::streamsize)(pos * sizeof(double)));
            if (dfs.gcount() != (std::streamsize)(pos * sizeof(double)))
                throw IOError{"FileIO error : could not read all data"};
        }
        for (auto &it : items) {
            if (!it.f.is_string) it.f.data.assign(data.begin() + it.begin, data.begin() + it.begin + it.size);
            io->fields[it.key] = std::move(it.f);
        }
    } catch (const IOError &e) {
        io->err = e.msg;
        return 1;
    }
    return 0;
}

// Output pair of FileIO::~FileIO (FileIO.hxx:96-132): the computed keys only, in key order.
int hfm_io_write_files(hfm_io *io, const char *prefix) {
    if (!io || !prefix) return 1;
    io->err.clear();
    const std::string pre(prefix);
    std::ofstream ofs(pre + "_Format.txt");
    if (!ofs.is_open()) {
        io->err = "FileIO error : output format file could not be opened " + pre + "_Format.txt";
        return 1;
    }
    std::vector<double> out;
    for (const auto &kv : io->fields) {
        const Field &f = kv.second;
        if (!f.computed) continue;
        ofs << kv.first << "\n";
        if (f.is_string) {
            ofs << "-1\n" << f.str << "\n\n";
        } else {
            ofs << f.dims.size() << "\n";
            for (int64_t d : f.dims) ofs << d << "\n";
            ofs << "\n";
            out.insert(out.end(), f.data.begin(), f.data.end());
        }
    }
    ofs.close();
    std::ofstream dfs(pre + "_Data.dat", std::ios::binary);
    if (!dfs.is_open()) {
        io->err = "FileIO error : output data file could not be opened " + pre + "_Data.dat";
        return 1;
    }
    dfs.write(reinterpret_cast<const char *>(out.data()), (std::streamsize)(out.size() * sizeof(double)));
    return dfs.good() ? 0 : 1;
}

int hfm_run(hfm_io *io) {
    if (!io) return 1;
    io->err.clear();
    io->log.clear();
    try {
        run_impl(io);
        usage_report(io);
    } catch (const IOError &e) {
        io->err = e.msg;
        return 1;
    } catch (const std::exception &e) {
        io->err = e.what();
        return 1;
    }
    return 0;
}

}  // extern "C"
