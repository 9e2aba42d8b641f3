// hfm_plan.hip -- host orchestration of the device-resident solver plan (C ABI hfm_plan_*).
//
// A plan owns every device buffer of one grid: per-point stencils (or the shared per-theta
// table and the speed field), the tiled values, the tile dependency graph, the two active-tile
// lists with their flags and counters.  Buffers are allocated once at plan creation (or when
// the dependency graph grows) and reused by every solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hfm_amd.h"
#include "hfm_internal.h"
#include "hfm_kernels.h"
#include "selling.h"

using namespace hfm;

namespace hfm {

thread_local std::string g_last_error;

void set_error(const std::string &m) { g_last_error = m; }

}  // namespace hfm

// speed = 1 / cost in the tiled layout (HFMInterface.hxx:131-150 DataSource_Inverse).
template <typename T>
__global__ void k_speed_from_cost(Geom g, const double *cost, int is_field, T *speed, long long np, RedSlot *red) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    double cv = 0., cc = 0.;
    if (!in_grid(g, c)) {
        speed[p] = T(0);
    } else {
        const long long lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
        cv = cost[is_field ? lin : 0];
        cc = 1.;
        speed[p] = (T)(1. / cv);
    }
    wave_add_tau(cv, cc, red);  // mean cost
}

namespace {

const double kPi = 3.141592653589793238462643383279502884L;

int model_from_name(const std::string &m) {
    if (m == "Riemann2") return MODEL_RIEMANN2;
    if (m == "Riemann3") return MODEL_RIEMANN3;
    if (m == "ReedsSheppForward2") return MODEL_RSF2;
    if (m == "Elastica2" || m == "Elastica2_5") return MODEL_ELASTICA2;
    return -1;
}

int model_ndiff(int model) {
    switch (model) {
        case MODEL_RIEMANN2: return 3;
        case MODEL_RIEMANN3: return 6;
        case MODEL_RSF2: return 4;
        case MODEL_ELASTICA2: return 27;
    }
    return 0;
}

}  // namespace


namespace {

void plan_free(hfm_plan *p) {
    dfree(p->u);
    dfree(p->offs);
    dfree(p->w);
    dfree(p->speed);
    dfree(p->tab_offs);
    dfree(p->tab_w);
    dfree(p->fixedmask);
    dfree(p->dep_cnt);
    dfree(p->dep_start);
    dfree(p->dep_fill);
    dfree(p->deps);
    dfree(p->lists[0]);

    dfree(p->counts);
    dfree(p->dep_src);
    dfree(p->ex_cnt);
    dfree(p->ex_start);
    dfree(p->ex_fill);
    dfree(p->ex_rcount);
    dfree(p->ex_scan_work);
    dfree(p->ex_roffs);
    dfree(p->ex_acc);
    dfree(p->ex_active);
    dfree(p->ex_quad);
    dfree(p->ex_heap);
    dfree(p->ex_out);
    dfree(p->ex_flags);
    dfree(p->ex_flags_u);
    dfree(p->fac_vals);
    dfree(p->fac_grads);
    dfree(p->fac_mask);
    dfree(p->dep_nsrc);
    dfree(p->scan_work);
    dfree(p->keys);
    dfree(p->tmax);
    dfree(p->nextmin);
    dfree(p->bounds);
    dfree(p->red);
    dfree(p->adj_lambda);
    dfree(p->adj_sens);
    dfree(p->geo_tips);
    dfree(p->geo_pts);
    dfree(p->geo_len);
    dfree(p->geo_st);
    dfree(p->ode_dist);
    dfree(p->ode_hkeys);
    dfree(p->ode_hcnt);
    if (p->geo_host) (void)hipHostFree(p->geo_host);
    dfree(p->adj_ecnt);
    dfree(p->adj_eidx);
    dfree(p->adj_var);
    dfree(p->adj_ew);
    dfree(p->adj_indeg);
    dfree(p->adj_counts);
    dfree(p->adj_list[0]);
    dfree(p->adj_list[1]);
    dfree(p->adj_seed_idx);
    dfree(p->adj_seed_w);
    dfree(p->adj_tp);
    dfree(p->adj_tw);
    dfree(p->walls);
    dfree(p->vis);
    dfree(p->wdist[0]);
    dfree(p->wdist[1]);
    dfree(p->fwd_pending);
    dfree(p->fwd_rcount);
    dfree(p->fwd_rstart);
    dfree(p->fwd_scan_work);
    dfree(p->fwd_rev);
    dfree(p->fwd_res);
    dfree(p->fwd_sv);
    dfree(p->fwd_sp);
    dfree(p->status);
    dfree(p->stats);
    dfree(p->seed_p);
    dfree(p->seed_v);
    if (p->h_count) (void)hipHostFree(p->h_count);
    if (p->h_red) (void)hipHostFree(p->h_red);
    for (hipEvent_t &e : p->stencil_ev)
        if (e) (void)hipEventDestroy(e);
    if (p->h_stats) (void)hipHostFree(p->h_stats);
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph_stream) (void)hipStreamDestroy(p->graph_stream);
}

// Build the tile dependency graph from the stencils now on the device.
template <typename T> void build_deps(hfm_plan *p, hipStream_t s) {
    const int nt = p->g.ntiles;
    p->exact_rev_dirty = true;
    HIPCHECK(hipMemsetAsync(p->dep_cnt, 0, sizeof(int) * nt, s));
    HIPCHECK(launch_tile_deps<T>(p->model, p->g, p->sd<T>(), 0, p->dep_cnt, nullptr, nullptr, nullptr, p->dep_src,
                                 p->dep_nsrc, s));
    HIPCHECK(launch_exclusive_scan(p->dep_cnt, p->dep_start, nt, s, p->scan_work));
    int total = 0;
    HIPCHECK(hipMemcpyAsync(p->h_count, p->dep_start + nt, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    total = *p->h_count;
    if (total > p->deps_cap) {
        if (p->workspace)
            throw HipFail{"tile dependency graph of " + std::to_string(total) + " entries exceeds the " +
                          std::to_string(p->deps_cap) + " of the allocation-free workspace"};
        dfree(p->deps);
        p->deps = nullptr;
        p->deps_cap = std::max<long long>(total, 1);
        p->deps = dalloc<int>(p->deps_cap);
    }
    p->ndeps = total;
    HIPCHECK(hipMemsetAsync(p->dep_fill, 0, sizeof(int) * nt, s));
    HIPCHECK(launch_tile_deps<T>(p->model, p->g, p->sd<T>(), 1, p->dep_cnt, p->dep_start, p->dep_fill, p->deps,
                                 p->dep_src, p->dep_nsrc, s));
}

// Voronoi1Vec (Headers/Specializations/CommonTraits.h:118-161) on the host: the shared
// curvature stencils are 64-odd small decompositions, computed once per plan.
template <int Dv> void voronoi1vec(const double *v, double eps, int (*off)[3], double *w) {
    constexpr int K = Dv * (Dv + 1) / 2;
    const double angleTol = 1. / std::sqrt(double(Dv - 1));
    const double weightTol = 1000 * 2.220446049250313080847e-16;
    SymPacked<Dv> relaxed;
    double sqn = 0.;
    for (int i = 0; i < Dv; ++i) sqn += v[i] * v[i];
    for (int i = 0; i < Dv; ++i)
        for (int j = 0; j <= i; ++j) {
            const double exact = v[i] * v[j];
            const double id = (i == j) ? 1. : 0.;
            relaxed.m[i * (i + 1) / 2 + j] = exact * (1. - eps * eps) + (sqn * id) * (eps * eps);
        }
    int o[K][Dv];
    double wr[K];
    if (selling_decompose<Dv>(relaxed, o, wr) != SELLING_OK)
        throw HipFail{"ObtuseSuperbase error : input must be positive definite"};
    const double wTol = sqn * weightTol;
    const double angTol = (1 + angleTol * angleTol) / sqn;
    for (int k = 0; k < K; ++k) {
        double scal = 0., offn = 0.;
        for (int c = 0; c < Dv; ++c) scal += v[c] * double(o[k][c]);
        for (int c = 0; c < Dv; ++c) offn += double(o[k][c]) * double(o[k][c]);
        w[k] = (wr[k] < wTol || (scal * scal) * angTol < offn) ? 0. : wr[k];
        off[k][0] = off[k][1] = off[k][2] = 0;
        for (int c = 0; c < Dv; ++c) off[k][c] = (int)(int8_t)(scal <= 0 ? o[k][c] : -o[k][c]);
    }
}

// Shared stencils, forward differences first then symmetric ones, like the reference layout
// (EulerianStencil.h:23-40): ReedsSheppForward2 (Curvature2.h:88-102) and Elastica2<5>
// (Curvature2.h:178-217, Fejer weights :233-234).
void curvature_table(hfm_plan *p, std::vector<uint32_t> &toff, std::vector<double> &tw, int &maxw) {
    const int nth = p->g.n[2];
    const int nd = p->ndiff;
    toff.assign((size_t)nth * nd, 0);
    tw.assign((size_t)nth * nd, 0.);
    maxw = 0;
    const double gs = p->h[0];
    for (int t = 0; t < nth; ++t) {
        int off[27][3];
        double w[27];
        if (p->model == MODEL_RSF2) {
            const double theta = t * p->dependScale;
            double v[2] = {std::cos(theta), std::sin(theta)};
            const double b = 1. / gs;
            v[0] *= b;
            v[1] *= b;
            voronoi1vec<2>(v, p->eps, off, w);
            off[3][0] = 0;
            off[3][1] = 0;
            off[3][2] = -1;
            const double xs = p->xi * p->dependScale;
            w[3] = 1. / (xs * xs);
        } else {
            const int nFejer = 5;
            const double fejer[5] = {0.167781, 0.525552, 0.613333, 0.525552, 0.167781};
            const double theta = (2. * kPi * t) / nth;
            for (int l = 0; l < nFejer / 2; ++l) {
                const double phi = kPi * (l + 0.5) / nFejer;
                const double v[3] = {std::sin(phi) * std::cos(theta) / gs, std::sin(phi) * std::sin(theta) / gs,
                                     std::cos(phi) / (p->xi * p->dependScale)};
                const int s0 = 2 * l, s1 = 2 * l + 1;
                voronoi1vec<3>(v, p->eps, &off[6 * s0], &w[6 * s0]);
                for (int i = 0; i < 6; ++i) {
                    w[6 * s0 + i] *= fejer[l];
                    w[6 * s1 + i] = w[6 * s0 + i];
                    off[6 * s1 + i][0] = off[6 * s0 + i][0];
                    off[6 * s1 + i][1] = off[6 * s0 + i][1];
                    off[6 * s1 + i][2] = (int)(int8_t)(-off[6 * s0 + i][2]);
                }
            }
            const double v2[2] = {std::cos(theta) / gs, std::sin(theta) / gs};
            const int s = nFejer - 1;
            voronoi1vec<2>(v2, p->eps, &off[6 * s], &w[6 * s]);
            for (int i = 0; i < 3; ++i) w[6 * s + i] *= fejer[nFejer / 2];
        }
        for (int k = 0; k < nd; ++k) {
            toff[(size_t)t * nd + k] = pack_offset(off[k][0], off[k][1], off[k][2]);
            tw[(size_t)t * nd + k] = w[k];
            if (w[k] != 0.) maxw = std::max(maxw, std::abs(off[k][0]) + std::abs(off[k][1]) + std::abs(off[k][2]));
        }
    }
}


// Seed placement (StaticFactoring.hxx:242-393): ADim, floor, out-of-range seeds skipped,
// the first value kept for a repeated index.  Returns tiled indices in row-major index order.
void place_seeds(const hfm_plan *p, int nseeds, const double *seeds, const double *sv,
                 std::vector<long long> &tp, std::vector<double> &vals, std::vector<long long> *inputs = nullptr) {
    std::map<long long, std::pair<long long, double>> bylin;
    if (inputs) inputs->assign(nseeds, -1);
    const int dim = p->g.dim;
    for (int s = 0; s < nseeds; ++s) {
        int c[3] = {0, 0, 0};
        bool ok = true;
        for (int i = 0; i < dim; ++i) {
            double q;
            if (p->shared && i == 2) q = (seeds[s * dim + i] - p->origin[i]) / p->dependScale;
            else q = (seeds[s * dim + i] - p->origin[i]) / p->h[i];
            const double f = std::floor(q);
            if (!(f > -2e9 && f < 2e9)) { ok = false; break; }
            int ci = (int)f;
            if (p->g.periodic >> i & 1) {
                ci %= p->g.n[i];
                if (ci < 0) ci += p->g.n[i];
            } else if (ci < 0 || ci >= p->g.n[i]) {
                ok = false;
                break;
            }
            c[i] = ci;
        }
        if (!ok) continue;
        const long long lin = ((long long)c[0] * p->g.n[1] + c[1]) * p->g.n[2] + c[2];
        const long long tpi = (long long)tile_of(p->g, c) * 64 + local_of(p->g, c);
        if (inputs) (*inputs)[s] = tpi;
        bylin.insert({lin, {tpi, sv ? sv[s] : 0.}});
    }
    tp.clear();
    vals.clear();
    for (const auto &kv : bylin) {
        tp.push_back(kv.second.first);
        vals.push_back(kv.second.second);
    }
}

// Visibility mask of the walls for the current stencils.
void build_visibility(hfm_plan *p, hipStream_t s) {
    if (!p->vis) {
        p->vis = dalloc<uint32_t>(p->np);
        p->wdist[0] = dalloc<int8_t>(p->np);
        p->wdist[1] = dalloc<int8_t>(p->np);
    }
    // maxDist = MaxStencilWidth as the reference's int8 ShortType (WallObstructionTest.h:47)
    const int maxDist = (int)(int8_t)p->max_stencil_width;
    if (p->precision == 32)
        HIPCHECK(launch_wall_visibility<float>(p->model, p->g, p->sd<float>(), p->walls, maxDist, p->walls_refine,
                                               p->wdist[0], p->wdist[1], p->vis, p->np, s));
    else
        HIPCHECK(launch_wall_visibility<double>(p->model, p->g, p->sd<double>(), p->walls, maxDist, p->walls_refine,
                                                p->wdist[0], p->wdist[1], p->vis, p->np, s));
    p->vis_dirty = false;
}

// Reversed offsets of the shared (per-theta) stencils, on the host like the stencils themselves
// (HFM_StencilShare.hxx:132-181): for each theta the stencil of the mid-domain point x at that
// theta is walked, and every nonzero-weight side x + s e lists s e under the theta of x + s e;
// each list is sorted (signed lexicographic order) without repetitions.
void shared_reversed_offsets(hfm_plan *p, std::vector<int> &start, std::vector<int> &count,
                             std::vector<uint32_t> &offs) {
    std::vector<uint32_t> toff;
    std::vector<double> tw;
    int maxw = 0;
    curvature_table(p, toff, tw, maxw);
    const int nth = p->g.n[2], nd = p->ndiff, nf = p->model == MODEL_RSF2 ? 3 : nd;
    std::vector<std::vector<std::array<int, 3>>> lists(nth);
    for (int t = 0; t < nth; ++t) {
        const int c[3] = {p->g.n[0] / 2, p->g.n[1] / 2, t};
        for (int k = 0; k < nd; ++k) {
            if (tw[(size_t)t * nd + k] == 0.) continue;
            int o[3];
            unpack_offset(toff[(size_t)t * nd + k], o[0], o[1], o[2]);
            for (int sg = 1; sg >= (k < nf ? 1 : -1); sg -= 2) {
                int r[3];
                if (!shift_coords(p->g, c, sg * o[0], sg * o[1], sg * o[2], r)) continue;
                lists[r[2]].push_back({sg * o[0], sg * o[1], sg * o[2]});
            }
        }
    }
    start.assign(nth + 1, 0);
    count.assign(nth, 0);
    offs.clear();
    for (int t = 0; t < nth; ++t) {
        auto &l = lists[t];
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        start[t] = (int)offs.size();
        count[t] = (int)l.size();
        for (const auto &e : l) offs.push_back(pack_offset(e[0], e[1], e[2]));
    }
    start[nth] = (int)offs.size();
}

// The serial acceptance loop (hfm_exact.hip) for order 2 / 3: reversed offsets (rebuilt after a
// change of stencils), the solver state, then one launch of k_fm_exact.
void run_exact(hfm_plan *p, hipStream_t s) {
    const std::string what = p->order > 1  ? "order " + std::to_string(p->order)
                             : p->fac.on ? std::string("static factoring")
                                         : std::string("a stopping criterion");
    p->exact_what = p->order > 1 ? "an order " + std::to_string(p->order) + " solve"
                    : p->fac.on  ? std::string("a factored solve")
                                 : std::string("a solve with stopping criteria");
    if (p->precision != 64)
        throw HipFail{what + " needs an FP64 plan (precision 64): the serial acceptance loop reproduces the "
                      "reference's FP64 values"};
    if (p->walls) throw HipFail{what + " with walls is not supported by this library"};
    if (p->workspace) throw HipFail{what + " is not available in the allocation-free workspace"};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, s));
    const long long np = p->np;
    const long long nkeys = p->shared ? p->g.n[2] : np;
    if (!p->ex_acc) {
        p->ex_acc = dalloc<unsigned char>(np);
        p->ex_active = dalloc<unsigned long long>(np);
        p->ex_quad = dalloc<double>(4 * np);
        p->ex_out = dalloc<long long>(4);
        p->ex_rcount = dalloc<int>(nkeys);
        p->ex_start = dalloc<int>(nkeys + 1);
        if (!p->shared) {
            p->ex_cnt = dalloc<int>(np);
            p->ex_fill = dalloc<int>(np);
            p->ex_scan_work = dalloc<int>(scan_work_ints(np));
        }
    }
    if (p->exact_rev_dirty) {
        long long total = 0;
        if (!p->shared) {
            if (np > (1LL << 30)) throw HipFail{"grid too large for the serial acceptance loop"};
            HIPCHECK(hipMemsetAsync(p->ex_cnt, 0, np * sizeof(int), s));
            HIPCHECK(launch_rev_build(p->model, p->g, p->sd<double>(), 0, p->ex_cnt, nullptr, nullptr, nullptr, np, s));
            HIPCHECK(launch_exclusive_scan(p->ex_cnt, p->ex_start, (int)np, s, p->ex_scan_work));
            HIPCHECK(hipMemcpyAsync(p->h_count, p->ex_start + np, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            total = p->h_count[0];
            if (total > p->ex_roffs_cap) {
                dfree(p->ex_roffs);
                p->ex_roffs = nullptr;
                p->ex_roffs_cap = std::max<long long>(total, 1);
                p->ex_roffs = dalloc<uint32_t>(p->ex_roffs_cap);
            }
            HIPCHECK(hipMemsetAsync(p->ex_fill, 0, np * sizeof(int), s));
            HIPCHECK(launch_rev_build(p->model, p->g, p->sd<double>(), 1, p->ex_cnt, p->ex_start, p->ex_fill,
                                      p->ex_roffs, np, s));
            HIPCHECK(launch_rev_sort(p->ex_cnt, p->ex_start, p->ex_roffs, p->ex_rcount, np, s));
            p->ex_pushes = total;
        } else {
            std::vector<int> st, cnt;
            std::vector<uint32_t> offs;
            shared_reversed_offsets(p, st, cnt, offs);
            total = (long long)offs.size();
            if (total > p->ex_roffs_cap) {
                dfree(p->ex_roffs);
                p->ex_roffs = nullptr;
                p->ex_roffs_cap = std::max<long long>(total, 1);
                p->ex_roffs = dalloc<uint32_t>(p->ex_roffs_cap);
            }
            if (total > 0)
                HIPCHECK(hipMemcpyAsync(p->ex_roffs, offs.data(), total * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            HIPCHECK(hipMemcpyAsync(p->ex_start, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice, s));
            HIPCHECK(hipMemcpyAsync(p->ex_rcount, cnt.data(), cnt.size() * sizeof(int), hipMemcpyHostToDevice, s));
            HIPCHECK(hipStreamSynchronize(s));  // the host vectors go out of scope
            long long per_layer = 0;
            for (int c : cnt) per_layer += c;
            p->ex_pushes = per_layer * p->g.n[0] * p->g.n[1];
        }
        p->exact_rev_dirty = false;
    }
    // every update pushes once at most: the seeds plus every point's reversed offsets bound the heap
    const int nseeds = (int)p->last_seeds.size();
    const long long cap = p->ex_pushes + nseeds + 1;
    if (cap > p->ex_heap_cap) {
        dfree(p->ex_heap);
        p->ex_heap = nullptr;
        p->ex_heap_cap = cap;
        p->ex_heap = dalloc<char>((size_t)cap * 16);
    }
    double *u = static_cast<double *>(p->u);
    HIPCHECK(launch_exact_init(u, p->ex_acc, p->ex_active, p->ex_quad, np, s));
    ExactArgs a;
    a.g = p->g;
    a.sd = p->sd<double>();
    a.u = u;
    a.accepted = p->ex_acc;
    a.active = p->ex_active;
    a.quad = p->ex_quad;
    a.rstart = p->ex_start;
    a.rcount = p->ex_rcount;
    a.roffs = p->ex_roffs;
    a.heap = p->ex_heap;
    a.heap_cap = p->ex_heap_cap;
    a.seed_p = p->seed_p;
    a.seed_v = p->seed_v;
    a.nseeds = nseeds;
    a.order = p->order;
    a.out = p->ex_out;
    // stopping criteria: target flags per tiled point / per (x, y), uploaded for this solve
    a.max_accepted = p->stop_max >= 0 ? p->stop_max : (long long)(~0ull >> 1);
    a.stop_distance = p->has_stopping ? p->stop_dist : __builtin_huge_val();
    a.stop_flags = nullptr;
    a.stop_flags_u = nullptr;
    a.all_left = 0;
    a.all_left_u = 0;
    a.fac = p->fac;
    p->ex_order = p->order;
    p->ex_fac = p->fac;
    std::vector<unsigned char> fl, flu;
    if (!p->stop_any.empty() || !p->stop_all.empty()) {
        fl.assign(np, 0);
        for (int pass = 0; pass < 2; ++pass)
            for (long long lin : pass == 0 ? p->stop_any : p->stop_all) {
                int c[3] = {0, 0, 0};
                long long r = lin;
                for (int i = p->g.dim - 1; i >= 0; --i) {
                    c[i] = (int)(r % p->g.n[i]);
                    r /= p->g.n[i];
                }
                unsigned char &f = fl[(long long)tile_of(p->g, c) * 64 + local_of(p->g, c)];
                if (pass == 1 && !(f & 2)) ++a.all_left;  // a set: each index once
                f |= pass == 0 ? 1 : 2;
            }
        if (!p->ex_flags) p->ex_flags = dalloc<unsigned char>(np);
        HIPCHECK(hipMemcpyAsync(p->ex_flags, fl.data(), np, hipMemcpyHostToDevice, s));
        a.stop_flags = p->ex_flags;
    }
    if (!p->stop_any_u.empty() || !p->stop_all_u.empty()) {
        const long long nxy = (long long)p->g.n[0] * p->g.n[1];
        flu.assign(nxy, 0);
        for (long long xy : p->stop_any_u) flu[xy] |= 1;
        for (long long xy : p->stop_all_u) {
            if (!(flu[xy] & 2)) ++a.all_left_u;
            flu[xy] |= 2;
        }
        if (!p->ex_flags_u) p->ex_flags_u = dalloc<unsigned char>(nxy);
        HIPCHECK(hipMemcpyAsync(p->ex_flags_u, flu.data(), nxy, hipMemcpyHostToDevice, s));
        a.stop_flags_u = p->ex_flags_u;
    }
    HIPCHECK(launch_fm_exact(p->model, a, s));
    HIPCHECK(hipEventRecord(e1, s));
    long long out[4] = {0, 0, 0, 0};
    HIPCHECK(hipMemcpyAsync(out, p->ex_out, sizeof(out), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (out[2] != 0) throw HipFail{"serial acceptance loop: heap capacity exceeded"};
    p->st.solve_ms = ms;
    p->st.iterations = out[0];  // points accepted
    p->st.n_accepted = out[3];  // pops seen by the stopping criteria (the reference's nAccepted)
    p->st.tile_visits = 0;
    p->st.sweeps = 0;
    p->st.fim_ms = 0.;
    p->st.fim_launches = 0;
}

template <typename T> void run_fim(hfm_plan *p, hipStream_t s) {
    // solve time from events, except in the allocation-free workspace (creates no event)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!p->workspace) {
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        HIPCHECK(hipEventRecord(e0, s));
    }
    const int nseeds = (int)p->last_seeds.size();
    HIPCHECK(launch_init<T>(p->g, static_cast<T *>(p->u), p->fixedmask, p->keys, p->tmax, p->np, p->seed_p,
                            p->seed_v, nseeds, s, 2));
    HIPCHECK(hipMemsetAsync(p->counts, 0, 4 * sizeof(int), s));
    HIPCHECK(hipMemsetAsync(p->stats, 0, STAT_WORDS * sizeof(unsigned long long), s));
    HIPCHECK(hipMemsetAsync(p->nextmin, 0xff, 4 * PEND_WORDS * sizeof(unsigned), s));
    const float delta0 = (float)(p->delta_factor > 0 ? p->delta_factor * p->tau : 0.);
    const float binit[5] = {0.f, -__builtin_huge_valf(), delta0, delta0, delta0};
    HIPCHECK(hipMemcpyAsync(p->bounds, binit, sizeof(binit), hipMemcpyHostToDevice, s));
    // Default round target: 1024 tiles for 2-D grids and 3-D grids of fewer than 4096 tiles,
    // whose rounds fit the update grid's waves, so wider bands cost no latency and save rounds;
    // 64 for larger 3-D grids, whose rounds hold several tiles per wave and where every redundant
    // visit lengthens a wave's chain (Riemann3 128^3: 2.26 ms at 64, 2.29 at 128; 64^3: 0.90 ms at
    // 64, 0.99 at 1024).
    const int target = p->round_target > 0 ? p->round_target : (p->g.dim == 3 && p->g.ntiles >= 4096) ? 64 : 1024;
    HIPCHECK(launch_seed_activate(p->seed_p, p->seed_v, nseeds, p->dep_start, p->deps, p->keys, p->tmax,
                                  &p->nextmin[0], s));

    FimArgs<T> a;
    std::memset(&a, 0, sizeof(a));  // the bytes of `a` key the captured graph
    a.g = p->g;
    a.sd = p->sd<T>();
    a.u = static_cast<T *>(p->u);
    a.fixedmask = p->fixedmask;
    a.dep_start = p->dep_start;
    a.deps = p->deps;
    a.keys = p->keys;
    a.tmax = p->tmax;
    a.list = p->lists[0];
    a.stats = p->stats;
    // 960 blocks of 4 waves (3.75 waves per SIMD; the kernel's 81 VGPRs allow 5): Riemann3 128^3
    // 2.33 / 2.26 / 2.27 ms at 640 / 960 / 1280 blocks, ReedsSheppForward2 256 x 256 x 64 2.91 /
    // 2.59 / 2.50 ms at 384 / 640 / 960; 1280 for 2-D grids scheduled by the listing kernel
    // (Riemann2 1024^2: 4.54 / 4.25 / 4.14 ms at 640 / 960 / 1280).  More waves used to be slower
    // while every wave ended with its own atomics on two shared cache lines.
    const bool fused = p->fused_schedule > 0 || (p->fused_schedule < 0 && p->g.dim == 2 && p->g.ntiles <= 4096);
    const int grid_default = (p->g.dim == 2 && !fused) ? 1280 : 960;
    a.grid = (unsigned)std::min<long long>((p->g.ntiles + 3) / 4, p->fim_grid > 0 ? p->fim_grid : grid_default);
    // At most 4 sweeps per visit on Riemann3 and Elastica2 (128^3: 2.73 against 2.79 ms at 8;
    // Elastica2 192 x 192 x 64: 23.9 against 28.1 ms; a tile still changing after 4 sweeps goes
    // back to the schedule); 8 on Riemann2 (256^2 FP32: 0.90 ms at 8 against 1.07 at 4, the
    // self-listing rounds needing more rounds at 4) and ReedsSheppForward2 (3.08 against 3.21 ms).
    a.max_sweeps = p->max_sweeps > 0                                           ? p->max_sweeps
                   : (p->model == MODEL_RIEMANN3 || p->model == MODEL_ELASTICA2) ? 4
                                                                                : 8;
    // Rounds that list their own tiles save the listing kernel and a kernel boundary per round,
    // but each block then visits only the tiles it owns, and a mark is taken one round later than
    // with the listing kernel (the keys are double-buffered).  That pays on small 2-D grids
    // (Riemann2 256^2 FP32 0.98 -> 0.90 ms, FP64 1.63 -> 1.52 ms) and not on larger ones, whose
    // rounds list a thousand tiles or more and leave the blocks' shares uneven (Riemann2 1024^2
    // 5.40 -> 5.70 ms, Riemann3 128^3 2.79 -> 3.67 ms).
    a.fused = fused ? 1 : 0;
    std::vector<hipEvent_t> tev;
    long long k = 0;
    const int batch = 8;  // even: every batch starts on an even round, so all batches are alike
    const long long max_rounds = 8LL * p->g.ntiles + 65536;
    // One batch of rounds k0 .. k0 + batch - 1 on stream st, then the read-back of the count
    // of its last round (the update kernel cleared only the other slot).
    auto enqueue_batch = [&](long long k0, hipStream_t st, bool timed) {
        for (long long kk = k0; kk < k0 + batch; ++kk) {
            if (a.fused) {
                // round kk lists and visits its own tiles: pend[kk % 4] in, pend[(kk + 1) % 4] and
                // counts[kk % 4] out, pend[(kk + 2) % 4] and counts[(kk + 1) % 4] cleared; it takes
                // from key buffer kk % 2 and marks into buffer (kk + 1) % 2
                a.keys = p->keys + (kk & 1) * (long long)p->g.ntiles;
                a.keys_out = p->keys + ((kk + 1) & 1) * (long long)p->g.ntiles;
                a.sr = SchedRound{&p->nextmin[(kk % 4) * PEND_WORDS],
                                  &p->nextmin[((kk + 1) % 4) * PEND_WORDS],
                                  &p->bounds[(kk + 1) & 1],
                                  &p->bounds[kk & 1],
                                  nullptr,
                                  &p->counts[kk % 4],
                                  &p->bounds[4],
                                  &p->bounds[2 + ((kk + 1) & 1)],
                                  &p->bounds[2 + (kk & 1)],
                                  &p->counts[(kk + 3) % 4],
                                  target};
                a.count_reset = &p->counts[(kk + 1) % 4];
                a.pend_reset = &p->nextmin[((kk + 2) % 4) * PEND_WORDS];
                if (timed) {
                    hipEvent_t ea, eb;
                    HIPCHECK(hipEventCreate(&ea));
                    HIPCHECK(hipEventCreate(&eb));
                    HIPCHECK(launch_fim<T>(p->model, a, st, ea, eb));
                    tev.push_back(ea);
                    tev.push_back(eb);
                } else {
                    HIPCHECK(launch_fim<T>(p->model, a, st));
                }
                continue;
            }
            // round kk: schedule lists tiles into lists[0] / counts[kk&1], update consumes them
            SchedRound sr{&p->nextmin[(kk & 1) * PEND_WORDS],
                          &p->nextmin[((kk + 1) & 1) * PEND_WORDS],
                          &p->bounds[(kk + 1) & 1],
                          &p->bounds[kk & 1],
                          p->lists[0],
                          &p->counts[kk & 1],
                          &p->bounds[4],
                          &p->bounds[2 + ((kk + 1) & 1)],
                          &p->bounds[2 + (kk & 1)],
                          &p->counts[(kk + 1) & 1],
                          target};
            HIPCHECK(launch_schedule(p->g.ntiles, p->keys, sr, st));
            a.count = &p->counts[kk & 1];
            a.count_reset = &p->counts[(kk + 1) & 1];
            a.pend_out = &p->nextmin[((kk + 1) & 1) * PEND_WORDS];
            a.pend_reset = &p->nextmin[(kk & 1) * PEND_WORDS];
            if (timed) {
                hipEvent_t ea, eb;
                HIPCHECK(hipEventCreate(&ea));
                HIPCHECK(hipEventCreate(&eb));
                HIPCHECK(launch_fim<T>(p->model, a, st, ea, eb));
                tev.push_back(ea);
                tev.push_back(eb);
            } else {
                HIPCHECK(launch_fim<T>(p->model, a, st));
            }
        }
        if (a.fused)  // the counts of the batch's last two rounds (counts[2], counts[3]: k0 % 4 == 0)
            HIPCHECK(hipMemcpyAsync(p->h_count, &p->counts[(k0 + batch - 2) % 4], 2 * sizeof(int),
                                    hipMemcpyDeviceToHost, st));
        else
            HIPCHECK(hipMemcpyAsync(p->h_count, &p->counts[(k0 + batch - 1) & 1], sizeof(int), hipMemcpyDeviceToHost,
                                    st));
    };
    const bool graph = p->use_graph && !p->timing && !p->workspace;
    if (graph) {
        // launch arguments of the batch: re-capture when any of them changed
        std::vector<char> key(sizeof(a) + sizeof(target));
        std::memcpy(key.data(), &a, sizeof(a));
        std::memcpy(key.data() + sizeof(a), &target, sizeof(target));
        if (!p->graph_exec || key != p->graph_key) {
            if (p->graph_exec) HIPCHECK(hipGraphExecDestroy(p->graph_exec));
            p->graph_exec = nullptr;
            if (!p->graph_stream) HIPCHECK(hipStreamCreateWithFlags(&p->graph_stream, hipStreamNonBlocking));
            hipGraph_t g;
            HIPCHECK(hipStreamBeginCapture(p->graph_stream, hipStreamCaptureModeThreadLocal));
            enqueue_batch(0, p->graph_stream, false);
            HIPCHECK(hipStreamEndCapture(p->graph_stream, &g));
            const hipError_t ie = hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHECK(ie);
            p->graph_key = key;
        }
    }
    // Batches enqueued before the first look at the count: three quarters of the last solve's
    // rounds on this plan (a plan solved again on a similar field needs about as many; rounds
    // after convergence list nothing and cost a few microseconds each), which saves a host
    // round trip per batch.
    long long ahead = graph ? p->prev_rounds * 3 / 4 / batch : 0;
    while (true) {
        const long long nb = ahead > 1 ? ahead : 1;
        for (long long b = 0; b < nb; ++b) {
            if (graph) HIPCHECK(hipGraphLaunch(p->graph_exec, s));
            else enqueue_batch(k, s, p->timing);
            k += batch;
        }
        ahead = 0;
        HIPCHECK(hipStreamSynchronize(s));
        // converged: the batch's last round listed nothing
        if (p->h_count[a.fused ? 1 : 0] == 0) break;
        if (k > max_rounds) throw HipFail{"block fast-iterative method did not converge"};
    }
    p->prev_rounds = k;
    if (e1) HIPCHECK(hipEventRecord(e1, s));
    HIPCHECK(hipMemcpyAsync(p->h_stats, p->stats, STAT_WORDS * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    float ms = 0;
    if (e0) HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    p->st.solve_ms = ms;
    p->st.iterations = k;
    p->st.tile_visits = 0;
    p->st.sweeps = 0;
    for (int i = 0; i < PEND_SLOTS; ++i) {
        p->st.tile_visits += (int64_t)p->h_stats[i * STAT_STRIDE];
        p->st.sweeps += (int64_t)p->h_stats[i * STAT_STRIDE + 1];
    }
    p->st.tiles_carried = 0;
    p->st.delta = delta0;
    double fim = 0;
    // Each pair brackets one launch; launches whose list was empty still count as launches.
    for (size_t i = 0; i + 1 < tev.size(); i += 2) {
        float d = 0;
        HIPCHECK(hipEventElapsedTime(&d, tev[i], tev[i + 1]));
        fim += d;
        (void)hipEventDestroy(tev[i]);
        (void)hipEventDestroy(tev[i + 1]);
    }
    p->st.fim_ms = p->timing ? fim : 0.;
    p->st.fim_launches = p->timing ? (int64_t)(tev.size() / 2) : 0;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
}

template <typename T> void set_metric_impl(hfm_plan *p, const double *d_metric, int is_field, int dual, hipStream_t s) {
    InvHH ih{};
    const int D = p->g.dim;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j <= i; ++j) ih.v[i * (i + 1) / 2 + j] = 1. / (p->h[i] * p->h[j]);
    p->stencil_ev_pending = false;
    if (!p->workspace) HIPCHECK(hipEventRecord(p->stencil_ev[0], s));
    int init[2] = {0, 0x7fffffff};
    HIPCHECK(hipMemcpyAsync(p->status, init, sizeof(init), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemsetAsync(p->red, 0, RED_SLOTS * sizeof(RedSlot), s));
    HIPCHECK(launch_riemann_stencil<T>(p->g, d_metric, is_field, dual, ih, p->offs, static_cast<T *>(p->w), p->np,
                                       p->status, p->red, s));
    // Status words and reduction slots come back to pinned memory without a synchronisation of
    // their own: the tile graph's synchronisation (its size decides the allocation) covers them.
    // A failed stencil only wastes that graph; the error is raised below.
    HIPCHECK(hipMemcpyAsync(&p->h_count[2], p->status, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(p->h_red, p->red, RED_SLOTS * sizeof(RedSlot), hipMemcpyDeviceToHost, s));
    build_deps<T>(p, s);
    const int hs[2] = {p->h_count[2], p->h_count[3]};
    const RedTotals tot = red_total(p->h_red);
    p->tau = tot.count > 0 ? tot.sum / tot.count : 0.;
    if (hs[0] == SELLING_NOT_SPD)
        throw HipFail{"ObtuseSuperbase error : input must be positive definite (grid point " + std::to_string(hs[1]) +
                      ")"};
    if (hs[0] == SELLING_NO_CONVERGENCE)
        throw HipFail{"ObtuseSuperbase error : did not terminate in 200 iterations (grid point " +
                      std::to_string(hs[1]) + ")"};
    p->max_stencil_width = tot.l1max;
    // the second tile-graph pass is still running: the phase time is read from the events later
    if (!p->workspace) {
        HIPCHECK(hipEventRecord(p->stencil_ev[1], s));
        p->stencil_ev_pending = true;
    }
}

// Model, grid geometry, tiling, grid scales and origin of a plan (no device memory).
void setup_geometry(hfm_plan *p, int m, int precision, const int64_t *dims, const double *gridScales,
                    const double *origin, double xi, double eps) {
    p->model = m;
    p->precision = precision;
    p->shared = (m == MODEL_RSF2 || m == MODEL_ELASTICA2);
    p->ndiff = model_ndiff(m);
    const int dim = (m == MODEL_RIEMANN2) ? 2 : 3;
    Geom &g = p->g;
    g.dim = dim;
    g.n[2] = 1;
    p->N = 1;
    for (int i = 0; i < dim; ++i) {
        if (dims[i] <= 0) throw HipFail{"BaseGrid error : domain size must be positive"};
        if (dims[i] > (1 << 24)) throw HipFail{"grid extent too large"};
        g.n[i] = (int)dims[i];
        p->N *= dims[i];
    }
    if (dim == 2) {
        g.sh[0] = 3; g.sh[1] = 3; g.sh[2] = 0;
    } else {
        g.sh[0] = 2; g.sh[1] = 2; g.sh[2] = 2;
    }
    long long nt = 1;
    for (int i = 0; i < 3; ++i) {
        g.nt[i] = (g.n[i] + (1 << g.sh[i]) - 1) >> g.sh[i];
        nt *= g.nt[i];
    }
    if (nt > (1LL << 26)) throw HipFail{"grid too large"};
    g.ntiles = (int)nt;
    g.periodic = p->shared ? 4 : 0;
    p->np = nt * 64;
    if (p->shared) {
        if (!(gridScales[0] > 0)) throw HipFail{"Parametrization error : gridscale should be positive"};
        p->h[0] = p->h[1] = gridScales[0];
        p->dependScale = 2 * kPi / g.n[2];
        p->h[2] = p->dependScale;
        p->origin[0] = origin ? origin[0] : 0.;
        p->origin[1] = origin ? origin[1] : 0.;
        p->origin[2] = -p->dependScale / 2.;
        p->xi = xi;
        p->eps = eps;
    } else {
        for (int i = 0; i < dim; ++i) {
            if (!(gridScales[i] > 0)) throw HipFail{"Parametrization error : gridscales should be positive"};
            p->h[i] = gridScales[i];
            p->origin[i] = origin ? origin[i] : 0.;
        }
    }
    p->st.n_points = p->N;
    p->st.n_tiles = nt;
}

}  // namespace

extern "C" {

const char *hfm_last_error(void) { return g_last_error.c_str(); }

int hfm_plan_create(hfm_plan **out, const char *model, const int64_t *dims, const double *gridScales,
                    const double *origin, int precision, double xi, double eps) {
    if (!out || !model || !dims || !gridScales) {
        set_error("hfm_plan_create: null argument");
        return 1;
    }
    *out = nullptr;
    const int m = model_from_name(model);
    if (m < 0) {
        set_error(std::string("Unrecognized model ") + model);
        return 1;
    }
    if (precision != 32 && precision != 64) {
        set_error("precision must be 32 or 64");
        return 1;
    }
    hfm_plan *p = new hfm_plan;
    try {
        setup_geometry(p, m, precision, dims, gridScales, origin, xi, eps);
        const Geom &g = p->g;
        const long long nt = g.ntiles;
        const size_t ts = precision == 32 ? 4 : 8;
        p->u = dalloc<char>(p->np * ts);
        if (!p->shared) {
            p->offs = dalloc<uint32_t>(p->np * p->ndiff);
            p->w = dalloc<char>(p->np * p->ndiff * ts);
        } else {
            p->speed = dalloc<char>(p->np * ts);
            p->tab_offs = dalloc<uint32_t>((size_t)g.n[2] * p->ndiff);
            p->tab_w = dalloc<double>((size_t)g.n[2] * p->ndiff);
        }
        p->fixedmask = dalloc<unsigned long long>(nt);
        p->dep_cnt = dalloc<int>(nt);
        p->dep_start = dalloc<int>(nt + 1);
        p->dep_fill = dalloc<int>(nt);
        p->lists[0] = dalloc<int>(nt);

        p->counts = dalloc<int>(4);
        p->dep_src = dalloc<int>((size_t)nt * 64);
        p->dep_nsrc = dalloc<int>(nt);
        p->scan_work = dalloc<int>(scan_work_ints(nt));
        p->keys = dalloc<unsigned>(2 * (size_t)nt);  // two buffers: rounds that list their own tiles alternate them
        p->tmax = dalloc<unsigned>(nt);
        p->nextmin = dalloc<unsigned>(4 * PEND_WORDS);
        p->bounds = dalloc<float>(5);  // 2 band bounds, 2 band widths, initial band width
        p->red = dalloc<RedSlot>(RED_SLOTS);
        HIPCHECK(hipHostMalloc((void **)&p->h_red, sizeof(RedSlot) * RED_SLOTS, hipHostMallocDefault));
        HIPCHECK(hipEventCreate(&p->stencil_ev[0]));
        HIPCHECK(hipEventCreate(&p->stencil_ev[1]));
        p->status = dalloc<int>(4);
        p->stats = dalloc<unsigned long long>(STAT_WORDS);
        HIPCHECK(hipHostMalloc((void **)&p->h_count, sizeof(int) * 4, hipHostMallocDefault));
        HIPCHECK(hipHostMalloc((void **)&p->h_stats, sizeof(unsigned long long) * STAT_WORDS, hipHostMallocDefault));
        p->st.n_points = p->N;
        p->st.n_tiles = nt;
    } catch (const HipFail &e) {
        set_error(e.msg);
        plan_free(p);
        delete p;
        return 1;
    }
    *out = p;
    return 0;
}

void hfm_plan_destroy(hfm_plan *plan) {
    if (!plan) return;
    plan_free(plan);
    delete plan;
}

int hfm_plan_set_metric(hfm_plan *p, const double *d_metric, int is_field, int dual, void *stream) {
    if (!p || !d_metric) {
        set_error("hfm_plan_set_metric: null argument");
        return 1;
    }
    if (p->shared) {
        set_error("hfm_plan_set_metric: model has no metric (use hfm_plan_set_cost)");
        return 1;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        p->stencil_ready = false;
        p->solved = false;
        if (p->precision == 32) set_metric_impl<float>(p, d_metric, is_field, dual, s);
        else set_metric_impl<double>(p, d_metric, is_field, dual, s);
        p->st.max_stencil_width = p->max_stencil_width;
        p->st.n_deps = p->ndeps;
        p->stencil_ready = true;
        p->vis_dirty = true;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

int hfm_plan_set_cost(hfm_plan *p, const double *d_cost, int is_field, void *stream) {
    if (!p || !d_cost) {
        set_error("hfm_plan_set_cost: null argument");
        return 1;
    }
    if (!p->shared) {
        set_error("hfm_plan_set_cost: model has no cost field (use hfm_plan_set_metric)");
        return 1;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        p->stencil_ready = false;
        p->solved = false;
        hipEvent_t e0, e1;
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        HIPCHECK(hipEventRecord(e0, s));
        std::vector<uint32_t> toff;
        std::vector<double> tw;
        int maxw = 0;
        curvature_table(p, toff, tw, maxw);
        HIPCHECK(hipMemcpyAsync(p->tab_offs, toff.data(), toff.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        HIPCHECK(hipMemcpyAsync(p->tab_w, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice, s));
        const unsigned nb = (unsigned)((p->np + 255) / 256);
        HIPCHECK(hipMemsetAsync(p->red, 0, RED_SLOTS * sizeof(RedSlot), s));
        if (p->precision == 32) {
            hipLaunchKernelGGL((k_speed_from_cost<float>), dim3(nb), dim3(256), 0, s, p->g, d_cost, is_field,
                               static_cast<float *>(p->speed), p->np, p->red);
            HIPCHECK(hipGetLastError());
            build_deps<float>(p, s);
        } else {
            hipLaunchKernelGGL((k_speed_from_cost<double>), dim3(nb), dim3(256), 0, s, p->g, d_cost, is_field,
                               static_cast<double *>(p->speed), p->np, p->red);
            HIPCHECK(hipGetLastError());
            build_deps<double>(p, s);
        }
        // one-step time ~ cost / sqrt(mean over theta of sum_k baseWeight_k)
        double mw = 0.;
        for (int t = 0; t < p->g.n[2]; ++t) {
            double sw = 0.;
            for (int k = 0; k < p->ndiff; ++k) sw += tw[(size_t)t * p->ndiff + k];
            mw += sw;
        }
        mw /= p->g.n[2];
        HIPCHECK(hipMemcpyAsync(p->h_red, p->red, RED_SLOTS * sizeof(RedSlot), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        const RedTotals tot = red_total(p->h_red);
        p->tau = (tot.count > 0 && mw > 0) ? (tot.sum / tot.count) / std::sqrt(mw) : 0.;
        HIPCHECK(hipEventRecord(e1, s));
        HIPCHECK(hipStreamSynchronize(s));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        p->st.stencil_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        p->max_stencil_width = maxw;
        p->st.max_stencil_width = maxw;
        p->st.n_deps = p->ndeps;
        p->stencil_ready = true;
        p->vis_dirty = true;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

int hfm_plan_solve(hfm_plan *p, int nseeds, const double *seeds, const double *seedValues, void *stream) {
    if (!p || (nseeds > 0 && !seeds)) {
        set_error("hfm_plan_solve: null argument");
        return 1;
    }
    if (!p->stencil_ready) {
        set_error("hfm_plan_solve: stencils not set (call hfm_plan_set_metric / hfm_plan_set_cost first)");
        return 1;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        std::vector<long long> tp, inputs;
        std::vector<double> vals;
        place_seeds(p, nseeds, seeds, seedValues, tp, vals, &inputs);
        if (tp.empty()) throw HipFail{nseeds > 0 ? "Error : seeds incorrectly set" : "No seeds to run fast marching solver"};
        if ((int)tp.size() > p->seed_cap) {
            if (p->workspace)
                throw HipFail{std::to_string(tp.size()) + " seed points exceed the " + std::to_string(p->seed_cap) +
                              " of the allocation-free workspace"};
            dfree(p->seed_p);
            dfree(p->seed_v);
            p->seed_p = nullptr;
            p->seed_v = nullptr;
            p->seed_cap = std::max<int>((int)tp.size(), 16);
            p->seed_p = dalloc<long long>(p->seed_cap);
            p->seed_v = dalloc<double>(p->seed_cap);
        }
        HIPCHECK(hipMemcpyAsync(p->seed_p, tp.data(), tp.size() * sizeof(long long), hipMemcpyHostToDevice, s));
        HIPCHECK(hipMemcpyAsync(p->seed_v, vals.data(), vals.size() * sizeof(double), hipMemcpyHostToDevice, s));
        p->last_seeds = tp;
        p->seed_inputs = inputs;
        if (p->walls && p->vis_dirty) build_visibility(p, s);
        p->exact_solved = false;
        p->st.n_accepted = -1;
        if (p->order > 1 || p->has_stopping || p->fac.on) {
            run_exact(p, s);
            p->exact_solved = true;
        } else if (p->precision == 32) {
            run_fim<float>(p, s);
        } else {
            run_fim<double>(p, s);
        }
        p->solved = true;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

int hfm_plan_load_values(hfm_plan *p, const double *d_values, const double *d_active, void *stream) {
    if (!p || !d_values || !d_active) {
        set_error("hfm_plan_load_values: null argument");
        return 1;
    }
    if (!p->stencil_ready) {
        set_error("hfm_plan_load_values: stencils not set (call hfm_plan_set_metric / hfm_plan_set_cost first)");
        return 1;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (p->walls && p->vis_dirty) build_visibility(p, s);
        if (p->precision == 32)
            HIPCHECK(launch_load_state<float>(p->g, d_values, d_active, static_cast<float *>(p->u), p->fixedmask,
                                              p->np, s));
        else
            HIPCHECK(launch_load_state<double>(p->g, d_values, d_active, static_cast<double *>(p->u), p->fixedmask,
                                               p->np, s));
        HIPCHECK(hipStreamSynchronize(s));
        p->exact_solved = false;
        p->last_seeds.clear();
        p->seed_inputs.clear();
        p->solved = true;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

int hfm_plan_values(hfm_plan *p, double *d_values, void *stream) {
    if (!p || !d_values) {
        set_error("hfm_plan_values: null argument");
        return 1;
    }
    if (!p->solved) {
        set_error("hfm_plan_values: no solve has run");
        return 1;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = p->precision == 32
                       ? launch_export_values<float>(p->g, static_cast<const float *>(p->u), d_values, p->np, s)
                       : launch_export_values<double>(p->g, static_cast<const double *>(p->u), d_values, p->np, s);
    if (e != hipSuccess) {
        set_error(std::string("hfm_plan_values: ") + hipGetErrorString(e));
        return 1;
    }
    return 0;
}

static int plan_flow(const char *who, hfm_plan *p, double *d_flow, double *d_active, double *d_aoffs, void *stream) {
    if (!p) {
        set_error(std::string(who) + ": null argument");
        return 1;
    }
    if (!p->solved) {
        set_error(std::string(who) + ": no solve has run");
        return 1;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    FlowScale fs{{p->h[0], p->h[1], p->h[2]}};
    if (p->exact_solved) {  // the flags of the serial loop and its order-aware recompute
        const hipError_t e = launch_flow_exact(p->model, p->g, p->sd<double>(), static_cast<const double *>(p->u),
                                               p->ex_acc, p->ex_active, p->ex_order, p->ex_fac, fs, d_flow, d_active,
                                               d_aoffs, p->np, s);
        if (e != hipSuccess) {
            set_error(std::string(who) + ": " + hipGetErrorString(e));
            return 1;
        }
        return 0;
    }
    hipError_t e = p->precision == 32
                       ? launch_flow<float>(p->model, p->g, p->sd<float>(), static_cast<const float *>(p->u),
                                            p->fixedmask, fs, d_flow, d_active, d_aoffs, p->np, s)
                       : launch_flow<double>(p->model, p->g, p->sd<double>(), static_cast<const double *>(p->u),
                                             p->fixedmask, fs, d_flow, d_active, d_aoffs, p->np, s);
    if (e != hipSuccess) {
        set_error(std::string(who) + ": " + hipGetErrorString(e));
        return 1;
    }
    return 0;
}

int hfm_plan_flow(hfm_plan *p, double *d_flow, double *d_active, void *stream) {
    return plan_flow("hfm_plan_flow", p, d_flow, d_active, nullptr, stream);
}

int hfm_plan_active_offsets(hfm_plan *p, double *d_offsets, void *stream) {
    if (!d_offsets) {
        set_error("hfm_plan_active_offsets: null argument");
        return 1;
    }
    return plan_flow("hfm_plan_active_offsets", p, nullptr, nullptr, d_offsets, stream);
}

int hfm_plan_n_active_slots(const hfm_plan *p) { return p ? p->ndiff : 0; }

int hfm_plan_residual(hfm_plan *p, double *max_rel, int64_t *n_mismatch, void *stream) {
    if (!p || !max_rel || !n_mismatch) {
        set_error("hfm_plan_residual: null argument");
        return 1;
    }
    if (!p->solved) {
        set_error("hfm_plan_residual: no solve has run");
        return 1;
    }
    if (p->exact_solved) {
        set_error("hfm_plan_residual: the residual check is the first-order scheme's; " + p->exact_what +
                  " does not satisfy it");
        return 1;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (p->walls && p->vis_dirty) build_visibility(p, s);
        ResSlot *slots = reinterpret_cast<ResSlot *>(p->red);
        HIPCHECK(hipMemsetAsync(slots, 0, RED_SLOTS * sizeof(ResSlot), s));
        if (p->precision == 32)
            HIPCHECK(launch_residual<float>(p->model, p->g, p->sd<float>(), static_cast<const float *>(p->u),
                                            p->fixedmask, p->np, slots, s));
        else
            HIPCHECK(launch_residual<double>(p->model, p->g, p->sd<double>(), static_cast<const double *>(p->u),
                                             p->fixedmask, p->np, slots, s));
        std::vector<ResSlot> h(RED_SLOTS);
        HIPCHECK(hipMemcpyAsync(h.data(), slots, RED_SLOTS * sizeof(ResSlot), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        unsigned long long mb = 0, mm = 0;
        for (const ResSlot &r : h) {
            mb = r.maxbits > mb ? r.maxbits : mb;
            mm += r.mismatch;
        }
        double d;
        std::memcpy(&d, &mb, sizeof(d));
        *max_rel = d;
        *n_mismatch = (int64_t)mm;
    } catch (const HipFail &f) {
        set_error("hfm_plan_residual: " + f.msg);
        return 1;
    }
    return 0;
}

int hfm_plan_geodesics(hfm_plan *p, int ntips, const double *tips, double step, double threshold,
                       double volumeBound, double *points, int64_t max_points, int64_t *lengths, int64_t *npoints,
                       int *nfailed, void *stream) {
    return hfm_plan_geodesics_impl(p, ntips, tips, false, step, threshold, volumeBound, points, max_points, lengths,
                                   npoints, nfailed, stream);
}

int hfm_plan_geodesics_ode(hfm_plan *p, int ntips, const double *tips, double step, double causality_tolerance,
                           double *points, int64_t max_points, int64_t *lengths, int64_t *npoints, int *tip_failed,
                           void *stream) {
    int nf = 0;
    return hfm_plan_geodesics_impl(p, ntips, tips, false, step, 0., 0., points, max_points, lengths, npoints, &nf,
                                   stream, 1, causality_tolerance, tip_failed);
}

}  // extern "C"

int hfm_plan_curvature_tables(const hfm_plan *p, std::vector<uint32_t> &offs, std::vector<double> &w) {
    if (!p || !p->shared || !p->stencil_ready) {
        set_error("hfm_plan_curvature_tables: no curvature stencil tables");
        return 1;
    }
    const size_t n = (size_t)p->g.n[2] * p->ndiff;
    offs.resize(n);
    w.resize(n);
    if (hipMemcpy(offs.data(), p->tab_offs, n * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(w.data(), p->tab_w, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("hfm_plan_curvature_tables: copy failed");
        return 1;
    }
    return 0;
}

int hfm_plan_geodesics_impl(hfm_plan *p, int ntips, const double *tips, bool tips_adim, double step,
                            double threshold, double volumeBound, double *points, int64_t max_points,
                            int64_t *lengths, int64_t *npoints, int *nfailed, void *stream, int solver,
                            double causality, int *tip_failed) {
    if (!p || (ntips > 0 && (!tips || !lengths)) || !npoints) {
        set_error("hfm_plan_geodesics: null argument");
        return 1;
    }
    if (!p->solved) {
        set_error("hfm_plan_geodesics: no solve has run");
        return 1;
    }
    if (solver == 1 && p->walls) {
        set_error("geodesic ODE solver: not available with walls");
        return 1;
    }
    *npoints = 0;
    if (nfailed) *nfailed = 0;
    if (ntips <= 0) return 0;
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        const int dim = p->g.dim;
        // ADim (HFM_ParamDefault.hxx:70-72 and :107-119), the same as the seeds' placement
        std::vector<double> tad((size_t)ntips * dim);
        GeoParamsHost h;
        for (int i = 0; i < dim; ++i) {
            h.scale[i] = (p->shared && i == 2) ? p->dependScale : p->h[i];
            h.origin[i] = p->origin[i];
        }
        for (int t = 0; t < ntips; ++t)
            for (int i = 0; i < dim; ++i) {
                const double x = tips[(size_t)t * dim + i];
                tad[(size_t)t * dim + i] = tips_adim ? x : (x - h.origin[i]) / h.scale[i];
            }
        h.step = step;
        h.threshold = threshold;
        h.volumeBound = volumeBound > 0 ? volumeBound : 5. * std::pow(1.3, dim);
        h.max_iter = 4 * p->np + 4096;
        long long cap = 1024;
        for (int i = 0; i < dim; ++i) cap += 8LL * p->g.n[i];
        if (ntips > p->geo_tip_cap) {
            dfree(p->geo_tips);
            dfree(p->geo_len);
            dfree(p->geo_st);
            p->geo_tips = nullptr;
            p->geo_len = nullptr;
            p->geo_st = nullptr;
            p->geo_tip_cap = 0;
            p->geo_tips = dalloc<double>((size_t)ntips * dim);
            p->geo_len = dalloc<long long>(ntips);
            p->geo_st = dalloc<int>(ntips);
            p->geo_tip_cap = ntips;
        }
        // pinned read-back layout: lengths, status, then the points of every tip (cap per tip)
        auto ensure_host = [&](size_t bytes) {
            if (bytes <= p->geo_host_cap) return;
            if (p->geo_host) HIPCHECK(hipHostFree(p->geo_host));
            p->geo_host = nullptr;
            p->geo_host_cap = 0;
            HIPCHECK(hipHostMalloc((void **)&p->geo_host, bytes, hipHostMallocDefault));
            p->geo_host_cap = bytes;
        };
        const size_t head = (((size_t)ntips * (sizeof(long long) + sizeof(int)) + 255) / 256) * 256;
        ensure_host(std::max(head, tad.size() * sizeof(double)));
        std::memcpy(p->geo_host, tad.data(), tad.size() * sizeof(double));
        HIPCHECK(hipMemcpyAsync(p->geo_tips, p->geo_host, tad.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHECK(hipStreamSynchronize(s));  // the pinned buffer is reused for the read-back below
        std::vector<long long> hlen(ntips);
        std::vector<int> hst(ntips);
        for (int attempt = 0;; ++attempt) {
            h.max_points = cap;
            const long long need = (long long)ntips * cap * dim;
            if (need > p->geo_pts_cap) {
                dfree(p->geo_pts);
                p->geo_pts = nullptr;
                p->geo_pts_cap = 0;
                p->geo_pts = dalloc<double>((size_t)need);
                p->geo_pts_cap = need;
            }
            // after a serial solve the flows are its recompute over its accepted set (FP64 plans only)
            ExactFlow xf;
            if (p->exact_solved) {
                xf.acc = p->ex_acc;
                xf.active = p->ex_active;
                xf.order = p->ex_order;
                xf.fac = p->ex_fac;
                xf.on = 1;
            }
            hipError_t e;
            if (solver == 1) {
                // the ODE solver: the seeds' L-infinity distance, then per-tip read-count tables
                if (!p->ode_dist) p->ode_dist = dalloc<int>(p->np);
                HIPCHECK(launch_ode_distance(p->model, p->g, p->ode_dist, p->seed_p, (int)p->last_seeds.size(),
                                             p->np, s));
                int hlog2 = 15;
                while (hlog2 > 10 && (double)ntips * (double)(1LL << hlog2) * 12. > 512. * 1024. * 1024.) --hlog2;
                const long long slots = (long long)ntips << hlog2;
                if (slots > p->ode_hcap) {
                    dfree(p->ode_hkeys);
                    dfree(p->ode_hcnt);
                    p->ode_hkeys = nullptr;
                    p->ode_hcnt = nullptr;
                    p->ode_hcap = 0;
                    p->ode_hkeys = dalloc<unsigned long long>(slots);
                    p->ode_hcnt = dalloc<int>(slots);
                    p->ode_hcap = slots;
                }
                HIPCHECK(hipMemsetAsync(p->ode_hkeys, 0, slots * sizeof(unsigned long long), s));
                HIPCHECK(hipMemsetAsync(p->ode_hcnt, 0, slots * sizeof(int), s));
                p->ode_hlog2 = hlog2;
                e = p->precision == 32
                        ? launch_geodesics_ode<float>(p->model, p->g, p->sd<float>(), static_cast<const float *>(p->u),
                                                      p->fixedmask, xf, h, causality, p->ode_dist, p->ode_hkeys,
                                                      p->ode_hcnt, hlog2, ntips, p->geo_tips, p->geo_pts, p->geo_len,
                                                      p->geo_st, s)
                        : launch_geodesics_ode<double>(p->model, p->g, p->sd<double>(),
                                                       static_cast<const double *>(p->u), p->fixedmask, xf, h,
                                                       causality, p->ode_dist, p->ode_hkeys, p->ode_hcnt, hlog2,
                                                       ntips, p->geo_tips, p->geo_pts, p->geo_len, p->geo_st, s);
            } else {
                e = p->precision == 32
                        ? launch_geodesics<float>(p->model, p->g, p->sd<float>(), static_cast<const float *>(p->u),
                                                  p->fixedmask, xf, h, ntips, p->geo_tips, p->geo_pts, p->geo_len,
                                                  p->geo_st, s)
                        : launch_geodesics<double>(p->model, p->g, p->sd<double>(),
                                                   static_cast<const double *>(p->u), p->fixedmask, xf, h, ntips,
                                                   p->geo_tips, p->geo_pts, p->geo_len, p->geo_st, s);
            }
            HIPCHECK(e);
            long long *hl = reinterpret_cast<long long *>(p->geo_host);
            int *hs = reinterpret_cast<int *>(p->geo_host + (size_t)ntips * sizeof(long long));
            HIPCHECK(hipMemcpyAsync(hl, p->geo_len, ntips * sizeof(long long), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipMemcpyAsync(hs, p->geo_st, ntips * sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            std::memcpy(hlen.data(), hl, ntips * sizeof(long long));
            std::memcpy(hst.data(), hs, ntips * sizeof(int));
            long long mx = 0;
            for (long long l : hlen) mx = std::max(mx, l);
            if (mx <= cap || attempt >= 2) break;
            cap = mx;
        }
        int failed = 0;
        std::string err;
        for (int t = 0; t < ntips; ++t) {
            if (tip_failed) tip_failed[t] = hst[t] == GEO_STATUS_FAILED;
            if (hst[t] == GEO_STATUS_FAILED) ++failed;
            else if (hst[t] == 6)
                err = "geodesic ODE solver: more than " + std::to_string(1LL << p->ode_hlog2) +
                      " distinct grid points read for one tip";
            else if (hst[t] == 7)
                err = "geodesic ODE solver: geodesicStep too small (more than 1024 steps of distance history)";
            else if (hst[t] == GEO_STATUS_SET_OVERFLOW)
                err = "geodesic extraction: the transported mass spread over more than 256 grid points";
            else if (hst[t] == GEO_STATUS_STEPS_OVERFLOW)
                err = "geodesic extraction: more than 256 steps between two geodesic points";
            else if (hst[t] == GEO_STATUS_NO_END)
                err = "geodesic extraction: mass transport did not end";
            else if (hst[t] == GEO_STATUS_INDEX_RANGE)
                err = "geodesic extraction: the transported mass left the range of grid indices (|index| < 2^20)";
            else if (hlen[t] > cap)
                err = "geodesic extraction: geodesic longer than the point buffer";
        }
        long long total = 0;
        if (err.empty()) {
            if (points && max_points > 0) {
                // the whole point buffer in one copy into pinned memory, then each tip's points
                const size_t bytes = (size_t)ntips * cap * dim * sizeof(double);
                ensure_host(bytes);
                HIPCHECK(hipMemcpyAsync(p->geo_host, p->geo_pts, bytes, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
            }
            const double *hp = reinterpret_cast<const double *>(p->geo_host);
            for (int t = 0; t < ntips; ++t) {
                lengths[t] = hlen[t];
                const long long n = hlen[t];
                if (points && total < max_points) {
                    const long long take = std::min<long long>(n, max_points - total);
                    std::memcpy(points + (size_t)total * dim, hp + (size_t)t * cap * dim,
                                (size_t)take * dim * sizeof(double));
                }
                total += n;
            }
        }
        if (!err.empty()) throw HipFail{err};
        *npoints = total;
        if (nfailed) *nfailed = failed;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

extern "C" {

int hfm_plan_get_stats(const hfm_plan *p, hfm_plan_stats *st) {
    if (!p || !st) {
        set_error("hfm_plan_get_stats: null argument");
        return 1;
    }
    *st = p->st;
    if (p->stencil_ev_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p->stencil_ev[1]) == hipSuccess &&
            hipEventElapsedTime(&ms, p->stencil_ev[0], p->stencil_ev[1]) == hipSuccess)
            st->stencil_ms = ms;
    }
    return 0;
}

int hfm_plan_get_stencils(hfm_plan *p, int8_t *h_offsets, double *h_weights) {
    if (!p || !h_offsets || !h_weights) {
        set_error("hfm_plan_get_stencils: null argument");
        return 1;
    }
    if (p->shared || !p->stencil_ready) {
        set_error("hfm_plan_get_stencils: no per-point stencils (Riemann plan after hfm_plan_set_metric)");
        return 1;
    }
    try {
        const int K = p->ndiff, D = p->g.dim;
        std::vector<uint32_t> offs((size_t)p->np * K);
        HIPCHECK(hipMemcpy(offs.data(), p->offs, offs.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        std::vector<double> wd((size_t)p->np * K);
        if (p->precision == 32) {
            std::vector<float> wf((size_t)p->np * K);
            HIPCHECK(hipMemcpy(wf.data(), p->w, wf.size() * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < wf.size(); ++i) wd[i] = wf[i];
        } else {
            HIPCHECK(hipMemcpy(wd.data(), p->w, wd.size() * sizeof(double), hipMemcpyDeviceToHost));
        }
        for (long long tp = 0; tp < p->np; ++tp) {
            int c[3];
            tiled_coords(p->g, (int)(tp >> 6), (int)(tp & 63), c);
            if (!in_grid(p->g, c)) continue;
            const long long lin = ((long long)c[0] * p->g.n[1] + c[1]) * p->g.n[2] + c[2];
            for (int k = 0; k < K; ++k) {
                int o[3];
                unpack_offset(offs[(size_t)k * p->np + tp], o[0], o[1], o[2]);
                for (int q = 0; q < D; ++q) h_offsets[(lin * K + k) * D + q] = (int8_t)o[q];
                h_weights[lin * K + k] = wd[(size_t)k * p->np + tp];
            }
        }
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

int hfm_plan_set_delta(hfm_plan *p, double factor) {
    if (!p) {
        set_error("hfm_plan_set_delta: null argument");
        return 1;
    }
    p->delta_factor = factor;
    return 0;
}

int hfm_plan_set_grid(hfm_plan *p, int blocks) {
    if (!p) {
        set_error("hfm_plan_set_grid: null argument");
        return 1;
    }
    p->fim_grid = blocks;
    return 0;
}

int hfm_plan_set_round_target(hfm_plan *p, int tiles) {
    if (!p) {
        set_error("hfm_plan_set_round_target: null argument");
        return 1;
    }
    p->round_target = tiles;
    return 0;
}

int hfm_plan_set_graph(hfm_plan *p, int enable) {
    if (!p) {
        set_error("hfm_plan_set_graph: null argument");
        return 1;
    }
    p->use_graph = enable != 0;
    return 0;
}

int hfm_plan_set_fused_schedule(hfm_plan *p, int enable) {
    if (!p) {
        set_error("hfm_plan_set_fused_schedule: null argument");
        return 1;
    }
    p->fused_schedule = enable < 0 ? -1 : enable != 0;
    return 0;
}

int hfm_plan_set_order(hfm_plan *p, int order) {
    if (!p) {
        set_error("hfm_plan_set_order: null argument");
        return 1;
    }
    p->order = order > 1 ? (order > 3 ? 3 : order) : 1;
    return 0;
}

int hfm_plan_set_stopping(hfm_plan *p, int64_t max_accepted, double stop_distance, int n_any, const int64_t *any_idx,
                          int n_all, const int64_t *all_idx, int n_any_u, const int64_t *any_u_idx, int n_all_u,
                          const int64_t *all_u_idx) {
    if (!p || (n_any > 0 && !any_idx) || (n_all > 0 && !all_idx) || (n_any_u > 0 && !any_u_idx) ||
        (n_all_u > 0 && !all_u_idx) || n_any < 0 || n_all < 0 || n_any_u < 0 || n_all_u < 0) {
        set_error("hfm_plan_set_stopping: null argument");
        return 1;
    }
    if ((n_any_u > 0 || n_all_u > 0) && !p->shared) {
        set_error("hfm_plan_set_stopping: unoriented targets apply to the curvature models only");
        return 1;
    }
    const long long npts = p->N, nxy = (long long)p->g.n[0] * p->g.n[1];
    auto take = [&](int n, const int64_t *idx, long long lim, std::vector<long long> &dst) {
        dst.clear();
        for (int i = 0; i < n; ++i) {
            if (idx[i] < 0 || idx[i] >= lim) return false;
            dst.push_back(idx[i]);
        }
        return true;
    };
    if (!take(n_any, any_idx, npts, p->stop_any) || !take(n_all, all_idx, npts, p->stop_all) ||
        !take(n_any_u, any_u_idx, nxy, p->stop_any_u) || !take(n_all_u, all_u_idx, nxy, p->stop_all_u)) {
        set_error("hfm_plan_set_stopping: target index out of the grid");
        return 1;
    }
    p->stop_max = max_accepted;
    p->stop_dist = stop_distance;
    p->has_stopping = max_accepted >= 0 || stop_distance < __builtin_huge_val() || n_any > 0 || n_all > 0 ||
                      n_any_u > 0 || n_all_u > 0;
    return 0;
}

int hfm_plan_set_factoring(hfm_plan *p, const int64_t *box_dims, const int64_t *shift, const double *values,
                           const double *gradients, const uint8_t *mask) {
    if (!p) {
        set_error("hfm_plan_set_factoring: null argument");
        return 1;
    }
    try {
        if (p->exact_solved) {  // that solve's flows would read the arrays freed here
            p->exact_solved = false;
            p->solved = false;
        }
        if (!values) {  // remove the factor
            dfree(p->fac_vals);
            dfree(p->fac_grads);
            dfree(p->fac_mask);
            p->fac_vals = p->fac_grads = nullptr;
            p->fac_mask = nullptr;
            p->fac = FactorDev();
            return 0;
        }
        if (!box_dims || !gradients) {
            set_error("hfm_plan_set_factoring: box_dims and gradients are required with values");
            return 1;
        }
        const int dim = p->g.dim;
        FactorDev f;
        long long n = 1;
        for (int i = 0; i < dim; ++i) {
            if (box_dims[i] <= 0 || box_dims[i] > (1 << 24)) {
                set_error("hfm_plan_set_factoring: box extents must be positive");
                return 1;
            }
            if (!shift && box_dims[i] != p->g.n[i]) {
                set_error("hfm_plan_set_factoring: without a shift the factor covers the whole grid "
                          "(Static factoring error : inconsistent data dimension.)");
                return 1;
            }
            if (shift && (shift[i] < -(1LL << 30) || shift[i] > (1LL << 30))) {
                set_error("hfm_plan_set_factoring: shift out of range");
                return 1;
            }
            f.bd[i] = (int)box_dims[i];
            f.shift[i] = shift ? (int)shift[i] : 0;
            n *= box_dims[i];
        }
        f.sub = shift ? 1 : 0;
        f.on = 1;
        std::vector<unsigned char> m(n, 1);
        if (mask)
            for (long long i = 0; i < n; ++i) m[i] = mask[i] != 0;
        dfree(p->fac_vals);
        dfree(p->fac_grads);
        dfree(p->fac_mask);
        p->fac_vals = p->fac_grads = nullptr;
        p->fac_mask = nullptr;
        p->fac = FactorDev();
        p->fac_vals = dalloc<double>(n);
        p->fac_grads = dalloc<double>(n * dim);
        p->fac_mask = dalloc<unsigned char>(n);
        HIPCHECK(hipMemcpy(p->fac_vals, values, n * sizeof(double), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(p->fac_grads, gradients, n * dim * sizeof(double), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(p->fac_mask, m.data(), n, hipMemcpyHostToDevice));
        f.vals = p->fac_vals;
        f.grads = p->fac_grads;
        f.mask = p->fac_mask;
        p->fac = f;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

int hfm_plan_set_walls(hfm_plan *p, const uint8_t *d_walls, int refine, void *stream) {
    if (!p) {
        set_error("hfm_plan_set_walls: null argument");
        return 1;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (!d_walls) {
            dfree(p->walls);
            p->walls = nullptr;
        } else {
            long long n = 1;
            for (int i = 0; i < p->g.dim; ++i) n *= p->g.n[i];
            if (!p->walls) p->walls = dalloc<uint8_t>(n);
            HIPCHECK(hipMemcpyAsync(p->walls, d_walls, (size_t)n, hipMemcpyDeviceToDevice, s));
        }
        p->walls_refine = refine != 0;
        p->vis_dirty = true;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

int hfm_plan_set_max_sweeps(hfm_plan *p, int sweeps) {
    if (!p) {
        set_error("hfm_plan_set_max_sweeps: null argument");
        return 1;
    }
    if (sweeps != 0 && sweeps != 4 && sweeps != 8 && sweeps != 16 && sweeps != 32) {
        set_error("hfm_plan_set_max_sweeps: sweeps must be 4, 8, 16 or 32, or 0 for the default (got " +
                  std::to_string(sweeps) + ")");
        return 1;
    }
    p->max_sweeps = sweeps;
    return 0;
}

int hfm_plan_set_timing(hfm_plan *p, int enable) {
    if (!p) {
        set_error("hfm_plan_set_timing: null argument");
        return 1;
    }
    p->timing = enable != 0;
    return 0;
}

int hfm_plan_adjoint(hfm_plan *p, int ntargets, const double *targets, const double *weights, double *d_sensitivity,
                     int64_t *seed_index, double *seed_weight, int max_seeds, int *nseed_out, void *stream) {
    return hfm_plan_adjoint_impl(p, ntargets, targets, weights, d_sensitivity, seed_index, seed_weight, max_seeds,
                                 nseed_out, stream);
}

int hfm_plan_forward_variation(hfm_plan *p, int nvar, const double *d_cost_variation, int nseeds,
                               const double *seed_value_variation, double *d_value_variation, void *stream) {
    return hfm_plan_forward_variation_impl(p, nvar, d_cost_variation, nseeds, seed_value_variation,
                                           d_value_variation, stream);
}

}  // extern "C"

int hfm::plan_count_reached(hfm_plan *p, int64_t *n, void *stream) {
    if (!p || !n || !p->solved) {
        set_error("plan_count_reached: no solve has run");
        return 1;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        HIPCHECK(hipMemsetAsync(p->red, 0, RED_SLOTS * sizeof(RedSlot), s));
        if (p->precision == 32)
            HIPCHECK(launch_count_reached<float>(p->g, static_cast<const float *>(p->u), p->np, p->red, s));
        else
            HIPCHECK(launch_count_reached<double>(p->g, static_cast<const double *>(p->u), p->np, p->red, s));
        HIPCHECK(hipMemcpyAsync(p->h_red, p->red, RED_SLOTS * sizeof(RedSlot), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        *n = (int64_t)red_total(p->h_red).count;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Allocation-free Riemann solve.  Every buffer of an FP32 Riemann plan is carved, per call, out
// of one static device array (`scratch`, HFM_NOALLOC_WORKSPACE_BYTES per GPU, held from the
// library's load to the end of the process).  In host code the name `scratch` is a host-side
// stand-in for the symbol, not a device address: the address is taken with
// hipGetSymbolAddress once per device, with that device current, and only that address is used.
// ---------------------------------------------------------------------------------------------
namespace {

__device__ __align__(256) char scratch[HFM_NOALLOC_WORKSPACE_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
constexpr int NOALLOC_MAX_SEEDS = 4096;
std::mutex g_ws_mutex[NOALLOC_MAX_DEVICES];  // one call at a time per device
void *g_ws_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of the workspace on the current device (called with g_ws_mutex[dev] held).
char *workspace_address(int dev) {
    if (!g_ws_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw HipFail{std::string("hipGetSymbolAddress of the allocation-free workspace failed: ") +
                          hipGetErrorString(e)};
        g_ws_ptr[dev] = ptr;
    }
    return static_cast<char *>(g_ws_ptr[dev]);
}

// Consecutive 256-byte-aligned pieces of the workspace.
struct Carver {
    char *base;
    size_t off = 0;
    template <typename T> T *take(size_t n) {
        off = (off + 255) & ~size_t(255);
        T *r = reinterpret_cast<T *>(base + off);
        off += (n ? n : 1) * sizeof(T);
        if (off > HFM_NOALLOC_WORKSPACE_BYTES) throw HipFail{"allocation-free workspace exhausted"};
        return r;
    }
};

}  // namespace

extern "C" int hfm_riemann_solve_noalloc(const char *model, const int64_t *dims, const double *gridScales,
                                         const double *origin, const double *d_metric, int metric_is_field, int dual,
                                         int nseeds, const double *seeds, const double *seedValues,
                                         double *d_values, void *stream) {
    if (!model || !dims || !gridScales || !d_metric || !d_values || (nseeds > 0 && !seeds)) {
        set_error("hfm_riemann_solve_noalloc: null argument");
        return 1;
    }
    const int m = model_from_name(model);
    if (m != MODEL_RIEMANN2 && m != MODEL_RIEMANN3) {
        set_error(std::string("hfm_riemann_solve_noalloc: model must be Riemann2 or Riemann3 (got ") + model + ")");
        return 1;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_error("hfm_riemann_solve_noalloc: no usable current device");
        return 1;
    }
    std::lock_guard<std::mutex> lock(g_ws_mutex[dev]);
    hfm_plan plan;  // host bookkeeping only; its device buffers point into the workspace
    hfm_plan *p = &plan;
    // host mirrors of the read-backs (pageable: no pinned allocation)
    struct HostMirror {
        int count[4];
        RedSlot red[RED_SLOTS];
        unsigned long long stats[STAT_WORDS];
    };
    static thread_local HostMirror hm;
    try {
        setup_geometry(p, m, 32, dims, gridScales, origin, 1.0, 0.1);
        const long long nt = p->g.ntiles;
        if (nt > HFM_NOALLOC_MAX_TILES)
            throw HipFail{"hfm_riemann_solve_noalloc: the grid has " + std::to_string(nt) +
                          " tiles of 64 points, more than the " + std::to_string(HFM_NOALLOC_MAX_TILES) +
                          " the allocation-free workspace holds (use hfm_plan_create)"};
        if (nseeds > NOALLOC_MAX_SEEDS)
            throw HipFail{"hfm_riemann_solve_noalloc: at most " + std::to_string(NOALLOC_MAX_SEEDS) + " seeds"};
        hipStream_t s = static_cast<hipStream_t>(stream);
        Carver cv{workspace_address(dev)};
        p->workspace = true;
        p->use_graph = false;
        p->u = cv.take<float>(p->np);
        p->offs = cv.take<uint32_t>(p->np * p->ndiff);
        p->w = cv.take<float>(p->np * p->ndiff);
        p->fixedmask = cv.take<unsigned long long>(nt);
        p->dep_cnt = cv.take<int>(nt);
        p->dep_start = cv.take<int>(nt + 1);
        p->dep_fill = cv.take<int>(nt);
        p->lists[0] = cv.take<int>(nt);
        p->counts = cv.take<int>(4);
        p->dep_src = cv.take<int>(nt * 64);
        p->dep_nsrc = cv.take<int>(nt);
        p->scan_work = cv.take<int>(scan_work_ints(nt));
        p->keys = cv.take<unsigned>(2 * nt);
        p->tmax = cv.take<unsigned>(nt);
        p->nextmin = cv.take<unsigned>(4 * PEND_WORDS);
        p->bounds = cv.take<float>(5);
        p->red = cv.take<RedSlot>(RED_SLOTS);
        p->status = cv.take<int>(4);
        p->stats = cv.take<unsigned long long>(STAT_WORDS);
        p->seed_p = cv.take<long long>(NOALLOC_MAX_SEEDS);
        p->seed_v = cv.take<double>(NOALLOC_MAX_SEEDS);
        p->seed_cap = NOALLOC_MAX_SEEDS;
        // the tile graph gets the rest
        cv.off = (cv.off + 255) & ~size_t(255);
        p->deps = reinterpret_cast<int *>(cv.base + cv.off);
        p->deps_cap = (long long)((HFM_NOALLOC_WORKSPACE_BYTES - cv.off) / sizeof(int));
        p->h_count = hm.count;
        p->h_red = hm.red;
        p->h_stats = hm.stats;
        set_metric_impl<float>(p, d_metric, metric_is_field, dual, s);
        p->stencil_ready = true;
    } catch (const HipFail &e) {
        set_error(e.msg);
        return 1;
    }
    if (hfm_plan_solve(p, nseeds, seeds, seedValues, stream)) return 1;
    if (hfm_plan_values(p, d_values, stream)) return 1;
    // the next call on this device may reuse the workspace as soon as this one returns
    const hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        set_error(std::string("hfm_riemann_solve_noalloc: ") + hipGetErrorString(e));
        return 1;
    }
    return 0;
}
