// hfm_internal.h -- host-side internals shared by the plan, adjoint and IO translation units.
#ifndef HFM_AMD_INTERNAL_H
#define HFM_AMD_INTERNAL_H

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/hfm_amd.h"
#include "hfm_device.h"

namespace hfm {

void set_error(const std::string &m);

struct HipFail {
    std::string msg;
};

#define HIPCHECK(expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw hfm::HipFail{std::string(#expr) + " failed: " + hipGetErrorString(e_)};        \
    } while (0)

template <typename T> T *dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    HIPCHECK(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T *>(p);
}

inline void dfree(void *p) {
    if (p) (void)hipFree(p);
}


}  // namespace hfm

struct hfm_plan {
    using Geom = hfm::Geom;
    template <typename T> using StencilData = hfm::StencilData<T>;
    int model = 0;
    int precision = 32;
    Geom g{};
    long long np = 0;  // tiled points
    long long N = 0;   // grid points
    double h[3] = {1, 1, 1};
    double origin[3] = {0, 0, 0};
    double dependScale = 1;
    double xi = 1, eps = 0.1;
    int ndiff = 0;
    bool shared = false;

    // device buffers
    void *u = nullptr;
    uint32_t *offs = nullptr;
    void *w = nullptr;
    void *speed = nullptr;
    uint32_t *tab_offs = nullptr;
    double *tab_w = nullptr;
    unsigned long long *fixedmask = nullptr;
    int *dep_cnt = nullptr, *dep_start = nullptr, *dep_fill = nullptr, *deps = nullptr;
    long long deps_cap = 0, ndeps = 0;
    int *lists[1] = {nullptr};     // tiles listed by the current round
    int *counts = nullptr;         // 4 rotating list counters
    long long prev_rounds = 0;     // rounds the last solve on this plan ran
    int *dep_src = nullptr;        // tile-graph build: each tile's sources (<= 64 kept, 64 per tile)
    int *dep_nsrc = nullptr;       // tile-graph build: number of sources per tile
    int *scan_work = nullptr;      // multi-block exclusive scan of the tile-graph counts
    unsigned *keys = nullptr;      // per-tile scheduling key (KEY_NONE: not pending)
    unsigned *tmax = nullptr;      // per-tile largest value (as key) after its last visit
    unsigned *nextmin = nullptr;   // 4 rotating smallest pending keys (PEND_SLOTS words each)
    float *bounds = nullptr;       // 2 alternating band bounds + 2 alternating band widths
    hfm::RedSlot *red = nullptr;   // spread launch-wide reductions (mean one-step time, widest stencil)
    hfm::RedSlot *h_red = nullptr;  // pinned host copy of the slots (asynchronous read-back)
    // stencil-phase events of the last set_metric, read when the statistics are asked for (the
    // phase ends without a host synchronisation)
    hipEvent_t stencil_ev[2] = {nullptr, nullptr};
    bool stencil_ev_pending = false;
    double tau = 0.;               // mean one-step time of the current stencils
    double delta_factor = 8.;      // initial band width in units of tau (<= 0: unordered)
    int fim_grid = 0;              // update-kernel blocks of 4 waves (0: 960, or 1280 for listed 2-D grids)
    int round_target = 0;          // tiles per round the band width is steered to (0: 64 for 3-D grids of >= 4096 tiles, else 1024)
    int max_sweeps = 0;            // sweeps per tile visit (4, 8, 16 or 32; 0: 4 for Riemann3 and Elastica2, else 8)
    int *status = nullptr;  // stencil status: worst Selling error, smallest failing row-major index
    unsigned long long *stats = nullptr;
    long long *seed_p = nullptr;
    double *seed_v = nullptr;
    int seed_cap = 0;
    // pinned host mirrors
    int *h_count = nullptr;
    unsigned long long *h_stats = nullptr;

    // adjoint scratch (allocated on the first hfm_plan_adjoint call)
    double *adj_lambda = nullptr, *adj_sens = nullptr;
    // each point's active edges, computed once per adjoint: count, var, and up to ADJ_EDGES
    // (neighbour tiled index, w'/S) pairs, slot-major (slot k of point p at k * np + p)
    int *adj_ecnt = nullptr;
    unsigned *adj_eidx = nullptr;
    double *adj_var = nullptr, *adj_ew = nullptr;
    int *adj_indeg = nullptr, *adj_counts = nullptr;
    long long *adj_list[2] = {nullptr, nullptr};
    long long *adj_seed_idx = nullptr, *adj_tp = nullptr;
    double *adj_seed_w = nullptr, *adj_tw = nullptr;
    int adj_seed_cap = 0, adj_t_cap = 0;
    // forward-variation scratch (first hfm_plan_forward_variation call; lists shared with the adjoint)
    int *fwd_pending = nullptr, *fwd_rcount = nullptr, *fwd_rstart = nullptr, *fwd_rev = nullptr;
    int *fwd_scan_work = nullptr;
    long long fwd_rev_cap = 0;
    double *fwd_res = nullptr, *fwd_sv = nullptr;
    long long fwd_res_cap = 0, fwd_sv_cap = 0;
    long long *fwd_sp = nullptr;
    int fwd_seed_cap = 0;

    // geodesic scratch (grown on demand, kept for the next call): tips, lengths and status on the
    // device, the points buffer (doubles), and one pinned host buffer for the read-back
    double *geo_tips = nullptr, *geo_pts = nullptr;
    long long *geo_len = nullptr;
    int *geo_st = nullptr;
    // the ODE geodesic solver: the seeds' L-infinity distance and the per-tip read-count tables
    int *ode_dist = nullptr;
    unsigned long long *ode_hkeys = nullptr;
    int *ode_hcnt = nullptr;
    long long ode_hcap = 0;
    int ode_hlog2 = 15;
    long long geo_tip_cap = 0, geo_pts_cap = 0;
    char *geo_host = nullptr;
    size_t geo_host_cap = 0;

    // walls (hfm_plan_set_walls): row-major bytes, the visibility mask derived from them and the
    // current stencils (rebuilt when either changes), scratch for the wall distances
    uint8_t *walls = nullptr;
    int walls_refine = 0;
    uint32_t *vis = nullptr;
    int8_t *wdist[2] = {nullptr, nullptr};
    bool vis_dirty = true;

    // Buffers carved out of the static workspace of hfm_riemann_solve_noalloc: no event, graph,
    // pinned buffer or device allocation is made, and a buffer that would have to grow (tile
    // graph, seeds) is an error instead.
    bool workspace = false;
    bool stencil_ready = false;
    bool solved = false;
    bool timing = false;
    // The update loop runs as a captured HIP graph of one batch of rounds (schedule + update
    // launches and the read-back of the last round's count), replayed until convergence.  The
    // graph is captured on graph_stream and re-captured when its launch arguments change.
    bool use_graph = true;
    int fused_schedule = -1;       // rounds list their own tiles (k_fim_sched) instead of k_schedule + k_fim: 1 on, 0 off, -1 on for 2-D grids
    // Serial acceptance loop (hfm_exact.hip), run instead of the block fast-iterative method when
    // order > 1: reversed offsets per key (tiled point, or theta index for shared stencils), the
    // accepted flags, activeNeighs bits and incremental quadratics of the last such solve.
    int order = 1;
    bool exact_solved = false;
    std::string exact_what;  // what made the last solve run the serial loop, for refusals
    bool exact_rev_dirty = true;
    int *ex_cnt = nullptr, *ex_start = nullptr, *ex_fill = nullptr, *ex_rcount = nullptr, *ex_scan_work = nullptr;
    uint32_t *ex_roffs = nullptr;
    long long ex_roffs_cap = 0;
    long long ex_pushes = 0;  // reversed offsets summed over every point: the heap's capacity less the seeds
    unsigned char *ex_acc = nullptr;
    unsigned long long *ex_active = nullptr;
    double *ex_quad = nullptr;
    void *ex_heap = nullptr;
    long long ex_heap_cap = 0;
    long long *ex_out = nullptr;
    // stopping criteria (hfm_plan_set_stopping), row-major grid indices; flags uploaded per solve
    bool has_stopping = false;
    long long stop_max = -1;
    double stop_dist = 0.;
    std::vector<long long> stop_any, stop_all, stop_any_u, stop_all_u;
    unsigned char *ex_flags = nullptr, *ex_flags_u = nullptr;
    // static factor (hfm_plan_set_factoring): device copies of values, gradients (grid units) and
    // mask over the box; fac.on = 0 for none
    hfm::FactorDev fac;
    // the order and factor the last serial solve ran with (its flows and geodesics use them); a
    // new factor invalidates that solve, whose factor arrays it frees
    int ex_order = 1;
    hfm::FactorDev ex_fac;
    double *fac_vals = nullptr, *fac_grads = nullptr;
    unsigned char *fac_mask = nullptr;
    hipStream_t graph_stream = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<char> graph_key;
    int max_stencil_width = 0;
    std::vector<long long> last_seeds;  // tiled indices of the last solve's seeds
    std::vector<long long> seed_inputs; // tiled index of each seed as passed (-1: out of the grid)
    hfm_plan_stats st{};

    template <typename T> StencilData<T> sd() const {
        StencilData<T> s;
        s.offs = offs;
        s.w = static_cast<const T *>(w);
        s.tab_offs = tab_offs;
        s.tab_w = tab_w;
        s.speed = static_cast<const T *>(speed);
        s.np = np;
        s.vis = walls ? vis : nullptr;
        return s;
    }
};

int hfm_plan_adjoint_impl(hfm_plan *p, int ntargets, const double *targets, const double *weights,
                          double *d_sensitivity, int64_t *seed_index, double *seed_weight, int max_seeds,
                          int *nseed_out, void *stream);

// Host copy of the per-theta stencil tables of a curvature plan (theta x ndiff, packed offsets
// and base weights).
int hfm_plan_curvature_tables(const hfm_plan *p, std::vector<uint32_t> &offs, std::vector<double> &w);

// hfm_plan_geodesics with tips already in grid units (ADim'd) when tips_adim.
int hfm_plan_geodesics_impl(hfm_plan *p, int ntips, const double *tips, bool tips_adim, double step,
                            double threshold, double volumeBound, double *points, int64_t max_points,
                            int64_t *lengths, int64_t *npoints, int *nfailed, void *stream, int solver = 0,
                            double causality = 4., int *tip_failed = nullptr);

namespace hfm {
// Number of reached grid points (finite values) of the last solve, counted on the device.
int plan_count_reached(hfm_plan *p, int64_t *n, void *stream);
}  // namespace hfm

int hfm_plan_forward_variation_impl(hfm_plan *p, int nvar, const double *d_cost_variation, int nseeds,
                                    const double *seed_value_variation, double *d_value_variation, void *stream);

#endif
