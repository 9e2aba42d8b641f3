// hfm_kernels.h -- host-side launchers of the HIP kernels (see hfm_kernels.hip).
#ifndef HFM_AMD_KERNELS_H
#define HFM_AMD_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hfm_device.h"

namespace hfm {

// 1/(h_i h_j) in packed symmetric order (Headers/Specializations/Riemannian.h:50-52).
struct InvHH {
    double v[6];
};

// Per-axis scale applied by ReDim to a vector (h_i, or the theta step on the bundle axis).
struct FlowScale {
    double h[3];
};

constexpr unsigned KEY_NONE = 0xffffffffu;
constexpr int PEND_SLOTS = 32;  // spread of the per-round minimum-key and statistics counters
// The single-solve kernels keep each slot on its own 128-byte line (PEND_STRIDE words for the
// pending minima, STAT_STRIDE 64-bit words for the statistics): atomics to different words of
// one line serialise like atomics to one word.
constexpr int PEND_STRIDE = 32;
constexpr int STAT_STRIDE = 16;
constexpr int PEND_WORDS = PEND_SLOTS * PEND_STRIDE;  // words of one round's pending minima
constexpr int STAT_WORDS = PEND_SLOTS * STAT_STRIDE;  // 64-bit words of the statistics

// One scheduling round (see k_schedule): reads the pending minimum and the previous bound,
// writes this round's bound, lists the tiles it takes, accumulates the keys it leaves.
struct SchedRound {
    const unsigned *pend_in;   // PEND_SLOTS words, PEND_STRIDE apart
    unsigned *pend_out;        // PEND_SLOTS words, PEND_STRIDE apart
    const float *bound_in;
    float *bound_out;
    int *list;
    int *count;
    const float *delta0;    // initial band width (device word); <= 0 lists every pending tile
    const float *delta_in;  // band width of the previous round
    float *delta_out;       // band width of this round
    const int *prev_count;  // tiles the previous round listed
    int target;             // tiles per round the band width is steered to
};

template <typename T> struct FimArgs {
    Geom g;
    StencilData<T> sd;
    T *u;
    const unsigned long long *fixedmask;
    const int *dep_start;
    const int *deps;
    const int *list;
    const int *count;
    int *count_reset;
    unsigned *keys;
    unsigned *keys_out;  // fused rounds: the key buffer this round marks into (keys: the one it takes from)
    unsigned *tmax;
    unsigned *pend_out;
    unsigned *pend_reset;
    unsigned long long *stats;  // PEND_SLOTS pairs (tiles processed, sweeps summed), STAT_STRIDE apart
    unsigned grid;
    int max_sweeps;  // sweeps per tile visit before the tile is handed back to the schedule: 4, 8, 16 or 32
    int fused;       // 1: k_fim_sched lists its own tiles with the round parameters in sr (no list)
    SchedRound sr;
};

// Serial acceptance loop of the options fixed by the heap's acceptance order (hfm_exact.hip):
// order 2 / 3 recompute.  Arrays are tiled; the reversed offsets are listed per key (the tiled
// point, or the theta index for shared stencils): rstart / rcount / roffs (packed int8 x 3).
struct ExactArgs {
    Geom g;
    StencilData<double> sd;
    double *u;
    unsigned char *accepted;
    unsigned long long *active;  // activeNeighs bits (side_bit layout)
    double *quad;                // per point: minVal, a, b, c of the incremental quadratic
    const int *rstart;
    const int *rcount;
    const uint32_t *roffs;
    void *heap;                  // heap_cap entries of 16 bytes
    long long heap_cap;
    const long long *seed_p;     // tiled seed points in row-major index order
    const double *seed_v;
    int nseeds;
    int order;
    long long *out;              // accepted, largest heap size, status (0 ok, 1 heap full), nAccepted
    // Stopping criteria (ExtraAlgorithms/CommonStoppingCriteria.h:57-92), checked when a point leaves the heap:
    // at most max_accepted pops; stop above stop_distance; stop_flags (per tiled point; null for
    // none): bit 0 stopWhenAnyAccepted, bit 1 stopWhenAllAccepted (cleared as it is accepted,
    // all_left of them to go); stop_flags_u (per (x, y) of the curvature models): the same for the
    // _Unoriented sets.
    long long max_accepted;
    double stop_distance;
    unsigned char *stop_flags;
    unsigned char *stop_flags_u;
    long long all_left, all_left_u;
    // Static factoring (StaticFactoring.hxx:11-55): a point in the factor's mask is recomputed at
    // acceptance at every order, its differences corrected.
    FactorDev fac;
};
hipError_t launch_exact_init(double *u, unsigned char *acc, unsigned long long *active, double *quad, long long np,
                             hipStream_t s);
hipError_t launch_rev_build(int model, const Geom &g, const StencilData<double> &sd, int pass, int *cnt,
                            const int *start, int *fill, uint32_t *roffs, long long np, hipStream_t s);
hipError_t launch_rev_sort(const int *cnt, const int *start, uint32_t *roffs, int *rcount, long long np, hipStream_t s);
hipError_t launch_fm_exact(int model, const ExactArgs &a, hipStream_t s);
hipError_t launch_flow_exact(int model, const Geom &g, const StencilData<double> &sd, const double *u,
                             const unsigned char *acc, const unsigned long long *active, int order,
                             const FactorDev &fac, const FlowScale &fs, double *flow, double *act_out, double *aoffs,
                             long long np, hipStream_t s);

// Geodesic backtracking (hfm_geodesic.hip).  Tips are ADim'd (grid-index units); points are
// written ReDim'd, max_points per tip, tip t at points[t * max_points * dim ...].
struct GeoParamsHost {
    double step = 0.25, threshold = 0.001, volumeBound = 0.;
    double scale[3] = {1., 1., 1.}, origin[3] = {0., 0., 0.};
    long long max_points = 0;  // per tip
    long long max_iter = 0;    // per pass
};
enum GeoStatusHost { GEO_STATUS_OK = 0, GEO_STATUS_FAILED = 1, GEO_STATUS_SET_OVERFLOW = 2,
                     GEO_STATUS_STEPS_OVERFLOW = 3, GEO_STATUS_NO_END = 4, GEO_STATUS_INDEX_RANGE = 5 };
// The flows of a serial solve for geodesic backtracking: the recompute over that loop's accepted
// set (order 2/3 combinations, the factor's corrections, HamiltonFastMarching.hxx:187-272), as
// the reference's solver takes them (GeodesicDiscreteSolver.hxx:131-132).  on = 0: a block solve,
// whose flows are recomputed from the values alone.
struct ExactFlow {
    const unsigned char *acc = nullptr;
    const unsigned long long *active = nullptr;
    int order = 1;
    FactorDev fac;
    int on = 0;
};
template <typename T>
hipError_t launch_geodesics(int model, const Geom &g, const StencilData<T> &sd, const T *u,
                            const unsigned long long *fm, const ExactFlow &xf, const GeoParamsHost &h, int ntips,
                            const double *tips, double *points, long long *lengths, int *status, hipStream_t s);
// The ODE geodesic solver (hfm_geodesic.hip): dist, the L-infinity grid distance to the seeds
// capped at 6 (launch_ode_distance), per-tip read-count tables of 2^hlog2 slots (hkeys / hcnt,
// zeroed by the caller).  Status GEO_STATUS_FAILED marks a stall; 6 a full count table; 7 a step
// too small for the distance history.
template <typename T>
hipError_t launch_geodesics_ode(int model, const Geom &g, const StencilData<T> &sd, const T *u,
                                const unsigned long long *fm, const ExactFlow &xf, const GeoParamsHost &h,
                                double causality, const int *dist, unsigned long long *hkeys, int *hcnt, int hlog2,
                                int ntips, const double *tips, double *points, long long *lengths, int *status,
                                hipStream_t s);
hipError_t launch_ode_distance(int model, const Geom &g, int *dist, const long long *seeds, int nseeds, long long np,
                               hipStream_t s);

template <typename T>
hipError_t launch_riemann_stencil(const Geom &g, const double *metric, int is_field, int dual, const InvHH &ih,
                                  uint32_t *offs, T *w, long long np, int *status, RedSlot *red, hipStream_t s);
template <typename T>
hipError_t launch_tile_deps(int model, const Geom &g, const StencilData<T> &sd, int pass, int *cnt, const int *start,
                            int *fill, int *deps, int *srcbuf, int *srcn, hipStream_t s);
// Solver state (values, fixed mask) from row-major FP64 arrival times and activeNeighs.
template <typename T>
hipError_t launch_load_state(const Geom &g, const double *values, const double *active, T *u,
                             unsigned long long *fixedmask, long long np, hipStream_t s);

// Walls (hfm_walls.hip): per-point visible-side bitmask (side_bit layout) from a row-major byte
// field of walls, the distance to the walls being capped at maxDist (MaxStencilWidth narrowed to
// int8); dist0 / dist1: np bytes of scratch each.
template <typename T>
hipError_t launch_wall_visibility(int model, const Geom &g, const StencilData<T> &sd, const uint8_t *walls,
                                  int maxDist, int refine, int8_t *dist0, int8_t *dist1, uint32_t *vis, long long np,
                                  hipStream_t s);

// Exclusive scan of n counts into start[0..n].  With work (scan_work_ints(n) ints of device
// memory) long arrays are scanned by one block per 4096 counts; without, by a single block.
long long scan_work_ints(long long n);
hipError_t launch_exclusive_scan(const int *cnt, int *start, int n, hipStream_t s, int *work = nullptr);
template <typename T>
hipError_t launch_init(const Geom &g, T *u, unsigned long long *fixedmask, unsigned *keys, unsigned *tmax,
                       long long np, const long long *seed_p, const double *seed_v, int nseeds, hipStream_t s,
                       int key_buffers = 1);
hipError_t launch_seed_activate(const long long *seed_p, const double *seed_v, int nseeds, const int *dep_start,
                                const int *deps, unsigned *keys, const unsigned *tmax, unsigned *pend, hipStream_t s);
hipError_t launch_schedule(int ntiles, unsigned *keys, const SchedRound &sr, hipStream_t s);
template <typename T>
hipError_t launch_fim(int model, const FimArgs<T> &a, hipStream_t s, hipEvent_t ev0 = nullptr,
                      hipEvent_t ev1 = nullptr);
template <typename T>
hipError_t launch_export_values(const Geom &g, const T *u, double *out, long long np, hipStream_t s);
// Reached (finite-valued) grid points, counted into the reduction slots' `count` fields.
template <typename T>
hipError_t launch_count_reached(const Geom &g, const T *u, long long np, RedSlot *red, hipStream_t s);
template <typename T>
hipError_t launch_flow(int model, const Geom &g, const StencilData<T> &sd, const T *u, const unsigned long long *fm,
                       const FlowScale &fs, double *flow, double *act, double *aoffs, long long np, hipStream_t s);
template <typename T>
hipError_t launch_residual(int model, const Geom &g, const StencilData<T> &sd, const T *u,
                           const unsigned long long *fm, long long np, ResSlot *out, hipStream_t s);

}  // namespace hfm

#endif
