// hfm_device.h -- data layout and per-point numerics shared by the HIP kernels.
//
// Layout in HBM.  A grid of dims (n0, n1[, n2]) is cut into tiles of 64 points, one wavefront
// per tile: 8x8 in 2-D, 4x4x4 in 3-D.  Every per-point array is stored tile-major
// ("tiled index" p = tile * 64 + local), so one wave reads a tile's 64 values of a field with a
// single coalesced 256 B (FP32) or 512 B (FP64) access.  Points of partial edge tiles that fall
// outside the grid are "padding": fixed at +inf and never updated.
//
// Stencils.  Riemann models keep one stencil per point (nDiff symmetric differences: offsets
// packed as int8x3 in a uint32, weights in the update precision).  The curvature models
// (ReedsSheppForward2, Elastica2) share one stencil per theta index (a small table) and keep
// a per-point speed (1/cost); the weight of a difference is baseWeight * speed^2
// (Headers/Base/EulerianStencil.h:84-97).
#ifndef HFM_AMD_DEVICE_H
#define HFM_AMD_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hfm {

enum ModelId { MODEL_RIEMANN2 = 0, MODEL_RIEMANN3 = 1, MODEL_RSF2 = 2, MODEL_ELASTICA2 = 3 };

constexpr int TILE_POINTS = 64;
constexpr int MAX_DIFF = 27;  // Elastica2<5>: 27 forward differences

// Grid geometry, passed by value to every kernel.
struct Geom {
    int dim;
    int n[3];      // grid extents (n[2] = 1 in 2-D)
    int sh[3];     // log2 of the tile extent per axis: (3,3,0) in 2-D, (2,2,2) in 3-D
    int nt[3];     // tiles per axis
    int ntiles;
    int periodic;  // bit k set: axis k is periodic (theta axis of R^2 x S^1)
};

__host__ __device__ inline int tile_extent(const Geom &g, int k) { return 1 << g.sh[k]; }

// Coordinates of the point at tiled index (tile, local).
__host__ __device__ inline void tiled_coords(const Geom &g, int tile, int local, int *c) {
    const int t2 = tile % g.nt[2];
    const int t01 = tile / g.nt[2];
    const int t1 = t01 % g.nt[1];
    const int t0 = t01 / g.nt[1];
    const int l2 = local & ((1 << g.sh[2]) - 1);
    const int l1 = (local >> g.sh[2]) & ((1 << g.sh[1]) - 1);
    const int l0 = local >> (g.sh[1] + g.sh[2]);
    c[0] = (t0 << g.sh[0]) + l0;
    c[1] = (t1 << g.sh[1]) + l1;
    c[2] = (t2 << g.sh[2]) + l2;
}

// Tile id and local index of in-grid coordinates.
__host__ __device__ inline int tile_of(const Geom &g, const int *c) {
    return ((c[0] >> g.sh[0]) * g.nt[1] + (c[1] >> g.sh[1])) * g.nt[2] + (c[2] >> g.sh[2]);
}
__host__ __device__ inline int local_of(const Geom &g, const int *c) {
    const int l0 = c[0] & ((1 << g.sh[0]) - 1);
    const int l1 = c[1] & ((1 << g.sh[1]) - 1);
    const int l2 = c[2] & ((1 << g.sh[2]) - 1);
    return (((l0 << g.sh[1]) | l1) << g.sh[2]) | l2;
}
__host__ __device__ inline bool in_grid(const Geom &g, const int *c) {
    return c[0] < g.n[0] && c[1] < g.n[1] && c[2] < g.n[2];
}

// Shift coordinates by s*offset with the reference's boundary rules: closed axes invalidate,
// periodic axes wrap (Headers/Base/Implementation/PeriodicGrid.hxx:120-140).
__host__ __device__ inline bool shift_coords(const Geom &g, const int *c, int o0, int o1, int o2, int *r) {
    r[0] = c[0] + o0;
    r[1] = c[1] + o1;
    r[2] = c[2] + o2;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (g.periodic >> k & 1) {
            int v = r[k] % g.n[k];
            r[k] = v < 0 ? v + g.n[k] : v;
        } else if (r[k] < 0 || r[k] >= g.n[k]) {
            return false;
        }
    }
    return true;
}

__host__ __device__ inline uint32_t pack_offset(int o0, int o1, int o2) {
    return (uint32_t)(uint8_t)(int8_t)o0 | ((uint32_t)(uint8_t)(int8_t)o1 << 8) |
           ((uint32_t)(uint8_t)(int8_t)o2 << 16);
}
__host__ __device__ inline void unpack_offset(uint32_t p, int &o0, int &o1, int &o2) {
    o0 = (int)(int8_t)(p & 0xff);
    o1 = (int)(int8_t)((p >> 8) & 0xff);
    o2 = (int)(int8_t)((p >> 16) & 0xff);
}

// Static description of a model's stencil shape.
template <int ModelT> struct ModelShape;
// PERIODIC: bitmask of periodic axes (the theta axis of R^2 x S^1, Curvature2.h:24-25).
template <> struct ModelShape<MODEL_RIEMANN2> {
    static constexpr int DIM = 2, NFWD = 0, NSYM = 3, SHARED = 0, PERIODIC = 0;
};
template <> struct ModelShape<MODEL_RIEMANN3> {
    static constexpr int DIM = 3, NFWD = 0, NSYM = 6, SHARED = 0, PERIODIC = 0;
};
template <> struct ModelShape<MODEL_RSF2> {
    static constexpr int DIM = 3, NFWD = 3, NSYM = 1, SHARED = 1, PERIODIC = 4;
};
template <> struct ModelShape<MODEL_ELASTICA2> {
    static constexpr int DIM = 3, NFWD = 27, NSYM = 0, SHARED = 1, PERIODIC = 4;
};

// shift_coords with the model's dimension and periodic axes known at compile time.
template <int ModelT>
__host__ __device__ inline bool shift_coords_m(const Geom &g, const int *c, int o0, int o1, int o2, int *r) {
    constexpr int DIM = ModelShape<ModelT>::DIM, PER = ModelShape<ModelT>::PERIODIC;
    const int o[3] = {o0, o1, o2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= DIM) {
            r[k] = c[k];
            continue;
        }
        int v = c[k] + o[k];
        if (PER >> k & 1) {
            const int n = g.n[k];
            if (v < 0) v += n;
            else if (v >= n) v -= n;
            if (v < 0 || v >= n) {
                v %= n;
                if (v < 0) v += n;
            }
        } else if (v < 0 || v >= g.n[k]) {
            return false;
        }
        r[k] = v;
    }
    return true;
}

// shift_coords_m without early exits, for gathers that issue every load before using any.
// Writes the shifted coordinates to r (wrapped on periodic axes) and returns whether the
// shifted point exists.  Coordinates off a closed axis are written unchanged (out of range):
// whatever is derived from them must not be used for a point that does not exist.
template <int ModelT>
__device__ inline bool shift_coords_flat(const Geom &g, const int *c, int o0, int o1, int o2, int *r) {
    constexpr int DIM = ModelShape<ModelT>::DIM, PER = ModelShape<ModelT>::PERIODIC;
    const int o[3] = {o0, o1, o2};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= DIM) {
            r[k] = c[k];
            continue;
        }
        const int n = g.n[k];
        int v = c[k] + o[k];
        if (PER >> k & 1) {
            v = v < 0 ? v + n : (v >= n ? v - n : v);
            if (v < 0 || v >= n) {  // |o| <= 127 may exceed n: full modulo (rare)
                v %= n;
                v = v < 0 ? v + n : v;
            }
        } else {
            ok = ok && (unsigned)v < (unsigned)n;  // one unsigned compare: 0 <= v < n
        }
        r[k] = v;
    }
    return ok;
}

// Device-side stencil data.
template <typename T> struct StencilData {
    // per-point (Riemann): offs[k * np + p], w[k * np + p]
    const uint32_t *__restrict__ offs;
    const T *__restrict__ w;
    // shared per theta (curvature): tab_offs[theta * ndiff + k], tab_w[theta * ndiff + k]
    const uint32_t *__restrict__ tab_offs;
    const double *__restrict__ tab_w;
    const T *__restrict__ speed;  // per point, tiled
    long long np;                 // number of tiled points (ntiles * 64)
    // Walls: per point (tiled), bit side_bit(k, s) set when side s of difference k is visible
    // (WallObstructionTest.h:96-150); null when there are no walls (everything visible).
    const uint32_t *__restrict__ vis = nullptr;
};

// Bit of side s (0: x + e_k, 1: x - e_k) of difference k, in the reference's activeNeighs layout
// (EulerianStencil.hxx:114-141): forward differences first, then two bits per symmetric one.
template <int NF> __host__ __device__ constexpr int side_bit(int k, int s) { return k < NF ? k : NF + 2 * (k - NF) + s; }

template <typename T> __device__ inline T dev_inf();
template <> __device__ inline float dev_inf<float>() { return __builtin_huge_valf(); }
template <> __device__ inline double dev_inf<double>() { return __builtin_huge_val(); }

// Local Hopf-Lax solve of sum_k w_k max(0, u - v_k)^2 = 1 over NC candidates.
// Candidates are taken in increasing v (ties: lower slot first) and accumulated while
// v < current root, in the reference's arithmetic (Headers/Base/Implementation/
// EulerianStencil.hxx:30-72): with m the smallest accepted value, a += w, b += w (v-m),
// c += (w (v-m)) (v-m) starting from c = -1, root = m + (b + sqrt(max(0, b^2 - a c))) / a.
// Zero-weight candidates are accepted like the others (the root stays +inf until a > 0).
// Returns the root (+inf if no finite candidate) and the bitmask of accumulated slots.
template <int NC, typename T>
__device__ inline T local_solve(const T (&v)[NC], const T (&w)[NC], uint32_t &used) {
    const T inf = dev_inf<T>();
    T r = inf, m = 0, a = 0, b = 0, c = -1;
    uint32_t sel = 0;
    for (int step = 0; step < NC; ++step) {
        T bv = inf, bw = 0;
        int bk = -1;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const bool ok = !((sel >> k) & 1u) && v[k] < bv;
            bv = ok ? v[k] : bv;
            bw = ok ? w[k] : bw;
            bk = ok ? k : bk;
        }
        if (bk < 0 || !(bv < r)) break;
        if (step == 0) m = bv;
        const T vr = bv - m;
        const T wv = bw * vr;
        const T wv2 = (bw * vr) * vr;
        a += bw;
        b += wv;
        c += wv2;
        const T d2 = b * b - a * c;
        const T d = sqrt(d2 > T(0) ? d2 : T(0));
        r = a > T(0) ? m + (b + d) / a : inf;  // no weight yet: +inf (QuadraticMax::Solve skips a == 0)
        sel |= 1u << bk;
    }
    used = sel;
    return r;
}

// Per-lane stencil access, shared by the tile kernels.
template <int ModelT, typename T> struct LaneStencil {
    using S = ModelShape<ModelT>;
    static constexpr int NC = S::NFWD + S::NSYM;
    uint32_t off[NC];
    T w[NC];
    uint32_t vis;  // visible sides (side_bit layout); all ones without walls

    __device__ bool visible(int k, int s) const { return (vis >> side_bit<S::NFWD>(k, s)) & 1u; }

    __device__ void load(const Geom &g, const StencilData<T> &sd, long long p, const int *c) {
        vis = sd.vis ? sd.vis[p] : 0xffffffffu;
        if (!S::SHARED) {
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                off[k] = sd.offs[k * sd.np + p];
                w[k] = sd.w[k * sd.np + p];
            }
        } else {
            int th = c[2] < g.n[2] ? c[2] : g.n[2] - 1;
            const T s = sd.speed[p];
            const T s2 = s * s;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                off[k] = sd.tab_offs[th * NC + k];
                w[k] = (T)sd.tab_w[th * NC + k] * s2;
            }
        }
    }
};

// Compare-exchange networks that sort NC values (checked exhaustively with the 0-1 principle and
// all permutations): the optimal 3-, 5- and 12-comparator networks for 3, 4 and 6 inputs, and
// odd-even transposition for other sizes up to 8.
template <int NC> struct SortNet {
    static constexpr int N = NC * (NC - 1) / 2;  // NC rounds of alternately floor(NC/2), floor((NC-1)/2) pairs
    __host__ __device__ static constexpr int pair(int q, int side) {
        // odd-even transposition: rounds r = 0..NC-1, pairs (j, j+1), j = r&1, r&1 + 2, ...
        int r = 0;
        while (true) {
            const int cnt = (NC - (r & 1)) / 2;
            if (q < cnt) return (r & 1) + 2 * q + side;
            q -= cnt;
            ++r;
        }
    }
};
template <> struct SortNet<3> {
    static constexpr int N = 3;
    __host__ __device__ static constexpr int pair(int q, int side) {
        constexpr int p[3][2] = {{1, 2}, {0, 2}, {0, 1}};
        return p[q][side];
    }
};
template <> struct SortNet<4> {
    static constexpr int N = 5;
    __host__ __device__ static constexpr int pair(int q, int side) {
        constexpr int p[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
        return p[q][side];
    }
};
template <> struct SortNet<6> {
    static constexpr int N = 12;
    __host__ __device__ static constexpr int pair(int q, int side) {
        constexpr int p[12][2] = {{0, 5}, {1, 3}, {2, 4}, {1, 2}, {3, 4}, {0, 3},
                                  {2, 5}, {0, 1}, {2, 3}, {4, 5}, {1, 2}, {3, 4}};
        return p[q][side];
    }
};

// Active differences of a local solve (the bitmask of accumulated slots), decided in FP64 --
// the reference's arithmetic -- whatever the storage precision of the values and weights.  An
// FP32 evaluation of local_solve decides "next value below the current root" against a root
// computed relative to the smallest value, which loses most of its digits when a low-weight
// candidate lies far below a heavy cluster just under the root.
template <int NC, typename T> __device__ inline uint32_t active_slots(const T (&v)[NC], const T (&w)[NC]) {
    double vd[NC], wd[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        vd[k] = (double)v[k];
        wd[k] = (double)w[k];
    }
    uint32_t used = 0;
    (void)local_solve<NC, double>(vd, wd, used);
    return used;
}

// Local solve used by the update kernel.  FP64: exactly local_solve (the reference's order of
// operations, so converged values reproduce the reference's FP64 values).  FP32: the same
// candidates in the same increasing order and the same root, but written relative to the
// largest accepted value p instead of the smallest, which removes every cancellation:
//   A = sum w_i, nB = sum w_i (p - v_i) >= 0, C = sum w_i (p - v_i)^2 < 1 over the accepted v_i;
//   the next value v (s = v - p >= 0) is below the current root iff
//     f = sum w_i (v - v_i)^2 = A s^2 + 2 nB s + C < 1,
//   and accepting it moves the pivot to v: C <- f, nB <- nB + A s, A <- A + w, p <- v;
//   the root is p + d with A d^2 + 2 nB d - (1 - C) = 0, i.e.
//     d = (1 - C) / (sqrt(nB^2 + A (1 - C)) + nB).
// Every sum adds terms of one sign.  The form relative to the smallest value m (the reference's)
// computes b^2 - a (c - 1) as the difference of two terms that are both about a^2 (p - m)^2: with
// a candidate of weight 4.8 at 0.005 below five of total weight 11575 at 0.4618 (Elastica2
// 192 x 192 x 64 beside the seed, across the periodic theta axis), FP32 loses 3e-5 of the value.
// One hardware square root and reciprocal per solve (1 ulp each).  Explicit fused multiply-adds:
// FP32 only, never on the FP64 path.
template <int NC, typename T> __device__ inline T local_solve_update(const T (&v)[NC], const T (&w)[NC]) {
    uint32_t used;
    return local_solve<NC, T>(v, w, used);
}
__device__ inline float root_from_pivot(float p, float A, float nB, float C) {
    const float e = 1.f - C;
    const float d2 = __builtin_fmaf(A, e, nB * nB);
    return p + e * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(d2) + nB);
}
template <int NC> __device__ inline float local_solve_update(const float (&v0)[NC], const float (&w0)[NC]) {
    const float inf = __builtin_huge_valf();
    float v[NC], w[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        v[k] = v0[k];
        w[k] = w0[k];
    }
    if constexpr (NC <= 8) {
        // sort (v, w) pairs by v with a compare-exchange network (SortNet<NC>)
#pragma unroll
        for (int q = 0; q < SortNet<NC>::N; ++q) {
            const int i = SortNet<NC>::pair(q, 0), j = SortNet<NC>::pair(q, 1);
            const float a0 = v[i], a1 = v[j], b0 = w[i], b1 = w[j];
            const bool sw = a1 < a0;
            v[i] = sw ? a1 : a0;
            v[j] = sw ? a0 : a1;
            w[i] = sw ? b1 : b0;
            w[j] = sw ? b0 : b1;
        }
        if (!(v[0] < inf)) return inf;
        float p = v[0], A = w[0], nB = 0.f, C = 0.f;
        bool go = true;
#pragma unroll
        for (int k = 1; k < NC; ++k) {
            const float s = v[k] - p;  // >= 0; +inf for a missing neighbour (then f = +inf)
            const float f = __builtin_fmaf(A * s, s, __builtin_fmaf(2.f * nB, s, C));
            go = go && f < 1.f;
            C = go ? f : C;
            nB = go ? __builtin_fmaf(A, s, nB) : nB;
            A = go ? A + w[k] : A;
            p = go ? v[k] : p;
        }
        return A > 0.f ? root_from_pivot(p, A, nB, C) : inf;  // no weight accepted: +inf
    } else {
        float p = 0.f, A = 0.f, nB = 0.f, C = 0.f;
        uint32_t sel = 0;
        bool any = false;
        for (int step = 0; step < NC; ++step) {
            float bv = inf, bw = 0.f;
            int bk = -1;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const bool ok = !((sel >> k) & 1u) && v[k] < bv;
                bv = ok ? v[k] : bv;
                bw = ok ? w[k] : bw;
                bk = ok ? k : bk;
            }
            if (bk < 0) break;
            if (step == 0) {
                p = bv;
                A = bw;
            } else {
                const float s = bv - p;
                const float f = __builtin_fmaf(A * s, s, __builtin_fmaf(2.f * nB, s, C));
                if (!(f < 1.f)) break;
                C = f;
                nB = __builtin_fmaf(A, s, nB);
                A += bw;
                p = bv;
            }
            sel |= 1u << bk;
            any = true;
        }
        return any && A > 0.f ? root_from_pivot(p, A, nB, C) : inf;
    }
}

// Static factor (StaticFactoring.h:20-63): values T, gradients of T in grid units (dim per point)
// and mask over a box of bd points whose first grid index is `shift` (sub), or over the whole grid
// (sub = 0, bd = the grid's extents), row-major over the box.  on = 0: no factoring.
struct FactorDev {
    const double *vals = nullptr;
    const double *grads = nullptr;
    const unsigned char *mask = nullptr;
    int bd[3] = {1, 1, 1}, shift[3] = {0, 0, 0};
    int sub = 0, on = 0;
};

// Launch-wide reductions (mean one-step time, widest stencil) go to RED_SLOTS accumulators, each
// on its own 256-byte line, chosen by the global wave index.  A single accumulator would
// serialise every wave of the launch on one address (tens of thousands of same-address atomics
// per 128^3 grid); the host sums the slots after the launch.
constexpr int RED_SLOTS = 256;
struct alignas(256) RedSlot {
    double sum;
    double count;
    int l1max;
    int pad_[59];
};
static_assert(sizeof(RedSlot) == 256, "RedSlot is one 256-byte line");

// Slot of the final residual check (same RED_SLOTS spread, same buffer): largest relative
// deviation (bits of a nonnegative double) and finite / infinite mismatches.
struct alignas(256) ResSlot {
    unsigned long long maxbits;
    unsigned long long mismatch;
    unsigned long long pad_[30];
};
static_assert(sizeof(ResSlot) == sizeof(RedSlot), "ResSlot reuses the RedSlot buffer");

#if defined(__HIPCC__)
__device__ inline RedSlot *red_slot(RedSlot *red) {
    const unsigned wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    return red + (wave & (RED_SLOTS - 1));
}

// Wave-reduced accumulation of (sum, count) into this wave's slot; every lane of the wave calls it.
__device__ inline void wave_add_tau(double v, double c, RedSlot *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o);
        c += __shfl_xor(c, o);
    }
    if ((threadIdx.x & 63) == 0 && c > 0.) {
        RedSlot *r = red_slot(red);
        atomicAdd(&r->sum, v);
        atomicAdd(&r->count, c);
    }
}

// Wave-reduced maximum into this wave's slot; every lane of the wave calls it.
__device__ inline void wave_max_l1(int v, RedSlot *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int t = __shfl_xor(v, o);
        v = t > v ? t : v;
    }
    if ((threadIdx.x & 63) == 0 && v > 0) atomicMax(&red_slot(red)->l1max, v);
}
#endif

// Host-side total of the slots (copied back after the launch).
struct RedTotals {
    double sum = 0., count = 0.;
    int l1max = 0;
};
inline RedTotals red_total(const RedSlot *h) {
    RedTotals t;
    for (int i = 0; i < RED_SLOTS; ++i) {
        t.sum += h[i].sum;
        t.count += h[i].count;
        t.l1max = h[i].l1max > t.l1max ? h[i].l1max : t.l1max;
    }
    return t;
}

}  // namespace hfm

#endif
