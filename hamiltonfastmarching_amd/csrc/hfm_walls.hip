// hfm_walls.hip -- obstacles: which differences of a point's stencil are visible.
//
// The reference's WallObstructionTest (Headers/ExtraAlgorithms/WallObstructionTest.h) keeps,
// per point, the grid L1 distance to the walls (unit steps along the axes, periodic axes
// wrapped) up to MaxStencilWidth, stored as its ShortType int8 with 127 meaning "farther"
// (:38-88); a difference from an accepted point y to the point x = y + v is visible unless the
// discrete segment from y to x meets a wall (Visible, :96-150), and an invisible difference is
// never used to update x (VisibleOffset, HamiltonFastMarching.hxx:160-170).
//
// Here visibility is a property of (x, difference, side) only, so it is computed once per
// stencil set as a per-point bitmask in the activeNeighs layout and read by every kernel that
// gathers neighbour values:
//   k_wall_init    distance 0 on walls, 127 elsewhere (tiled layout);
//   k_wall_relax   maxDist Jacobi passes d(x) = min(d(x), d(n) + 1 if <= maxDist) over the 2 d
//                  axis neighbours -- after maxDist passes this is the breadth-first distance;
//   k_wall_visibility  per point and side, the reference's line walk on the distances.
#include "hfm_internal.h"
#include "hfm_kernels.h"

namespace hfm {

constexpr int WALL_FAR = 127;  // std::numeric_limits<int_least8_t>::max()

__global__ __launch_bounds__(256) void k_wall_init(Geom g, const uint8_t *__restrict__ walls,
                                                   int8_t *__restrict__ dist, long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    int8_t d = (int8_t)WALL_FAR;
    if (in_grid(g, c)) {
        const long long lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
        if (walls[lin]) d = 0;
    }
    dist[p] = d;
}

__global__ __launch_bounds__(256) void k_wall_relax(Geom g, const int8_t *__restrict__ din,
                                                    int8_t *__restrict__ dout, int maxDist, long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    int d = din[p];
    if (in_grid(g, c)) {
        for (int i = 0; i < g.dim; ++i)
            for (int e = -1; e <= 1; e += 2) {
                int r[3];
                const int o[3] = {i == 0 ? e : 0, i == 1 ? e : 0, i == 2 ? e : 0};
                if (!shift_coords(g, c, o[0], o[1], o[2], r)) continue;
                const int dn = din[(long long)tile_of(g, r) * 64 + local_of(g, r)] + 1;
                if (dn <= maxDist && dn < d) d = dn;
            }
    }
    dout[p] = (int8_t)d;
}

__device__ inline int wall_dist_at(const Geom &g, const int8_t *dist, const int *c) {
    return dist[(long long)tile_of(g, c) * 64 + local_of(g, c)];
}

// Visible(base, v, sum) (WallObstructionTest.h:96-150): base = the accepted neighbour, sum =
// base + v (the updated point, in the domain).  Integer arithmetic as in the reference.
__device__ inline bool wall_visible(const Geom &g, const int8_t *dist, bool refine, const int *base, const int *v,
                                    const int *sum) {
    int wd = wall_dist_at(g, dist, sum);
    if (wd == WALL_FAR) return true;
    if (wd == 0) return false;
    if (refine && wall_dist_at(g, dist, base) == 1) return true;
    int sign[3] = {0, 0, 0};
    int vSum = 0, vProd = 1, vMax = 0;
    for (int i = 0; i < g.dim; ++i)
        if (v[i] != 0) {
            sign[i] = v[i] > 0 ? 1 : -1;
            const int a = v[i] > 0 ? v[i] : -v[i];
            vSum += a;
            vProd *= a;
            vMax = a > vMax ? a : vMax;
        }
    if (vSum <= wd) return true;
    // Straight discrete path from base towards sum: axis j steps when its next crossing
    // vProd / |v_j| * (2 m + 1) is reached by the threshold 2 i vProd / vMax.
    int pos[3] = {base[0], base[1], base[2]};
    long long next[3], inc[3];
    for (int i = 0; i < 3; ++i) {
        next[i] = (i < g.dim && v[i] != 0) ? vProd / (v[i] > 0 ? v[i] : -v[i]) : 0x7fffffffLL;
        inc[i] = 2 * next[i];
    }
    for (int i = 1; i < vMax; ++i) {
        const long long threshold = (2LL * i * vProd) / vMax;
        for (int j = 0; j < g.dim; ++j)
            if (next[j] <= threshold) {
                next[j] += inc[j];
                --vSum;
                pos[j] += sign[j];
            }
        int r[3];
        if (!shift_coords(g, pos, 0, 0, 0, r)) return true;  // not reached: the segment stays in the domain
        wd = wall_dist_at(g, dist, r);
        if (vSum <= wd) return true;
        if (wd == 0) return false;
    }
    return true;
}

template <int ModelT, typename T>
__global__ __launch_bounds__(256) void k_wall_visibility(Geom g, StencilData<T> sd, const int8_t *__restrict__ dist,
                                                         int refine, uint32_t *__restrict__ vis, long long np) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    if (!in_grid(g, c)) {
        vis[p] = 0u;
        return;
    }
    sd.vis = nullptr;
    LaneStencil<ModelT, T> st;
    st.load(g, sd, p, c);
    uint32_t m = 0;
    for (int k = 0; k < NC; ++k) {
        int o[3];
        unpack_offset(st.off[k], o[0], o[1], o[2]);
        for (int s = 0; s < (k >= NF ? 2 : 1); ++s) {
            const int sg = s == 0 ? 1 : -1;
            int y[3];
            if (!shift_coords_m<ModelT>(g, c, sg * o[0], sg * o[1], sg * o[2], y)) continue;
            const int v[3] = {-sg * o[0], -sg * o[1], -sg * o[2]};
            if (wall_visible(g, dist, refine != 0, y, v, c)) m |= 1u << side_bit<NF>(k, s);
        }
    }
    vis[p] = m;
}

template <int M, typename T>
static void visibility_one(const Geom &g, const StencilData<T> &sd, const int8_t *dist, int refine, uint32_t *vis,
                           long long np, hipStream_t s) {
    hipLaunchKernelGGL((k_wall_visibility<M, T>), dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, g, sd, dist,
                       refine, vis, np);
}

template <typename T>
hipError_t launch_wall_visibility(int model, const Geom &g, const StencilData<T> &sd, const uint8_t *walls,
                                  int maxDist, int refine, int8_t *dist0, int8_t *dist1, uint32_t *vis, long long np,
                                  hipStream_t s) {
    const unsigned nb = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_wall_init, dim3(nb), dim3(256), 0, s, g, walls, dist0, np);
    int8_t *a = dist0, *b = dist1;
    for (int it = 0; it < maxDist; ++it) {
        hipLaunchKernelGGL(k_wall_relax, dim3(nb), dim3(256), 0, s, g, a, b, maxDist, np);
        int8_t *t = a;
        a = b;
        b = t;
    }
    switch (model) {
        case MODEL_RIEMANN2: visibility_one<MODEL_RIEMANN2, T>(g, sd, a, refine, vis, np, s); break;
        case MODEL_RIEMANN3: visibility_one<MODEL_RIEMANN3, T>(g, sd, a, refine, vis, np, s); break;
        case MODEL_RSF2: visibility_one<MODEL_RSF2, T>(g, sd, a, refine, vis, np, s); break;
        case MODEL_ELASTICA2: visibility_one<MODEL_ELASTICA2, T>(g, sd, a, refine, vis, np, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_wall_visibility<float>(int, const Geom &, const StencilData<float> &, const uint8_t *, int,
                                                  int, int8_t *, int8_t *, uint32_t *, long long, hipStream_t);
template hipError_t launch_wall_visibility<double>(int, const Geom &, const StencilData<double> &, const uint8_t *,
                                                   int, int, int8_t *, int8_t *, uint32_t *, long long, hipStream_t);

}  // namespace hfm
