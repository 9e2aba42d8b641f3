// hfm_setup_dev.h -- device routines shared by the solve-setup kernels of a plan (hfm_kernels.hip)
// and by the server's setup launches that cover a group of requests at once (hfm_server.hip):
// the Riemann stencil of one point, the set of source tiles of one tile, and the chunked
// exclusive scan.
#pragma once

#include <hip/hip_runtime.h>

#include "hfm_device.h"
#include "hfm_kernels.h"
#include "selling.h"

namespace hfm {

// FP64 Selling stencil of the in-grid point with coordinates c (HFM_StencilRecompute.hxx via
// Riemannian.hxx: offsets of the obtuse superbase of D^{-1}, weights from its scalar products).
// lin receives the point's row-major index; returns the Selling status.
template <int D>
__device__ inline int riemann_point(const Geom &g, const double *__restrict__ metric, int metric_is_field, int dual,
                                    const InvHH &ih, const int *c, int8_t (*o8)[D], double *wd, long long &lin) {
    constexpr int K = D * (D + 1) / 2;
    lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
    double mc[K];
    const double *src = metric + (metric_is_field ? lin * K : 0);
    for (int k = 0; k < K; ++k) mc[k] = src[k];
    int mx = 0;
    return riemann_point_stencil<D>(mc, dual != 0, ih.v, o8, wd, &mx);
}

template <int D> __device__ inline uint32_t packed_offset(const int8_t (*o8)[D], int k) {
    return pack_offset(o8[k][0], o8[k][1], D > 2 ? o8[k][D > 2 ? 2 : 0] : 0);
}

// ---------------------------------------------------------------------------------------------
// Source tiles of a tile: the tiles other than T holding x +- e_k for the points x of T
// (nonzero weights only), gathered by the tile's wave in an LDS hash set of HASH_SLOTS ints
// (-1 = empty).  A tile has at most 64 points x 12 sides = 768 distinct sources, so 1024 slots
// always leave a free slot for the linear probing.
// ---------------------------------------------------------------------------------------------
constexpr int HASH_SLOTS = 1024;

__device__ inline void src_set_clear(int *set, int lane) {
    for (int i = lane; i < HASH_SLOTS; i += 64) set[i] = -1;
}

__device__ inline void src_set_add(int *set, int t) {
    unsigned h = ((unsigned)t * 2654435761u) & (HASH_SLOTS - 1);
    for (int probe = 0; probe < HASH_SLOTS; ++probe) {
        const int prev = atomicCAS(&set[h], -1, t);
        if (prev == -1 || prev == t) break;
        h = (h + 1) & (HASH_SLOTS - 1);
    }
}

// Adds the tiles of the in-domain points c + s*o (s = 1, and s = -1 when both sides count)
// other than `tile`.
template <int ModelT>
__device__ inline void src_set_add_sides(const Geom &g, const int *c, int tile, int o0, int o1, int o2, int nsides,
                                         int *set) {
    for (int s = 0; s < nsides; ++s) {
        const int sg = s == 0 ? 1 : -1;
        int r[3];
        if (!shift_coords_m<ModelT>(g, c, sg * o0, sg * o1, sg * o2, r)) continue;
        const int t = tile_of(g, r);
        if (t != tile) src_set_add(set, t);
    }
}

// Compacts the set to set[0..nsrc) and returns nsrc on every lane, so the global atomics that
// follow take ceil(sources / 64) passes with every lane busy.  Contains barriers: every thread
// of the (64-thread) block calls it.
__device__ inline int src_set_compact(int *set, int lane) {
    int mine[HASH_SLOTS / 64];
    int nmine = 0;
#pragma unroll
    for (int j = 0; j < HASH_SLOTS / 64; ++j) {
        const int v = set[j * 64 + lane];
        mine[j] = v;
        nmine += v >= 0;
    }
    int incl = nmine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    const int nsrc = __shfl(incl, 63);
    __syncthreads();
    int pos = incl - nmine;
#pragma unroll
    for (int j = 0; j < HASH_SLOTS / 64; ++j)
        if (mine[j] >= 0) set[pos++] = mine[j];
    __syncthreads();
    return nsrc;
}

// Source tiles of `tile` from a stored stencil, compacted to set[0..nsrc); returns nsrc.
// Contains barriers: every thread of the (64-thread) block calls it.
template <int ModelT, typename T>
__device__ inline int tile_sources(const Geom &g, const StencilData<T> &sd, int tile, int lane, int *set) {
    using S = ModelShape<ModelT>;
    constexpr int NC = S::NFWD + S::NSYM;
    src_set_clear(set, lane);
    __syncthreads();
    const long long p = (long long)tile * 64 + lane;
    int c[3];
    tiled_coords(g, tile, lane, c);
    if (in_grid(g, c)) {
        LaneStencil<ModelT, T> st;
        st.load(g, sd, p, c);
        for (int k = 0; k < NC; ++k) {
            if (st.w[k] == T(0)) continue;
            int o0, o1, o2;
            unpack_offset(st.off[k], o0, o1, o2);
            src_set_add_sides<ModelT>(g, c, tile, o0, o1, o2, (k >= S::NFWD) ? 2 : 1, set);
        }
    }
    __syncthreads();
    return src_set_compact(set, lane);
}

// ---------------------------------------------------------------------------------------------
// Exclusive scans of int counts.  The unit of work is a chunk of 4096 consecutive counts taken
// by a block of 1024 threads (four consecutive counts per thread, coalesced), scanned with wave
// shuffles plus one LDS pass over the 16 wave totals.
// ---------------------------------------------------------------------------------------------
constexpr int SCAN_CHUNK = 4096;

// Scans chunk [base, base + 4096) of cnt into start with the given carry; returns the chunk total
// (on every thread).  Contains barriers: every thread of the 1024-thread block calls it.
__device__ inline int scan_chunk(const int *__restrict__ cnt, int *__restrict__ start, int n, int base, int carry,
                                 int *wsum) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int i0 = base + tid * 4;
    int v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = i0 + k < n ? cnt[i0 + k] : 0;
    const int tsum = v[0] + v[1] + v[2] + v[3];
    int incl = tsum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int wpre = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
        const int x = wsum[w];
        wpre += w < wid ? x : 0;
        total += x;
    }
    if (start) {
        int run = carry + wpre + incl - tsum;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < n) start[i0 + k] = run;
            run += v[k];
        }
    }
    __syncthreads();  // wsum is reused by the next chunk
    return total;
}

}  // namespace hfm
