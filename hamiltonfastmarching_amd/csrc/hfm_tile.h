// hfm_tile.h -- per-tile update of the block fast-iterative method and its scheduling helpers,
// shared by the single-solve kernels (hfm_kernels.hip) and the multi-solve serving kernels
// (hfm_server.hip).
#ifndef HFM_AMD_TILE_H
#define HFM_AMD_TILE_H

#include "hfm_device.h"
#include "hfm_kernels.h"

namespace hfm {

// Scheduling keys: float values mapped to order-preserving uint32 so that atomicMin works for
// negative seed values too.  KEY_NONE (0xffffffff) means "nothing pending".
__device__ inline unsigned key_of(float f) {
    const unsigned b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float key_value(unsigned k) {
    if (k == KEY_NONE) return __builtin_huge_valf();
    const unsigned b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(b);
}
// Wave-wide minimum / maximum, returned to every lane; every lane of the wave must be active.
// DPP inclusive scan: rows of 16 lanes with row_shr 1, 2, 4, 8, then across rows with
// row_bcast 15 (into rows 1 and 3) and row_bcast 31 (into rows 2 and 3), after which lane 63
// holds the result.  Lanes without a source keep the identity.  Six dependent VALU steps
// instead of six LDS permutes, each waiting for the LDS.
template <bool MAX> __device__ inline unsigned wave_reduce_u32(unsigned v) {
    constexpr int id = MAX ? 0 : (int)0xffffffffu;
#define HFM_DPP_STEP(ctrl, rows)                                                                    \
    {                                                                                              \
        const unsigned w = (unsigned)__builtin_amdgcn_update_dpp(id, (int)v, ctrl, rows, 0xf, false); \
        v = MAX ? (w > v ? w : v) : (w < v ? w : v);                                               \
    }
    HFM_DPP_STEP(0x111, 0xf)  // row_shr:1
    HFM_DPP_STEP(0x112, 0xf)  // row_shr:2
    HFM_DPP_STEP(0x114, 0xf)  // row_shr:4
    HFM_DPP_STEP(0x118, 0xf)  // row_shr:8
    HFM_DPP_STEP(0x142, 0xa)  // row_bcast:15
    HFM_DPP_STEP(0x143, 0xc)  // row_bcast:31
#undef HFM_DPP_STEP
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ inline unsigned wave_min_u32(unsigned v) { return wave_reduce_u32<false>(v); }
__device__ inline unsigned wave_max_u32(unsigned v) { return wave_reduce_u32<true>(v); }
__device__ inline unsigned relaxed_load(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Mark tile d (d < 0: none) pending with key `key`: lowers keys[d] unless the tile's largest
// value is already <= key (then no point of d can use a value >= key).  Returns the smallest key
// marked by the wave (KEY_NONE if none); every lane of the wave calls it.
__device__ inline unsigned mark_pending(int d, unsigned key, unsigned *__restrict__ keys,
                                        const unsigned *__restrict__ tmax) {
    bool lowered = false;
    if (d >= 0 && relaxed_load(&tmax[d]) > key && relaxed_load(&keys[d]) > key) {
        atomicMin(&keys[d], key);
        lowered = true;
    }
    return wave_min_u32(lowered ? key : KEY_NONE);
}

// The smallest pending key is accumulated in PEND_SLOTS words, `stride` words apart (one per
// block or wave id modulo PEND_SLOTS), so that no single address takes an atomic from every wave.
__device__ inline unsigned read_pend(const unsigned *slots, int stride = 1) {
    unsigned m = KEY_NONE;
#pragma unroll
    for (int i = 0; i < PEND_SLOTS; ++i) {
        const unsigned v = slots[i * stride];
        m = v < m ? v : m;
    }
    return m;
}

// End of an update launch: the block's four waves fold their smallest marked key and their
// visit / sweep counts in LDS, and one thread adds them to the block's slot (one atomic per
// counter per block instead of per wave).  Every thread of the block calls it.
__device__ inline void block_flush(unsigned wmin, unsigned long long visits, unsigned long long sweeps,
                                   unsigned *__restrict__ pend_out, unsigned long long *__restrict__ stats) {
    __shared__ unsigned s_min[4];
    __shared__ unsigned long long s_vis[4], s_sw[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        s_min[wv] = wmin;
        s_vis[wv] = visits;
        s_sw[wv] = sweeps;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned m = KEY_NONE;
        unsigned long long v = 0, sw = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            m = s_min[w] < m ? s_min[w] : m;
            v += s_vis[w];
            sw += s_sw[w];
        }
        const int slot = blockIdx.x & (PEND_SLOTS - 1);
        if (m != KEY_NONE) atomicMin(&pend_out[slot * PEND_STRIDE], m);
        if (v) {
            atomicAdd(&stats[slot * STAT_STRIDE], v);
            atomicAdd(&stats[slot * STAT_STRIDE + 1], sw);
        }
    }
}

// Band bound of a round: a window sliding with the smallest pending key, never moving back --
// max(previous bound, smallest pending key + delta); +inf (every pending tile) when delta <= 0.
// A bound that waited at the previous value until every key below it was consumed spent many
// rounds on the last few tiles of each band (Riemann3 128^3: 104 rounds against 80).  The
// bound exceeds the smallest pending key, so a round lists at least one tile while any is
// pending.
__device__ inline float band_bound(float pend, float bprev, float delta) {
    float B = __builtin_huge_valf();
    if (delta > 0.f && pend < __builtin_huge_valf()) B = fmaxf(bprev, pend + delta);
    return B;
}

// Band width of this round, adapted to the number of tiles the previous round listed: x1.5
// below target/2, /1.5 above 2*target, kept within [dmin, dmax].  Every block of a round
// computes the same value from the same inputs.
__device__ inline float adapt_delta(float dprev, int prev_count, int target, float dmin, float dmax) {
    float d = dprev;
    if (prev_count < target / 2) d *= 1.5f;
    else if (prev_count > 2 * target) d *= (1.f / 1.5f);
    return fminf(fmaxf(d, dmin), dmax);
}

// Whether a tile with key k is taken by a round of bound B.
__device__ inline bool band_take(unsigned k, float B) {
    return k != KEY_NONE && (key_value(k) < B || !(B < __builtin_huge_valf()));
}

// Scheduling kernels run blocks of SCHED_BLOCK threads, one tile per thread.
constexpr int SCHED_BLOCK = 1024;

// Block-level aggregation of a scheduling pass (every thread of the block calls it once):
// `take` = the thread's tile is listed this round, `left` = the key it leaves pending
// (KEY_NONE if none).  One atomicAdd per block reserves the block's list entries (and adds them
// to *solve_count and *visits if given), one atomicMin folds the block's smallest left key into
// *pend_slot.  Returns the list position of a taken tile.  Per-wave atomics on the shared list
// counter would serialise thousands of waves on one address.
__device__ inline int block_take(bool take, unsigned left, int *__restrict__ list_count,
                                 int *__restrict__ solve_count, unsigned *__restrict__ pend_slot,
                                 unsigned long long *__restrict__ visits = nullptr) {
    constexpr int NW = SCHED_BLOCK / 64;
    __shared__ int wcnt[NW];
    __shared__ unsigned wmn[NW];
    __shared__ int base_s;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long mask = __ballot(take);
    const unsigned lmin = wave_min_u32(left);
    if (lane == 0) {
        wcnt[wid] = __popcll(mask);
        wmn[wid] = lmin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        unsigned m = KEY_NONE;
        for (int w = 0; w < NW; ++w) {
            tot += wcnt[w];
            m = wmn[w] < m ? wmn[w] : m;
        }
        base_s = tot ? atomicAdd(list_count, tot) : 0;
        if (tot && solve_count) atomicAdd(solve_count, tot);
        if (tot && visits) atomicAdd(visits, (unsigned long long)tot);
        if (m != KEY_NONE) atomicMin(pend_slot, m);
    }
    __syncthreads();
    int pre = 0;
    for (int w = 0; w < wid; ++w) pre += wcnt[w];
    const unsigned long long below = lane == 0 ? 0ull : (mask & (~0ull >> (64 - lane)));
    return base_s + pre + __popcll(below);
}

// LDS row of one wave: the tile's 64 values, then a +inf sentinel read by out-of-tile sides.
constexpr int SU_SENTINEL = 64;
constexpr int SU_ROW = 65;

// Minimum without NaN handling (no operand is NaN here): a compare and a select.
template <typename T> __device__ inline T min_nn(T a, T b) { return b < a ? b : a; }
__device__ inline float min_nn(float a, float b) { return __builtin_fminf(a, b); }  // one v_min_f32

template <typename T> struct TileCtx {
    Geom g;
    StencilData<T> sd;
    T *__restrict__ u;
    const unsigned long long *__restrict__ fixedmask;
    const int *__restrict__ dep_start;
    const int *__restrict__ deps;
    unsigned *__restrict__ keys;
    unsigned *__restrict__ tmax;
};

// One visit of `tile` by one wavefront (lane = 0..63, su = the wave's SU_ROW-entry LDS row).
//
// The tile's values are staged in LDS, the values of neighbours outside the tile are cached in
// registers, and Jacobi sweeps of the local solve run until no lane changes (wave ballot) or
// MAXSWEEP sweeps.  If the tile changed, its dependent tiles are marked pending with key =
// smallest lowered value (the tile itself too if it did not converge); wmin is lowered to the
// smallest key marked.  Values of other tiles may be read while another wave lowers them: any
// value read is an upper bound of the solution and the scheme is monotone, and the writer marks
// its dependents pending, so the fixed point is unaffected.  Returns the number of sweeps.
template <int ModelT, typename T, int MAXSWEEP>
__device__ inline int fim_tile(const TileCtx<T> &ctx, int tile_in, int lane, T *su, unsigned &wmin) {
    const int tile = __builtin_amdgcn_readfirstlane(tile_in);  // uniform: scalar address arithmetic
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NS = S::NSYM, NC = NF + NS;
    const T inf = dev_inf<T>();
    const Geom &g = ctx.g;
    T *__restrict__ u = ctx.u;
    const long long p = (long long)tile * 64 + lane;
    int c[3];
    tiled_coords(g, tile, lane, c);
    const bool fixed = (ctx.fixedmask[tile] >> lane) & 1ull;
    const int db = ctx.dep_start[tile], de = ctx.dep_start[tile + 1];
    // Loads in order of need, none waiting on another except where an address depends on a
    // value: own value and stencil; the first 64 dependent tiles; then (below) the neighbour
    // values, whose addresses come from the stencil, and the dependents' key and tile maximum,
    // whose addresses come from the dependent list.  The marks after the sweeps use the last
    // two: within one update launch keys and tile maxima only decrease, so values read early
    // are upper bounds and filter conservatively.
    const T u0 = u[p];
    LaneStencil<ModelT, T> st;
    st.load(g, ctx.sd, p, c);
    const bool has_d0 = db + lane < de;
    int d0raw = tile;
    if (de > db) d0raw = ctx.deps[has_d0 ? db + lane : de - 1];  // uniform branch
    // Neighbour slots: forward k -> one side, symmetric k -> two sides.  A side inside the tile
    // reads the wave's LDS row (ia/ib = its lane); a side outside reads the row's +inf sentinel
    // (SU_SENTINEL) and contributes its value, frozen for the visit, through hk = the smaller of
    // the outside values.  Each sweep then takes v_k = min(su[ia_k], su[ib_k], hk) without a
    // branch, which is the smaller neighbour value of difference k exactly.
    //
    // The gather is straight-line code: every side computes an address (its own point when the
    // side is absent, inside the tile or invisible), all loads are issued, and only then are the
    // loaded values selected.  Loads under per-side branches made the compiler wait for each
    // side's load before the next one, one memory latency per side.
    int ia[NC], ib[NC];
    T hk[NC];
    unsigned tmax0, key0;
    {
        // point indices fit 32 bits: at most 2^26 tiles of 64 points
        unsigned qa[NC], qb[NC];
        const unsigned ps = (unsigned)p;
        bool xa[NC], xb[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            int o0, o1, o2;
            unpack_offset(st.off[k], o0, o1, o2);
            const bool act = !fixed && st.w[k] != T(0);
            int r[3];
            const bool oka = shift_coords_flat<ModelT>(g, c, o0, o1, o2, r) && act && st.visible(k, 0);
            const int ta = tile_of(g, r), la = local_of(g, r);
            ia[k] = oka && ta == tile ? la : SU_SENTINEL;
            xa[k] = oka && ta != tile;
            qa[k] = xa[k] ? (unsigned)ta * 64u + (unsigned)la : ps;
            if (k >= NF) {
                const bool okb = shift_coords_flat<ModelT>(g, c, -o0, -o1, -o2, r) && act && st.visible(k, 1);
                const int tb = tile_of(g, r), lb = local_of(g, r);
                ib[k] = okb && tb == tile ? lb : SU_SENTINEL;
                xb[k] = okb && tb != tile;
                qb[k] = xb[k] ? (unsigned)tb * 64u + (unsigned)lb : ps;
            } else {
                ib[k] = SU_SENTINEL;
                xb[k] = false;
                qb[k] = ps;
            }
        }
        T va[NC], vb[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            va[k] = u[qa[k]];
            vb[k] = k >= NF ? u[qb[k]] : inf;
        }
        const int dq = has_d0 ? d0raw : tile;
        tmax0 = ctx.tmax[dq];
        key0 = ctx.keys[dq];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const T oa = xa[k] ? va[k] : inf, ob = xb[k] ? vb[k] : inf;
            hk[k] = oa < ob ? oa : ob;
        }
    }
    T ucur = u0;
    su[lane] = ucur;
    if (lane == 0) su[SU_SENTINEL] = inf;
    bool converged = false;
    int sweeps = 0;
    for (int it = 0; it < MAXSWEEP; ++it) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        T v[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const T va = su[ia[k]];
            const T vab = (k >= NF) ? min_nn(va, su[ib[k]]) : va;
            v[k] = min_nn(vab, hk[k]);
        }
        const T r = local_solve_update<NC>(v, st.w);
        const bool ch = !fixed && r < ucur;
        ucur = ch ? r : ucur;
        ++sweeps;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        su[lane] = ucur;
        if (__ballot(ch) == 0ull) {
            converged = true;
            break;
        }
    }
    const bool lowered = ucur < u0;
    if (lowered) u[p] = ucur;
    const bool tile_changed = __ballot(lowered) != 0ull;
    if (tile_changed) {
        const unsigned tm = wave_max_u32(key_of((float)ucur));
        if (lane == 0) ctx.tmax[tile] = tm;
        const unsigned kmin = wave_min_u32(lowered ? key_of((float)ucur) : KEY_NONE);
        {
            const bool lw = has_d0 && tmax0 > kmin && key0 > kmin;
            if (lw) atomicMin(&ctx.keys[d0raw], kmin);
            const unsigned m = wave_min_u32(lw ? kmin : KEY_NONE);
            wmin = m < wmin ? m : wmin;
        }
        for (int j0 = db + 64; j0 < de; j0 += 64) {
            const int j = j0 + lane;
            const unsigned m = mark_pending(j < de ? ctx.deps[j] : -1, kmin, ctx.keys, ctx.tmax);
            wmin = m < wmin ? m : wmin;
        }
        if (!converged) {
            const unsigned m = mark_pending(lane == 0 ? tile : -1, kmin, ctx.keys, ctx.tmax);
            wmin = m < wmin ? m : wmin;
        }
    }
    return sweeps;
}

}  // namespace hfm

#endif
