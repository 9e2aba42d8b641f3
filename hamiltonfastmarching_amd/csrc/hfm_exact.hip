// hfm_exact.hip -- the reference's serial acceptance loop on the device, for the options whose
// result is fixed by the order in which the heap accepts points: the second / third order
// recompute at acceptance (`order` = 2 or 3, Headers/Base/Implementation/HamiltonFastMarching.hxx
// :59-123 with Recompute :187-272 and EulerianStencil.hxx:155-233), static factoring (the same
// recompute with corrected differences, Headers/ExtraAlgorithms/Implementation/
// StaticFactoring.hxx:11-55) and the stopping criteria (ExtraAlgorithms/CommonStoppingCriteria.h).
//
// With order >= 2 a point's value is recomputed when it leaves the heap, from its active
// differences and, where the next points along an offset are already accepted and no larger,
// second (third) order combinations of their values.  The recomputed value replaces the one every
// later update reads, and the corrected values are not monotone in the neighbours: there is no
// order-independent fixed point for a parallel iteration to converge to, and pops whose key lies
// below the previous pop's (0.2 to 1.2 % of them on the test fields) change which neighbours
// count as accepted.  The loop therefore runs as the reference runs it: one wave pops a binary
// heap -- the push / pop steps of libstdc++'s std::priority_queue, so that equal keys leave in
// the same order -- updates the popped point's reversed neighbours with the incremental
// quadratic of HopfLaxUpdate, and recomputes each point at acceptance.  Every value is FP64, in
// the reference's order of operations.  It is correct rather than fast: the pops form one
// dependent chain that a single wave runs step by step.
#include "hfm_kernels.h"

#include "hfm_device.h"
#include "hfm_recompute.h"

namespace hfm {

namespace {

struct HeapEntry {
    double value;
    long long p;  // tiled point index
};

// The queue's order (QueueElement::operator<, HamiltonFastMarching.hxx:51-56): a larger value
// has the lower priority.
__device__ inline bool heap_less(const HeapEntry &a, const HeapEntry &b) { return a.value > b.value; }

// Tiled index of the point c + s * o, or -1 off a closed axis.
template <int ModelT> __device__ inline long long shifted(const Geom &g, const int *c, int o0, int o1, int o2, int s) {
    int r[3];
    if (!shift_coords_m<ModelT>(g, c, s * o0, s * o1, s * o2, r)) return -1;
    return (long long)tile_of(g, r) * 64 + local_of(g, r);
}

// Heap entries 0 .. HEAP_LDS - 1 (the top 13 levels) live in LDS, the rest in device memory:
// a pop walks from the root, so most of its dependent loads stay in LDS.  The LDS part is a
// file-scope LDS array, so that its accesses compile to LDS instructions (a pointer that may
// point to either memory compiles to flat accesses, several times slower).
constexpr int HEAP_LDS = 8192;
__shared__ HeapEntry s_heap[HEAP_LDS];
using GlobalHeapEntry = __attribute__((address_space(1))) HeapEntry;
struct HeapView {
    // the device-memory side through a global-address-space pointer: the compiler then cannot
    // fold the two sides into one flat access through a selected address
    GlobalHeapEntry *glob;
    __device__ HeapEntry get(long long i) const {
        if (i < HEAP_LDS) return s_heap[i];
        const GlobalHeapEntry &e = glob[i];
        return HeapEntry{e.value, e.p};
    }
    __device__ void set(long long i, const HeapEntry &e) const {
        if (i < HEAP_LDS) {
            s_heap[i] = e;
        } else {
            GlobalHeapEntry &d = glob[i];
            d.value = e.value;
            d.p = e.p;
        }
    }
};

// The heap's steps, run by the whole wave with wave-uniform arguments; they leave the heap
// exactly as libstdc++'s sequential steps do, because the same comparisons decide the same moves.

// __push_heap(first, hole, 0, v): the hole moves up while its parent has lower priority than v.
// Lane j compares the ancestor j + 1 levels up with v; the sequential loop stops at the first
// ancestor that does not have lower priority, so the moves are the ancestors below it (the
// trailing ones of the ballot), each written one level down by its lane, and v takes the last
// place freed.  Every read precedes every write.
__device__ inline void wave_sift_up(const HeapView &h, long long hole, const HeapEntry &v, int lane) {
    const int depth = 63 - __clzll((unsigned long long)(hole + 1));  // number of ancestors
    double ev = 0.;
    long long ep = 0;
    bool lower = false;
    if (lane < depth) {
        const HeapEntry e = h.get(((hole + 1) >> (lane + 1)) - 1);
        ev = e.value;
        ep = e.p;
        lower = ev > v.value;  // heap_less(e, v)
    }
    const int t = __builtin_ctzll(~__ballot(lower));
    if (lane < t) h.set(((hole + 1) >> lane) - 1, HeapEntry{ev, ep});
    if (lane == t) h.set(((hole + 1) >> t) - 1, v);
    if (hole >= HEAP_LDS) __syncthreads();  // device-memory writes complete before the next reads
}

// std::priority_queue::push: push_back then push_heap.
__device__ inline void wave_push(const HeapView &h, long long &n, const HeapEntry &v, int lane) {
    ++n;
    wave_sift_up(h, n - 1, v, lane);
}

// std::priority_queue::pop after top(): pop_heap, then pop_back.  The last element re-enters
// through __adjust_heap from the root: while the hole has two children it takes the one of
// higher priority (the right one on equality), then the element moves up.  The descent goes five
// levels at a time: lane j < 31 takes the j-th node of the next five levels below the hole (the
// nodes that can become holes) and, if that node would descend, reads its two children and
// picks one; the path then follows the picks from the top (scalar bit steps), and the lanes on
// it move their picked child up.  Every read of a chunk precedes its writes, and a chunk's reads
// lie below the nodes earlier chunks wrote.
__device__ inline HeapEntry wave_pop(const HeapView &h, long long &n, int lane) {
    const HeapEntry top = h.get(0);
    if (n > 1) {
        const long long len = n - 1;
        const HeapEntry v = h.get(len);
        const long long lim = (len - 1) / 2;  // a hole below lim has both children
        long long hole = 0;
        while (hole < lim) {
            long long node = 0;
            double cv = 0.;
            long long cp = 0;
            bool desc = false, left = false;
            if (lane < 31) {
                const int l = 31 - __clz(lane + 1);  // level within the chunk
                node = ((hole + 1) << l) - 1 + (lane + 1 - (1 << l));
                desc = node < lim;
                if (desc) {
                    const HeapEntry r = h.get(2 * node + 2), lf = h.get(2 * node + 1);
                    left = heap_less(r, lf);
                    cv = left ? lf.value : r.value;
                    cp = left ? lf.p : r.p;
                }
            }
            const unsigned long long dm = __ballot(desc), lm = __ballot(left);
            unsigned long long path = 0;
            int m = 0;
            for (int step = 0; step < 5 && (dm >> m & 1ull); ++step) {
                path |= 1ull << m;
                m = 2 * m + ((lm >> m & 1ull) ? 1 : 2);
            }
            if (path >> lane & 1ull) h.set(node, HeapEntry{cv, cp});
            const int l = 31 - __clz(m + 1);
            hole = ((hole + 1) << l) - 1 + (m + 1 - (1 << l));
        }
        if ((len & 1) == 0 && hole == (len - 2) / 2) {  // a last hole with a left child only
            const HeapEntry lf = h.get(2 * hole + 1);
            if (lane == 0) h.set(hole, lf);
            hole = 2 * hole + 1;
        }
        if (hole >= HEAP_LDS) __syncthreads();  // device-memory writes complete before the next reads
        wave_sift_up(h, hole, v, lane);
    }
    --n;
    return top;
}

// Slot k of the stencil of point p (coordinates c), as LaneStencil::load reads it.
template <int ModelT>
__device__ inline void stencil_slot(const Geom &g, const StencilData<double> &sd, long long p, const int *c, int k,
                                    uint32_t &off, double &w) {
    using S = ModelShape<ModelT>;
    constexpr int NC = S::NFWD + S::NSYM;
    if (!S::SHARED) {
        off = sd.offs[k * sd.np + p];
        w = sd.w[k * sd.np + p];
    } else {
        const int th = c[2] < g.n[2] ? c[2] : g.n[2] - 1;
        const double sp = sd.speed[p];
        const double s2 = sp * sp;
        off = sd.tab_offs[th * NC + k];
        w = (double)sd.tab_w[th * NC + k] * s2;
    }
}

// The serial loop (HamiltonFastMarching.hxx:59-169) on one wave.  The pops and pushes run in
// the reference's sequence, each spread over the wave (wave_pop, wave_push); so is the work of
// one pop: the recompute's differences (one stencil slot per lane, written to LDS at their place
// in stencil order, then summed by lane 0 in that order) and the conditional updates of the
// popped point's reversed neighbours (one neighbour per lane: the neighbours are distinct points,
// each update touches its own point only; they are then pushed in the reversed offsets' order).
// Every value is computed with the same operations, in the same order, as the single-threaded
// loop.
template <int ModelT>
__global__ __launch_bounds__(64) void k_fm_exact(ExactArgs a) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    __shared__ double s_rv[2 * NC], s_rw[2 * NC];
    __shared__ int s_ro[2 * NC];
    __shared__ long long s_px[64];
    __shared__ double s_pv[64];
    const int lane = threadIdx.x;
    const Geom &g = a.g;
    const HeapView heap{(GlobalHeapEntry *)a.heap};
    long long n = 0, maxn = 0;                   // heap size, the same in every lane
    long long accepted = 0, npost = 0;           // lane 0's
    long long all_left = a.all_left, all_left_u = a.all_left_u;
    int status = 0;
    // RunInit: the seeds in index order, values set and pushed
    for (int i = 0; i < a.nseeds; ++i) {
        if (lane == 0) a.u[a.seed_p[i]] = a.seed_v[i];
        wave_push(heap, n, HeapEntry{a.seed_v[i], a.seed_p[i]}, lane);
    }
    while (true) {
        // the next entry; an already accepted point's is skipped (lazy deletion)
        if (n == 0) break;
        maxn = n > maxn ? n : maxn;
        const HeapEntry t = wave_pop(heap, n, lane);
        const long long y = t.p;
        double key = t.value;
        int cy[3];
        tiled_coords(g, (int)(y >> 6), (int)(y & 63), cy);
        // what the pop needs from device memory, loaded at once before the accepted test
        const unsigned char acc_y = a.accepted[y];
        const unsigned long long act = a.active[y];
        const double u_y = a.u[y];
        const long long kk = S::SHARED ? (long long)cy[2] : y;
        const long long r0 = a.rstart[kk], r1 = r0 + a.rcount[kk];
        uint32_t slot_off = 0;
        double slot_w = 0.;
        if (lane < NC) stencil_slot<ModelT>(g, a.sd, y, cy, lane, slot_off, slot_w);
        asm volatile("" ::: "memory");
        if (acc_y) continue;
        // PostProcess (HamiltonFastMarching.hxx:86-99): the stopping criteria see the point with
        // its heap value, before the recompute (ExtraAlgorithms/CommonStoppingCriteria.h:57-92)
        int stop = 0;
        if (lane == 0) {
            ++npost;
            stop = npost >= a.max_accepted;
            if (a.stop_flags) {
                const unsigned char f = a.stop_flags[y];
                if (f & 1) stop = 1;
                if (f & 2) {
                    a.stop_flags[y] = f & (unsigned char)~2;
                    if (--all_left == 0) stop = 1;
                }
            }
            if (a.stop_flags_u) {
                const long long xy = (long long)cy[0] * g.n[1] + cy[1];
                const unsigned char f = a.stop_flags_u[xy];
                if (f & 1) stop = 1;
                if (f & 2) {
                    a.stop_flags_u[xy] = f & (unsigned char)~2;
                    if (--all_left_u == 0) stop = 1;
                }
            }
            if (u_y > a.stop_distance) stop = 1;
        }
        stop = __shfl(stop, 0);
        // the recompute (order >= 2, or the factor applies): a point without active differences
        // keeps its value
        int fcur[3];
        const bool fact = factor_at(g, a.fac, cy, fcur);
        if ((a.order > 1 || fact) && act) {
            if (lane < NC) {
                const double w = slot_w;
                int o0, o1, o2;
                unpack_offset(slot_off, o0, o1, o2);
                for (int sd = 0; sd < (lane < NF ? 1 : 2); ++sd) {
                    const int bit = side_bit<NF>(lane, sd);
                    if (!(act >> bit & 1ull)) continue;
                    const int sg = sd == 0 ? 1 : -1;
                    const int o[3] = {sg * o0, sg * o1, sg * o2};
                    const int pos = __popcll(act & ((1ull << bit) - 1ull));
                    int ord;
                    s_rv[pos] = value_corr<ModelT>(g, a.u, a.accepted, cy, o, a.order, a.fac, fact, fcur, ord);
                    s_rw[pos] = w;
                    s_ro[pos] = ord;
                }
            }
            __syncthreads();
            if (lane == 0) {
                const int ne = __popcll(act);
                Quad q{__builtin_huge_val(), 0., 0., -1.};
                for (int i = 0; i < ne; ++i) q.minVal = s_rv[i] < q.minVal ? s_rv[i] : q.minVal;
                constexpr double mult2[4] = {0., 1., 9. / 4., 121. / 36.};
                for (int i = 0; i < ne; ++i) q.add(s_rv[i], mult2[s_ro[i]] * s_rw[i]);
                key = q.solve();
                a.u[y] = key;
            }
            key = __shfl(key, 0);
            __syncthreads();
        } else {
            key = __shfl(key, 0);
        }
        if (stop) break;  // the point that triggered the stop is not accepted
        if (lane == 0) {
            a.accepted[y] = 1;
            ++accepted;
        }
        // every store of an iteration completes at a barrier before the next iteration reads
        // (the update chunks end with one; a point without reversed neighbours takes its own)
        if (r0 == r1) __syncthreads();
        for (long long base = r0; base < r1; base += 64) {
            const long long r = base + lane;
            long long px = -1;
            double pv = 0.;
            if (r < r1) {
                int e0, e1, e2;
                unpack_offset(a.roffs[r], e0, e1, e2);
                // ConditionalUpdate (HamiltonFastMarching.hxx:136-154)
                const long long x = shifted<ModelT>(g, cy, e0, e1, e2, -1);
                if (x >= 0) {
                    // the neighbour's state and stencil, loaded at once before the tests
                    const unsigned char accx = a.accepted[x];
                    unsigned long long ax = a.active[x];
                    const double ux = a.u[x];
                    int cx[3];
                    tiled_coords(g, (int)(x >> 6), (int)(x & 63), cx);
                    LaneStencil<ModelT, double> st;
                    st.load(g, a.sd, x, cx);
                    double *qs = a.quad + 4 * x;
                    Quad q{qs[0], qs[1], qs[2], qs[3]};
                    asm volatile("" ::: "memory");
                    // seeds are never updated; no update unless the value would decrease
                    if (!accx && !(ax == 0ull && ux != __builtin_huge_val()) && !(ux <= key)) {
                        // HopfLaxUpdate (EulerianStencil.hxx:104-153): activate the differences
                        // with this offset (a symmetric one only if its other side is not
                        // active), add them to the point's quadratic, solve
                        if (q.minVal == -__builtin_huge_val()) q.minVal = key;
                        const uint32_t eo = pack_offset(e0, e1, e2), meo = pack_offset(-e0, -e1, -e2);
                        for (int k = 0; k < NC; ++k) {
                            if (k < NF) {
                                if (st.off[k] == eo) {
                                    ax |= 1ull << side_bit<NF>(k, 0);
                                    q.add(key, st.w[k]);
                                }
                                continue;
                            }
                            const int b0 = side_bit<NF>(k, 0), b1 = side_bit<NF>(k, 1);
                            if (st.off[k] == eo && !(ax >> b1 & 1ull)) {
                                ax |= 1ull << b0;
                                q.add(key, st.w[k]);
                            }
                            if (st.off[k] == meo && !(ax >> b0 & 1ull)) {
                                ax |= 1ull << b1;
                                q.add(key, st.w[k]);
                            }
                        }
                        const double value = q.solve();
                        qs[0] = q.minVal;
                        qs[1] = q.a;
                        qs[2] = q.b;
                        qs[3] = q.c;
                        a.active[x] = ax;
                        a.u[x] = value;
                        px = x;
                        pv = value;
                    }
                }
            }
            s_px[lane] = px;
            s_pv[lane] = pv;
            __syncthreads();
            const int cnt = r1 - base < 64 ? (int)(r1 - base) : 64;
            for (int j = 0; j < cnt; ++j) {
                const long long px_j = s_px[j];
                if (px_j < 0) continue;
                if (n >= a.heap_cap) {  // cannot happen: the capacity is the number of pushes possible
                    status = 1;
                    break;
                }
                wave_push(heap, n, HeapEntry{s_pv[j], px_j}, lane);
            }
            __syncthreads();
            if (status) break;
        }
        if (status) break;
    }
    if (lane == 0) {
        a.out[0] = accepted;
        a.out[1] = maxn;
        a.out[2] = status;
        a.out[3] = npost;
    }
}

// Reversed offsets of the per-point stencils (HFM_StencilRecompute.hxx:81-124): every point y
// lists the offsets e for which some x = y - e has a nonzero-weight difference reaching y
// (x + e for a forward one, x +- e for a symmetric one), sorted and without repetitions.  Pass
// 0 counts, pass 1 fills (positions from an exclusive scan), then each list is sorted.
template <int ModelT>
__global__ __launch_bounds__(256) void k_rev_build(Geom g, StencilData<double> sd, int pass, int *__restrict__ cnt,
                                                  const int *__restrict__ start, int *__restrict__ fill,
                                                  uint32_t *__restrict__ roffs, long long np) {
    using S = ModelShape<ModelT>;
    constexpr int NF = S::NFWD, NC = NF + S::NSYM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    if (!in_grid(g, c)) return;
    LaneStencil<ModelT, double> st;
    st.load(g, sd, p, c);
    for (int k = 0; k < NC; ++k) {
        if (st.w[k] == 0.) continue;
        int o0, o1, o2;
        unpack_offset(st.off[k], o0, o1, o2);
        for (int s = 0; s < (k < NF ? 1 : 2); ++s) {
            const int sg = s == 0 ? 1 : -1;
            const long long y = shifted<ModelT>(g, c, o0, o1, o2, sg);
            if (y < 0) continue;
            if (pass == 0) {
                atomicAdd(&cnt[y], 1);
            } else {
                const int pos = atomicAdd(&fill[y], 1);
                roffs[start[y] + pos] = pack_offset(sg * o0, sg * o1, sg * o2);
            }
        }
    }
}

__device__ inline bool offset_less(uint32_t a, uint32_t b) {
    int a0, a1, a2, b0, b1, b2;
    unpack_offset(a, a0, a1, a2);
    unpack_offset(b, b0, b1, b2);
    if (a0 != b0) return a0 < b0;
    if (a1 != b1) return a1 < b1;
    return a2 < b2;
}

__global__ __launch_bounds__(256) void k_rev_sort(const int *__restrict__ cnt, const int *__restrict__ start,
                                                 uint32_t *__restrict__ roffs, int *__restrict__ rcount, long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int n = cnt[p];
    uint32_t *l = roffs + start[p];
    for (int i = 1; i < n; ++i) {  // insertion sort of a few dozen offsets at most
        const uint32_t v = l[i];
        int j = i - 1;
        while (j >= 0 && offset_less(v, l[j])) {
            l[j + 1] = l[j];
            --j;
        }
        l[j + 1] = v;
    }
    int m = 0;
    for (int i = 0; i < n; ++i)
        if (m == 0 || l[i] != l[m - 1]) l[m++] = l[i];
    rcount[p] = m;
}

// Exports of an exact solve (HFMInterface.hxx:526-577): activeNeighs are the flags the loop set;
// geodesicFlow is -ReDim(sum_i w_i max(0, r - v_i) m_i e_i) from the recompute with the final
// accepted set (m = 1, 3/2, 11/6 by order, EulerianStencil.hxx:218-231); activeOffsets are the
// recompute's offsets packed as by the direct path.
template <int ModelT>
__global__ __launch_bounds__(256) void k_flow_exact(Geom g, StencilData<double> sd, const double *__restrict__ u,
                                                   const unsigned char *__restrict__ acc,
                                                   const unsigned long long *__restrict__ active, int order,
                                                   FactorDev fac, FlowScale fs, double *__restrict__ flow,
                                                   double *__restrict__ active_out, double *__restrict__ aoffs_out,
                                                   long long np) {
    using S = ModelShape<ModelT>;
    constexpr int NC = S::NFWD + S::NSYM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
    tiled_coords(g, (int)(p >> 6), (int)(p & 63), c);
    if (!in_grid(g, c)) return;
    const long long lin = ((long long)c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
    const unsigned long long act = active[p];
    RecomputeEntry re[2 * NC];
    int n = 0;
    double value = 0.;
    if (act) n = recompute<ModelT>(g, sd, u, acc, p, c, act, order, fac, re, &value);
    double f[3] = {0., 0., 0.};
    for (int i = 0; i < n; ++i) {
        const double wpd = flow_weight(re[i], value);
        for (int q = 0; q < 3; ++q) f[q] += wpd * (double)re[i].o[q];
    }
    if (flow)
        for (int q = 0; q < g.dim; ++q) flow[lin * g.dim + q] = -(f[q] * fs.h[q]);
    if (active_out) active_out[lin] = (double)act;
    if (aoffs_out) {
        int j = 0;
        for (; j < n; ++j) {
            unsigned long long packed = 0;
            for (int q = 0; q < S::DIM; ++q)
                packed |= (unsigned long long)(uint8_t)(int8_t)re[j].o[q] << (8 * (S::DIM - 1 - q));
            aoffs_out[lin * NC + j] = (double)packed;
        }
        for (; j < NC; ++j) aoffs_out[lin * NC + j] = 0.;
    }
}

// Solver state before the loop: values +inf, nothing accepted or active, every point's quadratic
// empty (minVal = -inf until the first update, a = b = 0, c = -1).
__global__ __launch_bounds__(256) void k_exact_init(double *__restrict__ u, unsigned char *__restrict__ acc,
                                                   unsigned long long *__restrict__ active, double *__restrict__ quad,
                                                   long long np) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    u[p] = __builtin_huge_val();
    acc[p] = 0;
    active[p] = 0ull;
    quad[4 * p] = -__builtin_huge_val();
    quad[4 * p + 1] = 0.;
    quad[4 * p + 2] = 0.;
    quad[4 * p + 3] = -1.;
}

}  // namespace

static inline unsigned exact_blocks(long long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_rev_build(int model, const Geom &g, const StencilData<double> &sd, int pass, int *cnt,
                            const int *start, int *fill, uint32_t *roffs, long long np, hipStream_t s) {
    switch (model) {
        case MODEL_RIEMANN2:
            hipLaunchKernelGGL((k_rev_build<MODEL_RIEMANN2>), dim3(exact_blocks(np)), dim3(256), 0, s, g, sd, pass, cnt,
                               start, fill, roffs, np);
            break;
        case MODEL_RIEMANN3:
            hipLaunchKernelGGL((k_rev_build<MODEL_RIEMANN3>), dim3(exact_blocks(np)), dim3(256), 0, s, g, sd, pass, cnt,
                               start, fill, roffs, np);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_exact_init(double *u, unsigned char *acc, unsigned long long *active, double *quad, long long np,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_exact_init, dim3(exact_blocks(np)), dim3(256), 0, s, u, acc, active, quad, np);
    return hipGetLastError();
}

hipError_t launch_rev_sort(const int *cnt, const int *start, uint32_t *roffs, int *rcount, long long np, hipStream_t s) {
    hipLaunchKernelGGL(k_rev_sort, dim3(exact_blocks(np)), dim3(256), 0, s, cnt, start, roffs, rcount, np);
    return hipGetLastError();
}

hipError_t launch_fm_exact(int model, const ExactArgs &a, hipStream_t s) {
    switch (model) {
        case MODEL_RIEMANN2: hipLaunchKernelGGL((k_fm_exact<MODEL_RIEMANN2>), dim3(1), dim3(64), 0, s, a); break;
        case MODEL_RIEMANN3: hipLaunchKernelGGL((k_fm_exact<MODEL_RIEMANN3>), dim3(1), dim3(64), 0, s, a); break;
        case MODEL_RSF2: hipLaunchKernelGGL((k_fm_exact<MODEL_RSF2>), dim3(1), dim3(64), 0, s, a); break;
        case MODEL_ELASTICA2: hipLaunchKernelGGL((k_fm_exact<MODEL_ELASTICA2>), dim3(1), dim3(64), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_flow_exact(int model, const Geom &g, const StencilData<double> &sd, const double *u,
                             const unsigned char *acc, const unsigned long long *active, int order,
                             const FactorDev &fac, const FlowScale &fs, double *flow, double *act_out, double *aoffs,
                             long long np, hipStream_t s) {
    const dim3 gr(exact_blocks(np)), bl(256);
    switch (model) {
        case MODEL_RIEMANN2:
            hipLaunchKernelGGL((k_flow_exact<MODEL_RIEMANN2>), gr, bl, 0, s, g, sd, u, acc, active, order, fac, fs, flow,
                               act_out, aoffs, np);
            break;
        case MODEL_RIEMANN3:
            hipLaunchKernelGGL((k_flow_exact<MODEL_RIEMANN3>), gr, bl, 0, s, g, sd, u, acc, active, order, fac, fs, flow,
                               act_out, aoffs, np);
            break;
        case MODEL_RSF2:
            hipLaunchKernelGGL((k_flow_exact<MODEL_RSF2>), gr, bl, 0, s, g, sd, u, acc, active, order, fac, fs, flow,
                               act_out, aoffs, np);
            break;
        case MODEL_ELASTICA2:
            hipLaunchKernelGGL((k_flow_exact<MODEL_ELASTICA2>), gr, bl, 0, s, g, sd, u, acc, active, order, fac, fs, flow,
                               act_out, aoffs, np);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace hfm
