// selling.h -- FP64 Selling (obtuse superbase) reduction of 2x2 / 3x3 SPD tensors, usable on the
// host and in HIP device code.
//
// The decomposition D = sum_k w_k e_k e_k^T must reproduce the reference's integer offsets
// slot for slot, so every sign test sees the same FP64 numbers as the reference:
//   * packed symmetric storage xx, xy, yy[, xz, yz, zz]
//     (JMM_CPPLibs/JMM_CPPLibs/LinearAlgebra/Implementation/SymmetricMatrixType.hxx:11-16);
//   * <u, D v> summed as sum_i u_i (sum_j D_ij v_j) (SymmetricMatrixType.hxx:18-30);
//   * inverse = comatrix * (1 / determinant) with the reference's determinant expansion
//     (SymmetricMatrixType.hxx:199-243; reciprocal multiply from VectorType.h:37-42);
//   * flips visited in the order i < j, superbase updated in place, at most 200 sweeps
//     (BasisReduction.hpp:65-99, BasisReduction.h:72);
//   * output slots: 2-D k -> Perp(sb_k), weight -<sb_{k+1}, D sb_{k+2}>; 3-D (i,j) lexicographic
//     -> Cross(sb_k, sb_l) over the complement, weight -<sb_i, D sb_j> (BasisReduction.hpp:122-157).
// Callers compile this with -ffp-contract=off so no multiply-add is fused.
#ifndef HFM_AMD_SELLING_H
#define HFM_AMD_SELLING_H

#include <stdint.h>

#if defined(__HIPCC__)
#define HFM_HD __host__ __device__ inline
#else
#define HFM_HD inline
#endif

namespace hfm {

enum SellingStatus { SELLING_OK = 0, SELLING_NOT_SPD = 1, SELLING_NO_CONVERGENCE = 2 };

template <int D> struct SymPacked {
    static constexpr int K = D * (D + 1) / 2;
    double m[K];
    HFM_HD static int lin(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }
    HFM_HD double operator()(int i, int j) const { return m[lin(i, j)]; }
};

template <int D> HFM_HD double sym_trace(const SymPacked<D> &c) {
    double r = 0;
    for (int i = 0; i < D; ++i) r += c(i, i);
    return r;
}

template <int D> HFM_HD double sym_det(const SymPacked<D> &c) {
    if (D == 2) return c(0, 0) * c(1, 1) - c(1, 0) * c(0, 1);
    return c(0, 0) * (c(1, 1) * c(2, 2) - c(1, 2) * c(2, 1))
         + c(0, 1) * (2. * c(1, 2) * c(2, 0) - c(1, 0) * c(2, 2))
         - c(0, 2) * (c(1, 1) * c(2, 0));
}

template <int D> HFM_HD SymPacked<D> sym_inverse(const SymPacked<D> &c) {
    SymPacked<D> r;
    if (D == 2) {
        r.m[0] = c(1, 1);
        r.m[1] = -c(1, 0);
        r.m[2] = c(0, 0);
    } else {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j <= i; ++j) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                r.m[i * (i + 1) / 2 + j] = c(i1, j1) * c(i2, j2) - c(i1, j2) * c(i2, j1);
            }
    }
    const double b = 1. / sym_det<D>(c);
    for (int k = 0; k < SymPacked<D>::K; ++k) r.m[k] *= b;
    return r;
}

template <int D> HFM_HD double sym_dot(const SymPacked<D> &c, const int *u, const int *v) {
    double sum = 0.;
    for (int i = 0; i < D; ++i) {
        double sumi = 0.;
        for (int j = 0; j < D; ++j) sumi += c(i, j) * v[j];
        sum += u[i] * sumi;
    }
    return sum;
}

// D v for an integer vector v held in doubles, with the reference's summation order: entry i is
// sum_j D_ij v_j accumulated from 0 (the inner sum "sumi" of SymmetricMatrixType.hxx:18-30).
template <int D> HFM_HD void sym_apply(const SymPacked<D> &c, const double *v, double *out) {
    for (int i = 0; i < D; ++i) {
        double sumi = 0.;
        for (int j = 0; j < D; ++j) sumi += c(i, j) * v[j];
        out[i] = sumi;
    }
}

// <u, D v> = sum_i u_i (D v)_i accumulated from 0 -- the reference's outer sum.
template <int D> HFM_HD double dot_cached(const double *u, const double *Dv) {
    double sum = 0.;
    for (int i = 0; i < D; ++i) sum += u[i] * Dv[i];
    return sum;
}

// Tensor decomposition: offsets off[K][D] (int, before narrowing) and weights w[K].
// The superbase vectors are small integers held exactly in doubles, and D sb_j is cached per
// vector and refreshed when the vector changes; every scalar product therefore performs exactly
// the floating-point operations of the reference's ScalarProduct, in the same order (negating a
// vector negates its cached product exactly).
template <int D>
HFM_HD int selling_decompose(const SymPacked<D> &m, int (*off)[D], double *w, int maxIt = 200) {
    double sb[D + 1][D];
    double Dsb[D + 1][D];
    for (int i = 0; i <= D; ++i)
        for (int j = 0; j < D; ++j) sb[i][j] = i < D ? (i == j ? 1. : 0.) : -1.;
    if (sym_det<D>(m) <= 0 || sym_trace<D>(m) <= 0) return SELLING_NOT_SPD;
    for (int i = 0; i <= D; ++i) sym_apply<D>(m, sb[i], Dsb[i]);
    bool reduced = true;
    for (int it = 0; reduced && it < maxIt; ++it) {
        reduced = false;
        for (int i = 0; i < D; ++i)
            for (int j = i + 1; j <= D; ++j) {
                if (dot_cached<D>(sb[i], Dsb[j]) > 0) {
                    reduced = true;
                    if (D == 2) {
                        const int k = 3 - i - j;
                        for (int c = 0; c < D; ++c) sb[k][c] = sb[i][c] - sb[j][c];
                        sym_apply<D>(m, sb[k], Dsb[k]);
                    } else {
                        int k0 = -1, k1 = -1;
                        for (int q = 0; q <= D; ++q)
                            if (q != i && q != j) { if (k0 < 0) k0 = q; else k1 = q; }
                        for (int c = 0; c < D; ++c) { sb[k0][c] += sb[i][c]; sb[k1][c] += sb[i][c]; }
                        sym_apply<D>(m, sb[k0], Dsb[k0]);
                        sym_apply<D>(m, sb[k1], Dsb[k1]);
                    }
                    for (int c = 0; c < D; ++c) {
                        sb[i][c] = -sb[i][c];
                        Dsb[i][c] = -Dsb[i][c];
                    }
                }
            }
    }
    if (reduced) return SELLING_NO_CONVERGENCE;
    if (D == 2) {
        for (int i = 0; i < 3; ++i) {
            const int j = (i + 1) % 3, k = (i + 2) % 3;
            off[i][0] = -(int)sb[i][1];
            off[i][1] = (int)sb[i][0];
            w[i] = -dot_cached<D>(sb[j], Dsb[k]);
        }
    } else {
        int s = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) {
                int r0 = -1, r1 = -1;
                for (int q = 0; q < 4; ++q)
                    if (q != i && q != j) { if (r0 < 0) r0 = q; else r1 = q; }
                for (int c = 0; c < 3; ++c) {
                    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                    off[s][c] = (int)sb[r0][c1] * (int)sb[r1][c2] - (int)sb[r0][c2] * (int)sb[r1][c1];
                }
                w[s] = -dot_cached<D>(sb[i], Dsb[j]);
                ++s;
            }
    }
    return SELLING_OK;
}

// Riemannian stencil at one point (Riemannian.h:41-44,50-52,68-71 + CommonTraits.h:82-105):
// D = (dual ? M : M^-1) o [1/(h_i h_j)], Selling, weights below 1000 eps Tr(D) zeroed,
// offsets narrowed to int8 (the reference's int_least8_t).  invhh[k] = 1/(h_i h_j) packed.
template <int D>
HFM_HD int riemann_point_stencil(const double *mcomp, bool dual, const double *invhh, int8_t (*off8)[D],
                                 double *w, int *maxAbsOffset) {
    constexpr int K = D * (D + 1) / 2;
    SymPacked<D> m;
    for (int k = 0; k < K; ++k) m.m[k] = mcomp[k];
    if (!dual) m = sym_inverse<D>(m);
    for (int k = 0; k < K; ++k) m.m[k] = m.m[k] * invhh[k];
    const double atol = 1000 * 2.220446049250313080847e-16 * sym_trace<D>(m);
    int off[K][D];
    double wr[K];
    const int st = selling_decompose<D>(m, off, wr);
    int mx = 0;
    for (int k = 0; k < K; ++k) {
        w[k] = (st == SELLING_OK) ? (wr[k] < atol ? 0. : wr[k]) : 0.;
        for (int c = 0; c < D; ++c) {
            const int o = (st == SELLING_OK) ? off[k][c] : 0;
            off8[k][c] = (int8_t)o;
            const int a = o < 0 ? -o : o;
            if (w[k] != 0. && a > mx) mx = a;
        }
    }
    *maxAbsOffset = mx;
    return st;
}

}  // namespace hfm

#endif
