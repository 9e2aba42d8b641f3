"""Seeded synthetic inputs for the benchmark and parity workloads.

Every field is generated from numpy's PCG64 with an explicit seed and a separable Gaussian
filter written here, so that any machine with the same numpy produces the same bits.

Workloads (grid points are cell centred, x_i = origin + (i + 0.5) h, row-major arrays):

* ``riemann2``  -- 2-D Riemannian metric M = R(theta) diag(1, r^2) R(theta)^T with
  theta = pi sin(2 pi x) cos(2 pi y) + phi, r = rescale(smooth(sigma=8), 1, 10).
* ``riemann3``  -- 3-D metric R diag(1, a^2, b^2) R^T with R = Rz(alpha) Ry(beta) Rz(gamma),
  alpha, beta, gamma = pi smooth(sigma=6), a, b = rescale(smooth(sigma=6), 1, amax).
* ``rsf2``      -- ReedsSheppForward2 cost 1 + 4 sigmoid(smooth(sigma=10)) on (x, y, theta),
  theta periodic, 4 seed states and 32 tip states.
* ``elastica2`` -- Elastica2 cost exp(0.5 smooth(sigma=12)) on (x, y, theta), one seed at
  (0.5, 0.5, 0) and 32 sensitivity targets.

Metric components follow the reference's packed order xx, xy, yy[, xz, yz, zz]
(JMM_CPPLibs/JMM_CPPLibs/LinearAlgebra/Implementation/SymmetricMatrixType.hxx:11-16).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["smooth", "rescale", "riemann2", "riemann3", "rsf2", "elastica2", "riemann_request"]


def _gauss_kernel(sigma: float) -> np.ndarray:
    r = int(math.ceil(4.0 * sigma))
    t = np.arange(-r, r + 1, dtype=np.float64)
    g = np.exp(-0.5 * (t / sigma) ** 2)
    return g / g.sum()


def _filter_axis(a: np.ndarray, g: np.ndarray, axis: int, periodic: bool) -> np.ndarray:
    r = (len(g) - 1) // 2
    n = a.shape[axis]
    if periodic:
        idx = np.arange(-r, n + r) % n
    else:
        # symmetric reflection (d c b a | a b c d | d c b a), applied repeatedly for short axes
        idx = np.arange(-r, n + r)
        period = 2 * n
        idx = idx % period
        idx = np.where(idx >= n, period - 1 - idx, idx)
    padded = np.take(a, idx, axis=axis)
    out = np.zeros_like(a)
    sl = [slice(None)] * a.ndim
    for k in range(len(g)):
        sl[axis] = slice(k, k + n)
        out += g[k] * padded[tuple(sl)]
    return out


def smooth(shape, sigma: float, seed: int, periodic=()) -> np.ndarray:
    """Standardised (mean 0, std 1) Gaussian-filtered white noise, truncated at 4 sigma."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.standard_normal(tuple(shape))
    g = _gauss_kernel(sigma)
    for ax in range(a.ndim):
        a = _filter_axis(a, g, ax, ax in periodic)
    a -= a.mean()
    sd = a.std()
    return a / sd if sd > 0 else a


def rescale(n: np.ndarray, lo: float, hi: float) -> np.ndarray:
    """Affine map of n onto [lo, hi] (min -> lo, max -> hi)."""
    mn, mx = n.min(), n.max()
    if mx == mn:
        return np.full_like(n, lo)
    return lo + (hi - lo) * (n - mn) / (mx - mn)


def _cell_centres(n: int, h: float) -> np.ndarray:
    return (np.arange(n, dtype=np.float64) + 0.5) * h


def riemann2(n: int = 256, seed: int = 0, rmax: float = 10.0, sigma: float = 8.0):
    """Riemann2 workload: returns the HFM input dict (RowMajor arrays)."""
    h = 1.0 / n
    x = _cell_centres(n, h)
    X, Y = np.meshgrid(x, x, indexing="ij")
    phi = np.random.Generator(np.random.PCG64(seed + 1000)).uniform(0.0, 2 * math.pi)
    theta = math.pi * np.sin(2 * math.pi * X) * np.cos(2 * math.pi * Y) + phi
    r = rescale(smooth((n, n), sigma, seed), 1.0, rmax)
    c, s = np.cos(theta), np.sin(theta)
    l1, l2 = 1.0, r * r
    m = np.empty((n, n, 3))
    m[..., 0] = c * c * l1 + s * s * l2          # xx
    m[..., 1] = c * s * (l1 - l2)                # xy
    m[..., 2] = s * s * l1 + c * c * l2          # yy
    return {
        "model": "Riemann2", "dims": np.array([n, n], dtype=np.float64), "gridScale": h,
        "origin": np.zeros(2), "seeds": np.array([[0.5, 0.5]]), "exportValues": 1.0,
        "metric": m, "arrayOrdering": "RowMajor",
    }


def _rot_z(a):
    c, s = np.cos(a), np.sin(a)
    R = np.zeros(a.shape + (3, 3))
    R[..., 0, 0], R[..., 0, 1] = c, -s
    R[..., 1, 0], R[..., 1, 1] = s, c
    R[..., 2, 2] = 1.0
    return R


def _rot_y(a):
    c, s = np.cos(a), np.sin(a)
    R = np.zeros(a.shape + (3, 3))
    R[..., 0, 0], R[..., 0, 2] = c, s
    R[..., 1, 1] = 1.0
    R[..., 2, 0], R[..., 2, 2] = -s, c
    return R


def metric3_from_angles(alpha, beta, gamma, a, b) -> np.ndarray:
    """Packed (xx, xy, yy, xz, yz, zz) of R diag(1, a^2, b^2) R^T, R = Rz(alpha) Ry(beta) Rz(gamma)."""
    R = np.einsum("...ij,...jk->...ik", np.einsum("...ij,...jk->...ik", _rot_z(alpha), _rot_y(beta)),
                  _rot_z(gamma))
    lam = np.stack([np.ones_like(a), a * a, b * b], axis=-1)
    M = np.einsum("...ik,...k,...jk->...ij", R, lam, R)
    out = np.empty(alpha.shape + (6,))
    out[..., 0] = M[..., 0, 0]
    out[..., 1] = M[..., 1, 0]
    out[..., 2] = M[..., 1, 1]
    out[..., 3] = M[..., 2, 0]
    out[..., 4] = M[..., 2, 1]
    out[..., 5] = M[..., 2, 2]
    return out


def riemann3(n: int = 128, seed: int = 0, amax: float = 20.0, sigma: float = 6.0):
    """Riemann3 workload: returns the HFM input dict (RowMajor arrays)."""
    h = 1.0 / n
    shp = (n, n, n)
    alpha = math.pi * smooth(shp, sigma, seed)
    beta = math.pi * smooth(shp, sigma, seed + 1)
    gamma = math.pi * smooth(shp, sigma, seed + 2)
    a = rescale(smooth(shp, sigma, seed + 3), 1.0, amax)
    b = rescale(smooth(shp, sigma, seed + 4), 1.0, amax)
    m = metric3_from_angles(alpha, beta, gamma, a, b)
    return {
        "model": "Riemann3", "dims": np.array([n, n, n], dtype=np.float64), "gridScale": h,
        "origin": np.zeros(3), "seeds": np.array([[0.5, 0.5, 0.5]]), "exportValues": 1.0,
        "metric": m, "arrayOrdering": "RowMajor",
    }


def _sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def _states(rng, k: int) -> np.ndarray:
    xy = rng.uniform(0.1, 0.9, size=(k, 2))
    th = rng.uniform(0.0, 2 * math.pi, size=(k, 1))
    return np.concatenate([xy, th], axis=1)


def rsf2(n: int = 256, ntheta: int = 64, seed: int = 0, xi: float = 0.1, sigma: float = 10.0,
         nseeds: int = 4, ntips: int = 32):
    """ReedsSheppForward2 workload: returns the HFM input dict (RowMajor arrays)."""
    h = 1.0 / n
    cost = 1.0 + 4.0 * _sigmoid(smooth((n, n, ntheta), sigma, seed, periodic=(2,)))
    rng = np.random.Generator(np.random.PCG64(seed + 2000))
    return {
        "model": "ReedsSheppForward2", "dims": np.array([n, n, ntheta], dtype=np.float64),
        "gridScale": h, "origin": np.zeros(2), "xi": xi, "eps": 0.1, "cost": cost,
        "seeds": _states(rng, nseeds), "tips": _states(rng, ntips), "exportValues": 1.0,
        "arrayOrdering": "RowMajor",
    }


def elastica2(n: int = 192, ntheta: int = 64, seed: int = 0, xi: float = 0.2, sigma: float = 12.0,
              ntargets: int = 32):
    """Elastica2 workload with 32 sensitivity targets: returns the HFM input dict."""
    h = 1.0 / n
    cost = np.exp(0.5 * smooth((n, n, ntheta), sigma, seed, periodic=(2,)))
    rng = np.random.Generator(np.random.PCG64(seed + 3000))
    targets = _states(rng, ntargets)
    return {
        "model": "Elastica2", "dims": np.array([n, n, ntheta], dtype=np.float64), "gridScale": h,
        "origin": np.zeros(2), "xi": xi, "eps": 0.1, "cost": cost,
        "seeds": np.array([[0.5, 0.5, 0.0]]), "exportValues": 1.0,
        "inspectSensitivity": targets, "inspectSensitivityWeights": np.ones(ntargets),
        "inspectSensitivityLengths": np.array([float(ntargets)]), "arrayOrdering": "RowMajor",
    }


def riemann_request(rng: np.random.Generator, dim: int, side: int, kappa: float, nseeds: int,
                    sigma: float = 6.0):
    """One serving request: a smooth random SPD field of anisotropy up to kappa (sqrt of the
    eigenvalue ratio), 1-16 seeds uniform in [0.05, 0.95]^d."""
    seed = int(rng.integers(0, 2**31 - 1))
    h = 1.0 / side
    if dim == 2:
        x = _cell_centres(side, h)
        X, Y = np.meshgrid(x, x, indexing="ij")
        theta = math.pi * smooth((side, side), sigma, seed) + X * Y
        r = rescale(smooth((side, side), sigma, seed + 1), 1.0, kappa)
        c, s = np.cos(theta), np.sin(theta)
        m = np.empty((side, side, 3))
        m[..., 0] = c * c + s * s * r * r
        m[..., 1] = c * s * (1.0 - r * r)
        m[..., 2] = s * s + c * c * r * r
    else:
        shp = (side,) * 3
        al = math.pi * smooth(shp, sigma, seed)
        be = math.pi * smooth(shp, sigma, seed + 1)
        ga = math.pi * smooth(shp, sigma, seed + 2)
        a = rescale(smooth(shp, sigma, seed + 3), 1.0, kappa)
        b = rescale(smooth(shp, sigma, seed + 4), 1.0, kappa)
        m = metric3_from_angles(al, be, ga, a, b)
    seeds = rng.uniform(0.05, 0.95, size=(nseeds, dim))
    return {
        "model": "Riemann2" if dim == 2 else "Riemann3",
        "dims": np.array([side] * dim, dtype=np.float64), "gridScale": h,
        "origin": np.zeros(dim), "seeds": seeds, "exportValues": 1.0, "metric": m,
        "arrayOrdering": "RowMajor",
    }
