"""Checker: seed spreading and static factoring of the Riemann models, restated in Python floats
(IEEE FP64, one rounding per operation, the reference's order of operations).

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle/``.

* ``spread_seeds``  -- StaticFactoring::SetSeeds with a positive seedRadius
  (Headers/ExtraAlgorithms/Implementation/StaticFactoring.hxx:284-345).
* ``compute_factor`` -- StaticFactoring::ComputeFactor, choices "Seed" / "Key" and "Both"
  (StaticFactoring.hxx:116-239).
* ``import_factor`` -- StaticFactoring::ImportFactor (StaticFactoring.hxx:77-113).
* ``setup``          -- StaticFactoring::Setup + SetSeeds as HFMInterface.hxx:397 runs them: which
  of the above apply, in which order, the default seedRadius and the errors.

The distance guesses are GetGuess of the Riemannian stencil (Headers/Specializations/
Riemannian.h:62-71): the metric at a grid index, or the metric interpolated at a point over the
2^d cells around it (BaseGrid.h:80-104, PeriodicGrid.hxx:28-40, MapWeightedSum.h:21-36), made
primal and scaled by h_i h_j.
"""
from __future__ import annotations

import math

import numpy as np


def _lin(i, j):
    if i < j:
        i, j = j, i
    return i * (i + 1) // 2 + j


def _det(m, d):  # SymmetricMatrixType.hxx:199-209
    c = lambda i, j: m[_lin(i, j)]  # noqa: E731
    if d == 2:
        return c(0, 0) * c(1, 1) - c(1, 0) * c(0, 1)
    return (c(0, 0) * (c(1, 1) * c(2, 2) - c(1, 2) * c(2, 1))
            + c(0, 1) * (2. * c(1, 2) * c(2, 0) - c(1, 0) * c(2, 2))
            - c(0, 2) * (c(1, 1) * c(2, 0)))


def _inverse(m, d):  # Comatrix (SymmetricMatrixType.hxx:220-237) times 1 / Determinant (:239-243)
    c = lambda i, j: m[_lin(i, j)]  # noqa: E731
    r = [0.] * len(m)
    if d == 2:
        r[0], r[1], r[2] = c(1, 1), -c(1, 0), c(0, 0)
    else:
        for i in range(3):
            for j in range(i + 1):
                i1, i2, j1, j2 = (i + 1) % 3, (i + 2) % 3, (j + 1) % 3, (j + 2) % 3
                r[_lin(i, j)] = c(i1, j1) * c(i2, j2) - c(i1, j2) * c(i2, j1)
    b = 1. / _det(m, d)
    return [x * b for x in r]


class Guess:
    """A distance guess: the scaled primal metric M, Norm(v) = sqrt(<v, M v>) and
    Gradient(v) = (M v) / Norm(v) in the reference's order (SymmetricMatrixType.h:42-54,
    SymmetricMatrixType.hxx:18-44, VectorType.h:37-42)."""

    def __init__(self, m, d):
        self.m, self.d = m, d

    def mul(self, v):
        out = []
        for i in range(self.d):
            s = 0.
            for j in range(self.d):
                s += self.m[_lin(i, j)] * v[j]
            out.append(s)
        return out

    def norm(self, v):
        s = 0.
        for i in range(self.d):
            si = 0.
            for j in range(self.d):
                si += self.m[_lin(i, j)] * v[j]
            s += v[i] * si
        return math.sqrt(s)

    def gradient(self, v):
        mv = self.mul(v)
        b = 1. / self.norm(v)
        return [x * b for x in mv]


class Metric:
    """The metric field as given (`metric`, or `dualMetric` when dual), for the guesses."""

    def __init__(self, metric, dual, dims, h):
        self.d = len(dims)
        self.K = self.d * (self.d + 1) // 2
        self.field = np.asarray(metric, np.float64)
        self.is_field = self.field.ndim > 1
        self.dual = dual
        self.dims = [int(n) for n in dims]
        self.h = [float(x) for x in h]

    def raw(self, idx):
        if not self.is_field:
            return [float(x) for x in self.field.reshape(-1)[:self.K]]
        return [float(x) for x in self.field[tuple(idx)]]

    def rescale(self, m):  # Rescale<false> (Riemannian.h:68-71)
        if self.dual:
            m = _inverse(m, self.d)
        out = list(m)
        for i in range(self.d):
            for j in range(i + 1):
                out[_lin(i, j)] = out[_lin(i, j)] * (self.h[i] * self.h[j])
        return Guess(out, self.d)

    def inside(self, idx):
        return all(0 <= idx[i] < self.dims[i] for i in range(self.d))

    def guess_index(self, idx):
        return self.rescale(self.raw(idx))

    def guess_point(self, p):
        d = self.d
        nearest = [int(math.floor(p[i])) for i in range(d)]
        direction, weight = [], []
        for i in range(d):
            diff = p[i] - nearest[i] - 0.5
            direction.append(1 if diff > 0 else -1)
            weight.append(abs(diff))
        acc = None
        for c in range(1 << d):
            q = list(nearest)
            w = 1.
            for j in range(d):
                if (c >> j) & 1:
                    w *= weight[j]
                    q[j] += direction[j]
                else:
                    w *= 1. - weight[j]
            if not self.inside(q):
                w = 0.
            if c == 0:
                if not self.inside(q):
                    raise RuntimeError("Error : point has no valid neighbors in the grid domain")
                acc = [w * x for x in self.raw(q)]
            elif w != 0.:
                m = self.raw(q)
                acc = [a + w * x for a, x in zip(acc, m)]
        return self.rescale(acc)


def _euclid(v):
    s = 0.
    for x in v:
        s += x * x
    return math.sqrt(s)


def spread_seeds(metric: Metric, seeds_adim, values, radius):
    """SetSeeds' spreading (StaticFactoring.hxx:284-345) for radius > 0: every grid point within
    `radius` (in pixels) of each seed, with value seedValue + (|v|_p + |v|_q) / 2.  Returns the
    new ADim points and values, in the reference's order."""
    d = metric.d
    bd = 1 + int(math.ceil(radius))
    side = 1 + 2 * bd
    pts, vals = [], []
    for p, value in zip(seeds_adim, values):
        seen = set()
        distp = metric.guess_point(p)

        def insert(index, rad):
            q = [index[i] + 0.5 for i in range(d)]
            v = [p[i] - q[i] for i in range(d)]
            if _euclid(v) >= rad:
                return
            key = tuple(index)
            if key in seen:
                return
            seen.add(key)
            if not metric.inside(index):
                return
            distq = metric.guess_index(index)
            pts.append(q)
            vals.append(value + 0.5 * (distp.norm(v) + distq.norm(v)))

        pidx = [int(math.floor(p[i])) for i in range(d)]
        insert(pidx, math.inf)
        for b in range(side ** d):
            idx, r = [0] * d, b
            for i in range(d - 1, -1, -1):
                idx[i] = pidx[i] + (r % side) - bd
                r //= side
            insert(idx, radius)
    return pts, vals


def _box_index(i, bd):
    idx, r = [0] * len(bd), i
    for k in range(len(bd) - 1, -1, -1):
        idx[k] = r % bd[k]
        r //= bd[k]
    return idx


def compute_factor(metric: Metric, seed, factoring_radius, choice):
    """ComputeFactor (StaticFactoring.hxx:116-239): (sub, box extents, shift, values, gradients)
    with gradients in index units, row-major over the box; None when the box is empty (the
    factor is then disabled, as values.empty() disables it)."""
    d = metric.d
    R = int(factoring_radius)  # DiscreteType, truncated
    if R < 0:
        sub, bdims, shift = False, list(metric.dims), [0] * d
    else:
        sub = True
        bottom = [max(0, int(math.floor(seed[i] - R))) for i in range(d)]
        top = [min(metric.dims[i] - 1, int(math.ceil(seed[i] + R))) for i in range(d)]
        shift = bottom
        bdims = [top[i] - bottom[i] for i in range(d)]
    size = 1
    for n in bdims:
        size *= n
    if size <= 0:
        return None
    values = [math.nan] * size
    grads = [[0.] * d for _ in range(size)]
    seed_dist = metric.guess_point(seed)
    if choice in ("Seed", "Key"):
        for i in range(size):
            loc = _box_index(i, bdims)
            p = [loc[k] + shift[k] + 0.5 for k in range(d)]
            v = [p[k] - seed[k] for k in range(d)]
            if all(x == 0. for x in v):
                g = [0.] * d
            else:
                g = [-x for x in seed_dist.gradient([-x for x in v])]
            grads[i] = g
            s = 0.
            for k in range(d):
                s += g[k] * v[k]
            values[i] = s
    elif choice == "Both":
        def in_box(idx):
            return all(0 <= idx[k] < bdims[k] for k in range(d))

        def conv(idx):
            r = idx[0]
            for k in range(1, d):
                r = r * bdims[k] + idx[k]
            return r

        for pi in range(size):
            pidx = _box_index(pi, bdims)
            p = [pidx[k] + shift[k] + 0.5 for k in range(d)]
            v = [p[k] - seed[k] for k in range(d)]
            pdist = metric.guess_index(pidx)  # the box index, as the reference passes it
            null = all(x == 0. for x in v)
            seed_g = [0.] * d if null else [-x for x in seed_dist.gradient([-x for x in v])]
            p_g = [0.] * d if null else [-x for x in pdist.gradient([-x for x in v])]
            g = [0.5 * (p_g[k] + seed_g[k]) for k in range(d)]
            s = 0.
            for k in range(d):
                s += g[k] * v[k]
            values[pi] = s
            for k in range(d):
                grads[pi][k] += g[k]
            for k in range(d):
                for eps in (-1, 1):
                    q = list(pidx)
                    q[k] += eps
                    if in_box(q):
                        qp = [q[j] + shift[j] + 0.5 for j in range(d)]
                        qval = pdist.norm([-(qp[j] - seed[j]) for j in range(d)]) / 2.
                        qi = conv(q)
                        q[k] += eps
                        if in_box(q):
                            qval /= 2
                        grads[qi][k] -= eps * qval
                    else:
                        sp = 0.
                        for j in range(d):
                            sp += p_g[j] * v[j]
                        grads[pi][k] += eps * sp / 2.
    else:
        raise RuntimeError(f"StaticFactoring error : Unsupported factoringPointChoice {choice}.")
    return sub, bdims, shift, np.array(values), np.array(grads, np.float64).reshape(size * d)


def import_factor(params, dims, scales):
    """ImportFactor (StaticFactoring.hxx:77-113), RowMajor arrays: gradients are re-dimensioned
    (times the grid scale per axis) into index units."""
    d = len(dims)
    vals = np.asarray(params["factoringValues"], np.float64)
    if "factoringGradients" not in params:
        raise RuntimeError("StaticFactoring error : factoringGradients are not provided")
    grads = np.asarray(params["factoringGradients"], np.float64)
    mask = (np.asarray(params["factoringMask"], np.float64) != 0) if "factoringMask" in params \
        else np.ones(vals.shape, bool)
    sub = "factoringIndexShift" in params
    shift = [int(x) for x in np.asarray(params["factoringIndexShift"], np.float64).ravel()] if sub else [0] * d
    want = list(vals.shape) if sub else list(dims)
    if list(vals.shape) != want or list(grads.shape[:-1]) != want or list(mask.shape) != want:
        raise RuntimeError("Static factoring error : inconsistent data dimension.")
    g = grads.reshape(-1, d) * np.asarray(scales, np.float64)[None, :]
    return sub, want, shift, np.ascontiguousarray(vals.ravel()), np.ascontiguousarray(g.ravel()), \
        np.ascontiguousarray(mask.ravel(), np.uint8)


def setup(params, model, dims, h, origin, metric, dual, seeds, seed_values):
    """Setup + SetSeeds (StaticFactoring.hxx:58-75,242-393) for the given input dict: returns
    (seeds, seed values, factor or None, exports) with seeds as physical points.  The factor is
    (sub, box extents, shift, values, gradients, mask or None)."""
    riemann = model in ("Riemann2", "Riemann3")
    d = len(dims)
    has_import = "factoringValues" in params
    has_radius = "factoringRadius" in params
    compute = has_radius and float(params["factoringRadius"]) != 0
    if has_import and compute:
        raise RuntimeError("StaticFactoring error : factoringValues (import) and factoringRadius (compute) "
                           "both specified.")
    if riemann:
        scales = [float(x) for x in h]
        org = [0.] * d if origin is None else [float(x) for x in origin]
    else:  # the bundle axis: scale 2 pi / n_theta, origin minus half of it (HFM_ParamDefault.hxx:24-31)
        scales = [float(h[0]), float(h[0]), 2 * math.pi / dims[2]]
        org = ([0., 0.] if origin is None else [float(x) for x in origin][:2]) + [-scales[2] / 2.]
    factor = import_factor(params, dims, scales) if has_import else None
    disabled = factor is None and not has_radius
    exports = {}
    seeds = np.asarray(seeds, np.float64).reshape(-1, d)
    sv = np.zeros(len(seeds)) if seed_values is None else np.asarray(seed_values, np.float64).ravel()
    adim = [[(float(s[i]) - org[i]) / scales[i] for i in range(d)] for s in seeds]
    if has_radius and len(seeds) != 1:
        raise RuntimeError("Exactly one seed point must be specified for computing the factor, found "
                           f"{len(seeds)}.")
    seed0 = adim[0] if len(adim) else None
    radius = float(params.get("seedRadius", 0. if disabled else 2.))
    mobj = Metric(metric, dual, dims, h) if riemann else None
    if radius != 0 and len(seeds):
        if not riemann:
            raise RuntimeError("Equation factoring error : no guess")
        if radius < 0:
            raise NotImplementedError("negative seedRadius is not restated here")
        pts, vals = spread_seeds(mobj, adim, list(sv), abs(radius))
        seeds = np.array([[q[i] * scales[i] + org[i] for i in range(d)] for q in pts]).reshape(-1, d)
        sv = np.array(vals)
        exports["spreadedSeeds"], exports["spreadedSeedValues"] = seeds, sv
    if compute:
        if not riemann:
            raise RuntimeError("Equation factoring error : no guess")
        choice = params.get("factoringPointChoice", "Seed" if float(params.get("order", 1)) < 3 else "Both")
        f = compute_factor(mobj, seed0, float(params["factoringRadius"]), choice)
        if f is not None:
            sub, bdims, shift, vals, grads = f
            factor = (sub, bdims, shift, vals, grads, None)
            if float(params.get("exportFactoring", 0)):
                exports["factoringValues"] = vals.reshape(bdims)
                exports["factoringGradients"] = grads.reshape(tuple(bdims) + (d,))
                if sub:
                    exports["factoringIndexShift"] = np.array(shift, np.float64)
    return seeds, sv, factor, exports
