"""CPU tests of the oracle (the FP64 restatement that the GPU library is checked against).

Pins, in order of strength:
  * golden fixtures written by the reference's own FileHFM executables (tests/golden/): values,
    activeNeighs, geodesicFlow, activeOffsets, Selling offsets, shared curvature stencils,
    costSensitivity --
    all required bit-exact;
  * closed forms (identity metric, constant metrics), invariants (Selling reconstruction,
    non-negative weights, unimodular superbase; the discrete eikonal residual), symmetry and a
    brute-force Gauss-Seidel fixed point on tiny grids;
  * planted mistakes: perturbed outputs must fail the same checks.
"""
import os

import numpy as np
import pytest

import oracle
from hamiltonfastmarching_amd import synthetic as syn
from pack_utils import pack_active_offsets

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ------------------------------------------------------------------------------------------
# Golden fixtures from the reference
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("name,make", [("riemann2_64", lambda: syn.riemann2(64, seed=1)),
                                       ("riemann3_24", lambda: syn.riemann3(24, seed=2))])
def test_riemann_golden(name, make):
    g = load(name)
    p = make()
    assert sha(p["metric"]) == str(g["metric_sha256"]), "synthetic generator changed"
    out = oracle.solve(p, flow=True, active=True)
    assert np.array_equal(out["values"], g["values"])
    assert np.array_equal(out["activeNeighs"], g["activeNeighs"].astype(np.uint64))
    assert np.array_equal(out["geodesicFlow"], g["geodesicFlow"])
    dims = p["metric"].shape[:-1]
    off, w = oracle.riemann_stencils(dims, p["gridScale"], p["metric"])
    assert np.array_equal(off, g["offsets"])
    np.testing.assert_allclose(w, g["weights_printed"], rtol=1e-5, atol=1e-12)
    nz = w > 0
    l1 = np.abs(off.astype(int)).sum(-1)
    assert int(l1[nz].max()) == int(g["MaxStencilWidth"])


@pytest.mark.parametrize("key", ["metric", "dualMetric"])
def test_selling_ties_golden(key):
    g = load("ties3")
    t = g["tensors"].reshape(-1, 6)
    off, _ = oracle.riemann_stencils((4, 4, 4), 1.0, t, dual=(key == "dualMetric"))
    assert np.array_equal(off, g[key + "_offsets"])


@pytest.mark.parametrize("name,model,make", [
    ("rsf2_32x32x16", "ReedsSheppForward2", lambda: syn.rsf2(32, 16, seed=3)),
    ("elastica2_24x24x16", "Elastica2", lambda: syn.elastica2(24, 16, seed=4))])
def test_curvature_golden(name, model, make):
    g = load(name)
    p = make()
    assert sha(p["cost"]) == str(g["cost_sha256"])
    p.pop("tips", None)
    out = oracle.solve(p, flow=True, active=True)
    assert np.array_equal(out["values"], g["values"])
    assert np.array_equal(out["activeNeighs"], g["activeNeighs"].astype(np.uint64))
    assert np.array_equal(out["geodesicFlow"], g["geodesicFlow"])
    off, w, sym = oracle.curvature_stencils(model, 16, p["gridScale"], p["xi"], p["eps"])
    assert np.array_equal(off, g["table_offsets"])
    np.testing.assert_allclose(w, g["table_weights_printed"], rtol=1e-5, atol=1e-12)
    if "costSensitivity_0" in g:
        v, s, (si, sw) = oracle.solve_backward(p, p["inspectSensitivity"], p["inspectSensitivityWeights"])
        assert np.array_equal(s, g["costSensitivity_0"])
        ss = g["seedSensitivity_0"]
        assert np.array_equal(sw, ss[:, -1])


@pytest.mark.parametrize("name,make,model", [("riemann2_64", lambda: syn.riemann2(64, seed=1), "Riemann2"),
                                             ("riemann3_24", lambda: syn.riemann3(24, seed=2), "Riemann3"),
                                             ("rsf2_32x32x16", lambda: syn.rsf2(32, 16, seed=3),
                                              "ReedsSheppForward2"),
                                             ("elastica2_24x24x16", lambda: syn.elastica2(24, 16, seed=4),
                                              "Elastica2")])
def test_active_offsets_golden(name, make, model):
    """The reference's activeOffsets export equals the packing of the oracle's active set over
    the oracle's stencils (which are themselves pinned above)."""
    g = load(name)
    p = make()
    p.pop("tips", None)
    out = oracle.solve(p, active=True)
    if model.startswith("Riemann"):
        dims = p["metric"].shape[:-1]
        offs, _ = oracle.riemann_stencils(dims, p["gridScale"], p["metric"])
        nfwd, dim = 0, len(dims)
    else:
        dims = p["cost"].shape
        tab, _, _ = oracle.curvature_stencils(model, dims[2], p["gridScale"], p["xi"], p["eps"])
        offs = np.broadcast_to(tab[None, None], dims[:2] + tab.shape).reshape(-1, tab.shape[1], 3)
        nfwd, dim = (3 if model == "ReedsSheppForward2" else 27), 3
    mine = pack_active_offsets(out["activeNeighs"], offs, nfwd, dim)
    ref = g["activeOffsets"].reshape(mine.shape)
    assert np.array_equal(mine, ref)
    # planted mistakes: little-endian packing, or dropping the side sign, must fail
    rev = pack_active_offsets(out["activeNeighs"], offs[..., ::-1].copy(), nfwd, dim)
    assert not np.array_equal(rev, ref)
    if nfwd < offs.shape[1]:
        act = out["activeNeighs"].astype(np.uint64).ravel()
        b = np.uint64(nfwd + 1)
        flipped = act ^ (((act >> b) & np.uint64(1)) << b) ^ (((act >> b) & np.uint64(1)) << (b - np.uint64(1)))
        assert not np.array_equal(pack_active_offsets(flipped, offs, nfwd, dim), ref)


def test_riemann3_128_digest():
    """Benchmark-size values against the reference's digest and 4096 seeded samples."""
    g = load("riemann3_128")
    p = syn.riemann3(128, seed=0)
    assert sha(p["metric"]) == str(g["metric_sha256"])
    v = oracle.solve(p)["values"]
    assert np.array_equal(v.ravel()[g["sample_index"]], g["sample_values"])
    assert sha(v) == str(g["values_sha256"])


def test_riemann3_128_offsets_digest():
    """Benchmark-size Selling offsets: all 2,097,152 x 6 offsets against the digest of the
    reference's getStencils output (row-major points, slot order), 4096 seeded samples, and
    MaxStencilWidth.  A transposed component or a swapped slot changes the digest."""
    g = load("riemann3_128")
    p = syn.riemann3(128, seed=0)
    off, w = oracle.riemann_stencils((128,) * 3, p["gridScale"], p["metric"])
    assert np.array_equal(off[g["sample_index"]], g["sample_offsets"])
    assert sha(off) == str(g["offsets_sha256"])
    # MaxStencilWidth: largest L1 norm of a nonzero-weight side x +- e whose neighbour is in the grid
    c = np.stack(np.unravel_index(np.arange(128 ** 3), (128,) * 3), axis=1)[:, None, :]
    o = off.astype(np.int32)
    inside = np.all((c + o >= 0) & (c + o < 128), axis=2) | np.all((c - o >= 0) & (c - o < 128), axis=2)
    l1 = np.abs(o).sum(axis=2)
    assert int(np.max(np.where((w != 0) & inside, l1, 0))) == int(g["MaxStencilWidth"])
    assert int(np.max(np.where(w != 0, l1, 0))) > int(g["MaxStencilWidth"])  # the grid test matters here
    # planted mistakes: components reversed, or two slots swapped, change the digest
    assert sha(np.ascontiguousarray(off[..., ::-1])) != str(g["offsets_sha256"])
    assert sha(np.ascontiguousarray(off[:, [1, 0, 2, 3, 4, 5]])) != str(g["offsets_sha256"])


# ------------------------------------------------------------------------------------------
# Closed forms and invariants
# ------------------------------------------------------------------------------------------
def _sym3(M):
    return np.array([M[0, 0], M[1, 0], M[1, 1], M[2, 0], M[2, 1], M[2, 2]])


def _unsym(c, d):
    if d == 2:
        return np.array([[c[0], c[1]], [c[1], c[2]]])
    return np.array([[c[0], c[1], c[3]], [c[1], c[2], c[4]], [c[3], c[4], c[5]]])


def test_identity_2d_closed_form():
    """D = I: offsets (0,1), (-1,0) with weight 1 and (1,-1) with weight 0 (the 4-point scheme);
    values along the axes are exactly k h; at a diagonal neighbour (a+b+sqrt(2h^2-(a-b)^2))/2."""
    off, w = oracle.riemann_stencils((1, 1), 1.0, np.array([1.0, 0.0, 1.0]))
    assert sorted(map(tuple, off[0].tolist())) == [(-1, 0), (0, 1), (1, -1)]
    assert np.allclose(sorted(w[0]), [0, 1, 1])
    n, h = 21, 0.1
    p = {"model": "Riemann2", "dims": [n, n], "gridScale": h, "metric": np.array([1.0, 0.0, 1.0]),
         "seeds": np.array([[(n // 2 + 0.5) * h] * 2])}
    u = oracle.solve(p)["values"]
    c = n // 2
    for k in range(n - c):
        assert u[c + k, c] == pytest.approx(k * h, rel=1e-14, abs=1e-15)
        assert u[c, c - k] == pytest.approx(k * h, rel=1e-14, abs=1e-15)
    a = u[c + 1, c]
    assert u[c + 1, c + 1] == pytest.approx((a + a + np.sqrt(2 * h * h)) / 2, rel=1e-14)


def test_selling_invariants_random():
    rng = np.random.Generator(np.random.PCG64(5))
    for d in (2, 3):
        K = 3 if d == 2 else 6
        A = rng.standard_normal((500, d, d))
        M = A @ np.swapaxes(A, 1, 2) + 0.05 * np.eye(d)
        comps = np.array([[m[0, 0], m[1, 0], m[1, 1]] if d == 2 else _sym3(m) for m in M])
        off, w = oracle.riemann_stencils((500,) + (1,) * (d - 1), 1.0, comps, dual=True)
        assert (w >= 0).all()
        for i in range(500):
            D = _unsym(comps[i], d)
            R = sum(w[i, k] * np.outer(off[i, k], off[i, k]) for k in range(K))
            # reconstruction up to the zeroed weights (< 1000 eps Tr D each)
            assert np.allclose(R, D, atol=K * 1000 * 2.3e-16 * np.trace(D) + 1e-12 * np.abs(D).max())


def test_planted_selling_mistake_is_caught():
    """A transposed offset or a wrong sign in a weight breaks the reconstruction check."""
    comps = np.array([2.0, 0.7, 1.5, 0.3, -0.2, 1.1])
    off, w = oracle.riemann_stencils((1, 1, 1), 1.0, comps, dual=True)
    D = _unsym(comps, 3)
    good = sum(w[0, k] * np.outer(off[0, k], off[0, k]) for k in range(6))
    assert np.allclose(good, D)
    bad_off = off.copy()
    bad_off[0, 0] = bad_off[0, 0][::-1]
    bad = sum(w[0, k] * np.outer(bad_off[0, k], bad_off[0, k]) for k in range(6))
    w2 = w.copy()
    w2[0, 1] = -w2[0, 1]
    bad2 = sum(w2[0, k] * np.outer(off[0, k], off[0, k]) for k in range(6))
    assert not (np.allclose(bad, D) and np.allclose(bad2, D))


def residual(p, u):
    """max |sum_k w_k max(0, u - u(x+e), u - u(x-e))^2 - 1| over non-seed reachable points."""
    dims = u.shape
    d = len(dims)
    off, w = oracle.riemann_stencils(dims, p["gridScale"], p["metric"])
    N = u.size
    uf = u.ravel()
    idx = np.array(np.unravel_index(np.arange(N), dims)).T
    res = np.zeros(N)
    for k in range(off.shape[1]):
        best = np.zeros(N)
        for s in (1, -1):
            nb = idx + s * off[:, k, :].astype(int)
            ok = np.all((nb >= 0) & (nb < np.array(dims)), axis=1)
            vv = np.full(N, np.inf)
            vv[ok] = uf[np.ravel_multi_index(nb[ok].T, dims)]
            best = np.maximum(best, np.where(np.isfinite(vv), np.maximum(0, uf - vv), 0))
        res += w[:, k] * best ** 2
    return res


def test_residual_invariant_and_planted_perturbation():
    p = syn.riemann3(12, seed=3)
    u = oracle.solve(p)["values"]
    r = residual(p, u)
    seed_idx = np.ravel_multi_index((6, 6, 6), u.shape)
    fin = np.isfinite(u.ravel())
    mask = fin.copy()
    mask[seed_idx] = False
    assert np.max(np.abs(r[mask] - 1)) < 1e-9
    u2 = u.copy()
    u2[3, 4, 5] *= 1 + 1e-6
    r2 = residual(p, u2)
    assert np.max(np.abs(r2[mask] - 1)) > 1e-7


def test_bruteforce_gauss_seidel_fixed_point():
    """FP64 Gauss-Seidel sweeps of the local solve to a fixed point equal the FM values."""
    p = syn.riemann2(10, seed=4)
    u_fm = oracle.solve(p)["values"]
    dims = u_fm.shape
    off, w = oracle.riemann_stencils(dims, p["gridScale"], p["metric"])
    u = np.full(dims, np.inf)
    u[5, 5] = 0.0
    for _ in range(200):
        changed = False
        for i in range(dims[0]):
            for j in range(dims[1]):
                if (i, j) == (5, 5):
                    continue
                lin = i * dims[1] + j
                cand = []
                for k in range(3):
                    if w[lin, k] == 0:
                        continue
                    vs = []
                    for s in (1, -1):
                        a, b = i + s * off[lin, k, 0], j + s * off[lin, k, 1]
                        if 0 <= a < dims[0] and 0 <= b < dims[1]:
                            vs.append(u[a, b])
                    if vs:
                        cand.append((min(vs), w[lin, k]))
                cand.sort()
                r, A, B, C = np.inf, 0.0, 0.0, -1.0
                m = None
                for v, wk in cand:
                    if not v < r:
                        break
                    m = v if m is None else m
                    vr = v - m
                    A += wk
                    B += wk * vr
                    C += (wk * vr) * vr
                    r = m + (B + np.sqrt(max(0.0, B * B - A * C))) / A
                if r < u[i, j]:
                    u[i, j] = r
                    changed = True
        if not changed:
            break
    fin = np.isfinite(u_fm)
    assert np.array_equal(np.isfinite(u), fin)
    assert np.max(np.abs(u[fin] - u_fm[fin])) < 1e-13


def test_reflection_symmetry():
    """A metric field symmetric under x -> -x with a centred seed gives symmetric values."""
    n = 17
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h - 0.5
    X, Y = np.meshgrid(x, x, indexing="ij")
    a = 1 + 3 * np.exp(-((X ** 2 + 2 * Y ** 2) * 10))
    m = np.stack([a, 0.3 * X * Y * 0 + 0.0, 1 + Y ** 2], axis=-1)
    p = {"model": "Riemann2", "dims": [n, n], "gridScale": h, "metric": m, "seeds": np.array([[0.5, 0.5]])}
    u = oracle.solve(p)["values"]
    assert np.allclose(u, u[::-1, :], rtol=1e-13)
    assert np.allclose(u, u[:, ::-1], rtol=1e-13)


def test_reference_and_port_agree_where_reference_built():
    """When the reference binaries are present (development container), the restatement is
    bit-exact against a fresh reference run on an unseen input."""
    if not oracle.ref_available("Riemann3"):
        pytest.skip("reference binaries not built here")
    p = syn.riemann3(14, seed=99, amax=30.0)
    r = oracle.run_reference(dict(p, exportActiveNeighs=1.0))
    mine = oracle.solve(p, active=True)
    assert np.array_equal(r["values"], mine["values"])
    assert np.array_equal(r["activeNeighs"].astype(np.uint64), mine["activeNeighs"])


# ------------------------------------------------------------------------------------------
# Geodesic backtracking (discrete solver)
# ------------------------------------------------------------------------------------------
GEO_CASES = {
    "rsf2": lambda: syn.rsf2(32, 16, seed=3),
    "riemann2": lambda: syn.riemann2(64, seed=1),
    "riemann3": lambda: syn.riemann3(24, seed=2),
    "elastica2": lambda: syn.elastica2(24, 16, seed=4),
}


def _strip(p):
    return {k: v for k, v in p.items() if k not in ("tips", "inspectSensitivity", "inspectSensitivityWeights",
                                                    "inspectSensitivityLengths")}


@pytest.mark.parametrize("name", list(GEO_CASES))
def test_geodesics_golden(name):
    """geodesicPoints / geodesicLengths bit-exact against the reference's discrete solver."""
    g = load("geodesics")
    p = _strip(GEO_CASES[name]())
    pts, lens = oracle.geodesics(p, g[name + "_tips"])
    assert np.array_equal(lens, g[name + "_lengths"].astype(np.int64))
    assert np.array_equal(pts, g[name + "_points"])
    # the same solver on given values / active sets (the entry the GPU kernel is checked with)
    o = oracle.solve(p, active=True)
    pts2, lens2 = oracle.geodesics_given(p, o["values"], o["activeNeighs"], g[name + "_tips"])
    assert np.array_equal(pts2, pts) and np.array_equal(lens2, lens)
    # planted mistake: another geodesicStep gives another curve
    pts3, _ = oracle.geodesics(p, g[name + "_tips"], step=0.3)
    assert pts3.shape != pts.shape or not np.array_equal(pts3, pts)


def test_geodesics_isotropic_straight():
    """Identity metric, seed at the centre: each geodesic runs from its tip to the seed along the
    straight segment (the discrete mass transport stays within a cell of it)."""
    n = 64
    h = 1.0 / n
    p = {"model": "Riemann2", "dims": np.array([n, n], float), "gridScale": h, "origin": np.zeros(2),
         "seeds": np.array([[0.5, 0.5]]), "metric": np.array([1.0, 0.0, 1.0]), "arrayOrdering": "RowMajor"}
    rng = np.random.Generator(np.random.PCG64(3))
    tips = rng.uniform(0.1, 0.9, size=(12, 2))
    pts, lens = oracle.geodesics(p, tips)
    seed = np.array([0.5, 0.5]) + 0.5 * h  # centre of the seed cell
    start = 0
    for t, L in zip(tips, lens):
        c = pts[start:start + L]
        start += L
        assert np.max(np.abs(c[0] - t)) < 1e-12
        assert np.linalg.norm(c[-1] - seed) < 1.5 * h
        d = seed - t
        rel = c - t
        s = np.clip(rel @ d / (d @ d), 0, 1)
        dist = np.linalg.norm(rel - s[:, None] * d, axis=1)
        assert dist.max() < 1.0 * h, (t, dist.max() / h)


@pytest.mark.parametrize("name", ["riemann2", "riemann3", "rsf2"])
def test_geodesics_descend_arrival_times(name):
    """Along each geodesic the arrival time at the nearest cell falls from the tip towards a seed
    (cell-rounded values of the averaged positions fluctuate, so only the trend is checked)."""
    g = load("geodesics")
    p = _strip(GEO_CASES[name]())
    u = oracle.solve(p)["values"]
    dims = np.array(u.shape)
    h = np.array([p["gridScale"]] * len(dims), float)
    origin = np.zeros(len(dims))
    if name == "rsf2":
        h[2] = 2 * np.pi / dims[2]
        origin[2] = -h[2] / 2
    pts, lens = g[name + "_points"], g[name + "_lengths"].astype(int)
    start = 0
    for L in lens:
        c = pts[start:start + L]
        start += L
        idx = np.floor((c - origin) / h).astype(int)
        if name == "rsf2":
            idx[:, 2] %= dims[2]
        idx = np.clip(idx, 0, dims - 1)
        val = u[tuple(idx.T)]
        assert val[-1] < 0.5 * val[0]
        if L >= 8:
            assert np.corrcoef(np.arange(L), val)[0, 1] < -0.5


# ------------------------------------------------------------------------------------------
# Forward variation (tangent mode)
# ------------------------------------------------------------------------------------------
FWD = {
    "riemann2": lambda: syn.riemann2(32, seed=1),
    "riemann3": lambda: syn.riemann3(12, seed=2),
    "rsf2": lambda: syn.rsf2(16, 8, seed=3),
    "elastica2": lambda: syn.elastica2(16, 8, seed=4),
}


@pytest.mark.parametrize("name", list(FWD))
def test_forward_variation_golden(name):
    g = load("forward_variation")
    p = _strip(FWD[name]())
    _, vv = oracle.forward_variation(p, g[name + "_costVariation"], g[name + "_seedValueVariation"])
    assert np.array_equal(vv, g[name + "_valueVariation"])
    # planted mistake: the seed variations' layout transposed
    sv = g[name + "_seedValueVariation"]
    if sv.shape[1] > 1:
        _, bad = oracle.forward_variation(p, g[name + "_costVariation"], sv[::-1].copy())
        assert not np.array_equal(bad, g[name + "_valueVariation"])


def _scaled(p, eps, dc):
    """The input with cost -> cost (1 + eps dc) (curvature models), metric -> (1 + eps dc)^2 metric
    (Riemann models: arrival times scale like a cost)."""
    q = dict(p)
    if "cost" in p:
        q["cost"] = p["cost"] * (1 + eps * dc)
    else:
        q["metric"] = p["metric"] * ((1 + eps * dc) ** 2)[..., None]
    return q


@pytest.mark.parametrize("name", ["riemann2", "rsf2"])
def test_forward_variation_finite_differences(name):
    """valueVariation is the derivative of the arrival times along the cost / seed-value
    direction: central differences with eps = 1e-7 agree (a larger step crosses the kinks where
    active sets change)."""
    p = _strip(FWD[name]())
    dims = tuple(int(x) for x in p["dims"])
    rng = np.random.Generator(np.random.PCG64(11))
    dc = rng.uniform(-1, 1, size=dims)
    cost = p["cost"] if "cost" in p else np.ones(dims)
    _, vv = oracle.forward_variation(p, (dc * cost)[..., None])
    eps = 1e-7
    up = oracle.solve(_scaled(p, eps, dc))["values"]
    dn = oracle.solve(_scaled(p, -eps, dc))["values"]
    fd = (up - dn) / (2 * eps)
    fin = np.isfinite(fd) & np.isfinite(vv[..., 0])
    err = np.abs(fd[fin] - vv[..., 0][fin])
    scale = np.abs(vv[..., 0][fin]).max()
    assert np.quantile(err, 0.99) < 1e-6 * scale
    # seed values: a unit variation of every seed value shifts every reached value by 1
    ns = len(np.asarray(p["seeds"]).reshape(-1, len(dims)))
    _, vs = oracle.forward_variation(p, None, np.ones((1, ns)))
    reached = np.isfinite(oracle.solve(p)["values"])
    assert np.allclose(vs[..., 0][reached], 1.0, atol=1e-12)


# ------------------------------------------------------------------------------------------
# Walls (WallObstructionTest.h:38-150)
# ------------------------------------------------------------------------------------------
WALLS = {
    "riemann2": (lambda: syn.riemann2(48, seed=1), "riemann2", {}),
    "riemann2_refine": (lambda: syn.riemann2(48, seed=1), "riemann2", {"refineStencilAtWallBoundary": 1.0}),
    "riemann3": (lambda: syn.riemann3(20, seed=2), "riemann3", {}),
    "rsf2": (lambda: syn.rsf2(24, 16, seed=3), "rsf2", {}),
    "elastica2": (lambda: syn.elastica2(20, 16, seed=4), "elastica2", {}),
}
WALL_TIPS = {"riemann2": [[0.9, 0.5], [0.7, 0.3]], "riemann3": [[0.9, 0.5, 0.5], [0.95, 0.1, 0.9]],
             "rsf2": [[0.8, 0.2, 1.0], [0.5, 0.9, 4.0]], "elastica2": [[0.2, 0.8, 0.5], [0.85, 0.15, 2.0]]}


def wall_params(name):
    g = load("walls")
    make, base, extra = WALLS[name]
    return dict(_strip(make()), walls=g[name + "_walls"], **extra), g, base


@pytest.mark.parametrize("name", list(WALLS))
def test_walls_golden(name):
    """values, activeNeighs, geodesicFlow and geodesics with walls equal the reference's."""
    p, g, base = wall_params(name)
    r = oracle.solve(p, flow=True, active=True)
    assert np.array_equal(r["values"], g[name + "_values"])
    assert np.array_equal(r["activeNeighs"].astype(np.float64), g[name + "_activeNeighs"])
    assert np.array_equal(r["geodesicFlow"], g[name + "_geodesicFlow"])
    pts, lens = oracle.geodesics(p, WALL_TIPS[base])
    assert np.array_equal(lens, g[name + "_geodesicLengths"].astype(np.int64))
    assert np.array_equal(pts, g[name + "_geodesicPoints"])
    # the walls matter here: without them the values differ
    assert not np.array_equal(oracle.solve(_strip(WALLS[name][0]()))["values"], r["values"])


def test_walls_refine_changes_the_result():
    """refineStencilAtWallBoundary lets differences from points next to a wall through."""
    g = load("walls")
    assert not np.array_equal(g["riemann2_values"], g["riemann2_refine_values"])


def test_walls_identities():
    """No wall cells = no walls; wall cells are never reached; a closed box around the seed leaves
    everything outside unreached; a planted mistake (walls ignored beyond the first step of the
    line walk) changes the fixture's result."""
    p = _strip(syn.riemann2(32, seed=5))
    u0 = oracle.solve(p)["values"]
    assert np.array_equal(oracle.solve(dict(p, walls=np.zeros((32, 32))))["values"], u0)
    w = np.zeros((32, 32))
    w[8, 8:25] = w[24, 8:25] = w[8:25, 8] = w[8:25, 24] = 1
    u = oracle.solve(dict(p, walls=w))["values"]
    inside = np.zeros((32, 32), bool)
    inside[9:24, 9:24] = True
    assert np.all(np.isinf(u[w != 0]))
    assert np.all(np.isinf(u[~inside & (w == 0)]))
    assert np.all(np.isfinite(u[inside]))
    # planted mistake in the oracle: the visibility test without its line walk (only the end
    # point's distance to the walls) disagrees with the reference's fixtures
    for name in ("riemann2", "riemann3"):
        q, g, _ = wall_params(name)
        bad = oracle.solve(dict(q, _plant_no_wall_walk=True))["values"]
        assert not np.array_equal(bad, g[name + "_values"])
