"""GPU walls (WallObstructionTest.h): the visibility mask in the update, flow, geodesic, adjoint
and forward-variation kernels, against the fixtures the reference executables wrote with walls
(tests/golden/walls.npz) and against the oracle on larger grids.

Tolerances as for the solves without walls: values 1e-12 relative (FP64) / 1e-5 (FP32, floored
like the other value tests); activeNeighs equal (FP64); geodesicFlow 1e-10 (FP64); geodesic
points 1e-6 of a grid step (FP64); costSensitivity 1e-9; valueVariation 1e-11."""
import os

import numpy as np
import pytest

import oracle
import hamiltonfastmarching_amd as hfm
from hamiltonfastmarching_amd import synthetic as syn

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DROP = ("tips", "inspectSensitivity", "inspectSensitivityWeights", "inspectSensitivityLengths")
CASES = {
    "riemann2": (lambda: syn.riemann2(48, seed=1), "riemann2", {}),
    "riemann2_refine": (lambda: syn.riemann2(48, seed=1), "riemann2", {"refineStencilAtWallBoundary": 1.0}),
    "riemann3": (lambda: syn.riemann3(20, seed=2), "riemann3", {}),
    "rsf2": (lambda: syn.rsf2(24, 16, seed=3), "rsf2", {}),
    "elastica2": (lambda: syn.elastica2(20, 16, seed=4), "elastica2", {}),
}
TIPS = {"riemann2": [[0.9, 0.5], [0.7, 0.3]], "riemann3": [[0.9, 0.5, 0.5], [0.95, 0.1, 0.9]],
        "rsf2": [[0.8, 0.2, 1.0], [0.5, 0.9, 4.0]], "elastica2": [[0.2, 0.8, 0.5], [0.85, 0.15, 2.0]]}


def _strip(p):
    return {k: v for k, v in p.items() if k not in DROP}


def _floor_rel(u, ref):
    fin = np.isfinite(ref)
    pos = ref[fin & (ref > 0)]
    floor = pos.min() if pos.size else 0.0
    m = fin & (ref >= floor) & (ref > 0)
    return float(np.max(np.abs(u[m] - ref[m]) / ref[m])) if m.any() else 0.0


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", [64, 32])
def test_walls_vs_golden(name, precision):
    print(f"walls golden {name} precision {precision}", flush=True)
    g = np.load(os.path.join(GOLD, "walls.npz"))
    make, base, extra = CASES[name]
    p = dict(_strip(make()), walls=g[name + "_walls"], exportActiveNeighs=1.0, exportGeodesicFlow=1.0,
             tips=np.array(TIPS[base]), precision=float(precision), **extra)
    out = hfm.run(p)
    u, ref = out["values"], g[name + "_values"]
    assert np.array_equal(np.isfinite(u), np.isfinite(ref))
    err = _floor_rel(u, ref)
    act = float(np.mean(out["activeNeighs"] == g[name + "_activeNeighs"]))
    print(f"{name} fp{precision}: values rel err {err:.3e}, activeNeighs equal {act:.5f}", flush=True)
    if precision == 64:
        assert err <= 1e-12
        assert act == 1.0
        fl, fref = out["geodesicFlow"], g[name + "_geodesicFlow"]
        ferr = float(np.abs(fl - fref).max() / max(np.abs(fref).max(), 1e-300))
        pts, pref = out["geodesicPoints"], g[name + "_geodesicPoints"]
        assert np.array_equal(out["geodesicLengths"], g[name + "_geodesicLengths"])
        h = float(np.atleast_1d(make()["gridScale"])[0])
        perr = float(np.abs(pts - pref).max()) / h
        print(f"{name}: flow rel err {ferr:.3e}, geodesic points err {perr:.3e} h", flush=True)
        assert ferr <= 1e-10
        assert perr <= 1e-6
    else:
        assert err <= 1e-5
    assert out["MaxStencilWidth"] == g[name + "_MaxStencilWidth"]


def _random_walls(dims, seed, frac=0.05, keep=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    w = (rng.random(dims) < frac).astype(float)
    if keep is not None:
        w[keep] = 0
    return w


LARGE = [
    ("riemann3_32", lambda: syn.riemann3(32, seed=31), (slice(14, 18),) * 3),
    ("riemann2_96", lambda: syn.riemann2(96, seed=32), (slice(46, 50),) * 2),
    ("rsf2_40x40x16", lambda: syn.rsf2(40, 16, seed=33), None),
]


@pytest.mark.parametrize("name,make,keep", LARGE, ids=[c[0] for c in LARGE])
def test_walls_vs_oracle(name, make, keep):
    print(f"walls oracle {name}", flush=True)
    p = _strip(make())
    dims = tuple(int(x) for x in p["dims"])
    w = _random_walls(dims, 7, keep=keep)
    if keep is None:  # keep the seeds free
        for s in np.asarray(p["seeds"]).reshape(-1, 3):
            i, j = int(s[0] * dims[0]), int(s[1] * dims[1])
            w[max(i - 1, 0):i + 2, max(j - 1, 0):j + 2, :] = 0
    q = dict(p, walls=w)
    ref = oracle.solve(q, active=True)
    out = hfm.run(dict(q, exportActiveNeighs=1.0))
    err = _floor_rel(out["values"], ref["values"])
    act = float(np.mean(out["activeNeighs"] == ref["activeNeighs"].astype(np.float64)))
    print(f"{name}: values rel err {err:.3e}, activeNeighs equal {act:.5f}", flush=True)
    assert np.array_equal(np.isfinite(out["values"]), np.isfinite(ref["values"]))
    assert err <= 1e-12
    assert act == 1.0
    # the device plan with walls set directly gives the same field
    import torch
    plan = hfm.Plan(p["model"], dims, np.atleast_1d(p["gridScale"])[0], origin=p.get("origin"), precision=64,
                    xi=float(p.get("xi", 1.0)), eps=float(p.get("eps", 0.1)))
    if "metric" in p:
        plan.set_metric(torch.from_numpy(p["metric"]))
    else:
        plan.set_cost(torch.from_numpy(p["cost"]))
    plan.set_walls(torch.from_numpy(w))
    plan.solve(p["seeds"])
    assert np.array_equal(plan.values().cpu().numpy(), out["values"])
    # removing the walls restores the solve without walls
    plan.set_walls(None)
    plan.solve(p["seeds"])
    free = plan.values().cpu().numpy()
    assert _floor_rel(free, oracle.solve(p)["values"]) <= 1e-12


def test_walls_adjoint_and_forward():
    """Backward and forward differentiation see the walls through the active sets."""
    print("walls adjoint + forward", flush=True)
    g = np.load(os.path.join(GOLD, "walls.npz"))
    p = dict(_strip(syn.rsf2(24, 16, seed=3)), walls=g["rsf2_walls"])
    tg = np.array([[0.8, 0.2, 1.0], [0.5, 0.9, 4.0], [0.3, 0.7, 2.5]])
    tw = np.array([1.0, 0.5, 2.0])
    _, sref, (sidx, sw) = oracle.solve_backward(p, tg, tw)
    out = hfm.run(dict(p, inspectSensitivity=tg, inspectSensitivityWeights=tw,
                       inspectSensitivityLengths=np.array([3.0])))
    s = out["costSensitivity_0"]
    err = float(np.abs(s - sref).max() / np.abs(sref).max())
    dims = tuple(int(x) for x in p["dims"])
    rng = np.random.Generator(np.random.PCG64(4))
    cv = rng.normal(size=dims + (2,))
    nseeds = len(np.asarray(p["seeds"]).reshape(-1, 3))
    sv = rng.normal(size=(2, nseeds))
    _, vref = oracle.forward_variation(p, cv, sv)
    vv = hfm.run(dict(p, costVariation=cv, seedValueVariation=sv))["valueVariation"]
    ferr = float(np.abs(vv - vref).max() / np.abs(vref).max())
    print(f"walls: costSensitivity rel err {err:.3e}, valueVariation rel err {ferr:.3e}", flush=True)
    assert err <= 1e-9
    assert ferr <= 1e-11
