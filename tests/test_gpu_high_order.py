"""GPU: order 2 / 3 solves (the serial acceptance loop) against the reference's outputs and the
oracle, bit for bit in FP64, and the refusals of what an order 2 / 3 solve does not support
(FP32 plans, walls)."""
import os
import sys

import numpy as np
import pytest

import oracle
import hamiltonfastmarching_amd as hfm
from hamiltonfastmarching_amd import synthetic as syn

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

FIX = os.path.join(HERE, "golden", "high_order.npz")


def _params(name):
    p = mg.HIGH_ORDER_CASES[name]()
    return {k: v for k, v in p.items()
            if k not in ("tips", "inspectSensitivity", "inspectSensitivityWeights", "inspectSensitivityLengths")}


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("name", list(mg.HIGH_ORDER_CASES))
def test_matches_reference(name, order):
    print(f"case {name} order {order}", flush=True)
    d = np.load(FIX)
    q = dict(_params(name), order=float(order), exportActiveNeighs=1.0, exportGeodesicFlow=1.0,
             exportActiveOffsets=1.0)
    out = hfm.run(q)
    pre = f"{name}_order{order}_"
    assert np.array_equal(out["values"], d[pre + "values"])
    assert np.array_equal(out["activeNeighs"], d[pre + "activeNeighs"])
    assert np.array_equal(out["geodesicFlow"], d[pre + "geodesicFlow"])
    assert np.array_equal(out["activeOffsets"], d[pre + "activeOffsets"])
    assert out["nAccepted"] == float(d[pre + "nAccepted"])


@pytest.mark.parametrize("order", [2, 3])
def test_several_seeds_match_oracle(order):
    """Three seeds with seed values (pushed in index order, ties between them included)."""
    print(f"several seeds order {order}", flush=True)
    p = syn.riemann2(48, seed=21)
    p["seeds"] = np.array([[0.2, 0.3], [0.7, 0.8], [0.6, 0.25]])
    p["seedValues"] = np.array([0.0, 0.05, 0.0])
    ref = oracle.solve(dict(p, order=order), active=True)
    out = hfm.run(dict(p, order=float(order), exportActiveNeighs=1.0))
    assert np.array_equal(out["values"], ref["values"])
    assert np.array_equal(out["activeNeighs"], ref["activeNeighs"].astype(np.float64))


def test_plan_api_order():
    """Plan.set_order gives the same values as the IO path; a plan switched back to order 1 solves
    with the block fast-iterative method again."""
    import torch
    p = syn.riemann3(16, seed=22)
    plan = hfm.Plan("Riemann3", (16, 16, 16), p["gridScale"], precision=64)
    plan.set_metric(torch.from_numpy(p["metric"]).cuda())
    plan.set_order(2)
    plan.solve(p["seeds"])
    u2 = plan.values().cpu().numpy()
    assert np.array_equal(u2, oracle.solve(dict(p, order=2))["values"])
    plan.set_order(1)
    plan.solve(p["seeds"])
    u1 = plan.values().cpu().numpy()
    r1 = oracle.solve(p)["values"]
    fin = np.isfinite(r1)
    assert np.max(np.abs(u1[fin] - r1[fin]) / np.maximum(r1[fin], 1e-300)) < 1e-12
    assert not np.array_equal(u1, u2)


@pytest.mark.parametrize("extra,msg", [
    ({"precision": 32.0}, "precision 64"),
    ({"walls": "box"}, "walls"),
])
def test_refusals(extra, msg):
    p = syn.riemann2(24, seed=23)
    q = dict(p, order=2.0, **extra)
    if isinstance(extra.get("walls"), str):
        w = np.zeros((24, 24))
        w[5, 5:10] = 1
        q["walls"] = w
    with pytest.raises(Exception, match=msg):
        hfm.run(q)
