"""GPU: the stopping criteria (ExtraAlgorithms/CommonStoppingCriteria.h:96-165) on the serial acceptance loop,
against the reference's outputs bit for bit in FP64, through the IO path and the Plan API."""
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

FIX = os.path.join(HERE, "golden", "stopping.npz")
CASES = mg.stopping_params()


@pytest.mark.parametrize("name", list(CASES))
def test_matches_reference(name):
    print(f"stopping case {name}", flush=True)
    d = np.load(FIX)
    out = hfm.run(dict(CASES[name], exportActiveNeighs=1.0))
    assert np.array_equal(out["values"], d[f"{name}_values"])
    assert np.array_equal(out["activeNeighs"], d[f"{name}_activeNeighs"])
    assert out["nAccepted"] == float(d[f"{name}_nAccepted"])


def test_out_of_range_target_warns():
    io = hfm.HFMIO(default_ordering="RowMajor")
    for k, v in CASES["riemann2_all_o1"].items():
        io[k] = v
    io.run()
    assert "Warning : target {2,0.5} is out of range." in io.messages()
    io = hfm.HFMIO(default_ordering="RowMajor")
    for k, v in CASES["rsf2_anyU_o1"].items():
        io[k] = v
    io["stopWhenAnyAccepted_Unoriented"] = np.array([[0.125, 0.775], [3.0, 0.5]])
    io.run()
    assert 'Warning : unoriented target "p" -> {3,0.5}, is out of range.' in io.messages()


def test_plan_api_stopping():
    """Plan.set_stopping with row-major indices matches the oracle; removing the criteria gives a
    complete first-order solve again."""
    import torch
    p = syn.riemann3(16, seed=25)
    plan = hfm.Plan("Riemann3", (16, 16, 16), p["gridScale"], precision=64)
    plan.set_metric(torch.from_numpy(p["metric"]).cuda())
    tgt = [np.ravel_multi_index((3, 12, 5), (16, 16, 16)), np.ravel_multi_index((14, 2, 9), (16, 16, 16))]
    plan.set_stopping(all_idx=tgt)
    plan.solve(p["seeds"])
    u = plan.values().cpu().numpy()
    h = p["gridScale"]
    pts = np.array([[(3.5) * h, 12.5 * h, 5.5 * h], [14.5 * h, 2.5 * h, 9.5 * h]])
    ref = oracle.solve(dict(p, stopWhenAllAccepted=pts))
    assert np.array_equal(u, ref["values"])
    assert plan.stats()["n_accepted"] == ref["nAccepted"]
    plan.set_stopping()
    plan.solve(p["seeds"])
    u1 = plan.values().cpu().numpy()
    full = oracle.solve(p)["values"]  # points no stencil reaches stay at +inf in both
    fin = np.isfinite(full)
    assert np.array_equal(np.isfinite(u1), fin)
    assert np.max(np.abs(u1[fin] - full[fin]) / np.maximum(full[fin], 1e-300)) < 1e-12
    assert plan.stats()["n_accepted"] == -1


@pytest.mark.parametrize("extra,msg", [
    ({"precision": 32.0}, "precision 64"),
    ({"walls": "box"}, "walls"),
])
def test_refusals(extra, msg):
    p = dict(CASES["riemann2_max_o1"])
    if extra.get("walls") == "box":
        w = np.zeros((40, 40))
        w[10, 5:30] = 1
        extra = {"walls": w}
    with pytest.raises(hfm.PyExp, match=msg):
        hfm.run(dict(p, **extra))
