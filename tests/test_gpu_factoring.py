"""GPU: static factoring (StaticFactoring.hxx:11-393) on the serial acceptance loop, against the
reference's outputs bit for bit in FP64 -- computed factors ("Seed" and "Both", on a box and on
the whole grid, orders 1-3), imported ones (mask, index shift, periodic angle axis) and the
exported factor -- through the IO path and the Plan API, and the refusals."""
import os
import sys

import numpy as np
import pytest

import oracle
import hamiltonfastmarching_amd as hfm

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

FIX = os.path.join(HERE, "golden", "factoring.npz")
CASES = mg.factoring_params()


@pytest.mark.parametrize("name", list(CASES))
def test_matches_reference(name):
    print(f"factoring case {name}", flush=True)
    d = np.load(FIX)
    out = hfm.run(dict(CASES[name], exportActiveNeighs=1.0, exportGeodesicFlow=1.0, exportFactoring=1.0))
    assert np.array_equal(out["values"], d[f"{name}_values"])
    assert np.array_equal(out["activeNeighs"], d[f"{name}_activeNeighs"])
    assert np.array_equal(out["geodesicFlow"], d[f"{name}_geodesicFlow"])
    assert out["nAccepted"] == float(d[f"{name}_nAccepted"])
    for k in ("factoringValues", "factoringGradients", "factoringIndexShift", "spreadedSeeds", "spreadedSeedValues"):
        if f"{name}_{k}" in d.files:
            assert np.array_equal(np.asarray(out[k], np.float64), d[f"{name}_{k}"], equal_nan=True), k


def test_both_choice_warns():
    io = hfm.HFMIO(default_ordering="RowMajor")
    for k, v in CASES["r2_rad6_o3_both"].items():
        io[k] = v
    io.run()
    assert "factoringPointChoice = 'Both' yields a factor" in io.messages()


def test_plan_api_factoring():
    """Plan.set_factoring with the oracle's factor (whole grid, "Seed") and the spread seeds gives
    the oracle's values; removing the factor gives the block solver's first-order values again."""
    import torch
    from oracle import factoring as fac
    p = dict(CASES["r2_rad10_o1"])
    n = 40
    seeds, sv, f, _ = fac.setup(dict(p, factoringRadius=-1.0), "Riemann2", [n, n], [p["gridScale"]] * 2,
                                p["origin"], p["metric"], False, p["seeds"], None)
    sub, bd, shift, vals, grads, mask = f
    assert not sub
    plan = hfm.Plan("Riemann2", (n, n), p["gridScale"], precision=64)
    plan.set_metric(torch.from_numpy(p["metric"]).cuda())
    plan.set_factoring(vals.reshape(bd), grads.reshape(tuple(bd) + (2,)))
    plan.solve(seeds, sv)
    u = plan.values().cpu().numpy()
    ref = oracle.solve(dict(p, factoringRadius=-1.0))
    assert np.array_equal(u, ref["values"])
    assert plan.stats()["n_accepted"] == ref["nAccepted"]
    plan.set_factoring(None)
    plan.solve(seeds, sv)
    assert plan.stats()["n_accepted"] == -1


@pytest.mark.parametrize("extra,msg", [
    ({"precision": 32.0}, "precision 64"),
    ({"seeds": np.array([[0.3, 0.3], [0.6, 0.6]])}, "Exactly one seed"),
    ({"factoringPointChoice": "Bad"}, "Unsupported factoringPointChoice Bad"),
])
def test_refusals(extra, msg):
    with pytest.raises(hfm.PyExp, match=msg):
        hfm.run(dict(CASES["r2_rad10_o1"], **extra))


def test_import_and_compute_conflict():
    with pytest.raises(hfm.PyExp, match="both specified"):
        hfm.run(dict(CASES["r2_import_o1"], factoringRadius=3.0))


def test_curvature_compute_has_no_guess():
    """The curvature models have no distance guess (HamiltonFastMarching.h:203-204)."""
    p = {k: v for k, v in CASES["rsf2_import_o1"].items() if not k.startswith("factoring")}
    with pytest.raises(hfm.PyExp, match="no guess"):
        hfm.run(dict(p, factoringRadius=3.0))
