"""Final FP64 residual check (hfm_plan_residual): at every reached non-seed point the stored
value is the root of the local scheme sum_k w_k max(0, u - v_k)^2 = 1 over the smaller neighbour
value of each difference (EulerianStencil.hxx:30-72,104-153), recomputed in FP64.  Converged
solves of the four models in both precisions, the 128^3 benchmark field, and planted defects in
loaded values (a lowered value, a value made infinite) that the check must report."""
import numpy as np
import pytest

import hamiltonfastmarching_amd as hfm
from hamiltonfastmarching_amd import synthetic as syn

pytestmark = pytest.mark.gpu

SENS_KEYS = ("tips", "inspectSensitivity", "inspectSensitivityWeights", "inspectSensitivityLengths")
CASES = [
    ("riemann2_64", lambda: syn.riemann2(64, seed=1)),
    ("riemann3_24", lambda: syn.riemann3(24, seed=2)),
    ("rsf2_32x32x16", lambda: syn.rsf2(32, 16, seed=3)),
    ("elastica2_24x24x16", lambda: syn.elastica2(24, 16, seed=4)),
]
# FP64: the GPU's root in the reference's operation order against the same computation: equal
# up to the order-dependent last bits.  FP32: the FP32 local solve's rounding (hardware square
# root and reciprocal, one ulp each), measured up to 1.1e-6; bounded by the FP32 accuracy target.
TOL = {64: 1e-13, 32: 1e-5}


def _plan(p, precision):
    import torch
    dims = tuple(int(x) for x in p["dims"])
    h = float(np.atleast_1d(p["gridScale"])[0])
    if p["model"] in ("Riemann2", "Riemann3"):
        plan = hfm.Plan(p["model"], dims, h, precision=precision)
        plan.set_metric(torch.from_numpy(np.asarray(p["metric"], np.float64)).cuda())
    else:
        plan = hfm.Plan(p["model"], dims, h, origin=p.get("origin"), precision=precision, xi=float(p["xi"]),
                        eps=float(p.get("eps", 0.1)))
        plan.set_cost(torch.from_numpy(np.asarray(p["cost"], np.float64)).cuda())
    return plan


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("precision", [64, 32])
def test_converged_solves_have_small_residual(name, make, precision):
    print(f"residual {name} precision {precision}", flush=True)
    p = {k: v for k, v in make().items() if k not in SENS_KEYS}
    plan = _plan(p, precision)
    plan.solve(p["seeds"])
    dev, bad = plan.residual()
    print(f"  max relative deviation {dev:.2e}, mismatches {bad}", flush=True)
    assert bad == 0
    assert dev <= TOL[precision]


def test_riemann3_128_fp32_residual():
    import torch
    print("residual Riemann3 128^3 FP32", flush=True)
    p = syn.riemann3(128, seed=0)
    plan = hfm.Plan("Riemann3", (128, 128, 128), p["gridScale"], precision=32)
    plan.set_metric(torch.from_numpy(p["metric"]).cuda())
    plan.solve(p["seeds"])
    dev, bad = plan.residual()
    print(f"  max relative deviation {dev:.2e}, mismatches {bad}", flush=True)
    assert bad == 0 and dev <= TOL[32]


def test_planted_defects_are_reported():
    print("residual: planted defects in loaded values", flush=True)
    p = syn.riemann2(64, seed=1)
    out = hfm.run(dict(p, exportValues=1.0, exportActiveNeighs=1.0, precision=64.0))
    u, act = out["values"], out["activeNeighs"]
    plan = _plan(p, 64)
    plan.load_values(u, act)
    dev0, bad0 = plan.residual()
    assert bad0 == 0 and dev0 <= TOL[64]
    # a reached non-seed point in the interior, lowered by 1e-3 of its value
    i, j = 40, 22
    assert act[i, j] != 0 and np.isfinite(u[i, j])
    low = u.copy()
    low[i, j] *= 1 - 1e-3
    plan.load_values(low, act)
    dev1, bad1 = plan.residual()
    print(f"  lowered value: deviation {dev1:.2e}", flush=True)
    assert bad1 == 0 and dev1 >= 1e-4
    # the same point made unreachable while its neighbours are reached
    inf = u.copy()
    inf[i, j] = np.inf
    plan.load_values(inf, act)
    dev2, bad2 = plan.residual()
    print(f"  infinite value: mismatches {bad2}", flush=True)
    assert bad2 >= 1
