"""The block fast-iterative solver's schedule changes the work, never the fixed point: solves with
unordered rounds (band width 0), narrow and wide value bands, tiny and huge round targets, a
one-block and an odd-sized update grid, several sweep caps, and rounds that list their own tiles
or are listed by a separate kernel all agree with the oracle's FP64
Fast Marching (1e-12 relative for FP64 plans, 1e-5 for FP32) and with each other."""
import numpy as np
import pytest

import oracle
import hamiltonfastmarching_amd as hfm
from hamiltonfastmarching_amd import synthetic as syn

pytestmark = pytest.mark.gpu

# (band width in mean one-step times, tiles per round, update-kernel blocks, sweeps per visit,
#  rounds listing their own tiles)
SETTINGS = [
    (8.0, 0, 0, 8, None),       # defaults (rounds list their own tiles on 2-D grids)
    (8.0, 0, 0, 8, True),       # every round lists its own tiles
    (8.0, 0, 0, 8, False),      # a separate listing kernel per round
    (0.0, 0, 0, 8, True),       # unordered: every pending tile in every round
    (0.0, 0, 0, 8, False),
    (1.0, 16, 0, 8, True),      # narrow band, small rounds
    (1.0, 16, 0, 8, False),
    (64.0, 4096, 0, 32, True),  # wide band, large rounds, long visits
    (8.0, 0, 1, 16, True),      # one block of four waves (owning every tile)
    (8.0, 0, 1, 16, False),
    (8.0, 0, 7, 8, True),       # an odd grid
    (8.0, 0, 0, 4, True),       # at most 4 sweeps per visit
    (8.0, 0, 0, 4, False),
]
CASES = [
    ("riemann2_96", lambda: syn.riemann2(96, seed=21)),
    ("riemann3_32", lambda: syn.riemann3(32, seed=22)),
    ("rsf2_32x32x16", lambda: syn.rsf2(32, 16, seed=23)),
]
TOL = {64: 1e-12, 32: 1e-5}


def _floor_rel(u, ref):
    fin = np.isfinite(ref)
    pos = ref[fin & (ref > 0)]
    m = fin & (ref >= pos.min())
    return float(np.max(np.abs(u[m] - ref[m]) / ref[m]))


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("precision", [64, 32])
def test_fixed_point_independent_of_schedule(name, make, precision):
    import torch
    print(f"schedule settings {name} precision {precision}", flush=True)
    p = {k: v for k, v in make().items() if k not in ("tips",)}
    ref = oracle.solve(p)["values"]
    dims = tuple(int(x) for x in p["dims"])
    h = float(np.atleast_1d(p["gridScale"])[0])
    if p["model"].startswith("Riemann"):
        plan = hfm.Plan(p["model"], dims, h, precision=precision)
        plan.set_metric(torch.from_numpy(p["metric"]).cuda())
    else:
        plan = hfm.Plan(p["model"], dims, h, origin=p["origin"], precision=precision, xi=float(p["xi"]),
                        eps=float(p["eps"]))
        plan.set_cost(torch.from_numpy(p["cost"]).cuda())
    outs = []
    for delta, target, grid, sweeps, fused in SETTINGS:
        plan.set_delta(delta).set_round_target(target).set_grid(grid).set_max_sweeps(sweeps)
        plan.set_fused_schedule(fused)
        plan.solve(p["seeds"])
        u = plan.values().cpu().numpy()
        st = plan.stats()
        assert np.array_equal(np.isfinite(u), np.isfinite(ref))
        err = _floor_rel(u, ref)
        print(f"  delta {delta:g} target {target} grid {grid} sweeps {sweeps} fused {fused}: {st['iterations']} rounds, "
              f"{st['tile_visits']} visits, max rel err {err:.2e}", flush=True)
        assert err <= TOL[precision]
        outs.append(u)
    spread = max(_floor_rel(u, outs[0]) for u in outs[1:])
    print(f"  largest difference between settings {spread:.2e}", flush=True)
    assert spread <= TOL[precision]


@pytest.mark.parametrize("precision", [64, 32])
def test_self_listing_rounds_many_chunks(precision):
    """Rounds that list their own tiles, with several blocks each owning several 256-tile chunks
    (Riemann2 256^2 = 1024 tiles on 2 to 3 blocks): a block visits chunk after chunk while the other
    blocks, possibly on other XCDs, mark tiles it owns.  Values agree with the oracle and pass
    the FP64 residual check."""
    import torch
    p = syn.riemann2(256, seed=24)
    ref = oracle.solve(p)["values"]
    plan = hfm.Plan("Riemann2", (256, 256), float(p["gridScale"]), precision=precision)
    plan.set_metric(torch.from_numpy(p["metric"]).cuda())
    for grid in (2, 3, 0):
        print(f"self-listing rounds Riemann2 256^2 precision {precision} grid {grid}", flush=True)
        plan.set_fused_schedule(True).set_grid(grid)
        plan.solve(p["seeds"])
        u = plan.values().cpu().numpy()
        assert np.array_equal(np.isfinite(u), np.isfinite(ref))
        err = _floor_rel(u, ref)
        dev, bad = plan.residual()
        st = plan.stats()
        print(f"  {st['iterations']} rounds, {st['tile_visits']} visits, max rel err {err:.2e}, "
              f"residual {dev:.2e}", flush=True)
        assert err <= TOL[precision]
        assert bad == 0 and dev <= (1e-12 if precision == 64 else 2e-6)


def test_sweep_cap_values():
    """The sweep cap takes 4, 8, 16, 32 or 0 (the model's default); anything else is an error,
    not a silent rounding."""
    p = syn.riemann2(16, seed=25)
    plan = hfm.Plan("Riemann2", (16, 16), float(p["gridScale"]), precision=32)
    for ok in (0, 4, 8, 16, 32):
        plan.set_max_sweeps(ok)
    for bad in (1, 5, 12, 64, -1):
        with pytest.raises(hfm.PyExp, match="sweeps must be 4, 8, 16 or 32"):
            plan.set_max_sweeps(bad)
