"""The captured-graph update loop against individual launches: several graph replays per solve,
with eager solves, device work of the caller and synchronisations in between, on two metrics
(the second changes the initial band width, which the graph reads from device memory)."""
import numpy as np
import pytest

import oracle
import hamiltonfastmarching_amd as hfm
from hamiltonfastmarching_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _rel(u, ref):
    fin = np.isfinite(ref)
    m = fin & (ref > 0)
    assert np.array_equal(np.isfinite(u), fin)
    return float(np.max(np.abs(u[m] - ref[m]) / ref[m]))


@pytest.mark.parametrize("precision", [32, 64])
def test_graph_replays_match_eager(precision):
    import torch
    n = 40
    metrics = [syn.riemann3(n, seed=5), syn.riemann3(n, seed=6, amax=8.0)]
    plan = hfm.Plan("Riemann3", (n,) * 3, metrics[0]["gridScale"], precision=precision)
    for rep in range(3):
        for p in metrics:
            print(f"graph replay rep {rep} precision {precision}", flush=True)
            m = torch.from_numpy(p["metric"]).cuda()
            plan.set_metric(m)
            plan.set_graph(False)
            plan.solve(p["seeds"])
            eager = plan.values().cpu().numpy()
            eager_rounds = plan.stats()["iterations"]
            # caller's own device work between the runs
            junk = torch.randn(1 << 20, device="cuda").sum()
            torch.cuda.synchronize()
            plan.set_graph(True)
            for _ in range(2):
                plan.solve(p["seeds"])
                g = plan.values().cpu().numpy()
                assert plan.stats()["iterations"] % 8 == 0
                err = _rel(g, eager)
                print(f"  rounds eager {eager_rounds} graph {plan.stats()['iterations']}, "
                      f"rel diff {err:.2e}, exact {np.mean(g == eager):.4f}", flush=True)
                assert err <= (1e-12 if precision == 64 else 2e-6)
            del junk
    ref = oracle.solve(metrics[1])["values"]
    assert _rel(g, ref) <= (1e-12 if precision == 64 else 1e-5)
