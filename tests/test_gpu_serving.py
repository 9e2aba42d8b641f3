"""GPU tests of the continuous-batching server: each result is the same alone and in the mix,
matches a plain plan solve and the oracle, admission by HBM budget, error isolation."""
import numpy as np
import pytest

import oracle
import hamiltonfastmarching_amd as hfm

pytestmark = pytest.mark.gpu


def _rel(u, ref):
    fin = np.isfinite(ref)
    pos = ref[fin & (ref > 0)]
    floor = pos.min()
    m = fin & (ref >= floor)
    assert np.array_equal(np.isfinite(u), fin)
    return float(np.max(np.abs(u[m] - ref[m]) / ref[m]))


@pytest.fixture(scope="module")
def reqs():
    from hamiltonfastmarching_amd import serving
    return serving.make_requests(12, seed=3, side2d=(48, 200), side3d=(24, 48))


def test_alone_vs_mix_vs_plan(reqs):
    from hamiltonfastmarching_amd import serving
    print("serving alone vs mix", flush=True)
    srv = serving.Server(hbm_budget_bytes=8e9, max_resident=64)
    mix = serving.run_load(srv, reqs, qps=0, collect_values=True)
    assert mix["completed"] == len(reqs)
    alone = []
    for r in reqs:
        out = np.empty(int(np.prod(r["dims"])), np.float64)
        t = srv.submit(r["model"], r["dims"], r["gridScale"], r["metric"], r["seeds"], values_out=out)
        res = srv.wait(t)
        assert res["status"] == 0, res["error"]
        alone.append(out.reshape(r["dims"]))
    seq = serving.run_sequential(reqs, collect_values=True)
    srv.close()
    worst_mix, worst_plan = 0.0, 0.0
    for a, m, s in zip(alone, mix["values"], seq["values"]):
        worst_mix = max(worst_mix, _rel(m, a))
        worst_plan = max(worst_plan, _rel(m, s))
    print(f"alone vs mix {worst_mix:.2e}, mix vs plan {worst_plan:.2e}", flush=True)
    assert worst_mix <= 1e-6
    assert worst_plan <= 1e-6


def test_served_values_match_oracle(reqs):
    from hamiltonfastmarching_amd import serving
    print("serving vs oracle", flush=True)
    small = [r for r in reqs if np.prod(r["dims"]) <= 60000][:3]
    srv = serving.Server(hbm_budget_bytes=4e9, max_resident=8)
    res = serving.run_load(srv, small, qps=0, collect_values=True)
    srv.close()
    for r, u in zip(small, res["values"]):
        p = {"model": r["model"], "dims": list(r["dims"]), "gridScale": r["gridScale"],
             "metric": r["metric"].cpu().numpy(), "seeds": r["seeds"]}
        ref = oracle.solve(p)["values"]
        err = _rel(u, ref)
        print(f"{r['model']} {r['dims']} kappa {r['kappa']:.1f}: rel err {err:.2e}", flush=True)
        assert err <= 1e-5


def test_admission_by_budget(reqs):
    """A budget that fits one request at a time still serves every request."""
    from hamiltonfastmarching_amd import serving
    print("serving admission", flush=True)
    sub = reqs[:6]
    # a solve reserves ~40 B/point (2-D) to ~65 B/point (3-D): this budget holds about one
    # of the largest requests, so the others queue until memory is returned
    biggest = max(int(np.prod(r["dims"])) * (65 if len(r["dims"]) == 3 else 41) for r in sub)
    srv = serving.Server(hbm_budget_bytes=1.2 * biggest, max_resident=64)
    out = serving.run_load(srv, sub, qps=0)
    st = srv.stats()
    srv.close()
    print(f"admission: max resident {st['max_resident_seen']} of {len(sub)}", flush=True)
    assert out["completed"] == len(sub)
    assert st["max_resident_seen"] < len(sub)
    srv1 = serving.Server(hbm_budget_bytes=8e9, max_resident=1)
    out = serving.run_load(srv1, sub, qps=0)
    st = srv1.stats()
    srv1.close()
    assert out["completed"] == len(sub) and st["max_resident_seen"] == 1
    srv2 = serving.Server(hbm_budget_bytes=1e5, max_resident=4)
    r = sub[0]
    with pytest.raises(hfm.PyExp, match="budget"):
        srv2.submit(r["model"], r["dims"], r["gridScale"], r["metric"], r["seeds"])
    srv2.close()


def test_bad_request_fails_alone(reqs):
    """A non positive-definite metric fails its own request; the others complete."""
    import torch
    from hamiltonfastmarching_amd import serving
    print("serving error isolation", flush=True)
    srv = serving.Server(hbm_budget_bytes=4e9, max_resident=16)
    good = reqs[0]
    bad_metric = good["metric"].clone()
    bad_metric[..., 0] = -1.0
    tb = srv.submit(good["model"], good["dims"], good["gridScale"], bad_metric, good["seeds"])
    tg = srv.submit(good["model"], good["dims"], good["gridScale"], good["metric"], good["seeds"])
    rb, rg = srv.wait(tb), srv.wait(tg)
    srv.close()
    assert rb["status"] != 0 and "positive definite" in rb["error"]
    assert rg["status"] == 0
    torch.cuda.synchronize()
