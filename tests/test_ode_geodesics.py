"""The ODE geodesic solver (geodesicSolver = "ODE": Headers/Base/Implementation/
GeodesicODESolver.hxx:17-217, selected at HFMInterface.hxx:459-464): midpoint steps along the
flow interpolated from the cell corners, stopped at the seeds, by the seeds' L-infinity
distance, or by a stall that the reference reports as a failed tip.  The oracle against the
reference's outputs (CPU), and the GPU against them, bit for bit in FP64 (marked gpu)."""
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

FIX = os.path.join(HERE, "golden", "ode_geodesics.npz")
CASES = mg.ode_geodesic_params()


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_reference(name):
    d = np.load(FIX)
    p = {k: v for k, v in CASES[name].items() if k != "tips"}
    pts, lens = oracle.geodesics(p, CASES[name]["tips"])
    assert np.array_equal(np.asarray(lens), d[name + "_lengths"].astype(np.int64))
    assert np.array_equal(np.asarray(pts), d[name + "_points"])
    assert oracle.last_geodesic_failures() == int(d[name + "_failed"][0])


def test_oracle_geodesics_descend():
    """Along an ODE geodesic the arrival time decreases overall (first point above the last),
    and the walk ends within a few cells of the seed."""
    p = {k: v for k, v in CASES["r2"].items() if k != "tips"}
    pts, lens = oracle.geodesics(p, CASES["r2"]["tips"])
    vals = oracle.solve(p)["values"]
    n = vals.shape[0]
    seed = np.asarray(p["seeds"], np.float64).reshape(-1, 2)[0]
    start = 0
    for ln in np.asarray(lens, np.int64):
        g = pts[start:start + ln]
        start += ln
        i0 = np.clip((g[0] * n).astype(int), 0, n - 1)
        i1 = np.clip((g[-1] * n).astype(int), 0, n - 1)
        assert vals[tuple(i1)] < vals[tuple(i0)]
        assert np.abs(g[-1] - seed).max() < 6.0 / n


def test_oracle_discrete_after_ode_resets():
    """The solver choice does not leak into the next call: the discrete solver's output is the
    same before and after an ODE run."""
    p = {k: v for k, v in CASES["r2"].items() if k not in ("tips", "geodesicSolver")}
    a = oracle.geodesics(p, CASES["r2"]["tips"])
    oracle.geodesics(dict(p, geodesicSolver="ODE"), CASES["r2"]["tips"])
    b = oracle.geodesics(p, CASES["r2"]["tips"])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_oracle_given_matches_oracle():
    """The ODE solver on given values / activeNeighs (the form the GPU tests use) is the ODE
    solver of a full oracle run when those are the oracle's own."""
    for name in ("r2", "elastica2"):
        c = CASES[name]
        p = {k: v for k, v in c.items() if k != "tips"}
        o = oracle.solve(p, active=True)
        a = oracle.geodesics_given(p, o["values"], o["activeNeighs"], c["tips"])
        b = oracle.geodesics(p, c["tips"])
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_matches_reference(name):
    """Bit-exact against the oracle's ODE solver on the GPU's own FP64 values and active sets;
    against the reference's fixture bit-exact where those values are the reference's (the
    serial loop at order 2, and the block solves whose fixed point lands on the same bits), else
    the same number of points per tip within 1e-9 (the block solver's FP64 values are within
    1e-12 of the reference's)."""
    import hamiltonfastmarching_amd as hfm
    print(f"ode geodesics {name}", flush=True)
    d = np.load(FIX)
    io = hfm.HFMIO(default_ordering="RowMajor")
    for k, v in CASES[name].items():
        io[k] = v
    io["exportValues"] = 1.0
    io["exportActiveNeighs"] = 1.0
    io.run()
    pts, lens = np.asarray(io["geodesicPoints"]), np.asarray(io["geodesicLengths"])
    p = {k: v for k, v in CASES[name].items() if k != "tips"}
    assert np.array_equal(lens, d[name + "_lengths"])
    if p.get("order", 1.0) > 1:  # the serial loop: the reference's values, flows and points
        assert np.array_equal(pts, d[name + "_points"])
        return
    ref_pts, ref_lens = oracle.geodesics_given(p, io["values"], io["activeNeighs"], CASES[name]["tips"])
    assert np.array_equal(lens.astype(np.int64), ref_lens) and np.array_equal(pts, ref_pts)
    same_values = np.array_equal(io["values"], oracle.solve(p)["values"])
    print(f"  values bit-equal to the reference's: {same_values}", flush=True)
    if same_values:
        assert np.array_equal(pts, d[name + "_points"])
    else:
        assert np.abs(pts - d[name + "_points"]).max() < 1e-9
    msgs = [ln for ln in io.messages().splitlines() if "seems to have failed" in ln]
    assert msgs == [m for m in d[name + "_messages"].tolist() if m]


@pytest.mark.gpu
def test_gpu_plan_api_matches_io():
    """hfm_plan_geodesics_ode on a Plan gives the same points, lengths and failed tips as the
    input-dictionary path (tips in physical units, the stall case)."""
    import torch

    import hamiltonfastmarching_amd as hfm
    print("ode geodesics plan api", flush=True)
    d = np.load(FIX)
    c = CASES["r2_stall"]
    plan = hfm.Plan("Riemann2", c["metric"].shape[:2], c["gridScale"], origin=c["origin"], precision=64)
    plan.set_metric(torch.from_numpy(c["metric"]).cuda())
    plan.solve(np.asarray(c["seeds"], np.float64).reshape(-1, 2))
    pts, lens, failed = plan.geodesics_ode(c["tips"], step=0.25, causality_tolerance=0.0)
    assert np.array_equal(lens, d["r2_stall_lengths"].astype(np.int64))
    assert np.array_equal(pts, d["r2_stall_points"])
    assert int(failed.sum()) == int(d["r2_stall_failed"][0])


@pytest.mark.gpu
def test_gpu_unrecognized_solver():
    import hamiltonfastmarching_amd as hfm
    print("ode geodesics unrecognized solver", flush=True)
    with pytest.raises(RuntimeError, match="Unrecognized geodesic solver Foo"):
        hfm.run(dict(CASES["r2"], geodesicSolver="Foo"))


@pytest.mark.gpu
def test_gpu_long_walk_spills_read_counts():
    """A 4000-cell walk along a 4096 x 8 strip reads more distinct grid points than the LDS
    read-count table holds, so part of the counts go to the device-memory table: still
    bit-exact against the oracle's ODE solver on the GPU's values."""
    import hamiltonfastmarching_amd as hfm
    print("ode geodesics long walk", flush=True)
    n0, n1 = 4096, 8
    x = np.linspace(0.0, 1.0, n0)[:, None]
    metric = np.zeros((n0, n1, 3))
    metric[..., 0] = 1.0 + 0.5 * np.sin(7.0 * x)
    metric[..., 2] = 1.0 + 0.3 * np.cos(5.0 * x)
    metric[..., 1] = 0.1 * np.sin(3.0 * x)
    p = {"model": "Riemann2", "dims": np.array([n0, n1], np.float64), "gridScale": 1.0,
         "origin": np.zeros(2), "arrayOrdering": "RowMajor", "seeds": np.array([[0.5, 4.5]]),
         "metric": metric, "geodesicSolver": "ODE"}
    tips = np.array([[4090.3, 2.2], [4000.7, 6.1]])
    out = hfm.run(dict(p, tips=tips, exportValues=1.0, exportActiveNeighs=1.0))
    pts, lens = np.asarray(out["geodesicPoints"]), np.asarray(out["geodesicLengths"]).astype(np.int64)
    assert lens.min() > 4 * 3900  # four points per cell along the strip
    ref_pts, ref_lens = oracle.geodesics_given(p, out["values"], out["activeNeighs"], tips)
    assert np.array_equal(lens, ref_lens) and np.array_equal(pts, ref_pts)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["r2", "r3", "rsf2", "elastica2"])
def test_gpu_fp32_solve_close_to_reference(name):
    """The ODE solver after an FP32 solve (the configuration the benchmark times): every tip
    within one point of the reference's count, and the tips of equal count within 1e-4 of its
    points (FP32 values are within about 1e-7 of the reference's)."""
    import hamiltonfastmarching_amd as hfm
    print(f"ode geodesics fp32 {name}", flush=True)
    d = np.load(FIX)
    out = hfm.run(dict(CASES[name], precision=32.0))
    lens = np.asarray(out["geodesicLengths"]).astype(np.int64)
    ref_lens = d[name + "_lengths"].astype(np.int64)
    assert np.abs(lens - ref_lens).max() <= 1
    pts, ref_pts = np.asarray(out["geodesicPoints"]), d[name + "_points"]
    a = b = 0
    worst = 0.0
    for la, lb in zip(lens, ref_lens):
        if la == lb:
            worst = max(worst, float(np.abs(pts[a:a + la] - ref_pts[b:b + lb]).max()))
        a += la
        b += lb
    print(f"  lengths equal {np.mean(lens == ref_lens):.2f}, max point diff {worst:.2e}", flush=True)
    assert worst < 1e-4
