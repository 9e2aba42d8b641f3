"""Seed set-up variants against the fixtures the reference executables wrote
(tests/golden/seeds.npz, cases built by tests/golden/make_golden.py):

* seedRadius spreading (StaticFactoring.hxx:284-345): spreadedSeeds / spreadedSeedValues exactly,
  values 1e-12 relative, valueVariation of the given seeds 1e-11 (primal and dual metrics,
  2-D and 3-D; a negative radius also seeds one step of the reversed stencil from every point of
  the box);
* seeds_Unoriented / seedValues_Unoriented (StaticFactoring.hxx:351-380) and tips_Unoriented
  (HFMInterface.hxx:498-522) of the curvature models: values 1e-12, geodesicLengths_Unoriented
  equal, geodesicPoints_Unoriented within 1e-6 of a grid step;
* nAccepted (ExtraAlgorithms/CommonStoppingCriteria.h:162-165) equal.
FP64 solves."""
import os
import sys

import numpy as np
import pytest

import hamiltonfastmarching_amd as hfm

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import make_golden as mg  # noqa: E402


def _floor_rel(u, ref):
    fin = np.isfinite(ref)
    pos = ref[fin & (ref > 0)]
    floor = pos.min() if pos.size else 0.0
    m = fin & (ref >= floor) & (ref > 0)
    return float(np.max(np.abs(u[m] - ref[m]) / ref[m])) if m.any() else 0.0


@pytest.mark.parametrize("name", list(mg.SEED_CASES))
def test_seed_variants_vs_golden(name):
    print(f"seeds {name}", flush=True)
    g = np.load(os.path.join(GOLD, "seeds.npz"))
    p = dict(mg.seed_case_params(name), exportValues=1.0)
    out = hfm.run(p)
    u, ref = out["values"], g[name + "_values"]
    assert np.array_equal(np.isfinite(u), np.isfinite(ref))
    err = _floor_rel(u, ref)
    print(f"{name}: values rel err {err:.3e}", flush=True)
    assert err <= 1e-12
    assert out["nAccepted"] == float(g[name + "_nAccepted"])
    if name + "_spreadedSeeds" in g.files:
        assert np.array_equal(out["spreadedSeeds"], g[name + "_spreadedSeeds"])
        assert np.array_equal(out["spreadedSeedValues"], g[name + "_spreadedSeedValues"])
    if name + "_valueVariation" in g.files:
        vv, vref = out["valueVariation"], g[name + "_valueVariation"]
        verr = float(np.abs(vv - vref).max() / np.abs(vref).max())
        print(f"{name}: valueVariation rel err {verr:.3e}", flush=True)
        assert verr <= 1e-11
    if name + "_geodesicPoints_Unoriented" in g.files:
        assert np.array_equal(out["geodesicLengths_Unoriented"], g[name + "_geodesicLengths_Unoriented"])
        h = float(np.atleast_1d(p["gridScale"])[0])
        perr = float(np.abs(out["geodesicPoints_Unoriented"] - g[name + "_geodesicPoints_Unoriented"]).max()) / h
        print(f"{name}: unoriented geodesic points err {perr:.3e} h", flush=True)
        assert perr <= 1e-6


def test_seed_radius_on_curvature_model_is_the_reference_error():
    """The curvature models have no distance guess (HamiltonFastMarching.h:203-204)."""
    p = mg.seed_case_params("rsf2_unoriented")
    with pytest.raises(hfm.PyExp, match="no guess"):
        hfm.run(dict(p, seeds=np.array([[0.5, 0.5, 1.0]]), seedRadius=2.0))


@pytest.mark.parametrize("name", list(mg.BYPASS_CASES))
def test_bypass_with_cached_values_vs_golden(name):
    """values + activeNeighs given: the solver is bypassed (HFMInterface.hxx:423-426) and geodesics,
    geodesicFlow and costSensitivity come from the cached data, as the reference's do."""
    print(f"bypass {name}", flush=True)
    g = np.load(os.path.join(GOLD, "bypass.npz"))
    make, tips = mg.BYPASS_CASES[name]
    p = {k: v for k, v in make().items()
         if k not in ("tips", "inspectSensitivity", "inspectSensitivityWeights", "inspectSensitivityLengths")}
    q = dict(p, values=g[name + "_values_in"], activeNeighs=g[name + "_activeNeighs_in"], tips=np.array(tips),
             exportGeodesicFlow=1.0, inspectSensitivity=np.array(tips[:1]), verbosity=1.0)
    io = hfm.HFMIO(default_ordering="RowMajor")
    for k, v in q.items():
        io[k] = v
    io.run()
    out = {k: io[k] for k in io.computed_keys()}
    assert out["nAccepted"] == 0.0
    assert "FMCPUTime" not in out
    assert np.array_equal(out["geodesicLengths"], g[name + "_geodesicLengths"])
    h = float(np.atleast_1d(p["gridScale"])[0])
    perr = float(np.abs(out["geodesicPoints"] - g[name + "_geodesicPoints"]).max()) / h
    fref = g[name + "_geodesicFlow"]
    ferr = float(np.abs(out["geodesicFlow"] - fref).max() / np.abs(fref).max())
    sref = g[name + "_costSensitivity_0"]
    serr = float(np.abs(out["costSensitivity_0"] - sref).max() / np.abs(sref).max())
    print(f"{name}: geodesic points {perr:.3e} h, flow {ferr:.3e}, costSensitivity {serr:.3e}", flush=True)
    assert perr <= 1e-6
    assert ferr <= 1e-10
    assert serr <= 1e-9


@pytest.mark.parametrize("name", list(mg.STENCIL_CASES))
def test_get_stencils_string_equals_reference(name):
    """getStencils -> the 'stencils' string, character for character (HFMInterface.hxx:360-375)."""
    import json
    with open(os.path.join(GOLD, "stencils.json")) as f:
        ref = json.load(f)[name]
    out = hfm.run(mg.stencil_params(name))
    assert out["stencils"] == ref


@pytest.mark.parametrize("name", ["riemann2", "riemann3", "rsf2"])
def test_point_index_conversions_equal_reference(name):
    """pointToIndex -> indexFromPoint and indexToPoint -> pointFromIndex (HFMInterface.hxx:377-396)."""
    import json
    with open(os.path.join(GOLD, "stencils.json")) as f:
        ref = json.load(f)
    p = mg.stencil_params(name)
    dim = len(p["dims"])
    pts = p.pop("getStencils")
    out = hfm.run(dict(p, pointToIndex=pts, indexToPoint=np.array([[1, 2, 3][:dim], [7, 0, 5][:dim]], float)))
    assert np.array_equal(out["indexFromPoint"], np.array(ref[name + "_indexFromPoint"]))
    assert np.array_equal(out["pointFromIndex"], np.array(ref[name + "_pointFromIndex"]))
