"""Differentiation after a serial solve (order 2/3, static factoring, stopping criteria):
BackwardVariation and ValueVariation read the stored activeNeighs and values with first-order
weights (Headers/ExtraAlgorithms/FirstVariation.h:198-251,341-385,389-429), whatever the order.
The oracle against the reference's outputs (CPU, exact), and the GPU against them (marked gpu;
the reverse sweep accumulates in a different order: 1e-9 relative, and 1e-11 of the largest
magnitude for valueVariation)."""
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

FIX = os.path.join(HERE, "golden", "serial_variations.npz")
CASES = mg.serial_variation_params()
SENS = [n for n in CASES if "inspectSensitivity" in CASES[n]]
FWD = [n for n in CASES if "costVariation" in CASES[n]]
STRIP = ("inspectSensitivity", "inspectSensitivityWeights", "inspectSensitivityLengths", "costVariation",
         "seedValueVariation")


def _base(name):
    return {k: v for k, v in CASES[name].items() if k not in STRIP}


@pytest.mark.parametrize("name", SENS)
def test_oracle_sensitivity_matches_reference(name):
    d = np.load(FIX)
    p = CASES[name]
    v, s, (si, sw) = oracle.solve_backward(_base(name), p["inspectSensitivity"], p["inspectSensitivityWeights"])
    assert np.array_equal(v, d[name + "_values"])
    assert np.array_equal(s, d[name + "_costSensitivity_0"])
    assert np.array_equal(sw, d[name + "_seedSensitivity_0"][:, -1])


@pytest.mark.parametrize("name", FWD)
def test_oracle_value_variation_matches_reference(name):
    d = np.load(FIX)
    p = CASES[name]
    v, vv = oracle.forward_variation(_base(name), p["costVariation"], p["seedValueVariation"])
    assert np.array_equal(v, d[name + "_values"])
    assert np.array_equal(vv, d[name + "_valueVariation"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", SENS + FWD)
def test_gpu_matches_reference(name):
    import hamiltonfastmarching_amd as hfm
    print(f"serial variation {name}", flush=True)
    d = np.load(FIX)
    out = hfm.run(dict(CASES[name]))
    assert np.array_equal(out["values"], d[name + "_values"])
    if name in SENS:
        ref = d[name + "_costSensitivity_0"]
        assert np.max(np.abs(out["costSensitivity_0"] - ref)) <= 1e-9 * np.abs(ref).max()
        rs = d[name + "_seedSensitivity_0"]
        ss = out["seedSensitivity_0"]
        assert ss.shape == rs.shape
        assert np.array_equal(ss[:, :-1], rs[:, :-1])
        np.testing.assert_allclose(ss[:, -1], rs[:, -1], rtol=1e-9)
    else:
        ref = d[name + "_valueVariation"]
        assert np.max(np.abs(out["valueVariation"] - ref)) <= 1e-11 * np.abs(ref).max()
