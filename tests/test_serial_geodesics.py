"""Geodesic backtracking after a serial solve (order 2/3, static factoring, a stopping criterion):
the reference's discrete solver takes each point's flow from Recompute
(Headers/Base/Implementation/GeodesicDiscreteSolver.hxx:131-132, HamiltonFastMarching.hxx:187-272),
so the order-2/3 combinations, the factor's corrections and the loop's accepted set all enter
the transported mass.  The oracle against the reference's outputs (CPU), and the GPU against
them (marked gpu)."""
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

FIX = os.path.join(HERE, "golden", "serial_geodesics.npz")
CASES = mg.serial_geodesic_params()


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_reference(name):
    d = np.load(FIX)
    p = {k: v for k, v in CASES[name].items() if k != "tips"}
    assert np.array_equal(oracle.solve(p)["values"], d[name + "_values"])
    pts, lens = oracle.geodesics(p, CASES[name]["tips"])
    assert np.array_equal(np.asarray(lens), d[name + "_lengths"].astype(np.int64))
    assert np.array_equal(np.asarray(pts), d[name + "_points"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_matches_reference(name):
    import hamiltonfastmarching_amd as hfm
    print(f"serial geodesics {name}", flush=True)
    d = np.load(FIX)
    out = hfm.run(dict(CASES[name]))
    assert np.array_equal(out["values"], d[name + "_values"])
    assert np.array_equal(np.asarray(out["geodesicLengths"]), d[name + "_lengths"])
    assert np.array_equal(np.asarray(out["geodesicPoints"]), d[name + "_points"])
