"""CPU tests of the C ABI library and the host-side IO logic (no GPU calls)."""
import os
import re

import numpy as np
import pytest

import hamiltonfastmarching_amd as hfm
from hamiltonfastmarching_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "hfm_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hfm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/hfm_amd.h but not exported"
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_io_roundtrip():
    io = hfm.HFMIO(default_ordering="")
    io.set_string("model", "Riemann2")
    io.set_scalar("gridScale", 0.25)
    a = np.arange(24, dtype=float).reshape(2, 3, 4)
    io.set_array("metric", a)
    assert io.has_field("model") and not io.has_field("nope")
    assert io.get_string("model") == "Riemann2"
    assert io.get_scalar("gridScale") == 0.25
    assert np.array_equal(io.get_array("metric"), a)
    io.erase_field("metric")
    assert not io.has_field("metric")
    with pytest.raises(hfm.PyExp):
        io.get_string("gridScale")
    assert io.computed_keys() == []


@pytest.mark.parametrize("params,msg", [
    ({"dims": [4, 4], "gridScale": 1.0}, "model"),
    ({"model": "Seismic3", "dims": [4, 4], "gridScale": 1.0}, "Unrecognized model"),
    ({"model": "Riemann2", "dims": [4, 4], "gridScale": 1.0, "metric": np.ones((4, 4, 3)),
      "dualMetric": np.ones((4, 4, 3))}, "both primal and dual"),
    ({"model": "Riemann2", "dims": [4, 4], "gridScale": 1.0, "metric": np.ones((4, 5, 3))},
     "inconsistent dimensions"),
    ({"model": "Riemann2", "dims": [4, 4], "gridScale": 1.0, "metric": np.ones((4, 4, 2))},
     "does not match expected value 3"),
    ({"model": "Riemann2", "dims": [4, 4], "gridScale": 1.0, "metric": np.ones(3),
      "arrayOrdering": "Diagonal"}, "unrecognized array ordering"),
    ({"model": "Riemann2", "dims": [4, 4], "gridScale": 1.0, "metric": np.ones(3), "precision": 16.0},
     "precision"),
])
def test_run_rejects_bad_inputs_before_device_work(params, msg):
    """Input validation happens on the host, with the reference's messages, before any HIP call."""
    with pytest.raises(hfm.PyExp, match=msg):
        hfm.run(params)


def test_io_fp32_array_setter():
    """float32 arrays are stored widened to FP64 (the reference's only numerical type)."""
    io = hfm.HFMIO(default_ordering="")
    a = (np.arange(24, dtype=np.float32) / 7).reshape(2, 3, 4)
    io.set_array("metric", a)
    got = io.get_array("metric")
    assert got.dtype == np.float64 and got.shape == a.shape
    assert np.array_equal(got, a.astype(np.float64))


def test_last_error_empty_initially():
    assert isinstance(_lib.last_error(), str)
