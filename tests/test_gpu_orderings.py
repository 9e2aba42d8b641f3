"""Array orderings at the IO boundary (arrayOrdering: RowMajor, YXZ_RowMajor, ColumnMajor,
YXZ_ColumnMajor; JMM_CPPLibs/JMM_CPPLibs/Output/IO.hxx:68-110,153-170) against the reference's
own outputs on non-square grids (tests/golden/orderings.npz, written by the reference
executables): the field input (metric or cost) is given in each ordering, and every exported
array -- values, activeNeighs, geodesicFlow (components last), activeOffsets, costSensitivity --
comes back in that ordering with the reference's shape and entries.  The HFMIO object's default
ordering (YXZ_RowMajor, as the reference's Python interface, PythonIO.h:20) is checked on the
same fixture without setting the key."""
import os

import numpy as np
import pytest

import hamiltonfastmarching_amd as hfm
from pack_utils import ORDERINGS, ordering_params, to_ordering

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "orderings.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _check(name, ordering, out, gold):
    pre = f"{name}_{ordering}_"
    keys = [k[len(pre):] for k in gold.files if k.startswith(pre)]
    assert keys
    for k in keys:
        ref = gold[pre + k]
        got = np.asarray(out[k])
        assert got.shape == ref.shape, (k, got.shape, ref.shape)
        if k in ("activeNeighs", "activeOffsets", "geodesicLengths"):
            assert np.array_equal(got, ref), k
        elif k == "values":
            fin = np.isfinite(ref)
            assert np.array_equal(np.isfinite(got), fin)
            err = float(np.max(np.abs(got[fin] - ref[fin]) / np.maximum(np.abs(ref[fin]), 1e-300)))
            assert err <= 1e-12, (k, err)
        else:  # geodesicFlow, geodesicPoints, costSensitivity_0
            scale = max(float(np.abs(ref).max()), 1e-300)
            err = float(np.abs(got - ref).max()) / scale
            assert err <= 1e-10, (k, err)
    return keys


@pytest.mark.parametrize("ordering", ORDERINGS)
@pytest.mark.parametrize("name", ["riemann2", "riemann3", "rsf2"])
def test_ordering_vs_reference(name, ordering, gold):
    print(f"arrayOrdering {ordering} {name}", flush=True)
    base = ordering_params()[name]
    d = len(base["dims"])
    fkey = "metric" if "metric" in base else "cost"
    q = dict(base, arrayOrdering=ordering, precision=64.0)
    q[fkey] = to_ordering(base[fkey], ordering, d)
    out = hfm.run(q)
    keys = _check(name, ordering, out, gold)
    print(f"  {sorted(keys)} equal to the reference's", flush=True)


@pytest.mark.parametrize("name", ["riemann2", "riemann3", "rsf2"])
def test_hfmio_default_ordering(name, gold):
    """HFMIO() without an arrayOrdering key reads and writes YXZ_RowMajor."""
    print(f"HFMIO default ordering {name}", flush=True)
    base = ordering_params()[name]
    d = len(base["dims"])
    fkey = "metric" if "metric" in base else "cost"
    io = hfm.HFMIO()
    for k, v in base.items():
        io[k] = to_ordering(v, "YXZ_RowMajor", d) if k == fkey else v
    io["precision"] = 64.0
    io.run()
    out = {k: io[k] for k in io.computed_keys()}
    _check(name, "YXZ_RowMajor", out, gold)
