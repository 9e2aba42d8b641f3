"""Order 2 / 3 (the recompute at acceptance, Headers/Base/Implementation/HamiltonFastMarching.hxx
:86-123,187-272): the oracle against the reference's outputs, and what the mathematics fixes."""
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

FIX = os.path.join(HERE, "golden", "high_order.npz")


def _params(name):
    p = mg.HIGH_ORDER_CASES[name]()
    return {k: v for k, v in p.items()
            if k not in ("tips", "inspectSensitivity", "inspectSensitivityWeights", "inspectSensitivityLengths")}


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("name", list(mg.HIGH_ORDER_CASES))
def test_oracle_matches_reference(name, order):
    d = np.load(FIX)
    o = oracle.solve(dict(_params(name), order=order), flow=True, active=True)
    assert np.array_equal(o["values"], d[f"{name}_order{order}_values"])
    assert np.array_equal(o["activeNeighs"].astype(np.float64), d[f"{name}_order{order}_activeNeighs"])
    assert np.array_equal(o["geodesicFlow"], d[f"{name}_order{order}_geodesicFlow"])
    assert int(np.isfinite(o["values"]).sum()) == int(d[f"{name}_order{order}_nAccepted"])


@pytest.mark.parametrize("order", [2, 3])
def test_axis_values_exact(order):
    """Identity metric, seed on a grid point: along an axis the second / third order
    combinations of exact values k h are exact, so every value on the axis is k h."""
    p = _params("riemann2_iso")
    n = int(p["dims"][0])
    h = p["gridScale"]
    u = oracle.solve(dict(p, order=order))["values"]
    c = n // 2
    k = np.arange(n) - c
    np.testing.assert_allclose(u[c, :], np.abs(k) * h, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(u[:, c], np.abs(k) * h, rtol=1e-14, atol=1e-16)


def test_higher_order_is_more_accurate():
    """Identity metric: the exact distance is Euclidean; orders 2 and 3 must be closer to it than
    order 1 (the first order scheme overestimates along every oblique direction)."""
    p = _params("riemann2_iso")
    n = int(p["dims"][0])
    h = p["gridScale"]
    c = n // 2
    i, j = np.meshgrid(np.arange(n) - c, np.arange(n) - c, indexing="ij")
    exact = np.hypot(i, j) * h
    far = exact > 4 * h
    err = {}
    for order in (1, 2, 3):
        u = oracle.solve(dict(p, order=order))["values"]
        err[order] = float(np.mean(np.abs(u[far] - exact[far])))
    assert err[2] < 0.5 * err[1], err
    assert err[3] < 0.5 * err[1], err
