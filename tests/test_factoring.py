"""Static factoring (Headers/ExtraAlgorithms/Implementation/StaticFactoring.hxx:11-393, used by the
recompute of HamiltonFastMarching.hxx:86-123,187-272): the oracle against the reference's outputs,
and what the mathematics fixes -- with a constant metric the factor is the exact distance, which
the spread seeds take exactly and the corrected differences approach far better than plain ones."""
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

FIX = os.path.join(HERE, "golden", "factoring.npz")
CASES = mg.factoring_params()


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_reference(name):
    d = np.load(FIX)
    o = oracle.solve(dict(CASES[name], exportFactoring=1.0), flow=True, active=True)
    assert np.array_equal(o["values"], d[f"{name}_values"])
    assert np.array_equal(o["activeNeighs"].astype(np.float64), d[f"{name}_activeNeighs"])
    assert np.array_equal(o["geodesicFlow"], d[f"{name}_geodesicFlow"])
    assert o["nAccepted"] == int(d[f"{name}_nAccepted"])
    for k in ("factoringValues", "factoringGradients", "factoringIndexShift", "spreadedSeeds", "spreadedSeedValues"):
        if f"{name}_{k}" in d.files:
            assert np.array_equal(np.asarray(o[k], np.float64), d[f"{name}_{k}"], equal_nan=True), k


def _constant(n=25, dim=2):
    h = 1.0 / n
    m = np.array([2.0, 0.7, 1.0]) if dim == 2 else np.array([2.0, 0.3, 1.5, -0.2, 0.4, 1.2])
    return {"model": f"Riemann{dim}", "dims": np.array([n] * dim, float), "gridScale": h,
            "origin": np.zeros(dim), "seeds": np.full((1, dim), 0.43), "metric": m, "arrayOrdering": "RowMajor"}


def _exact(p):
    n = int(p["dims"][0])
    dim = len(p["dims"])
    h = p["gridScale"]
    m = p["metric"]
    M = np.array([[m[0], m[1]], [m[1], m[2]]]) if dim == 2 else \
        np.array([[m[0], m[1], m[3]], [m[1], m[2], m[4]], [m[3], m[4], m[5]]])
    x = (np.arange(n) + 0.5) * h
    X = np.stack(np.meshgrid(*([x] * dim), indexing="ij"), axis=-1) - p["seeds"][0]
    return np.sqrt(np.einsum("...i,ij,...j->...", X, M, X))


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dim", [2, 3])
def test_constant_metric_factored_is_more_accurate(dim, order):
    """Constant metric, factor |x - x0|_M over the whole grid: exact on the spread seeds, and the
    mean error over the grid at least 4x below the unfactored scheme's (first order: 20x).  The
    first pass stays unfactored, so points accepted before an upwind neighbour still carry an
    error: the factored scheme is not exact everywhere."""
    n = 25 if dim == 2 else 11
    p = _constant(n, dim)
    exact = _exact(p)
    o = oracle.solve(dict(p, order=float(order), factoringRadius=-1.0, factoringPointChoice="Seed"))
    u = o["values"]
    u0 = oracle.solve(dict(p, order=float(order), seedRadius=2.0))["values"]
    h = p["gridScale"]
    idx = np.floor(np.asarray(o["spreadedSeeds"]) / h).astype(int)  # cell centres (k + 1/2) h
    assert np.max(np.abs(u[tuple(idx.T)] - exact[tuple(idx.T)])) < 1e-14
    gain = np.mean(np.abs(u0 - exact)) / np.mean(np.abs(u - exact))
    assert gain > (20 if order == 1 else 4), gain


def test_factor_on_a_box_helps_near_the_seed():
    p = _constant(25, 2)
    exact = _exact(p)
    u = oracle.solve(dict(p, factoringRadius=6.0))["values"]
    u0 = oracle.solve(dict(p, seedRadius=2.0))["values"]
    c = 0.43 * 25
    i, j = np.meshgrid(np.arange(25) + 0.5, np.arange(25) + 0.5, indexing="ij")
    near = np.maximum(np.abs(i - c), np.abs(j - c)) < 5
    assert np.mean(np.abs(u - exact)[near]) * 10 < np.mean(np.abs(u0 - exact)[near])


def test_errors():
    p = CASES["r2_rad10_o1"]
    with pytest.raises(RuntimeError, match="Exactly one seed"):
        oracle.solve(dict(p, seeds=np.array([[0.3, 0.3], [0.6, 0.6]])))
    with pytest.raises(RuntimeError, match="both specified"):
        oracle.solve(dict(CASES["r2_import_o1"], factoringRadius=3.0))
