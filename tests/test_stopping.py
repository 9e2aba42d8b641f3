"""Stopping criteria (Headers/ExtraAlgorithms/CommonStoppingCriteria.h
:96-165, checked in the acceptance loop of HamiltonFastMarching.hxx:187-205): the oracle against
the reference's outputs, and the invariants the criteria fix."""
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

FIX = os.path.join(HERE, "golden", "stopping.npz")
CASES = mg.stopping_params()


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_reference(name):
    d = np.load(FIX)
    o = oracle.solve(CASES[name], active=True)
    assert np.array_equal(o["values"], d[f"{name}_values"])
    assert np.array_equal(o["activeNeighs"].astype(np.float64), d[f"{name}_activeNeighs"])
    assert o["nAccepted"] == int(d[f"{name}_nAccepted"])


def _base(name):
    return {k: v for k, v in CASES[name].items()
            if k not in ("stopAtDistance", "nMaxAccepted", "stopWhenAnyAccepted", "stopWhenAllAccepted",
                         "stopWhenAnyAccepted_Unoriented", "stopWhenAllAccepted_Unoriented")}


@pytest.mark.parametrize("name", ["riemann2_dist_o1", "elastica2_dist_o1"])
def test_distance_stop_keeps_final_values(name):
    """First order: every point whose value does not exceed the distance was accepted before the
    stop, and an accepted value is final, so those values equal the full solve's and no other
    point's does.  (At order 2 / 3 the recompute at acceptance can lower a value below the key it
    was popped with, so a point accepted after the stop may end below the distance.)"""
    dist = CASES[name]["stopAtDistance"]
    full = oracle.solve(_base(name))["values"]
    part = oracle.solve(CASES[name])["values"]
    inside = full <= dist
    assert inside.any() and not inside.all()
    assert np.array_equal(part[inside], full[inside])
    assert np.all(part[~inside] > dist)


def test_max_accepted_counts_exactly():
    o = oracle.solve(CASES["riemann2_max_o1"])
    assert o["nAccepted"] == 300
    # the 299 accepted points and the terminating one hold their final values
    full = oracle.solve(_base("riemann2_max_o1"))["values"]
    order = np.sort(full.ravel())
    assert np.array_equal(np.sort(o["values"].ravel())[:300], order[:300])


def test_unset_criteria_reset():
    """A solve without stopping keys after one with them runs to completion."""
    oracle.solve(CASES["riemann2_max_o1"])
    o = oracle.solve(_base("riemann2_max_o1"))
    assert o["nAccepted"] == 40 * 40 and np.isfinite(o["values"]).all()
