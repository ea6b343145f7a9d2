"""Host-side logic of the parameter sweep (no GPU): point sampling, the
dimensionless features, and the rule that counts an unsteady run as failed."""
import math

import numpy as np

import surrogate
import sweep


def test_sample_points_are_seeded_and_in_range():
    a = sweep.sample_points(200, seed=5)
    b = sweep.sample_points(200, seed=5)
    assert a == b
    n = np.array([p["n"] for p in a])
    V = np.array([p["V"] for p in a])
    tau = np.array([p["tau"] for p in a])
    k = np.array([p["n_colloids"] for p in a])
    assert n.min() >= 32 and n.max() <= 160
    assert V.min() >= 0.0 and V.max() <= 1.0
    assert tau.min() >= 0.1 and tau.max() <= 10.0
    assert k.min() >= 0 and k.max() <= 12


def test_features_closed_form():
    pt = dict(n=64, V=0.25, tau=2.0, n_colloids=3, omega=1.0, F0=1e-5, seed=1)
    wi, mn, phi, V = sweep.features(pt)
    gdot = 2 * 0.25 / 1.0
    assert math.isclose(wi, 2.0 * gdot)
    assert math.isclose(mn, gdot * sweep.R_COLL * (2 * sweep.R_COLL) ** 4 / 1e-5)
    assert math.isclose(phi, 3 * math.pi * sweep.R_COLL ** 2)
    assert V == 0.25
    X = surrogate.feature_matrix([pt])
    np.testing.assert_allclose(X[0], [math.log10(wi), math.log10(mn), phi, V])


def test_unsteady_run_rule():
    rng = np.random.default_rng(0)
    eta, dV, L1 = 1.0, 0.8, 1.0            # Newtonian wall stress 0.8
    steady = 0.8 * (1.0 + 0.03 * rng.standard_normal(20))
    assert not sweep.unsteady(steady, eta, dV, L1)
    # a run whose stress swings to either sign by several times its scale
    swinging = 0.8 * np.array([1.4, 0.9, -2.0, 3.5, -6.0, 1.1, -3.7, 0.2] * 2)
    assert sweep.unsteady(swinging, eta, dV, L1)
    # the threshold scales with the wall speed and the viscosity
    assert not sweep.unsteady(steady * 0.01, eta, dV * 0.01, L1)
    assert sweep.unsteady(steady, eta, dV, L1, factor=0.01)
    assert not sweep.unsteady(swinging, 10.0 * eta, dV, L1)
