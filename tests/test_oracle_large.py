"""The CPU oracle on models beyond every reference fixture: the whole gtinv
table (orders 2..6, gtinv_maxl [20, 10, 2, 1, 1], l up to 20, 20 Gaussians,
two species), linear and quadratic.  No reference output covers these, so
the oracle is pinned by what the mathematics fixes:

* every structural feature of every atom is invariant under a rigid rotation
  of the whole periodic system (a wrong Clebsch-Gordan term, sign, m index
  or Y_lm at large l breaks it);
* the forces are minus the gradient of the energy (central differences);
* the forces sum to zero;
and Y_lm up to l = 20 is pinned against scipy (test_oracle.py)."""
import math

import numpy as np
import pytest

import oracle_lib as O
from common import config_potential
from mlip_hip import workloads as W

MODELS = ["binary_order6_l20_linear", "binary_order6_l20_quadratic", "binary_order3_l10"]


def rotation():
    a, b = 0.61, -1.17
    R = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    return R @ np.array([[1, 0, 0], [0, math.cos(b), -math.sin(b)], [0, math.sin(b), math.cos(b)]])


def system(seed=2):
    return W.fcc(2, 4.05, 0.15, seed, type_fraction=0.4, type_seed=seed + 1)


def run(o, s, feats=False):
    return o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1, want_features=feats)


@pytest.mark.parametrize("name", MODELS)
def test_features_rotation_invariant(name):
    o = O.Oracle(config_potential(name))
    s = system()
    r = run(o, s, feats=True)
    R = rotation()
    rr = run(o, W.Structure(s.cell @ R.T, s.positions @ R.T, s.types), feats=True)
    X, Xr = r["features"], rr["features"]
    scale = np.abs(X).max(axis=0) + 1e-300
    err = (np.abs(Xr - X) / scale).max()
    print("%s: %d features per atom, max rotation change %.2e (relative to each feature's scale)"
          % (name, X.shape[1], err))
    assert err < 1e-11
    assert rr["energy"] == pytest.approx(r["energy"], rel=1e-11)
    np.testing.assert_allclose(rr["forces"], r["forces"] @ R.T, atol=1e-9 * np.abs(r["forces"]).max())


@pytest.mark.parametrize("name", MODELS)
def test_forces_minus_gradient(name):
    o = O.Oracle(config_potential(name))
    s = system(seed=4)
    r = run(o, s)
    np.testing.assert_allclose(r["forces"].sum(0), 0.0, atol=1e-10 * np.abs(r["forces"]).max())
    h = 1e-4
    rng = np.random.default_rng(1)
    for i in rng.choice(s.natoms, 2, replace=False):
        for k in range(3):
            sp = W.Structure(s.cell, s.positions.copy(), s.types)
            sp.positions[i, k] += h
            sm = W.Structure(s.cell, s.positions.copy(), s.types)
            sm.positions[i, k] -= h
            fd = -(run(o, sp)["energy"] - run(o, sm)["energy"]) / (2 * h)
            assert r["forces"][i, k] == pytest.approx(fd, rel=1e-6, abs=1e-7 * np.abs(r["forces"]).max())
