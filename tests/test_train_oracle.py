"""CPU oracle of the training outputs (MLIPModel::prepare_features,
src/USER-MLIP/mlipkk.cpp:461-470, 992-1281), pinned without golden vectors
(the reference ships none for this path) by what the mathematics fixes:

* identities against the evaluation path, itself pinned by the regression
  logs and dumps: sum_p w_p X_p = E, sum_p w_p FF[k][p] = F_k and
  sum_p w_p SF[p] = virial (components xx, yy, zz, yz, zx, xy);
* every force feature is minus the gradient of its energy feature
  X_p = sum_i P[i][p] (central differences, linear, quadratic and cubic
  coefficients);
* the stress features are minus the strain derivatives of X_p;
* every structural feature of every atom is rotation invariant (a dropped
  or mis-signed CG term breaks this).
A plausible slip -- a wrong sign on djri, a transposed stress component, a
missing parity factor -- fails one of these."""
import numpy as np
import pytest

import oracle_lib as O
from common import ref
from mlip_hip import workloads as W

POTS = ["gtinv-197.mlip", "gtinv-434.mlip", "gtinv-820.mlip"]


def perturbed_structure(seed=3, scale=0.05):
    s = W.read_lammps_data(ref("structure.lammps"))
    rng = np.random.default_rng(seed)
    return W.Structure(s.cell.copy(), s.positions + rng.normal(scale=scale, size=s.positions.shape), s.types)


def train(o, s, want=("X", "P", "FF", "SF")):
    return o.train(W.build_neighbor_list(s, o.cutoff), s.types - 1, want=want)


def vmap(v):
    # oracle virial (xx, yy, zz, xy, xz, yz) -> (xx, yy, zz, yz, zx, xy)
    return np.array([v[0], v[1], v[2], v[5], v[4], v[3]])


@pytest.mark.parametrize("name", POTS)
def test_weighted_sums_reproduce_energy_forces_virial(name):
    o = O.GtinvOracle(ref(name))
    s = perturbed_structure()
    tr = train(o, s)
    r = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    w = o.weights()
    assert tr["X"].shape == (s.natoms, o.info["n_des"])
    assert tr["P"].shape == (s.natoms, o.info["n_coeff"])
    e = w @ tr["P"].sum(axis=0)
    f = np.einsum("p,kpc->kc", w, tr["FF"])
    v = w @ tr["SF"]
    assert abs(e - r["energy"]) <= 1e-10 * abs(r["energy"])
    assert np.abs(f - r["forces"]).max() <= 1e-10 * np.abs(r["forces"]).max()
    vm = vmap(r["virial"])
    assert np.abs(v - vm).max() <= 1e-9 * np.abs(vm).max()
    # linear polynomial features are the structural features
    np.testing.assert_array_equal(tr["P"][:, :o.info["n_des"]], tr["X"])


def energy_features(o, s):
    return train(o, s, want=("P",))["P"].sum(axis=0)


@pytest.mark.parametrize("name", ["gtinv-434.mlip", "gtinv-820.mlip"])
def test_force_features_are_minus_gradients(name):
    o = O.GtinvOracle(ref(name))
    s = perturbed_structure(seed=5)
    tr = train(o, s, want=("FF",))
    nd, nc = o.info["n_des"], o.info["n_coeff"]
    n2 = o.info["n_comb2"]
    # coefficients: linear, quadratic and (gtinv-820) cubic, the largest of each
    mag = np.abs(tr["FF"]).max(axis=(0, 2))
    picks = [int(np.argmax(mag[:nd]))]
    if n2:
        picks.append(nd + int(np.argmax(mag[nd:nd + n2])))
    if nc > nd + n2:
        picks.append(nd + n2 + int(np.argmax(mag[nd + n2:])))
    h = 1e-5
    for k, c in [(0, 0), (7, 2), (19, 1)]:
        sp, sm = s.positions.copy(), s.positions.copy()
        sp[k, c] += h
        sm[k, c] -= h
        xp = energy_features(o, W.Structure(s.cell, sp, s.types))
        xm = energy_features(o, W.Structure(s.cell, sm, s.types))
        fd = -(xp - xm) / (2 * h)
        for p in picks:
            scale = max(np.abs(tr["FF"][:, p]).max(), 1e-12)
            assert abs(fd[p] - tr["FF"][k, p, c]) <= 1e-6 * scale, (name, p, k, c, fd[p], tr["FF"][k, p, c])


@pytest.mark.parametrize("comp,ab", [(0, (0, 0)), (3, (1, 2)), (4, (2, 0)), (5, (0, 1))])
def test_stress_features_are_strain_derivatives(comp, ab):
    o = O.GtinvOracle(ref("gtinv-820.mlip"))
    s = perturbed_structure(seed=7)
    tr = train(o, s, want=("SF",))
    a, b = ab
    h = 1e-6

    def deformed(eps):
        cell, x = s.cell.copy(), s.positions.copy()
        cell[:, b] += eps * cell[:, a]
        x[:, b] += eps * x[:, a]
        return W.Structure(cell, x, s.types)

    fd = -(energy_features(o, deformed(h)) - energy_features(o, deformed(-h))) / (2 * h)
    scale = np.abs(tr["SF"][:, comp]).max()
    assert np.abs(fd - tr["SF"][:, comp]).max() <= 1e-6 * scale


def test_structural_features_rotation_invariant():
    o = O.GtinvOracle(ref("gtinv-820.mlip"))
    s = perturbed_structure(seed=9)
    q, _ = np.linalg.qr(np.random.default_rng(4).normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    r = W.Structure(s.cell @ q.T, s.positions @ q.T, s.types)
    xa = train(o, s, want=("X",))["X"]
    xb = train(o, r, want=("X",))["X"]
    assert np.abs(xa - xb).max() <= 1e-10 * np.abs(xa).max()
