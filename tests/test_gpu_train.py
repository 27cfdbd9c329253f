"""Training outputs on the device (mlip_train_features) against the CPU
oracle of MLIPModel::prepare_features (oracle_train), FP64.

Bar: structural and polynomial features within 1e-10 of the largest
magnitude of their array (per structure); force and stress features within
1e-8 of the coefficient's largest force (stress) feature, that scale floored
at 1e-12 of the structure's largest one -- coefficients that vanish by exact
cancellation in the oracle's half list (an atom that only sees its own
images) come out at the 1e-18 roundoff level from the full list's two image
directions -- and at 1e-3 (1e-11 absolute) where everything vanishes by
symmetry (a one-atom cell); the weighted sums reproduce the device's own
energy, forces and virial (1e-10 / 1e-9 relative)."""
import numpy as np
import pytest

import oracle_lib as O
from common import ref
from mlip_hip import BatchSystem, PairMLIPGtinv, ResidentSystem
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu


def gpu_pair(pot):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, "Ti", "Al")
    return p


def perturbed_structure(seed=3, scale=0.05):
    s = W.read_lammps_data(ref("structure.lammps"))
    rng = np.random.default_rng(seed)
    return W.Structure(s.cell.copy(), s.positions + rng.normal(scale=scale, size=s.positions.shape), s.types)


def compare(g, o, tag):
    """g: device outputs of one structure (X, P, FF [n, nc, 3], SF [nc, 6])."""
    worst = {}
    for k in ("X", "P"):
        sc = max(np.abs(o[k]).max(), 1e-300)
        worst[k] = np.abs(g[k] - o[k]).max() / sc
        assert worst[k] <= 1e-10, (tag, k, worst[k])
    mx = np.abs(o["FF"]).max(axis=(0, 2))
    sc = np.maximum(mx, max(1e-12 * mx.max(), 1e-3))
    err = np.abs(g["FF"] - o["FF"]).max(axis=(0, 2))
    worst["FF"] = (err / sc).max()
    if worst["FF"] > 1e-8:
        for p in np.argsort(err / sc)[-3:]:
            print("%s: coefficient %d err %.3e max|FF_p| %.3e structure max %.3e" % (tag, p, err[p], mx[p], mx.max()),
                  flush=True)
    assert worst["FF"] <= 1e-8, (tag, "FF", worst["FF"])
    ms = np.abs(o["SF"]).max(axis=1)
    sc = np.maximum(ms, max(1e-12 * ms.max(), 1e-3))
    worst["SF"] = (np.abs(g["SF"] - o["SF"]).max(axis=1) / sc).max()
    assert worst["SF"] <= 1e-8, (tag, "SF", worst["SF"])
    return worst


@pytest.mark.parametrize("name", ["gtinv-197.mlip", "gtinv-434.mlip", "gtinv-820.mlip"])
def test_train_features_regression_structure(name):
    pot = ref(name)
    o = O.GtinvOracle(pot)
    s = perturbed_structure()
    ro = o.train(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    p = gpu_pair(pot)
    rs = ResidentSystem.from_cell(p, s.cell, s.positions, s.types)
    g = rs.train_features()
    worst = compare({"X": g["X"], "P": g["P"], "FF": g["FF"], "SF": g["SF"][0]}, ro, name)
    # weighted sums against the device's own evaluation
    rs.compute()
    rs.sync()
    ev = rs.get()
    w = o.weights()
    e = w @ g["P"].sum(axis=0)
    f = np.einsum("p,kpc->kc", w, g["FF"])
    v = w @ g["SF"][0]
    vm = np.array([ev["virial"][0], ev["virial"][1], ev["virial"][2], ev["virial"][5], ev["virial"][4],
                   ev["virial"][3]])
    assert abs(e - ev["energy"]) <= 1e-10 * abs(ev["energy"])
    assert np.abs(f - ev["forces"]).max() <= 1e-10 * np.abs(ev["forces"]).max()
    assert np.abs(v - vm).max() <= 1e-9 * np.abs(vm).max()
    print("%s: n_des=%d n_coeff=%d worst %s" % (name, o.info["n_des"], o.info["n_coeff"],
                                                 " ".join("%s=%.1e" % kv for kv in worst.items())), flush=True)


def test_train_features_batch():
    """Per-structure outputs of a batch (random cells, one empty and one
    one-atom structure) equal the oracle's structure by structure."""
    pot = ref("gtinv-820.mlip")
    o = O.GtinvOracle(pot)
    structs = W.random_structures(6, 4, 40, seed=21)
    structs.insert(2, W.Structure(np.diag([5.0, 5.0, 5.0]), np.zeros((0, 3)), np.zeros(0, dtype=np.int32)))
    structs.insert(4, W.Structure(np.diag([3.1, 3.3, 3.6]), np.array([[0.2, 0.1, 0.3]]), np.array([2], dtype=np.int32)))
    p = gpu_pair(pot)
    b = BatchSystem(p, structs)
    g = b.train_features()
    a0 = 0
    for k, s in enumerate(structs):
        n = s.natoms
        if n == 0:
            assert np.all(g["SF"][k] == 0.0)
            continue
        ro = o.train(W.build_neighbor_list(s, o.cutoff), s.types - 1)
        compare({"X": g["X"][a0:a0 + n], "P": g["P"][a0:a0 + n], "FF": g["FF"][a0:a0 + n], "SF": g["SF"][k]}, ro,
                "batch %d (%d atoms)" % (k, n))
        a0 += n
    print("batch: %d structures, %d atoms" % (len(structs), a0), flush=True)


def test_train_features_pair_style_is_an_error():
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", ref("pair-17.mlip"), "Ti", "Al")
    s = perturbed_structure()
    rs = ResidentSystem.from_cell(p, s.cell, s.positions, s.types)
    with pytest.raises(Exception):
        rs.train_features()
