"""CPU oracle of the radial-only style (pair_style mlip_pair): model
enumeration against the potential files' coefficient counts, step-0
energy and pressure against the regression logs (8 printed digits),
F = -grad E by central differences, sum F = 0, translation invariance."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from common import GOLD, ref
from mlip_hip import workloads as W
from test_md_oracle import digits_tol

NKTV2P = 1.6021765e6


def goldens():
    return json.load(open(os.path.join(GOLD, "regression_goldens.json")))


@pytest.mark.parametrize("name,n_coeff,n_des,nc2,nc3", [("pair-17", 90, 30, 30, 30),
                                                         ("pair-60", 22910, 60, 1430, 21420)])
def test_pair_model_enumeration(name, n_coeff, n_des, nc2, nc3):
    o = O.Oracle(ref(name + ".mlip"))
    assert (o.info["n_coeff"], o.info["n_des"], o.info["n_comb2"], o.info["n_comb3"]) == (n_coeff, n_des, nc2, nc3)
    if name == "pair-60":
        # model_type 2: every pair of features whose type combinations a
        # single centre type sees (never Ti-Ti with Al-Al)
        c2 = o.combs(2)
        tc = c2 // o.info["n_fn"]
        assert not np.any((tc[:, 0] == 0) & (tc[:, 1] == 2))
        assert np.all(c2[:, 0] <= c2[:, 1])


@pytest.mark.parametrize("name", ["pair-17", "pair-60"])
def test_pair_step0_goldens(name):
    g = goldens()[name]
    s = W.read_lammps_data(ref(g["structure"]))
    o = O.Oracle(ref(g["potential"]))
    r = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    p = r["virial"][:3].sum() / 3 / s.volume * NKTV2P
    assert abs(r["energy"] - g["step0"]["E_pair"]) <= digits_tol(g["step0"]["E_pair"])
    assert abs(p - g["step0"]["Press"]) <= digits_tol(g["step0"]["Press"])
    assert np.abs(r["forces"].sum(axis=0)).max() < 1e-10


def test_pair_forces_finite_difference():
    s = W.read_lammps_data(ref("structure.lammps"))
    o = O.Oracle(ref("pair-60.mlip"))
    r = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    h = 1e-4
    rng = np.random.default_rng(0)
    for _ in range(4):
        i, k = int(rng.integers(s.natoms)), int(rng.integers(3))
        e = []
        for sg in (1, -1):
            x = s.positions.copy()
            x[i, k] += sg * h
            t = W.Structure(s.cell, x, s.types)
            e.append(o.compute(W.build_neighbor_list(t, o.cutoff), s.types - 1)["energy"])
        fd = -(e[0] - e[1]) / (2 * h)
        assert abs(fd - r["forces"][i, k]) < 1e-6 * max(1.0, abs(fd))


def test_pair_translation_invariance():
    s = W.read_lammps_data(ref("structure.lammps"))
    o = O.Oracle(ref("pair-17.mlip"))
    a = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    t = W.Structure(s.cell, s.positions + np.array([0.37, -1.1, 2.9]), s.types)
    b = o.compute(W.build_neighbor_list(t, o.cutoff), s.types - 1)
    assert abs(a["energy"] - b["energy"]) < 1e-10
    assert np.abs(a["forces"] - b["forces"]).max() < 1e-10
