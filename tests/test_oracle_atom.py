"""Per-atom energy and virial of the CPU oracle (ev_tally_xyz split of the
reference CPU style, pair_mlip_gtinv.cpp:127-159,240-253; pair_mlip_pair.cpp
:214-236), pinned without the reference:

* sums: sum eatom = E and sum vatom = virial;
* a three-atom chain 0-1-2 (0-2 beyond the cutoff, isolated in a large box):
  atom 0 has one neighbor, so its site energy equals its one pair share by
  Euler's theorem once every term of dE/da is divided by its degree, and the
  chain forces eatom[1] = E / 2 exactly; a missing or wrong degree weight
  breaks it;
* the chain's end atoms have a single pair, so vatom[0] = 1/2 del_01 (x) F_0
  and vatom[2] = 1/2 del_21 (x) F_2 with the (pinned) total forces;
* a dimer splits its energy equally.
"""
import numpy as np
import pytest

import oracle_lib as O
from common import ref
from mlip_hip import workloads as W

POTS = ["gtinv-197.mlip", "gtinv-434.mlip", "gtinv-820.mlip", "pair-17.mlip", "pair-60.mlip"]


def chain(rc, types=(1, 2, 1)):
    """0-1 and 1-2 at about 0.56 rc, 0-2 beyond rc, in a box with no image
    within reach."""
    u = np.array([[0.0, 0.0, 0.0], [0.54, 0.12, 0.06], [1.07, 0.31, 0.1]]) * rc
    return W.Structure(np.eye(3) * 6.0 * rc, u + 2.0 * rc, np.array(types))


def vtensor(d, f):
    return np.array([d[0] * f[0], d[1] * f[1], d[2] * f[2], d[0] * f[1], d[0] * f[2], d[1] * f[2]])


@pytest.mark.parametrize("pot", POTS)
def test_chain_split(pot):
    o = O.Oracle(ref(pot))
    s = chain(o.cutoff)
    nl = W.build_neighbor_list(s, o.cutoff, full=False)
    assert nl.npairs == 2
    r = o.compute(nl, s.types - 1, eatom=True, vatom=True)
    E = r["energy"]
    assert abs(E) > 1e-6
    ea, va, f = r["eatom"], r["vatom"], r["forces"]
    assert abs(ea.sum() - E) <= 1e-12 * abs(E)
    assert np.allclose(va.sum(axis=0), r["virial"], rtol=0, atol=1e-12 * np.abs(r["virial"]).max())
    assert abs(ea[1] - 0.5 * E) <= 1e-12 * abs(E), (ea, E)
    x = s.positions
    sc = max(np.abs(va).max(), 1e-300)
    assert np.abs(va[0] - 0.5 * vtensor(x[0] - x[1], f[0])).max() <= 1e-12 * sc
    assert np.abs(va[2] - 0.5 * vtensor(x[2] - x[1], f[2])).max() <= 1e-12 * sc


@pytest.mark.parametrize("pot", ["gtinv-197.mlip", "gtinv-820.mlip", "pair-60.mlip"])
def test_dimer_split(pot):
    o = O.Oracle(ref(pot))
    s = W.Structure(np.eye(3) * 40.0, np.array([[10.0, 10.0, 10.0], [12.1, 13.0, 11.2]]), np.array([1, 2]))
    nl = W.build_neighbor_list(s, o.cutoff, full=False)
    r = o.compute(nl, s.types - 1, eatom=True, vatom=True)
    assert abs(r["eatom"][0] - r["eatom"][1]) <= 1e-14 * abs(r["energy"])
    assert np.allclose(r["vatom"][0], r["vatom"][1], rtol=1e-14, atol=0)


@pytest.mark.parametrize("pot", POTS)
def test_periodic_sums(pot):
    o = O.Oracle(ref(pot))
    s = W.read_lammps_data(ref("structure.lammps"))
    nl = W.build_neighbor_list(s, o.cutoff, full=False)
    r = o.compute(nl, s.types - 1, eatom=True, vatom=True)
    E = r["energy"]
    assert abs(r["eatom"].sum() - E) <= 1e-11 * abs(E)
    assert np.abs(r["vatom"].sum(axis=0) - r["virial"]).max() <= 1e-11 * np.abs(r["virial"]).max()
