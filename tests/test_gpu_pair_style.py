"""Radial-only style (pair_style mlip_pair) on the device through the C ABI.

Bars: energy within 1e-10 eV/atom, forces within 1e-8 of the largest
component, virial within 1e-8 relative, against the CPU oracle of the
reference style (half and full neighbor lists, regression structure and
random cells); 100-step NVE against the regression logs (8 printed digits)
and dumps (1e-8 A)."""
import json
import os

import numpy as np
import pytest

import md_oracle as MD
import oracle_lib as O
from common import GOLD, ref, rel_force_err
from mlip_hip import BatchSystem, PairMLIPGtinv, RequestPool
from mlip_hip import workloads as W
from test_md_oracle import check_thermo

pytestmark = pytest.mark.gpu


def gpu_pair(pot):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, "Ti", "Al")
    return p


def compare(r_e, r_f, r_v, o, natoms, tag):
    de = abs(r_e - o["energy"]) / natoms
    fe = rel_force_err(r_f, o["forces"])
    ve = np.abs(r_v - o["virial"]).max() / np.abs(o["virial"]).max()
    print("%s: dE/atom=%.2e relF=%.2e relV=%.2e" % (tag, de, fe, ve), flush=True)
    assert de <= 1e-10 and fe <= 1e-8 and ve <= 1e-8, tag


@pytest.mark.parametrize("name", ["pair-17", "pair-60"])
@pytest.mark.parametrize("full", [False, True])
def test_pair_style_regression_structure(name, full):
    s = W.read_lammps_data(ref("structure.lammps"))
    o = O.Oracle(ref(name + ".mlip"))
    ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    p = gpu_pair(ref(name + ".mlip"))
    nl = W.build_neighbor_list(s, p.init_one(), full=full)
    r = p.compute(nl, nl.types)
    compare(r["energy"], r["forces"], r["virial"], ro, s.natoms, "%s full=%d" % (name, full))


def test_pair_style_batch_random_cells():
    pot = ref("pair-60.mlip")
    structs = W.random_structures(6, 8, 64, seed=31)
    o = O.Oracle(pot)
    p = gpu_pair(pot)
    b = BatchSystem(p, structs)
    b.compute()
    b.sync()
    r = b.get()
    for k, s in enumerate(structs):
        ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
        compare(r["energies"][k], r["forces"][k], r["virials"][k], ro, s.natoms, "batch %d" % k)


@pytest.mark.parametrize("name", ["pair-17", "pair-60"])
def test_pair_style_nve_regression_runs(name):
    g = json.load(open(os.path.join(GOLD, "regression_goldens.json")))[name]
    s = W.read_lammps_data(ref(g["structure"]))
    p = gpu_pair(ref(g["potential"]))
    pool = RequestPool(p, 1000, 400000, 4)
    h = pool.admit(s, nsteps=100, dt=0.001, skin=2.0)
    res = None
    while pool.tick():
        pool.sync()
        for hh in pool.done():
            res = pool.result(hh)
    assert res is not None and res["thermo"].shape == (101, 13)
    th = lambda row: MD.thermo(s.natoms, s.volume, row[0], row[1:7], row[7:13])  # noqa: E731
    check_thermo(th(res["thermo"][0]), {k: g["step0"][k] for k in ("E_pair", "Press")}, name + " step 0")
    check_thermo(th(res["thermo"][100]), g["step100_dt0.001"], name + " step 100")
    d = MD.periodic_diff(res["positions"], MD.read_dump(ref(g["dump100_dt0.001"])), s.cell)
    print("%s: pool NVE max |x - dump| = %.2e A" % (name, np.abs(d).max()), flush=True)
    assert np.abs(d).max() < 1e-8
