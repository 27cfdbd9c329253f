"""Mixed precision (FP32 per-pair geometry, radial functions, Y_lm and their
gradients; FP64 a_nlm accumulation, invariants, adjoint and reductions)
against the FP64 oracle through the C ABI.

Bar: energy within 1e-6 eV/atom and every force component within 1e-4 eV/A
of the FP64 oracle; the errors are printed for every case."""
import numpy as np
import pytest

import oracle_lib as O
from common import config_potential, ref
from mlip_hip import BatchSystem, PairMLIPGtinv, ResidentSystem, periodic_full_list
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu

E_TOL = 1e-6   # eV/atom
F_TOL = 1e-4   # eV/A, absolute per component


def gpu_pair(pot, elems, precision):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, *elems)
    p.set_precision(precision)
    return p


def report(tag, r_e, r_f, o_e, o_f, natoms):
    de = abs(r_e - o_e) / natoms
    df = np.abs(r_f - o_f)
    print("%s: dE/atom=%.2e  F max=%.2e MAE=%.2e eV/A (|F|max %.2f)"
          % (tag, de, df.max(), df.mean(), np.abs(o_f).max()), flush=True)
    assert de <= E_TOL, tag
    assert df.max() <= F_TOL, tag


@pytest.mark.parametrize("pot", ["gtinv-197.mlip", "gtinv-434.mlip", "gtinv-820.mlip"])
def test_regression_structure_mixed(pot):
    s = W.read_lammps_data(ref("structure.lammps"))
    o = O.Oracle(ref(pot))
    ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    p = gpu_pair(ref(pot), ["Ti", "Al"], "mixed")
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    r = p.compute(nl, nl.types)
    report("mixed " + pot, r["energy"], r["forces"], ro["energy"], ro["forces"], s.natoms)


def test_benchmark_models_mixed():
    for name, s, elems in [
            ("fcc256_order3_linear", W.fcc(4, 4.05, 0.05, 1), ["X"]),
            ("binary_order4_quadratic", W.fcc(3, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4), ["Ti", "Al"])]:
        pot = config_potential(name)
        o = O.Oracle(pot)
        ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
        p = gpu_pair(pot, elems, "mixed")
        off, idx, sh = periodic_full_list(s.cell, s.positions, p.init_one())
        rs = ResidentSystem(p, s.positions, s.types, off, idx, sh)
        rs.compute()
        rs.sync()
        r = rs.get()
        report("mixed " + name, r["energy"], r["forces"], ro["energy"], ro["forces"], s.natoms)
        # repeatable bit for bit
        rs.compute()
        rs.sync()
        r2 = rs.get()
        assert r2["energy"] == r["energy"] and np.array_equal(r2["forces"], r["forces"])


def test_precision_switch_restores_fp64():
    pot = ref("gtinv-820.mlip")
    s = W.fcc(3, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
    p = gpu_pair(pot, ["Ti", "Al"], "fp64")
    off, idx, sh = periodic_full_list(s.cell, s.positions, p.init_one())
    rs = ResidentSystem(p, s.positions, s.types, off, idx, sh)
    rs.compute()
    rs.sync()
    a = rs.get()
    p.set_precision("mixed")
    rs.compute()
    rs.sync()
    b = rs.get()
    assert b["energy"] != a["energy"]
    p.set_precision("fp64")
    rs.compute()
    rs.sync()
    c = rs.get()
    assert c["energy"] == a["energy"] and np.array_equal(c["forces"], a["forces"])


def test_batch_mixed():
    pot = config_potential("binary_order4_quadratic")
    structs = W.random_structures(6, 8, 48, seed=6)
    o = O.Oracle(pot)
    p = gpu_pair(pot, ["Ti", "Al"], "mixed")
    b = BatchSystem(p, structs)
    b.compute()
    b.sync()
    r = b.get()
    for k, s in enumerate(structs):
        ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
        report("mixed batch %d (%d atoms)" % (k, s.natoms), r["energies"][k], r["forces"][k], ro["energy"],
               ro["forces"], s.natoms)
