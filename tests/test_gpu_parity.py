"""HIP path against the CPU oracle (both FP64) through the C ABI.

Bar (stated per test): energy within 1e-10 eV/atom, forces within 1e-8 of
the largest force component (relative), virial within 1e-8 relative to its
largest component.  Neighbor lists are the reference's own shapes: the half
list of the CPU style and the full list of the Kokkos style."""
import numpy as np
import pytest

import oracle_lib as O
from common import config_potential, ref, rel_force_err
from mlip_hip import PairMLIPGtinv, ResidentSystem
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu

E_TOL = 1e-10   # eV/atom
F_TOL = 1e-8    # relative to max |F|
V_TOL = 1e-8    # relative to max |virial|


def gpu_pair(pot, elems):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, *elems)
    return p


def compare(r, o, natoms, tag):
    de = abs(r["energy"] - o["energy"]) / natoms
    fe = rel_force_err(r["forces"], o["forces"])
    ve = np.abs(r["virial"] - o["virial"]).max() / max(np.abs(o["virial"]).max(), 1e-300)
    print("%s: dE/atom=%.2e relF=%.2e relV=%.2e" % (tag, de, fe, ve), flush=True)
    assert de <= E_TOL, tag
    assert fe <= F_TOL, tag
    assert ve <= V_TOL, tag


REG = [("gtinv-197.mlip", "structure.lammps"), ("gtinv-434.mlip", "structure.lammps"),
       ("gtinv-820.mlip", "structure.lammps")]


@pytest.mark.parametrize("pot,struct", REG)
@pytest.mark.parametrize("full", [False, True])
def test_regression_structures(pot, struct, full):
    s = W.read_lammps_data(ref(struct))
    o = O.Oracle(ref(pot))
    nl_half = W.build_neighbor_list(s, o.cutoff, full=False)
    ro = o.compute(nl_half, s.types - 1)
    p = gpu_pair(ref(pot), ["Ti", "Al"])
    nl = W.build_neighbor_list(s, p.init_one(), full=full)
    r = p.compute(nl, nl.types)
    compare(r, ro, s.natoms, "%s full=%d" % (pot, full))


def test_fcc256_linear_order3():
    pot = config_potential("fcc256_order3_linear")
    s = W.fcc(4, 4.05, 0.05, 1)
    o = O.Oracle(pot)
    ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    p = gpu_pair(pot, ["X"])
    r = p.compute(*_full(s, p))
    compare(r, ro, s.natoms, "fcc256")


def _full(s, p):
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    return nl, nl.types


def test_binary_order4_quadratic_small():
    pot = config_potential("binary_order4_quadratic")
    s = W.fcc(3, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
    o = O.Oracle(pot)
    ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    p = gpu_pair(pot, ["Ti", "Al"])
    r = p.compute(*_full(s, p))
    compare(r, ro, s.natoms, "binary108")


def test_resident_matches_boundary_and_is_deterministic():
    pot = ref("gtinv-820.mlip")
    s = W.fcc(3, 4.05, 0.1, 9, type_fraction=0.3, type_seed=3)
    p = gpu_pair(pot, ["Ti", "Al"])
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    rb = p.compute(nl, nl.types)
    off, idx, sh = W.resident_lists(nl)
    rs = ResidentSystem(p, nl.x[:nl.nlocal], s.types, off, idx, sh)
    rs.compute()
    rs.sync()
    a = rs.get()
    rs.compute()
    rs.sync()
    b = rs.get()
    assert a["energy"] == b["energy"]
    assert np.array_equal(a["forces"], b["forces"])
    assert abs(a["energy"] - rb["energy"]) / s.natoms < 1e-12
    assert rel_force_err(a["forces"], rb["forces"]) < 1e-12


def test_conventional_cell_goldens():
    # perfect 64-atom lattice: E and pressure against the reference log
    # (log.lammps.gtinv-197.conv64:35); forces vanish by symmetry
    s = W.read_lammps_data(ref("structure_00064_conventional.lammps"))
    p = gpu_pair(ref("gtinv-197.mlip"), ["Ti", "Al"])
    r = p.compute(*_full(s, p))
    assert r["energy"] == pytest.approx(-273.09596, abs=1e-5)
    press = r["virial"][:3].sum() / 3 / s.volume * 1.6021765e6
    assert press == pytest.approx(-18477.159, abs=2e-3)
    assert np.abs(r["forces"]).max() < 1e-10
