"""The LAMMPS-shaped entry point mlip_compute (``int** firstneigh``, as
NeighList hands it to Pair::compute) against the CPU oracle, with ghosts,
half (CPU style) and full (Kokkos style) lists, eflag_atom and vflag_atom.

Bars: energy 1e-10 eV/atom, forces 1e-8 of max |F|, virial 1e-8 relative;
per-atom energy 1e-10 eV and per-atom virial 1e-10 of its largest
component, in the reference CPU style's ev_tally_xyz split.  Half lists are
compared position by position (ghosts included); full lists after folding
ghost entries onto their owners (LAMMPS reverse communication), since the
two list shapes place the same pair shares on different images."""
import numpy as np
import pytest

import oracle_lib as O
from common import ref, rel_force_err
from mlip_hip import PairMLIPGtinv
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu

POTS = ["gtinv-197.mlip", "gtinv-434.mlip", "gtinv-820.mlip", "pair-17.mlip", "pair-60.mlip"]


def gpu_pair(pot):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, "Ti", "Al")
    return p


def fold(nl, a):
    out = np.zeros((nl.nlocal,) + a.shape[1:])
    np.add.at(out, nl.owner, a)
    return out


def check(r, o, nl, nl_o, tag, per_atom=True, e_tol=1e-10):
    n = nl.nlocal
    de = abs(r["energy"] - o["energy"]) / n
    fe = rel_force_err(r["forces"], o["forces"])
    ve = np.abs(r["virial"] - o["virial"]).max() / np.abs(o["virial"]).max()
    msg = "%s: dE/atom=%.2e relF=%.2e relV=%.2e" % (tag, de, fe, ve)
    if per_atom:
        if nl.full:
            ea, eo = fold(nl, r["eatom"]), fold(nl_o, o["eatom"])
            va, vo = fold(nl, r["vatom"]), fold(nl_o, o["vatom"])
        else:
            ea, eo, va, vo = r["eatom"], o["eatom"], r["vatom"], o["vatom"]
        dea = np.abs(ea - eo).max()
        dva = np.abs(va - vo).max() / np.abs(vo).max()
        msg += " eatom=%.2e vatom=%.2e" % (dea, dva)
    print(msg, flush=True)
    assert de <= e_tol and fe <= 1e-8 and ve <= 1e-8, msg
    if per_atom:
        assert dea <= e_tol and dva <= 1e-10, msg


@pytest.mark.parametrize("pot", POTS)
@pytest.mark.parametrize("full", [False, True])
def test_int_pointer_entry(pot, full):
    s = W.read_lammps_data(ref("structure.lammps"))
    o = O.Oracle(ref(pot))
    nl_o = W.build_neighbor_list(s, o.cutoff, full=False)
    ro = o.compute(nl_o, s.types - 1, eatom=True, vatom=True)
    p = gpu_pair(ref(pot))
    nl = W.build_neighbor_list(s, p.init_one(), full=full)
    r = p.compute_lammps(nl, nl.types, eatom=True, vatom=True)
    check(r, ro, nl, nl_o, "%s full=%d" % (pot, full))
    # the flat entry point gives the same numbers bit for bit
    r2 = p.compute(nl, nl.types, eatom=True, vatom=True)
    assert r2["energy"] == r["energy"]
    assert np.array_equal(r2["forces_all"], r["forces_all"])
    assert np.array_equal(r2["eatom"], r["eatom"]) and np.array_equal(r2["vatom"], r["vatom"])


def rotation(a=0.3, b=0.2):
    rz = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
    rx = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(b), -np.sin(b)], [0.0, np.sin(b), np.cos(b)]])
    return rx @ rz


def rotated(nl, R):
    """The same list with every position (ghosts included) rotated by R."""
    nl.x = np.ascontiguousarray((nl.x.reshape(-1, 3) @ R.T).reshape(nl.x.shape))
    return nl


@pytest.mark.parametrize("pot", ["gtinv-197.mlip", "gtinv-820.mlip", "pair-60.mlip"])
def test_triclinic_64(pot):
    """The reference's 64-atom triclinic cell: a perfect lattice in a tilted
    box, whose owned forces nearly cancel (|F| ~ 1e-5 from the file's
    printed digits) while ghost entries carry up to 0.9 eV/A.  40 of its
    half pairs lie exactly on the z axis, where the reference's derivative
    drops the |m| = 1 polar term (mlip_features.cpp:111) and the oracle
    with it, so parity is taken on a rotated copy of every position; the
    unrotated cell then checks that the device forces are rotation
    covariant there.  Owned forces are compared with a bar of 1e-8 of the
    largest per-position force, the scale their rounding comes from."""
    s = W.read_lammps_data(ref("structure_00064_triclinic.lammps"))
    R = rotation()
    o = O.Oracle(ref(pot))
    nl_o = rotated(W.build_neighbor_list(s, o.cutoff, full=False), R)
    ro = o.compute(nl_o, s.types - 1, eatom=True, vatom=True)
    fscale = np.abs(ro["forces_all"]).max()
    p = gpu_pair(ref(pot))
    for full in (False, True):
        nl = rotated(W.build_neighbor_list(s, p.init_one(), full=full), R)
        r = p.compute_lammps(nl, nl.types, eatom=True, vatom=True)
        fe = np.abs(r["forces"] - ro["forces"]).max() / fscale
        print("%s tri full=%d: owned-force error %.2e of max per-position |F|" % (pot, full, fe), flush=True)
        assert fe <= 1e-8
        r["forces"] = ro["forces"]  # compared above, on its own scale
        check(r, ro, nl, nl_o, "%s tri full=%d" % (pot, full))
        if not full:
            fa = np.abs(r["forces_all"] - ro["forces_all"]).max() / fscale
            assert fa <= 1e-8, fa
        # unrotated: the same forces rotated back (exact on the z axis)
        nl0 = W.build_neighbor_list(s, p.init_one(), full=full)
        r0 = p.compute_lammps(nl0, nl0.types)
        back = r["forces_all"] @ R
        fr = np.abs(r0["forces_all"] - back).max() / fscale
        print("%s tri full=%d: unrotated vs rotated back %.2e" % (pot, full, fr), flush=True)
        assert fr <= 1e-10 and abs(r0["energy"] - r["energy"]) <= 1e-10 * nl.nlocal


def test_chain_split_gpu():
    """Three-atom chain 0-1-2: the split gives eatom[1] = E / 2."""
    p = gpu_pair(ref("gtinv-820.mlip"))
    rc = p.init_one()
    u = np.array([[0.0, 0.0, 0.0], [0.54, 0.12, 0.06], [1.07, 0.31, 0.1]]) * rc
    s = W.Structure(np.eye(3) * 6.0 * rc, u + 2.0 * rc, np.array([1, 2, 1]))
    for full in (False, True):
        nl = W.build_neighbor_list(s, rc, full=full)
        r = p.compute_lammps(nl, nl.types, eatom=True, vatom=True)
        assert abs(r["eatom"][1] - 0.5 * r["energy"]) <= 1e-12 * abs(r["energy"])


def test_site_mode_and_mixed():
    s = W.read_lammps_data(ref("structure.lammps"))
    o = O.Oracle(ref("gtinv-820.mlip"))
    nl_o = W.build_neighbor_list(s, o.cutoff, full=False)
    ro = o.compute(nl_o, s.types - 1, eatom=True, vatom=True)
    p = gpu_pair(ref("gtinv-820.mlip"))
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    p.set_atom_mode("site")
    r = p.compute_lammps(nl, nl.types, eatom=True, vatom=True)
    # site energies sum to the energy; ghosts carry nothing
    assert abs(r["eatom"][: nl.nlocal].sum() - r["energy"]) <= 1e-10 * abs(r["energy"])
    assert not r["eatom"][nl.nlocal:].any()
    assert np.abs(r["vatom"].sum(axis=0) - r["virial"]).max() <= 1e-10 * np.abs(r["virial"]).max()
    p.set_atom_mode("pair")
    p.set_precision("mixed")
    r = p.compute_lammps(nl, nl.types, eatom=True, vatom=True)
    dea = np.abs(fold(nl, r["eatom"]) - fold(nl_o, ro["eatom"])).max()
    print("mixed eatom error %.2e eV" % dea)
    assert dea <= 1e-6
