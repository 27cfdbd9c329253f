"""Device-side neighbor and force-gather lists (mlip_resident_set_cell,
mlip_batch_set with the device builder) against the host builder and the
brute-force image search.

Bar: bit-exact.  Per atom the same neighbor count as the brute-force
numpy search over lattice images (workloads.build_neighbor_list, strict
r < rc, self-images kept); the same list as the host builder entry for
entry (neighbor index and the image shift's bits); gather lists equal to a
numpy restatement of "for position k, -(p+1) for pairs centred on k and
p+1 for pairs with neighbor k, in increasing p, centre entry first"; and
energies, forces and virials from device-built lists identical to those
from host-built lists."""
import numpy as np
import pytest

from common import config_potential, ref
from mlip_hip import BatchSystem, PairMLIPGtinv, ResidentSystem, periodic_full_list
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu


def gpu_pair(pot, elems):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, *elems)
    return p


def gather_expected(off, idx):
    n = off.size - 1
    np_ = idx.size
    cen = np.repeat(np.arange(n), np.diff(off))
    p = np.arange(np_)
    pos = np.concatenate([cen, idx])
    ent = np.concatenate([-(p + 1), p + 1]).astype(np.int32)
    pp = np.concatenate([p, p])
    side = np.concatenate([np.zeros(np_, dtype=np.int8), np.ones(np_, dtype=np.int8)])
    order = np.lexsort((side, pp, pos))
    goff = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(pos, minlength=n), out=goff[1:])
    return goff, ent[order]


def brute_counts(s, rc):
    nl = W.build_neighbor_list(s, rc, full=True)
    cnt = np.zeros(s.natoms, dtype=np.int64)
    for i in range(s.natoms):
        js = nl.idx[nl.off[i]:nl.off[i + 1]]
        r = np.linalg.norm(nl.x[i] - nl.x[js], axis=1)
        cnt[i] = int((r < rc).sum())
    return cnt


def check_lists(lists, off, idx, sh, tag):
    assert np.array_equal(lists["off"], off), tag
    assert np.array_equal(lists["idx"], idx), tag
    # image shifts bit for bit
    assert np.array_equal(lists["shift"].view(np.uint64), np.ascontiguousarray(sh).view(np.uint64)), tag
    goff, gent = gather_expected(off, idx)
    assert np.array_equal(lists["goff"], goff), tag
    assert np.array_equal(lists["gent"], gent), tag


CASES = ["fcc108", "tiny_triclinic", "reference_triclinic", "regression32", "one_atom"]


def case_structure(case):
    if case == "fcc108":
        return W.fcc(3, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
    if case == "tiny_triclinic":
        # 9 atoms: every atom sees several images of itself and of the others
        return W.random_cell_structure(9, 0.06, np.random.default_rng(6), type_fraction=0.5)
    if case == "reference_triclinic":
        return W.read_lammps_data(ref("structure_00064_triclinic.lammps"))
    if case == "regression32":
        return W.read_lammps_data(ref("structure.lammps"))
    return W.Structure(np.diag([3.1, 3.3, 3.6]), np.array([[0.2, 0.1, 0.3]]), np.array([2], dtype=np.int32))


@pytest.mark.parametrize("case", CASES)
def test_device_list_matches_host_and_brute_force(case):
    pot = config_potential("fcc256_order3_linear")
    p = gpu_pair(pot, ["X"])
    s = case_structure(case)
    types = np.ones(s.natoms, dtype=np.int32)
    rc = p.init_one()
    rs = ResidentSystem.from_cell(p, s.cell, s.positions, types)
    L = rs.lists()
    off, idx, sh = periodic_full_list(s.cell, s.positions, rc)
    check_lists(L, off, idx, sh, case)
    assert np.array_equal(np.diff(L["off"]), brute_counts(s, rc)), case
    # evaluation from device-built lists equals host-built lists bit for bit
    rs.compute()
    rs.sync()
    a = rs.get(eatom=True)
    rh = ResidentSystem(p, s.positions, types, off, idx, sh)
    rh.compute()
    rh.sync()
    b = rh.get(eatom=True)
    assert a["energy"] == b["energy"], case
    assert np.array_equal(a["forces"], b["forces"]), case
    assert np.array_equal(a["virial"], b["virial"]), case
    assert np.array_equal(a["eatom"], b["eatom"]), case
    print("%s: %d atoms, %d pairs, identical lists and results" % (case, s.natoms, idx.size), flush=True)


def test_device_batch_equals_host_batch():
    pot = ref("gtinv-197.mlip")
    structs = W.random_structures(48, 8, 160, seed=11)
    structs.insert(5, W.Structure(np.diag([5.0, 5.0, 5.0]), np.zeros((0, 3)), np.zeros(0, dtype=np.int32)))
    structs.insert(9, case_structure("one_atom"))
    out = {}
    for builder in ("host", "device"):
        p = gpu_pair(pot, ["Ti", "Al"])
        b = BatchSystem(p, structs, builder=builder)
        b.compute()
        b.sync()
        out[builder] = (b.get(), b.lists())
    (rh, lh), (rd, ld) = out["host"], out["device"]
    for k in ("off", "idx", "goff", "gent"):
        assert np.array_equal(lh[k], ld[k]), k
    assert np.array_equal(lh["shift"].view(np.uint64), ld["shift"].view(np.uint64))
    assert np.array_equal(rh["energies"], rd["energies"])
    assert np.array_equal(rh["virials"], rd["virials"])
    for fh, fd in zip(rh["forces"], rd["forces"]):
        assert np.array_equal(fh, fd)
    # per-structure lists equal the single-cell host builder's, offset by
    # the structure's first atom
    rc = gpu_pair(pot, ["Ti", "Al"]).init_one()
    a0 = 0
    for s in structs:
        if s.natoms:
            off, idx, sh = periodic_full_list(s.cell, s.positions, rc)
            lo, hi = ld["off"][a0], ld["off"][a0 + s.natoms]
            assert np.array_equal(ld["off"][a0:a0 + s.natoms + 1] - lo, off)
            assert np.array_equal(ld["idx"][lo:hi] - a0, idx)
        a0 += s.natoms
    print("batch: %d structures, %d pairs, host and device identical" % (len(structs), ld["idx"].size), flush=True)


def test_device_list_full_size():
    """The 256,000-atom benchmark structure (binary fcc, 40^3 cells, cutoff
    8): identical list and gather to the host builder."""
    pot = config_potential("binary_order4_quadratic")
    p = gpu_pair(pot, ["Ti", "Al"])
    s = W.fcc(40, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
    rs = ResidentSystem.from_cell(p, s.cell, s.positions, s.types)
    L = rs.lists()
    off, idx, sh = periodic_full_list(s.cell, s.positions, p.init_one())
    assert L["idx"].size == idx.size
    check_lists(L, off, idx, sh, "fcc256k")
    print("fcc256k: %d pairs identical to the host builder" % idx.size, flush=True)


def test_device_builder_rejects_bad_input():
    pot = config_potential("fcc256_order3_linear")
    p = gpu_pair(pot, ["X"])
    s = case_structure("fcc108")
    t = np.ones(s.natoms, dtype=np.int32)
    x = s.positions.copy()
    x[3, 1] = np.nan
    with pytest.raises(Exception):
        ResidentSystem.from_cell(p, s.cell, x, t)
    with pytest.raises(Exception):
        ResidentSystem.from_cell(p, np.zeros((3, 3)), s.positions, t)
    # the context still works afterwards
    rs = ResidentSystem.from_cell(p, s.cell, s.positions, t)
    rs.compute()
    rs.sync()
    assert np.isfinite(rs.get()["energy"])
