"""mlip_compute_flat_noalloc (every device buffer in the library's fixed
scratch array, no allocation) against mlip_compute_flat, through the C ABI:
two calls in a row, the second on a smaller input, energies, forces,
virials and per-atom energies/virials equal to the allocating entry point's
(the same kernels on the same data: bit for bit), and the oracle's bar
(1e-10 eV/atom, 1e-8 relative force) on the way."""
import numpy as np
import pytest

import oracle_lib as O
from common import config_potential, ref, rel_force_err
from mlip_hip import PairMLIPGtinv
from mlip_hip import workloads as W
from mlip_hip._lib import MlipError

pytestmark = pytest.mark.gpu


def pair_of(pot, elems=("Ti", "Al")):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, *elems)
    return p


def same(a, b, tag):
    assert a["energy"] == b["energy"], tag
    assert np.array_equal(a["forces_all"], b["forces_all"]), tag
    assert np.array_equal(a["virial"], b["virial"]), tag
    for k in ("eatom", "vatom"):
        if k in a:
            assert np.array_equal(a[k], b[k]), (tag, k)


def test_noalloc_matches_allocating_entry():
    # first call: 108-atom cell of the 256k benchmark model (two species,
    # orders up to 4, quadratic polynomial), full list, per-atom outputs
    pot = config_potential("binary_order4_quadratic")
    p = pair_of(pot)
    s = W.fcc(3, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    a = p.compute(nl, nl.types, eatom=True, vatom=True)
    b = p.compute(nl, nl.types, eatom=True, vatom=True, noalloc=True)
    same(a, b, "binary108")
    o = O.Oracle(pot)
    ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
    assert abs(b["energy"] - ro["energy"]) / s.natoms <= 1e-10
    assert rel_force_err(b["forces"], ro["forces"]) <= 1e-8
    # second call, smaller input: 32-atom regression structure, half list,
    # another model on another context (the scratch array keeps the first
    # call's contents)
    q = pair_of(ref("gtinv-820.mlip"))
    s2 = W.read_lammps_data(ref("structure.lammps"))
    nl2 = W.build_neighbor_list(s2, q.init_one(), full=False)
    c = q.compute(nl2, nl2.types, eatom=True, vatom=True)
    d = q.compute(nl2, nl2.types, eatom=True, vatom=True, noalloc=True)
    same(c, d, "gtinv-820 half")
    # and the first structure again after the smaller one
    e = p.compute(nl, nl.types, noalloc=True)
    assert e["energy"] == a["energy"] and np.array_equal(e["forces_all"], a["forces_all"])


def test_noalloc_rejects_oversize():
    p = pair_of(config_potential("binary_order4_quadratic"))
    s = W.fcc(8, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)   # 2048 atoms: beyond the array
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    with pytest.raises(MlipError) as ei:
        p.compute(nl, nl.types, noalloc=True)
    assert "scratch" in str(ei.value)
