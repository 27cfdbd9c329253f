"""The LAMMPS adapter (lammps-mlip-package_amd/lammps/pair_mlip_gtinv_hip.cpp)
driven through the calls of a LAMMPS run by tests/lammps_harness, with the
flags LAMMPS passes with newton on and per-atom computes active
(eflag = ENERGY_GLOBAL | ENERGY_ATOM, vflag = VIRIAL_FDOTR | VIRIAL_ATOM),
against the CPU oracle on the reference's structure with ghost images.

The adapter must request a full list, keep newton on, keep the global
virial (no_virial_fdotr_compute, so ev_init leaves vflag_global set), set the
masses from the potential file and return the cutoff from init_one.  Bars as
for the C ABI: energy 1e-10 eV/atom, forces (ghosts folded onto owners, as
LAMMPS's reverse communication does) 1e-8 of max |F|, virial 1e-8
relative, per-atom energy 1e-10 eV and per-atom virial 1e-10 of its largest
component after the same folding.  The adapter's numbers are those of the
C ABI called directly, bit for bit, and a second step repeats the first."""
import numpy as np
import pytest

import adapter_harness as H
import oracle_lib as O
from common import ref, rel_force_err
from mlip_hip import PairMLIPGtinv
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu


def fold(nl, a):
    out = np.zeros((nl.nlocal,) + a.shape[1:])
    np.add.at(out, nl.owner, a)
    return out


@pytest.mark.parametrize("pot", ["gtinv-197.mlip", "gtinv-820.mlip", "pair-60.mlip"])
def test_adapter_lammps_run(pot, tmp_path):
    s = W.read_lammps_data(ref("structure.lammps"))
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", ref(pot), "Ti", "Al")
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    print("case: adapter %s, %d owned + %d ghosts, %d list entries" % (pot, nl.nlocal, nl.nall - nl.nlocal,
                                                                      nl.npairs), flush=True)
    rc, err, out = H.run(tmp_path, "mlip_gtinv/hip", ref(pot), ["Ti", "Al"], nl,
                         H.ENERGY_GLOBAL | H.ENERGY_ATOM, H.VIRIAL_FDOTR | H.VIRIAL_ATOM, nsteps=2)
    assert rc == 0, err
    # what the adapter told LAMMPS
    assert out["request_full"] == 1 and out["request_half"] == 0
    assert out["newton_pair"] == 1
    assert out["vflag_fdotr"] == 0 and out["vflag_global"] == H.VIRIAL_FDOTR
    assert out["eflag_atom"] and out["vflag_atom"]
    assert out["cutoff"] == p.init_one()
    assert np.array_equal(out["mass"], np.asarray(p.masses()))
    st = out["steps"][0]
    # against the oracle (half list, the CPU style's own shape)
    o = O.Oracle(ref(pot))
    nl_o = W.build_neighbor_list(s, o.cutoff, full=False)
    ro = o.compute(nl_o, s.types - 1, eatom=True, vatom=True)
    n = nl.nlocal
    de = abs(st["energy"] - ro["energy"]) / n
    fe = rel_force_err(fold(nl, st["forces_all"]), ro["forces"])
    ve = np.abs(st["virial"] - ro["virial"]).max() / np.abs(ro["virial"]).max()
    dea = np.abs(fold(nl, st["eatom"]) - fold(nl_o, ro["eatom"])).max()
    vo = fold(nl_o, ro["vatom"])
    dva = np.abs(fold(nl, st["vatom"]) - vo).max() / np.abs(vo).max()
    msg = "%s: dE/atom=%.2e relF=%.2e relV=%.2e eatom=%.2e vatom=%.2e" % (pot, de, fe, ve, dea, dva)
    print(msg, flush=True)
    assert de <= 1e-10 and fe <= 1e-8 and ve <= 1e-8 and dea <= 1e-10 and dva <= 1e-10, msg
    # the C ABI called directly gives the same numbers bit for bit
    r = p.compute_lammps(nl, nl.types, eatom=True, vatom=True)
    assert st["energy"] == r["energy"]
    assert np.array_equal(st["virial"], r["virial"])
    assert np.array_equal(st["forces_all"], r["forces_all"])
    assert np.array_equal(st["eatom"], r["eatom"]) and np.array_equal(st["vatom"], r["vatom"])
    # a second step (force_clear, ev_init zeroing) repeats the first
    s2 = out["steps"][1]
    assert s2["energy"] == st["energy"] and np.array_equal(s2["forces_all"], st["forces_all"])
    assert np.array_equal(s2["virial"], st["virial"]) and np.array_equal(s2["vatom"], st["vatom"])


def test_adapter_energy_only_step(tmp_path):
    """eflag = ENERGY_GLOBAL, vflag = 0 (a step where no compute asks for the
    virial): energy and forces as with the virial, no per-atom arrays."""
    pot = "gtinv-820.mlip"
    s = W.read_lammps_data(ref("structure.lammps"))
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", ref(pot), "Ti", "Al")
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    print("case: adapter energy-only step %s" % pot, flush=True)
    rc, err, out = H.run(tmp_path, "mlip_gtinv/hip", ref(pot), ["Ti", "Al"], nl, H.ENERGY_GLOBAL, 0)
    assert rc == 0, err
    st = out["steps"][0]
    assert out["vflag_global"] == 0 and not out["eflag_atom"] and not out["vflag_atom"]
    r = p.compute_lammps(nl, nl.types)
    assert st["energy"] == r["energy"] and np.array_equal(st["forces_all"], r["forces_all"])
    assert np.array_equal(st["virial"], np.zeros(6))
