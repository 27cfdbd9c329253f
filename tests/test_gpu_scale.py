"""Parity at the workload sizes of the benchmark configurations, against the
CPU oracle where it finishes in seconds and through size-independent
properties where it does not.

* batched: 12 structures of 300-512 atoms (the top of the batched
  workload's 8-512 range) of the benchmark model in one batch, FP64 and mixed;
* served: pool requests of 1,000 to 4,096 atoms (the served workload's range)
  evaluated together, against the oracle;
* the 256,000-atom structure: forces sum to zero (1e-9 of the largest force),
  a rigid translation leaves energy and forces unchanged, and a repeated
  evaluation is bit-identical.

Bars: FP64 1e-10 eV/atom, forces 1e-8 of the structure's largest force;
mixed 1e-6 eV/atom and 1e-4 eV/A."""
import numpy as np
import pytest

import oracle_lib as O
from common import config_potential, rel_force_err
from mlip_hip import BatchSystem, PairMLIPGtinv, RequestPool, ResidentSystem
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu

POT = "binary_order4_quadratic"


def gpu_pair(precision="fp64"):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", config_potential(POT), "Ti", "Al")
    p.set_precision(precision)
    return p


def test_batch_of_large_structures():
    structs = W.random_structures(12, 300, 512, seed=41)
    assert min(s.natoms for s in structs) >= 300
    o = O.Oracle(config_potential(POT))
    ref = [o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1) for s in structs]
    for precision in ("fp64", "mixed"):
        p = gpu_pair(precision)
        b = BatchSystem(p, structs)
        b.compute()
        b.sync()
        r = b.get()
        worst_e = worst_f = 0.0
        for k, (s, ro) in enumerate(zip(structs, ref)):
            de = abs(r["energies"][k] - ro["energy"]) / s.natoms
            df = np.abs(r["forces"][k] - ro["forces"]).max()
            fr = df / np.abs(ro["forces"]).max()
            worst_e, worst_f = max(worst_e, de), max(worst_f, df if precision == "mixed" else fr)
            if precision == "fp64":
                assert de <= 1e-10 and fr <= 1e-8, (k, s.natoms, de, fr)
            else:
                assert de <= 1e-6 and df <= 1e-4, (k, s.natoms, de, df)
        print("batch of 12 structures (%d-%d atoms) %s: worst dE/atom %.2e, worst dF %.2e"
              % (min(s.natoms for s in structs), max(s.natoms for s in structs), precision, worst_e, worst_f),
              flush=True)
        del b, p


def test_pool_large_requests():
    sizes = [1000, 2500, 4096]
    structs = [W.random_structures(1, n, n, seed=50 + k)[0] for k, n in enumerate(sizes)]
    assert [s.natoms for s in structs] == sizes
    o = O.Oracle(config_potential(POT))
    p = gpu_pair("fp64")
    pool = RequestPool(p, 9000, 9000 * 200, 8)
    hs = [pool.admit(s) for s in structs]
    assert all(h is not None for h in hs)
    res = {}
    while pool.tick():
        pool.sync()
        for h in pool.done():
            res[h] = pool.result(h)
    for h, s in zip(hs, structs):
        ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
        de = abs(res[h]["energy"] - ro["energy"]) / s.natoms
        fr = rel_force_err(res[h]["forces"], ro["forces"])
        vr = np.abs(res[h]["virial"] - ro["virial"]).max() / np.abs(ro["virial"]).max()
        print("pool request of %d atoms: dE/atom %.2e relF %.2e relV %.2e" % (s.natoms, de, fr, vr), flush=True)
        assert de <= 1e-10 and fr <= 1e-8 and vr <= 1e-8


def test_256k_structure_properties():
    p = gpu_pair("fp64")
    s = W.fcc(40, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
    rs = ResidentSystem.from_cell(p, s.cell, s.positions, s.types)
    rs.compute()
    rs.sync()
    a = rs.get()
    rs.compute()
    rs.sync()
    b = rs.get()
    assert a["energy"] == b["energy"] and np.array_equal(a["forces"], b["forces"])  # bitwise repeat
    fmax = np.abs(a["forces"]).max()
    net = np.abs(a["forces"].sum(axis=0)).max()
    print("256k: E/atom %.10f, |F|max %.3f, |sum F| %.2e" % (a["energy"] / s.natoms, fmax, net), flush=True)
    assert net <= 1e-9 * fmax
    del rs
    shift = np.array([0.37, -1.13, 2.29])
    rt = ResidentSystem.from_cell(p, s.cell, s.positions + shift, s.types)
    rt.compute()
    rt.sync()
    c = rt.get()
    de = abs(c["energy"] - a["energy"]) / s.natoms
    df = np.abs(c["forces"] - a["forces"]).max() / fmax
    print("256k rigid translation: dE/atom %.2e, relF %.2e" % (de, df), flush=True)
    assert de <= 1e-12 and df <= 1e-9
