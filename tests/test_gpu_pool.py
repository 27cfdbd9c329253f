"""Request pool (continuous batching, device-resident NVE) through the C ABI.

Bars: NVE trajectories against the reference's regression logs (8 printed
digits) and dumps (1e-8 A) and against the CPU-oracle integrator (1e-9 A);
single evaluations equal the same structure evaluated alone to 1e-12
(relative); a request gives the same numbers alone and in a mixed batch to
1e-12 (relative); mixed precision within 1e-6 eV/atom and 1e-4 eV/A of the
FP64 oracle."""
import json
import os

import numpy as np
import pytest

import md_oracle as MD
import oracle_lib as O
from common import GOLD, config_potential, ref, rel_force_err
from mlip_hip import BatchSystem, PairMLIPGtinv, RequestPool
from mlip_hip import md as MDP
from mlip_hip import workloads as W
from test_md_oracle import check_thermo, potential_masses

pytestmark = pytest.mark.gpu


def gpu_pair(pot, elems, precision="fp64"):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, *elems)
    p.set_precision(precision)
    return p


def goldens():
    return json.load(open(os.path.join(GOLD, "regression_goldens.json")))


def drain(pool, max_ticks=10000):
    out = {}
    for _ in range(max_ticks):
        if pool.tick() == 0:
            break
        pool.sync()
        for h in pool.done():
            out[h] = pool.result(h)
    return out


def thermo_of(row, s):
    # the package's thermo output (mlip_hip.md), checked here against the
    # reference's logs and the oracle integrator
    return MDP.thermo(s.natoms, s.volume, row)


@pytest.mark.parametrize("name", ["gtinv-197", "gtinv-434", "gtinv-820"])
def test_pool_nve_regression_runs(name, tmp_path):
    g = goldens()[name]
    s = W.read_lammps_data(ref(g["structure"]))
    p = gpu_pair(ref(g["potential"]), g["elements"])
    pool = RequestPool(p, 1000, 200000, 4)
    h = pool.admit(s, nsteps=100, dt=0.001, skin=2.0)
    res = drain(pool)[h]
    th = res["thermo"]
    assert th.shape == (101, 13)
    disp = MD.periodic_diff(res["positions"], s.positions, s.cell)
    assert np.linalg.norm(disp, axis=1).max() < 1.0  # within half the skin
    check_thermo(thermo_of(th[0], s), {k: g["step0"][k] for k in ("E_pair", "Press")}, name + " step 0")
    check_thermo(thermo_of(th[100], s), g["step100_dt0.001"], name + " step 100")
    d = MD.periodic_diff(res["positions"], MD.read_dump(ref(g["dump100_dt0.001"])), s.cell)
    print("%s: pool NVE max |x - dump| = %.2e A" % (name, np.abs(d).max()), flush=True)
    assert np.abs(d).max() < 1e-8
    # the package's dump writer reproduces the reference's dump file
    out = os.path.join(str(tmp_path), "dump.atom")
    MDP.write_dump(out, 100, s, res["positions"])
    mine, theirs = open(out).read().split("\n"), open(ref(g["dump100_dt0.001"])).read().split("\n")
    assert mine[:9] == theirs[:9]
    assert np.abs(MD.read_dump(out) - MD.read_dump(ref(g["dump100_dt0.001"]))).max() < 1e-8
    log = MDP.thermo_log(th, s, every=50)
    assert len(log.strip().split("\n")) == 4
    # and against the oracle integrator, step by step in energy
    o = O.Oracle(ref(g["potential"]))
    xo, vo, tho = MD.run_nve(o, s, potential_masses(ref(g["potential"])), 0.001, 100, record=(25, 100))
    assert np.abs(MD.periodic_diff(res["positions"], xo, s.cell)).max() < 1e-9
    assert np.abs(res["velocities"] - vo).max() < 1e-8
    assert abs(thermo_of(th[25], s)["E_pair"] - tho[25]["E_pair"]) < 1e-9


def test_pool_nve_stress_run():
    g = goldens()["gtinv-197"]
    s = W.read_lammps_data(ref(g["structure"]))
    p = gpu_pair(ref(g["potential"]), g["elements"])
    pool = RequestPool(p, 1000, 400000, 4)
    h = pool.admit(s, nsteps=100, dt=0.005, skin=3.0)
    res = drain(pool)[h]
    disp = MD.periodic_diff(res["positions"], s.positions, s.cell)
    assert np.linalg.norm(disp, axis=1).max() < 1.5
    check_thermo(thermo_of(res["thermo"][50], s), g["step50_dt0.005"], "step 50")
    check_thermo(thermo_of(res["thermo"][100], s), g["step100_dt0.005"], "step 100")


def test_pool_single_requests_equal_alone():
    pot = ref("gtinv-820.mlip")
    structs = W.random_structures(10, 8, 160, seed=17)
    p = gpu_pair(pot, ["Ti", "Al"])
    pool = RequestPool(p, 4000, 2000000, 32)
    hs = [pool.admit(s) for s in structs]
    assert all(h is not None for h in hs)
    res = drain(pool)
    o = O.Oracle(pot)
    for k, (h, s) in enumerate(zip(hs, structs)):
        r = res[h]
        q = gpu_pair(pot, ["Ti", "Al"])
        b = BatchSystem(q, [s])
        b.compute()
        b.sync()
        a = b.get()
        e_rel = abs(r["energy"] - a["energies"][0]) / abs(a["energies"][0])
        f_rel = rel_force_err(r["forces"], a["forces"][0])
        v_rel = np.abs(r["virial"] - a["virials"][0]).max() / np.abs(a["virials"][0]).max()
        assert e_rel <= 1e-12 and f_rel <= 1e-12 and v_rel <= 1e-12, (k, e_rel, f_rel, v_rel)
        if k < 3:
            ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
            assert abs(r["energy"] - ro["energy"]) / s.natoms <= 1e-10
            assert rel_force_err(r["forces"], ro["forces"]) <= 1e-8


def test_pool_request_alone_equals_in_mix():
    pot = config_potential("binary_order4_quadratic")
    structs = W.random_structures(9, 16, 120, seed=7)
    target = structs[0]
    rng = np.random.default_rng(3)
    v0 = rng.normal(0.0, 5.0, target.positions.shape)
    # alone
    p = gpu_pair(pot, ["Ti", "Al"])
    pool = RequestPool(p, 4000, 2000000, 32)
    h = pool.admit(target, nsteps=10, dt=0.001, skin=1.0, velocities=v0)
    alone = drain(pool)[h]
    # in a mix: others admitted before and while the target runs
    p2 = gpu_pair(pot, ["Ti", "Al"])
    pool2 = RequestPool(p2, 4000, 2000000, 32)
    pool2.admit(structs[1], nsteps=3, dt=0.001, skin=1.0)
    pool2.admit(structs[2])
    pool2.tick()
    pool2.sync()
    h2 = pool2.admit(target, nsteps=10, dt=0.001, skin=1.0, velocities=v0)
    res = {}
    for k in range(40):
        if k < 6:
            pool2.admit(structs[3 + k], nsteps=(k % 3) * 2)
        if pool2.tick() == 0:
            break
        pool2.sync()
        for hh in pool2.done():
            res[hh] = pool2.result(hh)
    mix = res[h2]
    assert mix["thermo"].shape == alone["thermo"].shape == (11, 13)
    scale = np.abs(alone["thermo"]).max(axis=0) + 1e-300
    assert (np.abs(mix["thermo"] - alone["thermo"]) / scale).max() <= 1e-12
    assert np.abs(mix["positions"] - alone["positions"]).max() <= 1e-12
    assert rel_force_err(mix["forces"], alone["forces"]) <= 1e-12


def test_pool_admission_by_capacity():
    pot = ref("gtinv-434.mlip")
    structs = W.random_structures(3, 40, 60, seed=5)
    p = gpu_pair(pot, ["Ti", "Al"])
    cap = structs[0].natoms + structs[1].natoms - 1
    pool = RequestPool(p, cap, 1000000, 8)
    assert RequestPool.bytes_for(p, cap, 1000000, 8) > 0
    h0 = pool.admit(structs[0], nsteps=2)
    assert h0 is not None
    assert pool.admit(structs[1]) is None  # does not fit next to the first
    st = pool.stats()
    assert st["atoms_used"] == structs[0].natoms and st["running"] == 1
    res = drain(pool)
    assert h0 in res and pool.stats()["atoms_used"] == 0
    h1 = pool.admit(structs[1])
    assert h1 is not None
    assert h1 in drain(pool)


def test_pool_mixed_precision():
    """Served-workload model (order 4, quadratic, coefficients N(0, 1e-4)) on
    random cells: the stated bar.  A fitted potential (gtinv-820) on the same
    kind of cells, whose 1.8 A contacts give forces of tens of eV/A: energy
    within 1e-6 eV/atom, forces within 1e-5 of the largest component."""
    structs = W.random_structures(5, 16, 100, seed=23)
    for pot, elems, fixed in [(config_potential("binary_order4_quadratic"), ["Ti", "Al"], True),
                              (ref("gtinv-820.mlip"), ["Ti", "Al"], False)]:
        p = gpu_pair(pot, elems, "mixed")
        pool = RequestPool(p, 4000, 2000000, 16)
        hs = [pool.admit(s) for s in structs]
        res = drain(pool)
        o = O.Oracle(pot)
        for h, s in zip(hs, structs):
            ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
            de = abs(res[h]["energy"] - ro["energy"]) / s.natoms
            df = np.abs(res[h]["forces"] - ro["forces"]).max()
            fmax = np.abs(ro["forces"]).max()
            print("mixed pool %s %d atoms: dE/atom %.2e F max err %.2e (|F|max %.1f)"
                  % (os.path.basename(pot), s.natoms, de, df, fmax), flush=True)
            assert de <= 1e-6
            assert df <= (1e-4 if fixed else 1e-5 * fmax)


def test_pool_admit_many_prefix_and_results():
    pot = ref("gtinv-434.mlip")
    structs = W.random_structures(8, 20, 80, seed=29)
    p = gpu_pair(pot, ["Ti", "Al"])
    cap = sum(s.natoms for s in structs[:5])
    pool = RequestPool(p, cap, 2000000, 32)
    hs = pool.admit_many(structs, [0] * 8, nthreads=4)
    assert len(hs) == 5  # FIFO prefix that fits
    res = drain(pool)
    q = gpu_pair(pot, ["Ti", "Al"])
    pool1 = RequestPool(q, cap, 2000000, 32)
    for h, s in zip(hs, structs):
        h1 = pool1.admit(s)
        r1 = drain(pool1)[h1]
        assert res[h]["energy"] == pytest.approx(r1["energy"], rel=1e-12)
        assert rel_force_err(res[h]["forces"], r1["forces"]) <= 1e-12


def test_serving_loop_small_stream():
    from mlip_hip import serve as SV
    pot = config_potential("binary_order4_quadratic")
    p = gpu_pair(pot, ["Ti", "Al"], "mixed")
    stream = SV.make_stream(200, 0.25, 7, [47.88, 26.98], nmin=16, nmax=128)
    pool = RequestPool(p, 20000, 4000000, 256)
    out = SV.serve(pool, stream)
    print("serve:", out, flush=True)
    assert out["completed"] == len(stream) and out["unfinished"] == 0
    assert out["latency_p99_ms"] >= out["latency_p50_ms"] > 0


def test_pool_nve_rebuilds_lists():
    """200 NVE steps at 300 K (Maxwell draw, seed 3) with a 0.3 A skin: the
    pool rebuilds the request's list whenever an atom would pass skin / 2 and
    matches the oracle integrator, whose list is rebuilt exactly every step,
    to 1e-9 A; an NVE request without a skin is refused."""
    from mlip_hip import serve as SV
    from mlip_hip._lib import MlipError
    g = goldens()["gtinv-820"]
    pot = ref(g["potential"])
    s = W.read_lammps_data(ref(g["structure"]))
    masses = potential_masses(pot)
    v0 = SV.maxwell(s, masses, 300.0, np.random.default_rng(3))
    p = gpu_pair(pot, g["elements"])
    pool = RequestPool(p, 1000, 200000, 4)
    with pytest.raises(MlipError):
        pool.admit(s, nsteps=10, dt=0.001, skin=0.0, velocities=v0)
    h = pool.admit(s, nsteps=200, dt=0.001, skin=0.3, velocities=v0)
    res = drain(pool)[h]
    nreb = pool.rebuilds()
    o = O.Oracle(pot)
    xo, vo, tho = MD.run_nve(o, s, masses, 0.001, 200, record=(50, 100, 200), v0=v0)
    dx = np.abs(MD.periodic_diff(res["positions"], xo, s.cell)).max()
    moved = np.linalg.norm(MD.periodic_diff(res["positions"], s.positions, s.cell), axis=1).max()
    print("pool NVE 200 steps, skin 0.3 A: %d list rebuilds, largest move %.3f A, max |x - oracle| = %.2e A"
          % (nreb, moved, dx), flush=True)
    assert nreb >= 1
    assert dx < 1e-9
    assert np.abs(res["velocities"] - vo).max() < 1e-8
    for k in (50, 100, 200):
        assert abs(thermo_of(res["thermo"][k], s)["E_pair"] - tho[k]["E_pair"]) < 1e-9


def test_cfg1_nve_300k_against_oracle(tmp_path):
    """The 256-atom fcc configuration (a = 4.05 A, sigma = 0.05 A, seed 1;
    order-3 linear model) run 100 NVE steps at 1 fs from a 300 K Maxwell
    draw (seed 3) in the pool, against the oracle integrator (lists rebuilt
    every step): positions within 1e-9 A, thermo within 1e-9 eV; the run's
    LAMMPS-format thermo block and dump come from the package."""
    from mlip_hip import serve as SV
    pot = config_potential("fcc256_order3_linear")
    s = W.fcc(4, 4.05, 0.05, 1)
    masses = potential_masses(pot)
    v0 = SV.maxwell(s, masses, 300.0, np.random.default_rng(3))
    p = gpu_pair(pot, ["X"])
    pool = RequestPool(p, 1000, 200000, 4)
    h = pool.admit(s, nsteps=100, dt=0.001, skin=1.0, velocities=v0)
    res = drain(pool)[h]
    o = O.Oracle(pot)
    xo, vo, tho = MD.run_nve(o, s, masses, 0.001, 100, record=(0, 50, 100), v0=v0)
    dx = np.abs(MD.periodic_diff(res["positions"], xo, s.cell)).max()
    print("cfg1 NVE 100 steps from 300 K: max |x - oracle| = %.2e A, %d list rebuilds" % (dx, pool.rebuilds()))
    assert dx < 1e-9
    for k in (0, 50, 100):
        t = thermo_of(res["thermo"][k], s)
        assert abs(t["E_pair"] - tho[k]["E_pair"]) < 1e-9 and abs(t["TotEng"] - tho[k]["TotEng"]) < 1e-9
        assert abs(t["Temp"] - tho[k]["Temp"]) < 1e-6 * max(1.0, tho[k]["Temp"])
    log = MDP.thermo_log(res["thermo"], s, every=50)
    print(log)
    MDP.write_dump(os.path.join(str(tmp_path), "dump.atom"), 100, s, res["positions"])
    assert np.abs(MD.read_dump(os.path.join(str(tmp_path), "dump.atom")) - res["positions"]).max() < 1e-12
