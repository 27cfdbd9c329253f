#!/usr/bin/env python3
"""Benchmark of the gtinv energy/force path on one MI355X.

Workload (the 256k-atom configuration of the headline metric): binary random
substitutional fcc, 40x40x40 conventional cells = 256,000 atoms, a = 4.0 A,
Gaussian displacements sigma = 0.08 A (seed 4), species 50/50 (seed 4);
potential r_c = 8 A, 12 Gaussians (beta 1, centres 0..7 A), gtinv order 4
with maxl [12, 8, 2], model_type 4 / max_p 2 (quadratic in the order<=2
invariants, 239,184 coefficients, N(0, 1e-4), seed 5).  One step = one full
energy + force + virial evaluation of the whole structure with positions
and the neighbor list resident in HBM.  value = atom-evaluations per second.

Also reported on the same line:
* roofline: the dominant kernel's algorithmic FP64 rate (FLOP model of
  mlip_flops in the engine) over its average duration, measured with HIP
  events on the engine's stream during the timed steps, against the
  MI355X FP64 peak;
* cpu_baseline: the CPU oracle (FP64 restatement of the reference CPU pair
  style) timed on a bounded sub-cell of the same generator on this host.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lammps-mlip-package_amd"))

from mlip_hip import BatchSystem, PairMLIPGtinv, ResidentSystem, inspect_model, periodic_full_list  # noqa: E402
from mlip_hip import serve as SV  # noqa: E402
from mlip_hip import workloads as W  # noqa: E402

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_workload(ncell, tmpdir):
    pot = os.path.join(tmpdir, "binary_order4_quadratic.mlip")
    W.config_potential(pot, "binary_order4_quadratic", lambda probe: inspect_model(probe, 4))
    s = W.fcc(ncell, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
    return pot, s


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(pot, budget_s):
    """Oracle (FP64 restatement of the reference CPU pair style) on sub-cells
    of the same generator: a 10x10x10-cell (4,000-atom) sub-cell on all host
    threads, and a 4x4x4-cell (256-atom) sub-cell on one thread (the
    reference's CPU style scales poorly with OpenMP).  The per-atom cost does
    not depend on the cell size (a fixed neighbor count per atom)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    o = oracle_lib.Oracle(pot)

    def timed(ncell, budget):
        s = W.fcc(ncell, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
        nl = W.build_neighbor_list(s, o.cutoff, full=False)
        n_eval, t_tot = 0, 0.0
        while t_tot < budget or n_eval == 0:
            t0 = time.perf_counter()
            o.compute(nl, s.types - 1)
            t_tot += time.perf_counter() - t0
            n_eval += 1
            if n_eval >= 50:
                break
        return s.natoms, n_eval, t_tot

    L = oracle_lib.lib()
    threads = L.oracle_num_threads()
    na, n_eval, t_tot = timed(10, budget_s)
    L.oracle_set_num_threads(1)
    n1a, n1, t1 = timed(4, budget_s / 2)
    L.oracle_set_num_threads(threads)
    return {"value": na * n_eval / t_tot, "unit": "atom-evals/s", "cores": int(threads),
            "kind": "port", "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": "%d evaluation(s) of a %d-atom (10x10x10-cell) sub-cell of the same generator and model, "
                      "%.1f s of CPU work on %d threads; the per-atom cost is size-independent"
                      % (n_eval, na, t_tot, threads),
            "single_thread": {"value": n1a * n1 / t1, "cores": 1,
                              "sample": "%d evaluation(s) of a %d-atom (4x4x4-cell) sub-cell, %.1f s"
                                        % (n1, n1a, t1)}}


def cfg1_section(tmpdir, nsteps=100):
    """The 256-atom fcc configuration (4x4x4 cells, a = 4.05 A, sigma = 0.05 A,
    seed 1; order-3 [8,4] linear model, 10 Gaussians, r_c 6 A): one E/F/virial
    evaluation and a 100-step NVE run at 1 fs from a 300 K Maxwell draw
    (seed 3), on the host's cores with the oracle (neighbor list rebuilt each
    step, as the oracle's own integrator does) and on the device (request
    pool, skin 1 A, lists rebuilt when an atom moves half the skin); final
    positions and step-100 energies compared."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import md_oracle as MD
    import oracle_lib
    pot = os.path.join(tmpdir, "fcc256_order3_linear.mlip")
    W.config_potential(pot, "fcc256_order3_linear", lambda probe: inspect_model(probe, 4))
    s = W.fcc(4, 4.05, 0.05, 1)
    mass = [26.98]
    v0 = SV.maxwell(s, mass, 300.0, np.random.default_rng(3))
    o = oracle_lib.Oracle(pot)
    nl = W.build_neighbor_list(s, o.cutoff, full=False)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        o.compute(nl, s.types - 1)
        ts.append(time.perf_counter() - t0)
    t_eval_cpu = float(np.median(ts))
    t0 = time.perf_counter()
    xo, _, tho = MD.run_nve(o, s, mass, 0.001, nsteps, record=(nsteps,), v0=v0)
    t_nve_cpu = time.perf_counter() - t0
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, *W.CONFIG_MODELS["fcc256_order3_linear"]["elements"])
    rs = ResidentSystem.from_cell(p, s.cell, s.positions, s.types)
    for _ in range(3):
        rs.compute()
    rs.sync()
    t0 = time.perf_counter()
    for _ in range(20):
        rs.compute()
    rs.sync()
    t_eval_gpu = (time.perf_counter() - t0) / 20
    pool = SV.RequestPool(p, 1000, 200000, 4)
    t0 = time.perf_counter()
    h = pool.admit(s, nsteps=nsteps, dt=0.001, skin=1.0, velocities=v0)
    while pool.tick():
        pool.sync()
    t_nve_gpu = time.perf_counter() - t0
    res = pool.result(h, max_evals=nsteps + 8)
    dx = float(np.abs(MD.periodic_diff(res["positions"], xo, s.cell)).max())
    th = res["thermo"][nsteps]
    return {"workload": "256-atom fcc (4x4x4 cells, a=4.05 A, sigma=0.05 A, seed 1), order-3 [8,4] linear model, "
                        "10 Gaussians, r_c 6 A; 100 NVE steps at 1 fs from 300 K (seed 3)",
            "cpu_oracle": {"eval_s": t_eval_cpu, "atom_evals_per_s": s.natoms / t_eval_cpu, "nve_100_s": t_nve_cpu,
                           "cores": int(oracle_lib.lib().oracle_num_threads()),
                           "note": "median of 5 evaluations; the NVE time includes a host neighbor build per step"},
            "gpu": {"eval_s": t_eval_gpu, "atom_evals_per_s": s.natoms / t_eval_gpu, "nve_100_s": t_nve_gpu,
                    "list_rebuilds": pool.rebuilds(), "note": "a 256-atom evaluation is launch-latency bound"},
            "final_positions_max_diff_A": dx,
            "epair_step100_diff_per_atom": abs(th[0] - tho[nsteps]["E_pair"]) / s.natoms}


def nve_drift(pot, precision, nsteps=200, ncell=10):
    """Total-energy drift of a device-resident NVE run (4000-atom binary fcc
    of the benchmark generator, 300 K Maxwell start, 1 fs) in the request
    pool; returns the TotEng series per atom."""
    p = make_pair(pot, precision)
    s = W.fcc(ncell, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
    v = SV.maxwell(s, [47.88, 26.98], 300.0, np.random.default_rng(8))
    pool = SV.RequestPool(p, s.natoms + 64, (s.natoms + 64) * 260, 2)
    h = pool.admit(s, nsteps=nsteps, dt=0.001, skin=1.0, velocities=v)
    t0 = time.perf_counter()
    while pool.tick():
        pool.sync()
    dt = time.perf_counter() - t0
    th = pool.result(h, max_evals=nsteps + 8)["thermo"]
    etot = (th[:, 0] + 0.5 * (th[:, 7] + th[:, 8] + th[:, 9])) / s.natoms
    return etot, th[:, 0] / s.natoms, dt


def drift_section(pot, nsteps=200):
    e64, p64, t64 = nve_drift(pot, "fp64", nsteps)
    emx, pmx, tmx = nve_drift(pot, "mixed", nsteps)
    out = {"workload": "4000-atom binary fcc (10^3 cells, seed 4), same model, Maxwell 300 K (seed 8), "
                       "velocity Verlet 1 fs, %d steps in the request pool" % nsteps,
           "fp64_totE_drift_per_atom": float(e64[-1] - e64[0]),
           "fp64_totE_maxdev_per_atom": float(np.abs(e64 - e64[0]).max()),
           "mixed_totE_drift_per_atom": float(emx[-1] - emx[0]),
           "mixed_totE_maxdev_per_atom": float(np.abs(emx - emx[0]).max()),
           "mixed_vs_fp64_epair_maxdev_per_atom": float(np.abs(pmx - p64).max()),
           "unit": "eV/atom", "run_s": {"fp64": t64, "mixed": tmx}}
    log("NVE drift over %d steps: FP64 %.2e, mixed %.2e eV/atom (max dev %.2e / %.2e); mixed vs FP64 E_pair "
        "max %.2e eV/atom" % (nsteps, out["fp64_totE_drift_per_atom"], out["mixed_totE_drift_per_atom"],
                               out["fp64_totE_maxdev_per_atom"], out["mixed_totE_maxdev_per_atom"],
                               out["mixed_vs_fp64_epair_maxdev_per_atom"]))
    return out


def timed_steps(rs, warmup, steps, tag):
    """Warm up, pick the force kernel's launch shape, the a_nlm kernel and the
    order-3/4 workgroup plan with the autotuners, then time `steps`
    evaluations (HIP events per kernel on the engine's stream)."""
    for _ in range(max(warmup, 1)):
        rs.compute()
    rs.sync()
    nt = rs.autotune(reps=2)
    plan = rs.autotune_q34(reps=2)
    rs.compute()
    rs.sync()
    rs.tuning = {"force_threads": nt, "anlm_kernel": rs.pair.anlm_kernel(), "q34_plan_work_model": list(plan[:2]), "q34_paired_rows_lmax": plan[2]}
    log("%s: warmup done, force kernel %d threads per workgroup, a_nlm kernel %s, order-3/4 plan %s"
        % (tag, nt, rs.pair.anlm_kernel(), plan))
    rs.profile(True)
    rs.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        rs.compute()
    rs.sync()
    dt = time.perf_counter() - t0
    prof = rs.read_profile()
    rs.profile(False)
    return dt, prof, rs.get(), nt


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the calibrated FETCH_SIZE /
    WRITE_SIZE counter passes summarised in profiles/ (None if absent)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_256k_fp64.json")
    if not os.path.exists(path):
        return None
    data = json.load(open(path))
    k = data.get("kernels", {}).get(kernel)
    if not k:
        return None
    return {"bytes_per_launch": k["bytes"], "source": "profiles/hbm_traffic_256k_fp64.json"}


def measured_flops(kernel):
    """Executed FP64 FLOPs per launch of `kernel` from the SQ_INSTS_VALU_
    {FMA,MUL,ADD,TRANS}_F64 counter pass summarised in profiles/ (64 lanes per
    instruction, 2 per FMA), or None if absent."""
    path = os.path.join(ROOT, "profiles", "valu_flops_256k_fp64.json")
    if not os.path.exists(path):
        return None
    k = json.load(open(path)).get("kernels", {}).get(kernel)
    if not k:
        return None
    return {"flops_per_launch": k["flops"], "source": "profiles/valu_flops_256k_fp64.json"}


def make_pair(pot, precision):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, "Ti", "Al")
    p.set_precision(precision)
    return p


def batched_section(pot, steps):
    """4096 independent structures (log-uniform 8-512 atoms, seed 6) in one
    batched evaluation, replayed as a captured HIP graph; force launch shape
    from the autotuner.  Mixed precision, and the same workload in FP64 as
    its baseline; per-structure error of the mixed result against FP64."""
    structs = W.random_structures(4096, 8, 512, seed=6)
    natoms = sum(s.natoms for s in structs)
    nat = np.array([s.natoms for s in structs])
    runs = {}
    for precision in ("mixed", "fp64"):
        p = make_pair(pot, precision)
        t0 = time.time()
        b = BatchSystem(p, structs)
        t_setup = time.time() - t0
        b.compute()
        b.sync()
        nt = b.autotune(reps=2)
        b.graph(True)
        for _ in range(2):  # eager after the change, then the capture
            b.compute()
        b.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.compute()
        b.sync()
        dt = time.perf_counter() - t0
        r = b.get(forces=True)
        if not np.all(np.isfinite(r["energies"])):
            raise SystemExit("non-finite batched result")
        runs[precision] = {"dt": dt, "nt": nt, "setup": t_setup, "res": r}
        log("batched %s: %d structures, %d atoms, %.2f ms/eval, %.4g atom-evals/s (setup %.1fs)"
            % (precision, len(structs), natoms, 1e3 * dt / steps, natoms * steps / dt, t_setup))
        del b, p
    rm, r64 = runs["mixed"]["res"], runs["fp64"]["res"]
    de = np.abs(rm["energies"] - r64["energies"]) / nat
    df = np.array([np.abs(a - c).max() if a.size else 0.0 for a, c in zip(rm["forces"], r64["forces"])])
    dv = np.abs(rm["virials"] - r64["virials"]).max(axis=1)
    dm = runs["mixed"]["dt"]
    out = {"value": natoms * steps / dm, "unit": "atom-evals/s", "ms_per_eval": 1e3 * dm / steps,
           "nstruct": len(structs), "natoms": natoms, "precision": "mixed", "graph": True,
           "force_threads": runs["mixed"]["nt"], "host_setup_s": runs["mixed"]["setup"],
           "fp64": {"value": natoms * steps / runs["fp64"]["dt"], "ms_per_eval": 1e3 * runs["fp64"]["dt"] / steps,
                    "force_threads": runs["fp64"]["nt"]},
           "speedup_mixed_vs_fp64": runs["fp64"]["dt"] / dm,
           "mixed_vs_fp64_error": {
               "energy_per_atom_max": float(de.max()), "energy_per_atom_median": float(np.median(de)),
               "force_max": float(df.max()), "force_median_of_structure_max": float(np.median(df)),
               "virial_max": float(dv.max()), "structures": len(structs), "unit": "eV/atom, eV/A, eV",
               "worst_structure": int(np.argmax(df))},
           "workload": "4096 random triclinic cells, 8-512 atoms log-uniform (seed 6), 0.06 atoms/A^3, "
                       "1.8 A rejection, two species with random fraction; same model"}
    log("batched mixed vs FP64: %.2fx; per structure max dE/atom %.2e, max dF %.2e eV/A"
        % (out["speedup_mixed_vs_fp64"], de.max(), df.max()))
    return out


def trajectory_section(pot, s, nsteps=200, dt=0.001, skin=1.0, temperature=300.0):
    """The 256k-atom structure run 200 NVE steps (velocity Verlet, 1 fs,
    Maxwell start at 300 K, seed 3) twice: driven by FP64 forces, with the
    mixed-precision evaluation of every step's positions compared against
    FP64 (energy per atom, force MAE and max), and driven by mixed-precision
    forces; total-energy drift per atom of both.  Positions stay resident
    (the list is built once at r_c + skin and rebuilt if an atom moves more
    than skin / 2)."""
    p = make_pair(pot, "fp64")
    mass_t = p.masses()
    m = mass_t[s.types - 1][:, None]
    ftm2v = 1.0 / 1.0364269e-4
    v0 = SV.maxwell(s, mass_t, temperature, np.random.default_rng(3))

    def build(x):
        off, idx, sh = periodic_full_list(s.cell, x, p.init_one() + skin)
        return ResidentSystem(p, x, s.types, off, idx, sh)

    def run(drive, compare):
        x = s.positions.copy()
        v = v0.copy()
        rs = build(x)
        xref = x.copy()
        rebuilds = 0

        def evaluate(prec):
            p.set_precision(prec)
            rs.update_positions(x)
            rs.compute()
            rs.sync()
            return rs.get()

        r = evaluate(drive)
        rows, etot = [], []
        for step in range(nsteps + 1):
            if compare:
                rc = evaluate("mixed")
                df = np.abs(rc["forces"] - r["forces"])
                rows.append((abs(rc["energy"] - r["energy"]) / s.natoms, float(df.mean()), float(df.max())))
            ke = 0.5 * float((m * v * v).sum()) / ftm2v
            etot.append((r["energy"] + ke) / s.natoms)
            if step == nsteps:
                break
            v += 0.5 * dt * ftm2v * r["forces"] / m
            x += dt * v
            if np.sqrt(((x - xref) ** 2).sum(axis=1).max()) > 0.5 * skin:
                rs = build(x)
                xref = x.copy()
                rebuilds += 1
            r = evaluate(drive)
            v += 0.5 * dt * ftm2v * r["forces"] / m
        p.set_precision("fp64")
        return np.array(rows), np.array(etot), rebuilds

    t0 = time.perf_counter()
    rows, e64, rb64 = run("fp64", True)
    t1 = time.perf_counter()
    _, emx, rbmx = run("mixed", False)
    t2 = time.perf_counter()
    out = {"workload": "the 256,000-atom structure, velocity Verlet %d steps at %.0f fs from a %.0f K Maxwell draw "
                       "(seed 3), neighbor list at r_c + %.1f A" % (nsteps, dt * 1e3, temperature, skin),
           "mixed_vs_fp64_per_step": {"energy_per_atom_max": float(rows[:, 0].max()),
                                      "force_mae_max": float(rows[:, 1].max()),
                                      "force_max": float(rows[:, 2].max()),
                                      "energy_per_atom": [float("%.3g" % v) for v in rows[:, 0]],
                                      "force_mae": [float("%.3g" % v) for v in rows[:, 1]],
                                      "force_max_by_step": [float("%.3g" % v) for v in rows[:, 2]],
                                      "unit": "eV/atom, eV/A"},
           "fp64_totE_drift_per_atom": float(e64[-1] - e64[0]),
           "fp64_totE_maxdev_per_atom": float(np.abs(e64 - e64[0]).max()),
           "mixed_totE_drift_per_atom": float(emx[-1] - emx[0]),
           "mixed_totE_maxdev_per_atom": float(np.abs(emx - emx[0]).max()),
           "list_rebuilds": {"fp64": rb64, "mixed": rbmx}, "run_s": {"fp64_with_mixed_checks": t1 - t0,
                                                                  "mixed": t2 - t1}}
    log("256k trajectory: mixed vs FP64 over %d steps: dE/atom max %.2e, F MAE max %.2e, F max %.2e eV/A; "
        "drift FP64 %.2e, mixed %.2e eV/atom" % (nsteps, rows[:, 0].max(), rows[:, 1].max(), rows[:, 2].max(),
                                                  out["fp64_totE_drift_per_atom"], out["mixed_totE_drift_per_atom"]))
    del p
    return out


def train_section(nstruct, reps=3):
    """Training outputs (structural, polynomial, force and stress features)
    of a batch of small random cells with the fitted order-4 cubic potential
    of the reference's regression tests (gtinv-820: 3470 features, 6695
    coefficients), copied back to host; the CPU oracle on a few of the same
    structures for scale."""
    pot = os.path.join(ROOT, "tests", "golden", "reference", "gtinv-820.mlip")
    p = make_pair(pot, "fp64")
    structs = W.random_structures(nstruct, 8, 48, seed=13)
    natoms = sum(s.natoms for s in structs)
    b = BatchSystem(p, structs)
    b.train_features()
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        g = b.train_features()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    if not np.all(np.isfinite(g["FF"])):
        raise SystemExit("non-finite training features")
    out_bytes = sum(v.nbytes for v in g.values() if v is not None)
    del g, b
    cpu = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        o = oracle_lib.Oracle(pot)
        t0 = time.perf_counter()
        k = 0
        while k < min(8, nstruct) and time.perf_counter() - t0 < 10.0:
            s = structs[k]
            o.train(W.build_neighbor_list(s, o.cutoff), s.types - 1)
            k += 1
        cpu = {"structures_per_s": k / (time.perf_counter() - t0), "sample": "first %d structures" % k,
               "cores": oracle_lib.lib().oracle_num_threads()}
    except Exception as e:  # the oracle is test infrastructure; report its absence
        cpu = {"error": str(e)}
    out = {"structures_per_s": nstruct / best, "atoms_per_s": natoms / best, "ms_per_batch": 1e3 * best,
           "nstruct": nstruct, "natoms": natoms, "output_gb": out_bytes / 1e9, "includes_copy_to_host": True,
           "workload": "%d random triclinic cells, 8-48 atoms log-uniform (seed 13); gtinv-820 potential "
                       "(order 4, model_type 3, max_p 3)" % nstruct,
           "cpu_baseline": cpu}
    log("train: %d structures (%d atoms) in %.1f ms -> %.4g structures/s (%.2f GB of features to host); "
        "CPU oracle %s" % (nstruct, natoms, 1e3 * best, out["structures_per_s"], out["output_gb"],
                           cpu.get("structures_per_s") if cpu else None))
    return out


def served_section(pot, rates, windows, budget_gb):
    """Poisson request streams (seed 7) through the continuous-batching pool
    in mixed precision, and the same streams one request at a time in FP64."""
    masses = [47.88, 26.98]
    p = make_pair(pot, "mixed")
    pool, cap = SV.pool_for_budget(p, int(budget_gb * (1 << 30)), max_requests=8192)
    SV.serve(pool, SV.make_stream(1000, 0.05, 1, masses))  # warm-up stream
    q = make_pair(pot, "fp64")
    seq_pool = SV.RequestPool(q, 8192, 8192 * 200, 4)
    SV.serve_sequential(seq_pool, SV.make_stream(100, 0.05, 1, masses))
    out = {"pool_atoms": cap, "hbm_budget_gb": budget_gb, "precision": "mixed", "rates": {}}
    for rate, win in zip(rates, windows):
        stream = SV.make_stream(rate, win, 7, masses)
        r = SV.serve(pool, stream)
        seq_stream = SV.make_stream(rate, win, 7, masses)
        seq = SV.serve_sequential(seq_pool, seq_stream)
        r["arrival_window_s"] = win
        r["fp64_sequential"] = seq
        r["mixed_vs_fp64_error"] = SV.stream_errors(stream, seq_stream)
        # offered load: the stream's atom-evaluations over its arrival window;
        # a makespan well past the window means the queue grew (saturation)
        r["offered_atom_evals_per_s"] = r["atom_evals"] / win
        r["regime"] = "saturated" if r["makespan_s"] > 1.2 * win else "below saturation"
        out["rates"][str(rate)] = r
        log("served %d req/s (%d requests): p50 %.1f ms p99 %.1f ms, %.4g atom-evals/s; "
            "FP64 one-at-a-time: p50 %.1f ms p99 %.1f ms, %.4g atom-evals/s"
            % (rate, len(stream), r["latency_p50_ms"], r["latency_p99_ms"], r["atom_evals_per_s"],
               seq["latency_p50_ms"], seq["latency_p99_ms"], seq["atom_evals_per_s"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ncell", type=int, default=40, help="fcc cells per edge (40 -> 256,000 atoms)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-oracle work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision timing")
    ap.add_argument("--no-batched", action="store_true", help="skip the 4096-structure batched evaluation")
    ap.add_argument("--no-served", action="store_true", help="skip the served request streams")
    ap.add_argument("--no-train", action="store_true", help="skip the training-feature batch")
    ap.add_argument("--no-trajectory", action="store_true", help="skip the 256k-atom NVE error trajectory")
    ap.add_argument("--trajectory-steps", type=int, default=200)
    ap.add_argument("--train-structures", type=int, default=256)
    ap.add_argument("--serve-windows", type=str, default="10.0,1.0,0.5",
                    help="arrival windows (s) for the 200/1000/5000 req/s streams (200 req/s is below saturation: "
                         "a steady state; the others saturate the device)")
    ap.add_argument("--pool-gb", type=float, default=16.0, help="HBM budget of the request pool")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also launch a plain copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: evaluation runs on one GPU; --gpus %d is not supported "
              "(a multi-GPU run would shard the structure by spatial domain)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    tmp = tempfile.mkdtemp(prefix="mlip_bench_")
    pot, s = make_workload(args.ncell, tmp)
    pair = PairMLIPGtinv()
    pair.settings()
    pair.coeff("*", "*", pot, "Ti", "Al")
    pair.set_precision(args.precision)
    log("model %s" % pair.model_info())
    t0 = time.perf_counter()
    off, idx, sh = periodic_full_list(s.cell, s.positions, pair.init_one())
    t_host = time.perf_counter() - t0
    host_pairs = idx.size
    del off, idx, sh
    # the resident structure's lists come from the device builder (identical
    # to the host builder's); the second build reuses the allocations, as a
    # re-neighboring step of a run would
    t0 = time.perf_counter()
    rs = ResidentSystem.from_cell(pair, s.cell, s.positions, s.types)
    t_dev_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    rs = ResidentSystem.from_cell(pair, s.cell, s.positions, s.types)
    t_dev = time.perf_counter() - t0
    if rs.npairs != host_pairs:
        raise SystemExit("device and host neighbor lists differ in size")
    neighbor = {"pairs": rs.npairs, "pairs_per_atom": rs.npairs / s.natoms, "device_build_s": t_dev,
                "device_first_build_s": t_dev_first, "host_build_s": t_host, "host_threads": 1,
                "note": "device time: positions upload, bin sort, count/scan/fill, gather-list sort and merge, "
                        "slot layout; identical lists to the host builder (tests/test_gpu_neighbor.py)"}
    log("neighbor list: %d atoms, %d pairs (%.1f per atom): device %.3fs (first %.2fs), host %.2fs"
        % (s.natoms, rs.npairs, rs.npairs / s.natoms, t_dev, t_dev_first, t_host))
    dt, prof, res, nt = timed_steps(rs, args.warmup, args.steps, args.precision)
    tuning = dict(rs.tuning)
    if not np.isfinite(res["energy"]) or not np.all(np.isfinite(res["forces"])):
        raise SystemExit("non-finite result")
    flops = rs.flops()
    flops_ref = rs.flops_reference()
    mixed = None
    if args.precision == "fp64" and not args.no_mixed:
        pair.set_precision("mixed")
        dtm, profm, resm, ntm = timed_steps(rs, args.warmup, args.steps, "mixed")
        pair.set_precision("fp64")
        df = np.abs(resm["forces"] - res["forces"])
        mixed = {"value": s.natoms * args.steps / dtm, "ms_per_step": 1e3 * dtm / args.steps,
                 "speedup_vs_fp64": dt / dtm, "tuning": dict(rs.tuning),
                 "energy_err_per_atom": abs(resm["energy"] - res["energy"]) / s.natoms,
                 "force_err_max": float(df.max()), "force_mae": float(df.mean()),
                 "reference": "FP64 GPU result of the same step (itself within 1e-10 eV/atom of the CPU oracle)",
                 "kernels_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in profm.items()}}
        log("mixed: %.1f ms/step (%.2fx FP64), dE/atom %.2e, F max %.2e MAE %.2e eV/A"
            % (mixed["ms_per_step"], mixed["speedup_vs_fp64"], mixed["energy_err_per_atom"],
               mixed["force_err_max"], mixed["force_mae"]))
        mixed["nve_drift"] = drift_section(pot)
    if args.calibrate_copy_mib > 0:
        rs.calibration_copy(args.calibrate_copy_mib << 20, reps=3)

    npairs = rs.npairs
    del rs
    if mixed is not None and not args.no_trajectory:
        mixed["trajectory_256k"] = trajectory_section(pot, s, nsteps=args.trajectory_steps)
    batched = None if args.no_batched else batched_section(pot, max(args.steps, 3))
    train = None if args.no_train else train_section(args.train_structures)
    served = None
    if not args.no_served:
        served = served_section(pot, [200, 1000, 5000], [float(w) for w in args.serve_windows.split(",")],
                                args.pool_gb)

    natoms = s.natoms
    value = natoms * args.steps / dt
    kernels = {k: {"avg_ms": v[0] / max(v[1], 1), "launches": v[1]} for k, v in prof.items()}
    dom = max(prof.items(), key=lambda kv: kv[1][0])[0]
    fl_key = {"q34_adj": "q34_adj", "q34_feat": "q34_feat", "anlm": "anlm", "force": "force",
              "poly_gemm": "poly", "q12_adj": "q12", "q12_feat": "q12"}.get(dom)
    dom_ms = kernels[dom]["avg_ms"]
    ref_fl = flops_ref[fl_key] if fl_key else None
    exe_fl = flops[fl_key] if fl_key else None
    achieved = exe_fl / (dom_ms * 1e-3) / 1e12 if fl_key else None
    traffic = measured_traffic(dom)
    counted = measured_flops(dom)
    total_rate = flops["total"] * args.steps / dt / 1e12
    roofline = {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": (achieved / FP64_PEAK_TFLOPS) if achieved else None,
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "kernel": dom, "kernel_avg_ms": dom_ms,
                "executed_flops_per_launch": exe_fl,
                "counter_flops_per_launch": counted["flops_per_launch"] if counted else None,
                "counter_vs_model": (counted["flops_per_launch"] / exe_fl) if (counted and exe_fl) else None,
                "flops_source": counted["source"] if counted else None,
                "reference_rule_flops_per_launch": ref_fl,
                "reference_rule_tflops": ref_fl / (dom_ms * 1e-3) / 1e12 if fl_key else None,
                "step_executed_tflops": total_rate, "step_frac": total_rate / FP64_PEAK_TFLOPS,
                "traffic_source": traffic["source"] if traffic else None,
                "rule": "achieved = FP64 operations the kernel executes (engine count mlip_flops: add/mul 1, "
                        "FMA 2; checked against the SQ_INSTS_VALU_*_F64 counters of the same launch) / its "
                        "average HIP-event duration; the reference-formulation count (per half pair 40 + "
                        "12 n_fn + 38 n_lmh + 52 n_fn n_lmh; per atom n_fn sum T(6(q-1)+3) + n_fn sum "
                        "T q(6(q-1)+4) + 2 sum|p|) is %.1fx larger because the kernel pairs m with -m, shares "
                        "partial products, merges factor positions with identical signatures and takes the "
                        "energy of linear families from the adjoint by Euler's theorem"
                        % ((ref_fl / exe_fl) if (fl_key and exe_fl) else float("nan")),
                "note": "FP64 VALU kernel (the MI355X FP64 peak is 78.6 TFLOP/s for vector and matrix units)"}
    log("steps %d in %.3fs -> %.1f ms/step, %.4g atom-evals/s; whole-step FP64 rate %.2f TFLOP/s"
        % (args.steps, dt, 1e3 * dt / args.steps, value, total_rate))
    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["avg_ms"]):
        log("  %-14s %9.3f ms/launch" % (k, v["avg_ms"]))
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU oracle (bounded sample)")
        cpu = cpu_baseline(pot, args.cpu_budget)
    cfg1 = None
    if not args.no_cpu_baseline:
        log("256-atom order-3 configuration: CPU oracle and device, one evaluation and 100 NVE steps")
        cfg1 = cfg1_section(tmp)
        log("  oracle %.3g atom-evals/s (%d threads), NVE 100 steps %.1f s; device %.3g atom-evals/s, NVE %.2f s "
            "(%d list rebuilds); final positions differ by %.2e A"
            % (cfg1["cpu_oracle"]["atom_evals_per_s"], cfg1["cpu_oracle"]["cores"], cfg1["cpu_oracle"]["nve_100_s"],
               cfg1["gpu"]["atom_evals_per_s"], cfg1["gpu"]["nve_100_s"], cfg1["gpu"]["list_rebuilds"],
               cfg1["final_positions_max_diff_A"]))
    line = {
        "metric": "atom-evals/s (E+F) FP64 vs mixed at 256k atoms; p50/p99 latency served",
        "value": value,
        "unit": "atom-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32/f64 mixed",
        "data": "synthetic (seeded fcc structure and random potential coefficients)",
        "config": {"workload": "binary fcc %dx%dx%d cells = %d atoms, a=4.0 A, sigma=0.08 A; gtinv order 4 "
                               "[12,8,2], model_type 4 max_p 2 (239,184 coeffs), r_c 8 A, 12 Gaussians; "
                               "one E+F+virial evaluation per step, inputs resident in HBM"
                               % (args.ncell, args.ncell, args.ncell, natoms),
                   "natoms": natoms, "pairs_per_atom": npairs / natoms, "precision": args.precision,
                   "tuning": tuning},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fcc256_order3": cfg1,
        "mixed": mixed,
        "batched": batched,
        "served": served,
        "neighbor_build": neighbor,
        "train": train,
        "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in kernels.items()},
        "flops_per_step": flops,
        "flops_reference_per_step": flops_ref,
        "step_tflops": total_rate,
        "energy_per_atom": res["energy"] / natoms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
