"""Serving loop over the device request pool: continuous batching of
Poisson-arriving requests (single energy/force/stress evaluations and short
NVE runs), admission by pool capacity, latency statistics.

One loop iteration: enqueue one batched tick of everything resident, then
move arrived requests into the FIFO queue and admit from its head while the
pool has room (host work overlaps the device; they join the next tick), wait
for the tick and retire finished requests.  An idle pool admits first.  Latency = completion time minus
arrival time, both on the host clock."""
import time
from collections import deque

import numpy as np

from . import workloads as W
from .pair import RequestPool

BOLTZ = 8.617343e-5
MVV2E = 1.0364269e-4


class Request:
    __slots__ = ("t_arrival", "structure", "nsteps", "v0", "handle", "t_done", "evals", "energy", "forces")

    def __init__(self, t_arrival, structure, nsteps, v0):
        self.t_arrival = t_arrival
        self.structure = structure
        self.nsteps = nsteps
        self.v0 = v0
        self.handle = None
        self.t_done = None
        self.evals = 0
        self.energy = None
        self.forces = None


def maxwell(s, mass_of_type, temperature, rng):
    """Velocities (A/ps) from a Maxwell distribution at `temperature`, zero
    net momentum."""
    m = np.asarray(mass_of_type, dtype=np.float64)[s.types - 1]
    sig = np.sqrt(BOLTZ * temperature / (m * MVV2E))
    v = rng.normal(size=(s.natoms, 3)) * sig[:, None]
    v -= (m[:, None] * v).sum(axis=0) / m.sum()
    return v


def make_stream(rate, duration, seed, mass_of_type, nmin=16, nmax=4096, frac_md=0.3, nsteps=10,
                temperature=300.0):
    """Poisson arrivals at `rate` per second over `duration` seconds; sizes
    log-uniform in [nmin, nmax] from the batched-workload generator (random
    triclinic cells at 0.06 atoms/A^3, 1.8 A rejection, two species);
    a fraction `frac_md` are `nsteps`-step NVE runs from a Maxwell draw."""
    rng = np.random.default_rng(seed)
    t, times = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        times.append(t)
    structs = W.random_structures(len(times), nmin, nmax, seed) if times else []
    out = []
    for ta, s in zip(times, structs):
        md = rng.random() < frac_md
        v0 = maxwell(s, mass_of_type, temperature, rng) if md else None
        out.append(Request(ta, s, nsteps if md else 0, v0))
    return out


def serve(pool, stream, dt=0.001, skin=0.5, drain_timeout=120.0, nthreads=16, max_admit=512, device_time=True):
    """Runs the stream through the pool in real time.  Returns per-request
    latencies (s) and the run's totals; with device_time, the evaluation
    kernels' device time (HIP events around every launch) and its share of
    the makespan."""
    if device_time:
        pool.profile(True)
    t0 = time.perf_counter()
    pending = deque(sorted(stream, key=lambda r: r.t_arrival))
    queue = deque()
    by_handle = {}
    done = []
    atom_evals = 0
    ticks = 0
    batch_sizes = []
    host = {"admit_s": 0.0, "tick_s": 0.0, "sync_s": 0.0, "results_s": 0.0}

    def now():
        return time.perf_counter() - t0

    def admit():
        ta = time.perf_counter()
        _admit()
        host["admit_s"] += time.perf_counter() - ta

    def _admit():
        tn = now()
        while pending and pending[0].t_arrival <= tn:
            queue.append(pending.popleft())
        if not queue:
            return
        # only the FIFO prefix that fits the free atom capacity is prepared
        st = pool.stats()
        free = st["atom_capacity"] - st["atoms_used"]
        cand = []
        for r in queue:
            if len(cand) >= max_admit or r.structure.natoms > free:
                break
            free -= r.structure.natoms
            cand.append(r)
        if not cand:
            return
        hs = pool.admit_many([r.structure for r in cand], [r.nsteps for r in cand], dt=dt,
                             skins=[skin if r.nsteps > 0 else 0.0 for r in cand],
                             velocities=[r.v0 for r in cand], nthreads=nthreads)
        for r, h in zip(cand, hs):
            r.handle = h
            by_handle[h] = r
            queue.popleft()

    while pending or queue or by_handle:
        if now() > drain_timeout:
            break
        if not by_handle:
            # idle device: admit what has arrived, or wait for the next arrival
            admit()
            if not by_handle:
                if pending:
                    wait = pending[0].t_arrival - now()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
        tt = time.perf_counter()
        n = pool.tick()
        host["tick_s"] += time.perf_counter() - tt
        if n == 0:
            continue
        ticks += 1
        batch_sizes.append(n)
        admit()  # overlaps the tick on the device
        ts = time.perf_counter()
        pool.sync()
        tn = now()
        tr = time.perf_counter()
        host["sync_s"] += tr - ts
        for h in pool.done():
            res = pool.result(h)
            r = by_handle.pop(h)
            r.t_done = tn
            r.evals = res["thermo"].shape[0]
            r.energy = res["energy"]
            r.forces = res["forces"]
            atom_evals += r.evals * r.structure.natoms
            done.append(r)
        host["results_s"] += time.perf_counter() - tr
    makespan = now()
    dev_ms = None
    if device_time:
        dev_ms = sum(v[0] for v in pool.read_profile().values())
        pool.profile(False)
    lat = np.array([r.t_done - r.t_arrival for r in done]) if done else np.zeros(0)
    return {
        "completed": len(done),
        "offered": len(stream),
        "unfinished": len(stream) - len(done),
        "device_kernel_s": dev_ms / 1e3 if dev_ms is not None else None,
        "device_busy_fraction": dev_ms / 1e3 / makespan if (dev_ms is not None and makespan > 0) else None,
        "latency_p90_ms": float(np.percentile(lat, 90) * 1e3) if lat.size else None,
        "latency_p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
        "latency_p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
        "latency_mean_ms": float(lat.mean() * 1e3) if lat.size else None,
        "atom_evals": int(atom_evals),
        "makespan_s": makespan,
        "atom_evals_per_s": atom_evals / makespan if makespan > 0 else 0.0,
        "ticks": ticks,
        "mean_batch_requests": float(np.mean(batch_sizes)) if batch_sizes else 0.0,
        "host_time_s": host,
    }


def serve_sequential(pool, stream, dt=0.001, skin=0.5, drain_timeout=120.0):
    """The same stream without batching: one request at a time in arrival
    order, each evaluated alone (the plain single-structure path)."""
    t0 = time.perf_counter()
    done = []
    atom_evals = 0
    for r in sorted(stream, key=lambda q: q.t_arrival):
        while time.perf_counter() - t0 < r.t_arrival:
            time.sleep(min(r.t_arrival - (time.perf_counter() - t0), 0.01))
        if time.perf_counter() - t0 > drain_timeout:
            break
        h = pool.admit(r.structure, nsteps=r.nsteps, dt=dt, skin=skin if r.nsteps > 0 else 0.0, velocities=r.v0)
        if h is None:
            raise RuntimeError("pool too small for a single request")
        while pool.tick():
            pool.sync()
        for hh in pool.done():
            res = pool.result(hh)
            r.t_done = time.perf_counter() - t0
            r.evals = res["thermo"].shape[0]
            r.energy = res["energy"]
            r.forces = res["forces"]
            atom_evals += r.evals * r.structure.natoms
            done.append(r)
    makespan = time.perf_counter() - t0
    lat = np.array([r.t_done - r.t_arrival for r in done]) if done else np.zeros(0)
    return {
        "completed": len(done),
        "offered": len(stream),
        "latency_p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
        "latency_p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
        "atom_evals": int(atom_evals),
        "makespan_s": makespan,
        "atom_evals_per_s": atom_evals / makespan if makespan > 0 else 0.0,
    }


def stream_errors(stream, reference):
    """Per-request error of a served stream's results against the same
    requests evaluated in another run (the FP64 one-at-a-time stream):
    energy per atom and largest force difference of the last evaluation,
    separately for single evaluations and NVE runs (whose final state also
    carries the trajectories' divergence)."""
    out = {}
    for kind, sel in (("single", lambda r: r.nsteps == 0), ("nve", lambda r: r.nsteps > 0)):
        de, df = [], []
        for a, b in zip(stream, reference):
            if not sel(a) or a.energy is None or b.energy is None:
                continue
            de.append(abs(a.energy - b.energy) / a.structure.natoms)
            df.append(float(np.abs(a.forces - b.forces).max()) if a.forces.size else 0.0)
        out[kind] = {"requests": len(de),
                     "energy_per_atom_max": float(max(de)) if de else None,
                     "energy_per_atom_median": float(np.median(de)) if de else None,
                     "force_max": float(max(df)) if df else None,
                     "force_median": float(np.median(df)) if df else None}
    out["unit"] = "eV/atom, eV/A"
    return out


def pool_for_budget(pair, hbm_bytes, max_requests=4096, pairs_per_atom=160):
    """Largest pool (atoms, pairs = atoms * pairs_per_atom) within hbm_bytes."""
    lo, hi = 1024, 1 << 26
    while hi - lo > 1024:
        mid = (lo + hi) // 2
        if RequestPool.bytes_for(pair, mid, mid * pairs_per_atom, max_requests) <= hbm_bytes:
            lo = mid
        else:
            hi = mid
    return RequestPool(pair, lo, lo * pairs_per_atom, max_requests), lo
