"""Runs the LAMMPS adapter (lammps-mlip-package_amd/lammps/pair_mlip_gtinv_hip.cpp)
through tests/lammps_harness: the calls of a LAMMPS run (settings, coeff,
init_style, init_one, then force_clear + compute per step) on stand-in
LAMMPS classes, with the inputs of a workloads.NeighborList."""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS_DIR = os.path.join(HERE, "lammps_harness")
HARNESS = os.path.join(HARNESS_DIR, "_build", "lammps_harness")

# eflag / vflag bits as LAMMPS's Integrate::ev_set passes them with newton on
ENERGY_GLOBAL, ENERGY_ATOM = 1, 2
VIRIAL_PAIR, VIRIAL_FDOTR, VIRIAL_ATOM = 1, 2, 4


def hip_device_count():
    """Devices the HIP runtime sees (0 where there is none or no runtime)."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return 0
    n = ctypes.c_int(0)
    return n.value if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 else 0


def ensure_built():
    """The harness binary (g++ only; build() makes it with the engine)."""
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", HARNESS_DIR], check=True, capture_output=True)
    return HARNESS


def write_input(path, nl, ntypes, eflag, vflag, nsteps=1, tags=None):
    off = np.asarray(nl.off, dtype=np.int64)
    counts = np.asarray(nl.numneigh, dtype=np.int64)[np.asarray(nl.ilist)]
    assert np.array_equal(off[: nl.inum], np.concatenate([[0], np.cumsum(counts)[:-1]])), "lists in ilist order"
    tag = np.asarray(nl.tag if tags is None else tags)
    with open(path, "wb") as fp:
        np.array([nl.nlocal, nl.nall - nl.nlocal, ntypes, nl.inum, nl.npairs, eflag, vflag, nsteps],
                 dtype=np.int32).tofile(fp)
        np.ascontiguousarray(nl.x, dtype=np.float64).tofile(fp)
        np.ascontiguousarray(nl.types, dtype=np.int32).tofile(fp)
        np.ascontiguousarray(tag, dtype=np.int32).tofile(fp)
        np.ascontiguousarray(nl.ilist, dtype=np.int32).tofile(fp)
        np.ascontiguousarray(np.asarray(nl.numneigh)[: nl.nlocal], dtype=np.int32).tofile(fp)
        np.ascontiguousarray(nl.idx, dtype=np.int32).tofile(fp)


def read_output(path, nall, ntypes, nsteps):
    raw = open(path, "rb").read()
    flags = np.frombuffer(raw[:32], dtype=np.int32)
    vals = np.frombuffer(raw[32:], dtype=np.float64)
    out = {"request_full": int(flags[0]), "request_half": int(flags[1]), "newton_pair": int(flags[2]),
           "vflag_global": int(flags[3]), "vflag_fdotr": int(flags[4]), "eflag_atom": int(flags[5]),
           "vflag_atom": int(flags[6]), "cutoff": float(vals[0]), "mass": vals[1:1 + ntypes].copy(), "steps": []}
    k = 1 + ntypes
    per = 1 + 6 + 3 * nall + nall + 6 * nall
    assert vals.size == k + nsteps * per, (vals.size, k, per)
    for _ in range(nsteps):
        v = vals[k:k + per]
        out["steps"].append({"energy": float(v[0]), "virial": v[1:7].copy(),
                             "forces_all": v[7:7 + 3 * nall].reshape(nall, 3).copy(),
                             "eatom": v[7 + 3 * nall:7 + 4 * nall].copy(),
                             "vatom": v[7 + 4 * nall:].reshape(nall, 6).copy()})
        k += per
    return out


def run(tmpdir, style, potfile, elements, nl, eflag, vflag, nsteps=1, tags=None, timeout=120):
    """Returns (returncode, stderr text, parsed output or None)."""
    ensure_built()
    fin = os.path.join(str(tmpdir), "harness_in.bin")
    fout = os.path.join(str(tmpdir), "harness_out.bin")
    write_input(fin, nl, len(elements), eflag, vflag, nsteps, tags)
    p = subprocess.run([HARNESS, style, potfile, fin, fout] + list(elements), capture_output=True, text=True,
                       timeout=timeout)
    out = read_output(fout, nl.nall, len(elements), nsteps) if p.returncode == 0 else None
    return p.returncode, p.stderr, out
