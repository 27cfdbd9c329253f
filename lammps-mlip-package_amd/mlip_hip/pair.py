"""Python mirror of the reference pair style ``mlip_gtinv``.

``PairMLIPGtinv`` follows src/USER-MLIP/pair_mlip_gtinv.{h,cpp}: ``settings``
(:545-549) takes no argument, ``coeff`` (:555-593) reads ``* * potfile
elem...``, ``init_one`` (:599-603) returns the cutoff and ``compute``
(:63-125) evaluates energy, forces and virial for a LAMMPS-shaped neighbor
list.  All arithmetic happens in the HIP library.
"""
import ctypes

import numpy as np

from ._lib import MLIP_ERR_FULL, MLIP_OK, MlipError, library

INFO_KEYS = ["n_type", "n_fn", "n_gtinv", "n_des", "n_coeff", "maxl", "lmax3", "n_poly_features", "n_irreps",
             "gtinv_order", "model_type", "maxp"]


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _null(ct=ctypes.c_double):
    return ctypes.POINTER(ct)()


class PairMLIPGtinv:
    single_enable = 0
    restartinfo = 0
    one_coeff = 1
    manybody_flag = 1

    def __init__(self):
        self.ctx = None
        self.lib = library()

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.mlip_destroy(self.ctx)
            self.ctx = None

    def _check(self, rc):
        if rc != MLIP_OK:
            raise MlipError(rc, self.lib.mlip_last_error(self.ctx).decode())

    def settings(self, *args):
        if self.lib.mlip_settings(len(args)) != MLIP_OK:
            raise MlipError(1, "Illegal pair_style command")

    def coeff(self, *args, ntypes=None):
        """``coeff('*', '*', potfile, 'Ti', 'Al', ...)``"""
        if len(args) < 4 or args[0] != "*" or args[1] != "*":
            raise MlipError(1, "Incorrect args for pair coefficients")
        potfile, elems = args[2], list(args[3:])
        if ntypes is not None and len(elems) != ntypes:
            raise MlipError(1, "Incorrect args for pair coefficients")
        arr = (ctypes.c_char_p * len(elems))(*[e.encode() for e in elems])
        err = ctypes.create_string_buffer(1024)
        ctx = ctypes.c_void_p()
        rc = self.lib.mlip_create(potfile.encode(), len(elems), arr, ctypes.byref(ctx), err, 1024)
        if rc != MLIP_OK:
            raise MlipError(rc, err.value.decode())
        self.close()
        self.ctx = ctx
        self.ntypes = len(elems)

    def init_one(self, i=1, j=1):
        return self.lib.mlip_cutoff(self.ctx)

    def masses(self):
        out = np.zeros(self.ntypes)
        self._check(self.lib.mlip_masses(self.ctx, _p(out, ctypes.c_double)))
        return out

    def model_info(self):
        v = np.zeros(12, dtype=np.int32)
        self.lib.mlip_model_info(self.ctx, _p(v, ctypes.c_int))
        return dict(zip(INFO_KEYS, v.tolist()))

    def set_anlm_kernel(self, kind):
        """a_nlm kernel: "registers" or "mfma" (matrix cores)."""
        self._check(self.lib.mlip_set_anlm_kernel(self.ctx, {"registers": 0, "mfma": 1}[kind]))

    def anlm_kernel(self):
        return {0: "registers", 1: "mfma"}.get(self.lib.mlip_get_anlm_kernel(self.ctx))

    def set_precision(self, mode):
        self._check(self.lib.mlip_set_precision(self.ctx, {"fp64": 0, "mixed": 1}[mode]))

    def set_atom_mode(self, mode):
        """Per-atom energy/virial: "pair" (the reference CPU style's
        ev_tally_xyz split, default) or "site" (site energies E_i)."""
        self._check(self.lib.mlip_set_atom_mode(self.ctx, {"pair": 0, "site": 1}[mode]))

    def compute_lammps(self, nl, types, eflag=1, vflag=1, eatom=False, vatom=False, tag_dtype=np.int64):
        """The LAMMPS-shaped entry point ``mlip_compute``: ``firstneigh`` as
        an array of per-atom pointers into the list (``int**``), exactly as
        ``NeighList::firstneigh`` hands it to ``Pair::compute``."""
        f = np.zeros((nl.nall, 3))
        eng = np.zeros(1)
        vir = np.zeros(6)
        ea = np.zeros(nl.nall) if eatom else None
        va = np.zeros((nl.nall, 6)) if vatom else None
        typ = np.ascontiguousarray(types, dtype=np.int32)
        tag = np.ascontiguousarray(np.asarray(nl.tag).astype(tag_dtype).astype(np.int64))
        idx = np.ascontiguousarray(nl.idx, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        rows = (ip * max(nl.nlocal, 1))()
        base = idx.ctypes.data
        for ii in range(nl.inum):
            rows[int(nl.ilist[ii])] = ctypes.cast(base + 4 * int(nl.off[ii]), ip)
        rc = self.lib.mlip_compute(
            self.ctx, nl.nlocal, nl.nall - nl.nlocal, _p(nl.x, ctypes.c_double), _p(typ, ctypes.c_int),
            _p(tag, ctypes.c_int64), nl.inum, _p(nl.ilist, ctypes.c_int), _p(nl.numneigh, ctypes.c_int),
            ctypes.cast(rows, ctypes.POINTER(ip)), 1 if nl.full else 0, eflag, vflag, _p(f, ctypes.c_double),
            _p(eng, ctypes.c_double), _p(vir, ctypes.c_double), _p(ea, ctypes.c_double) if eatom else _null(),
            _p(va, ctypes.c_double) if vatom else _null())
        self._check(rc)
        out = {"energy": float(eng[0]), "forces_all": f, "virial": vir}
        if nl.owner is not None:
            fo = np.zeros((nl.nlocal, 3))
            np.add.at(fo, nl.owner, f)
            out["forces"] = fo
        if eatom:
            out["eatom"] = ea
        if vatom:
            out["vatom"] = va
        return out

    def compute(self, nl, types, eflag=1, vflag=1, eatom=False, vatom=False, noalloc=False):
        """Evaluate on a ``workloads.NeighborList``; ``types`` are the
        1-based LAMMPS types of all nall positions.  Returns a dict with the
        energy, forces on all positions (ghost entries unreduced), forces
        folded onto owned atoms, and the virial.  ``noalloc`` selects the
        entry point that allocates no device memory (mlip_compute_flat_noalloc)."""
        f = np.zeros((nl.nall, 3))
        eng = np.zeros(1)
        vir = np.zeros(6)
        ea = np.zeros(nl.nall) if eatom else None
        va = np.zeros((nl.nall, 6)) if vatom else None
        typ = np.ascontiguousarray(types, dtype=np.int32)
        tag = np.ascontiguousarray(nl.tag, dtype=np.int64)
        fn = self.lib.mlip_compute_flat_noalloc if noalloc else self.lib.mlip_compute_flat
        rc = fn(
            self.ctx, nl.nlocal, nl.nall - nl.nlocal, _p(nl.x, ctypes.c_double), _p(typ, ctypes.c_int),
            _p(tag, ctypes.c_int64), nl.inum, _p(nl.ilist, ctypes.c_int), _p(nl.numneigh, ctypes.c_int),
            _p(nl.off, ctypes.c_int), _p(nl.idx, ctypes.c_int), 1 if nl.full else 0, eflag, vflag,
            _p(f, ctypes.c_double), _p(eng, ctypes.c_double), _p(vir, ctypes.c_double),
            _p(ea, ctypes.c_double) if eatom else _null(), _p(va, ctypes.c_double) if vatom else _null())
        self._check(rc)
        out = {"energy": float(eng[0]), "forces_all": f, "virial": vir}
        if nl.owner is not None:
            fo = np.zeros((nl.nlocal, 3))
            np.add.at(fo, nl.owner, f)
            out["forces"] = fo
        if eatom:
            out["eatom"] = ea
        if vatom:
            out["vatom"] = va
        return out

    def debug_fetch(self, what, n):
        buf = np.zeros(n)
        got = self.lib.mlip_debug_fetch(self.ctx, what, _p(buf, ctypes.c_double), n)
        if got < 0:
            raise MlipError(5, "debug fetch failed")
        return buf[:got]


class _ResidentBase:
    def compute(self):
        self.pair._check(self.lib.mlip_resident_compute(self.pair.ctx))

    def sync(self):
        self.pair._check(self.lib.mlip_resident_sync(self.pair.ctx))

    def graph(self, on=True):
        self.pair._check(self.lib.mlip_graph_mode(self.pair.ctx, 1 if on else 0))

    def autotune(self, reps=3):
        v = ctypes.c_int(0)
        self.pair._check(self.lib.mlip_autotune(self.pair.ctx, int(reps), ctypes.byref(v)))
        return v.value

    def autotune_q34(self, reps=2):
        """Order-3/4 workgroup plan timed on the device; returns its two
        work-model constants and the largest L of species-paired rows."""
        v = np.zeros(3, dtype=np.int32)
        self.pair._check(self.lib.mlip_autotune_q34(self.pair.ctx, int(reps), _p(v, ctypes.c_int)))
        return tuple(int(x) for x in v)

    def set_q34_plan(self, region_cost, visit_cost, pair_lmax):
        """Order-3/4 workgroup plan by hand (what autotune_q34 returns)."""
        self.pair._check(self.lib.mlip_set_q34_plan(self.pair.ctx, int(region_cost), int(visit_cost),
                                                    int(pair_lmax)))

    def set_force_threads(self, threads):
        self.pair._check(self.lib.mlip_set_force_threads(self.pair.ctx, int(threads)))

    def profile(self, on=True):
        self.lib.mlip_profile_enable(self.pair.ctx, 1 if on else 0)

    def read_profile(self):
        """{kernel: (total_ms, launches)} since profile(True)."""
        names = (ctypes.c_char_p * 64)()
        ms = np.zeros(64)
        cnt = np.zeros(64, dtype=np.int32)
        n = self.lib.mlip_profile_read(self.pair.ctx, names, _p(ms, ctypes.c_double), _p(cnt, ctypes.c_int), 64)
        return {names[k].decode(): (float(ms[k]), int(cnt[k])) for k in range(n)}

    def flops(self):
        out = np.zeros(7)
        self.pair._check(self.lib.mlip_flops(self.pair.ctx, _p(out, ctypes.c_double)))
        keys = ["anlm", "q12", "q34_feat", "q34_adj", "poly", "force", "total"]
        return dict(zip(keys, out.tolist()))

    def flops_reference(self):
        """Algorithmic FLOPs by the reference-formulation rule (see
        mlip_flops_reference)."""
        out = np.zeros(7)
        self.pair._check(self.lib.mlip_flops_reference(self.pair.ctx, _p(out, ctypes.c_double)))
        keys = ["anlm", "q12", "q34_feat", "q34_adj", "poly", "force", "total"]
        return dict(zip(keys, out.tolist()))

    def train_features(self, want=("X", "P", "FF", "SF")):
        """Training outputs (mlip_train_features): X [natoms, n_des],
        P [natoms, n_coeff], FF [natoms, n_coeff, 3] (= -dX_p/dr_k),
        SF [nstruct, n_coeff, 6] (xx, yy, zz, yz, zx, xy)."""
        info = self.pair.model_info()
        n = self.n if hasattr(self, "n") else self.natoms
        ns = getattr(self, "nstruct", 1)
        nd, nc = info["n_des"], info["n_coeff"]
        out = {"X": np.zeros((n, nd)) if "X" in want else None,
               "P": np.zeros((n, nc)) if "P" in want else None,
               "FF": np.zeros((n, nc, 3)) if "FF" in want else None,
               "SF": np.zeros((ns, nc, 6)) if "SF" in want else None}
        q = lambda a: _p(a, ctypes.c_double) if a is not None else _null()  # noqa: E731
        self.pair._check(self.lib.mlip_train_features(self.pair.ctx, q(out["X"]), q(out["P"]), q(out["FF"]),
                                                      q(out["SF"])))
        return out

    def lists(self):
        """The resident neighbor and force-gather lists as built on the
        device: off, idx, shift (del = x_i - x_j + shift), goff, gent."""
        n = self.n if hasattr(self, "n") else self.natoms
        np_ = int(self.lib.mlip_resident_npairs(self.pair.ctx))
        off = np.zeros(n + 1, dtype=np.int64)
        idx = np.zeros(max(np_, 1), dtype=np.int32)
        sh = np.zeros((max(np_, 1), 3))
        goff = np.zeros(n + 1, dtype=np.int32)
        gent = np.zeros(max(2 * np_, 1), dtype=np.int32)
        self.pair._check(self.lib.mlip_resident_lists(self.pair.ctx, _p(off, ctypes.c_int64), _p(idx, ctypes.c_int),
                                                      _p(sh, ctypes.c_double), _p(goff, ctypes.c_int),
                                                      _p(gent, ctypes.c_int)))
        return {"off": off, "idx": idx[:np_], "shift": sh[:np_], "goff": goff, "gent": gent[:2 * np_]}

    def calibration_copy(self, nbytes, reps=1):
        self.pair._check(self.lib.mlip_calibration_copy(self.pair.ctx, int(nbytes), int(reps)))


class BatchSystem(_ResidentBase):
    """Batch of independent periodic structures evaluated together on the
    device (structures: list of workloads.Structure)."""

    def __init__(self, pair, structures, builder="device"):
        """builder: where the per-structure neighbor lists are built,
        "device" (HIP kernels) or "host" (threads; identical lists)."""
        self.pair = pair
        self.lib = pair.lib
        if builder not in ("device", "host"):
            raise ValueError("builder must be 'device' or 'host'")
        pair._check(self.lib.mlip_set_neighbor_builder(pair.ctx, 1 if builder == "device" else 0))
        self.nstruct = len(structures)
        nat = np.array([s.natoms for s in structures], dtype=np.int32)
        cells = np.ascontiguousarray(np.concatenate([s.cell.reshape(9) for s in structures]), dtype=np.float64)
        x = np.ascontiguousarray(np.concatenate([s.positions for s in structures]), dtype=np.float64)
        t = np.ascontiguousarray(np.concatenate([s.types for s in structures]), dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(nat)])
        self.natoms = int(self.offsets[-1])
        pair._check(self.lib.mlip_batch_set(pair.ctx, self.nstruct, _p(nat, ctypes.c_int), _p(cells, ctypes.c_double),
                                            _p(x, ctypes.c_double), _p(t, ctypes.c_int)))

    def get(self, forces=True):
        e = np.zeros(self.nstruct)
        v = np.zeros((self.nstruct, 6))
        f = np.zeros((self.natoms, 3)) if forces else None
        self.pair._check(self.lib.mlip_batch_get(self.pair.ctx, _p(e, ctypes.c_double), _p(v, ctypes.c_double),
                                                 _p(f, ctypes.c_double) if forces else _null()))
        out = {"energies": e, "virials": v}
        if forces:
            out["forces"] = [f[self.offsets[k]:self.offsets[k + 1]] for k in range(self.nstruct)]
        return out


class ResidentSystem(_ResidentBase):
    """Device-resident structure: positions, types and a full neighbor list
    of owned atoms with image shifts stay in HBM; compute() only enqueues.
    ResidentSystem.from_cell builds the list on the device."""

    @classmethod
    def from_cell(cls, pair, cell, x, types):
        """Periodic cell (rows a, b, c): neighbor list, gather lists and
        layout built on the device (mlip_resident_set_cell)."""
        self = cls.__new__(cls)
        self.pair = pair
        self.lib = pair.lib
        x = np.ascontiguousarray(x, dtype=np.float64)
        self.n = int(x.shape[0])
        cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
        t = np.ascontiguousarray(types, dtype=np.int32)
        pair._check(self.lib.mlip_resident_set_cell(pair.ctx, self.n, _p(cell, ctypes.c_double),
                                                    _p(x, ctypes.c_double), _p(t, ctypes.c_int)))
        self.npairs = int(self.lib.mlip_resident_npairs(pair.ctx))
        return self

    def __init__(self, pair, x, types, nbr_off, nbr_idx, nbr_shift=None):
        self.pair = pair
        self.lib = pair.lib
        self.n = int(x.shape[0])
        x = np.ascontiguousarray(x, dtype=np.float64)
        t = np.ascontiguousarray(types, dtype=np.int32)
        off = np.ascontiguousarray(nbr_off, dtype=np.int64)
        idx = np.ascontiguousarray(nbr_idx, dtype=np.int32)
        sh = None if nbr_shift is None else np.ascontiguousarray(nbr_shift, dtype=np.float64)
        rc = self.lib.mlip_resident_set(pair.ctx, self.n, _p(x, ctypes.c_double), _p(t, ctypes.c_int),
                                        int(idx.size), _p(off, ctypes.c_int64), _p(idx, ctypes.c_int),
                                        _p(sh, ctypes.c_double) if sh is not None else _null())
        pair._check(rc)
        self.npairs = int(idx.size)

    def update_positions(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        self.pair._check(self.lib.mlip_resident_update_positions(self.pair.ctx, _p(x, ctypes.c_double)))

    def get(self, eatom=False):
        f = np.zeros((self.n, 3))
        e = np.zeros(1)
        v = np.zeros(6)
        ea = np.zeros(self.n) if eatom else None
        self.pair._check(self.lib.mlip_resident_get(self.pair.ctx, _p(f, ctypes.c_double), _p(e, ctypes.c_double),
                                                    _p(v, ctypes.c_double),
                                                    _p(ea, ctypes.c_double) if eatom else _null()))
        out = {"energy": float(e[0]), "forces": f, "virial": v}
        if eatom:
            out["eatom"] = ea
        return out



class RequestPool:
    """Device-resident request pool with continuous batching (see the
    mlip_pool_* entry points in include/mlip_hip.h).  Structures are
    workloads.Structure objects; results come back per request."""

    THERMO = 13  # energy, virial[6], kinetic-energy tensor[6] per evaluation

    def __init__(self, pair, max_atoms, max_pairs, max_requests):
        self.pair = pair
        self.lib = pair.lib
        self.natoms = {}
        pair._check(self.lib.mlip_pool_init(pair.ctx, int(max_atoms), int(max_pairs), int(max_requests)))

    @staticmethod
    def bytes_for(pair, max_atoms, max_pairs, max_requests):
        return int(pair.lib.mlip_pool_bytes(pair.ctx, int(max_atoms), int(max_pairs), int(max_requests)))

    DEFAULT_SKIN = 1.0  # A, for NVE requests (their lists are rebuilt when an atom moves skin / 2)

    def admit(self, s, nsteps=0, dt=0.001, skin=None, velocities=None):
        """Returns a handle, or None when the pool has no room now.  skin
        defaults to 0 for single evaluations and DEFAULT_SKIN for NVE."""
        if skin is None:
            skin = self.DEFAULT_SKIN if nsteps > 0 else 0.0
        cell = np.ascontiguousarray(s.cell, dtype=np.float64).reshape(9)
        x = np.ascontiguousarray(s.positions, dtype=np.float64)
        t = np.ascontiguousarray(s.types, dtype=np.int32)
        v = None if velocities is None else np.ascontiguousarray(velocities, dtype=np.float64)
        h = ctypes.c_int(0)
        rc = self.lib.mlip_pool_admit(self.pair.ctx, int(s.natoms), _p(cell, ctypes.c_double), _p(x, ctypes.c_double),
                                      _p(t, ctypes.c_int), _p(v, ctypes.c_double) if v is not None else _null(),
                                      int(nsteps), float(dt), float(skin), ctypes.byref(h))
        if rc == MLIP_ERR_FULL:
            return None
        self.pair._check(rc)
        self.natoms[h.value] = int(s.natoms)
        return h.value

    def admit_many(self, structures, nsteps, dt=0.001, skins=None, velocities=None, nthreads=16):
        """FIFO admission of several structures (their neighbor lists are
        built in one batch on the device; `nthreads` is kept for callers of
        the earlier host builder and not used); returns the handles of those
        admitted (a prefix of the list)."""
        n = len(structures)
        if n == 0:
            return []
        nat = np.array([s.natoms for s in structures], dtype=np.int32)
        cells = np.ascontiguousarray(np.concatenate([s.cell.reshape(9) for s in structures]), dtype=np.float64)
        x = np.ascontiguousarray(np.concatenate([s.positions for s in structures]), dtype=np.float64)
        t = np.ascontiguousarray(np.concatenate([s.types for s in structures]), dtype=np.int32)
        ns = np.ascontiguousarray(nsteps, dtype=np.int32)
        if skins is None:
            skins = [self.DEFAULT_SKIN if k > 0 else 0.0 for k in nsteps]
        sk = np.ascontiguousarray(skins, dtype=np.float64)
        v = None
        if velocities is not None:
            v = np.ascontiguousarray(np.concatenate([np.zeros((s.natoms, 3)) if vv is None else vv
                                                     for s, vv in zip(structures, velocities)]), dtype=np.float64)
        hs = np.zeros(n, dtype=np.int32)
        k = self.lib.mlip_pool_admit_many(self.pair.ctx, n, _p(nat, ctypes.c_int), _p(cells, ctypes.c_double),
                                          _p(x, ctypes.c_double), _p(t, ctypes.c_int),
                                          _p(v, ctypes.c_double) if v is not None else _null(), _p(ns, ctypes.c_int),
                                          float(dt), _p(sk, ctypes.c_double), int(nthreads), _p(hs, ctypes.c_int))
        if k < 0:
            self.pair._check(-k)
        for j in range(k):
            self.natoms[int(hs[j])] = int(nat[j])
        return hs[:k].tolist()

    def tick(self):
        """Enqueue one batched evaluation of every resident request; returns
        the batch size (0: nothing resident)."""
        n = ctypes.c_int(0)
        self.pair._check(self.lib.mlip_pool_tick(self.pair.ctx, ctypes.byref(n)))
        return n.value

    def sync(self):
        n = ctypes.c_int(0)
        self.pair._check(self.lib.mlip_pool_sync(self.pair.ctx, ctypes.byref(n)))
        return n.value

    def done(self, cap=4096):
        ids = np.zeros(cap, dtype=np.int32)
        n = self.lib.mlip_pool_done(self.pair.ctx, _p(ids, ctypes.c_int), cap)
        if n < 0:
            self.pair._check(-n)
        return ids[:n].tolist()

    def result(self, handle, max_evals=4096):
        n = self.natoms.pop(handle)
        e = np.zeros(1)
        vir = np.zeros(6)
        f = np.zeros((n, 3))
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        th = np.zeros(max_evals * self.THERMO)
        nev = self.lib.mlip_pool_result(self.pair.ctx, int(handle), _p(e, ctypes.c_double), _p(vir, ctypes.c_double),
                                        _p(f, ctypes.c_double), _p(x, ctypes.c_double), _p(v, ctypes.c_double),
                                        _p(th, ctypes.c_double), th.size)
        if nev < 0:
            self.pair._check(-nev)
        return {"energy": float(e[0]), "virial": vir, "forces": f, "positions": x, "velocities": v,
                "thermo": th[:nev * self.THERMO].reshape(nev, self.THERMO)}

    def rebuilds(self):
        """Neighbor-list rebuilds of NVE requests so far."""
        return int(self.lib.mlip_pool_rebuilds(self.pair.ctx))

    def stats(self):
        out = np.zeros(6, dtype=np.int64)
        self.pair._check(self.lib.mlip_pool_stats(self.pair.ctx, _p(out, ctypes.c_int64)))
        return dict(zip(["atoms_used", "pairs_used", "atom_capacity", "pair_capacity", "requests", "running"],
                        out.tolist()))

    def profile(self, on=True):
        self.lib.mlip_profile_enable(self.pair.ctx, 1 if on else 0)

    def read_profile(self):
        """{kernel: (total_ms, launches)} since profile(True)."""
        names = (ctypes.c_char_p * 64)()
        ms = np.zeros(64)
        cnt = np.zeros(64, dtype=np.int32)
        n = self.lib.mlip_profile_read(self.pair.ctx, names, _p(ms, ctypes.c_double), _p(cnt, ctypes.c_int), 64)
        return {names[k].decode(): (float(ms[k]), int(cnt[k])) for k in range(n)}


def inspect_model(potfile, which, cap=1 << 26):
    """Device-free model tables (see mlip_model_inspect in include/mlip_hip.h)."""
    lib = library()
    out = np.zeros(cap, dtype=np.int32)
    err = ctypes.create_string_buffer(1024)
    n = lib.mlip_model_inspect(potfile.encode(), which, _p(out, ctypes.c_int), cap, err, 1024)
    if n < 0:
        raise MlipError(3, err.value.decode())
    out = out[:n]
    if which == 0:
        return dict(zip(INFO_KEYS, out.tolist()))
    if which == 4:
        return int(out[0])
    if which == 1:
        return out.reshape(-1, 9)
    return out.reshape(-1, which)


def periodic_full_list(cell, x, rc):
    """Binned periodic full list (native builder): returns (off int64[n+1],
    idx int32[npairs], shift float64[npairs,3]) with del = x_i - x_j + shift."""
    lib = library()
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    x = np.ascontiguousarray(x, dtype=np.float64)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    rc_ = lib.mlip_nlist_build(_p(cell, ctypes.c_double), int(x.shape[0]), _p(x, ctypes.c_double), float(rc),
                               ctypes.byref(h), err, 512)
    if rc_ != MLIP_OK:
        raise MlipError(rc_, err.value.decode())
    try:
        npairs = lib.mlip_nlist_npairs(h)
        off = np.zeros(x.shape[0] + 1, dtype=np.int64)
        idx = np.zeros(max(npairs, 1), dtype=np.int32)
        sh = np.zeros((max(npairs, 1), 3))
        lib.mlip_nlist_copy(h, _p(off, ctypes.c_int64), _p(idx, ctypes.c_int), _p(sh, ctypes.c_double))
    finally:
        lib.mlip_nlist_free(h)
    return off, idx[:npairs], sh[:npairs]
