"""Structures, potential files and neighbor lists for the gtinv evaluation path.

Everything here is host-side input preparation, the work LAMMPS does before it
calls ``Pair::compute``:

* reading a LAMMPS ``atomic``-style data file (orthogonal or triclinic box);
* synthetic structures of the benchmark workloads (fcc lattices with Gaussian
  displacements, random triclinic cells with minimum-distance rejection);
* writing a potential file in the reference format
  (``src/USER-MLIP/mlip_model.h:45-101``) with seeded random coefficients;
* periodic neighbor lists in LAMMPS form: local atoms followed by ghost
  images, ``ilist``/``numneigh``/``firstneigh`` as a half list (newton on,
  each physical pair once, as ``half/bin/newton`` builds for the CPU style)
  or a full list (as the Kokkos style requests with ``neigh full``).

The device-resident neighbor builder is native code; this module is the
reference-shaped path the tests and the boundary use.
"""
import math
import os

import numpy as np

# ----------------------------------------------------------------------------
# LAMMPS data files


class Structure:
    """Periodic cell: rows of ``cell`` are the a, b, c edge vectors (LAMMPS
    triclinic convention a=(lx,0,0), b=(xy,ly,0), c=(xz,yz,lz))."""

    def __init__(self, cell, positions, types, masses=None, origin=None):
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        self.positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        self.types = np.ascontiguousarray(types, dtype=np.int32)  # 1-based LAMMPS types
        self.masses = None if masses is None else np.asarray(masses, dtype=np.float64)
        self.origin = np.zeros(3) if origin is None else np.asarray(origin, dtype=np.float64)

    @property
    def natoms(self):
        return self.positions.shape[0]

    @property
    def volume(self):
        return abs(np.linalg.det(self.cell))

    def wrapped(self):
        """Positions mapped into the cell (fractional coordinates in [0,1))."""
        frac = np.linalg.solve(self.cell.T, (self.positions - self.origin).T).T
        frac -= np.floor(frac)
        return frac @ self.cell + self.origin


def read_lammps_data(path):
    lines = open(path).read().splitlines()
    natoms = ntypes = None
    xlo = xhi = ylo = yhi = zlo = zhi = 0.0
    xy = xz = yz = 0.0
    masses = {}
    atoms = []
    i = 1
    while i < len(lines):
        raw = lines[i].split("#")[0].strip()
        toks = raw.split()
        if not toks:
            i += 1
            continue
        if len(toks) == 2 and toks[1] == "atoms":
            natoms = int(toks[0])
        elif len(toks) == 3 and toks[1:] == ["atom", "types"]:
            ntypes = int(toks[0])
        elif toks[-2:] == ["xlo", "xhi"]:
            xlo, xhi = float(toks[0]), float(toks[1])
        elif toks[-2:] == ["ylo", "yhi"]:
            ylo, yhi = float(toks[0]), float(toks[1])
        elif toks[-2:] == ["zlo", "zhi"]:
            zlo, zhi = float(toks[0]), float(toks[1])
        elif toks[-3:] == ["xy", "xz", "yz"]:
            xy, xz, yz = float(toks[0]), float(toks[1]), float(toks[2])
        elif toks[0] == "Masses":
            i += 1
            while i < len(lines) and not lines[i].strip():
                i += 1
            while i < len(lines) and lines[i].strip():
                t = lines[i].split()
                masses[int(t[0])] = float(t[1])
                i += 1
            continue
        elif toks[0] == "Atoms":
            i += 1
            while i < len(lines) and not lines[i].strip():
                i += 1
            while i < len(lines) and lines[i].strip():
                t = lines[i].split()
                atoms.append((int(t[0]), int(t[1]), float(t[2]), float(t[3]), float(t[4])))
                i += 1
            continue
        i += 1
    if natoms is None or len(atoms) != natoms:
        raise ValueError("bad LAMMPS data file %s" % path)
    atoms.sort()
    cell = np.array([[xhi - xlo, 0.0, 0.0], [xy, yhi - ylo, 0.0], [xz, yz, zhi - zlo]])
    pos = np.array([[a[2], a[3], a[4]] for a in atoms])
    typ = np.array([a[1] for a in atoms], dtype=np.int32)
    m = None
    if masses:
        m = np.array([masses[k] for k in sorted(masses)])
    return Structure(cell, pos, typ, m, origin=[xlo, ylo, zlo])


# ----------------------------------------------------------------------------
# Synthetic structures


def fcc(ncell, a, sigma, seed, types=None, type_fraction=None, type_seed=None):
    """fcc supercell of ncell^3 conventional cells with N(0, sigma^2)
    displacements drawn from PCG64(seed).  ``type_fraction`` assigns type 2
    to a random fraction of sites drawn with PCG64(type_seed)."""
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    grid = np.stack(np.meshgrid(np.arange(ncell), np.arange(ncell), np.arange(ncell),
                                indexing="ij"), -1).reshape(-1, 3)
    frac = (grid[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    pos = frac * a
    rng = np.random.default_rng(seed)
    pos = pos + rng.normal(0.0, sigma, size=pos.shape)
    n = pos.shape[0]
    if types is not None:
        typ = np.asarray(types, dtype=np.int32)
    elif type_fraction is not None:
        trng = np.random.default_rng(type_seed)
        nb = int(round(type_fraction * n))
        typ = np.ones(n, dtype=np.int32)
        typ[trng.permutation(n)[:nb]] = 2
    else:
        typ = np.ones(n, dtype=np.int32)
    cell = np.eye(3) * (ncell * a)
    return Structure(cell, pos, typ)


def random_cell_structure(natoms, density, rng, rmin=1.8, type_fraction=None, max_tries=200000):
    """Random triclinic cell at the given number density with positions drawn
    uniformly under a minimum-distance rejection (periodic images included)."""
    vol = natoms / density
    while True:
        lengths = rng.uniform(0.8, 1.25, size=3)
        angles = np.radians(rng.uniform(70.0, 110.0, size=3))
        ca, cb, cg = np.cos(angles)
        sg = np.sin(angles[2])
        v_unit = math.sqrt(max(1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg, 1e-6))
        scale = (vol / (lengths.prod() * v_unit)) ** (1.0 / 3.0)
        la, lb, lc = lengths * scale
        ax = la
        bx, by = lb * cg, lb * sg
        cx = lc * cb
        cy = lc * (ca - cb * cg) / sg
        cz = math.sqrt(max(lc * lc - cx * cx - cy * cy, 1e-12))
        cell = np.array([[ax, 0, 0], [bx, by, 0], [cx, cy, cz]])
        if abs(np.linalg.det(cell)) > 0.5 * vol:
            break
    inv = np.linalg.inv(cell)
    pts = []
    tries = 0
    r2 = rmin * rmin
    shifts = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)])
    while len(pts) < natoms:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place atoms with rmin=%g" % rmin)
        f = rng.random(3)
        if pts:
            fp = np.array(pts)
            d = f[None, :] - fp
            d -= np.round(d)
            dd = (d[:, None, :] + shifts[None, :, :]) @ cell
            if np.min(np.einsum("ijk,ijk->ij", dd, dd)) < r2:
                continue
        pts.append(f)
    pos = np.array(pts) @ cell
    typ = np.ones(natoms, dtype=np.int32)
    if type_fraction is not None:
        nb = int(round(type_fraction * natoms))
        typ[rng.permutation(natoms)[:nb]] = 2
    del inv
    return Structure(cell, pos, typ)


# ----------------------------------------------------------------------------
# Potential files


def write_potential(path, elements, cutoff, model_type, maxp, maxl, gtinv_order, gtinv_maxl,
                    gtinv_sym, coeffs, params, masses, scales=None):
    """Write a potential in the reference format (mlip_model.h:45-101)."""
    coeffs = np.asarray(coeffs, dtype=np.float64)
    if scales is None:
        scales = np.ones_like(coeffs)
    with open(path, "w") as fh:
        fh.write("%s # element\n" % " ".join(elements))
        fh.write("%r # cutoff\n" % float(cutoff))
        fh.write("gaussian # pair_type\n")
        fh.write("gtinv # des_type\n")
        fh.write("%d # model_type\n" % model_type)
        fh.write("%d # max_p\n" % maxp)
        fh.write("%d # max_l\n" % maxl)
        fh.write("%d # gtinv_order\n" % gtinv_order)
        fh.write("%s  # gtinv_maxl\n" % " ".join(str(v) for v in gtinv_maxl))
        fh.write("%s  # gtinv_sym\n" % " ".join(str(int(v)) for v in gtinv_sym))
        fh.write("%d # number of regression coefficients\n" % coeffs.size)
        fh.write(" ".join("%.17g" % v for v in coeffs) + "\n")
        fh.write(" ".join("%.17g" % v for v in scales) + "\n")
        fh.write("%d # number of parameters\n" % len(params))
        for b, m in params:
            fh.write("%.17g %.17g # pair func. params\n" % (b, m))
        fh.write(" ".join("%.17g" % v for v in masses) + " # atomic mass\n")


# ----------------------------------------------------------------------------
# Neighbor lists


class NeighborList:
    """LAMMPS-shaped neighbor data: x holds nlocal owned atoms then ghosts;
    ``off``/``idx`` flatten ``firstneigh`` per ``ilist`` entry."""

    def __init__(self, x, tag, types, nlocal, ilist, numneigh, off, idx, full, image=None,
                 owner=None, shift=None):
        self.x = x
        self.tag = tag
        self.types = types
        self.nlocal = nlocal
        self.nall = x.shape[0]
        self.ilist = ilist
        self.numneigh = numneigh
        self.off = off
        self.idx = idx
        self.full = full
        self.owner = owner    # owned index of every slot (ghost -> its atom)
        self.shift = shift    # per-slot displacement from its owner

    @property
    def inum(self):
        return self.ilist.size

    @property
    def npairs(self):
        return int(self.idx.size)


def _lex_positive(n):
    return (n[:, 0] > 0) | ((n[:, 0] == 0) & ((n[:, 1] > 0) | ((n[:, 1] == 0) & (n[:, 2] > 0))))


def build_neighbor_list(struct, rcut, full=False, skin=0.0):
    """Periodic neighbor list with ghost images.

    Pairs with r < rcut + skin are listed (the pair style applies its own
    strict r < cutoff test).  Half lists keep each physical pair once: local
    pairs with i < j and, for periodic images, the image with a
    lexicographically positive lattice shift.  Full lists keep both
    directions."""
    pos = struct.wrapped()
    n = pos.shape[0]
    cell = struct.cell
    rl = rcut + skin
    inv = np.linalg.inv(cell)
    # number of images needed along each lattice vector: distance between
    # opposite faces is V / |b x c| etc.
    vol = abs(np.linalg.det(cell))
    widths = np.array([vol / np.linalg.norm(np.cross(cell[1], cell[2])),
                       vol / np.linalg.norm(np.cross(cell[2], cell[0])),
                       vol / np.linalg.norm(np.cross(cell[0], cell[1]))])
    nrep = np.ceil(rl / widths).astype(int)
    shifts = np.array([[i, j, k] for i in range(-nrep[0], nrep[0] + 1)
                       for j in range(-nrep[1], nrep[1] + 1)
                       for k in range(-nrep[2], nrep[2] + 1)])
    del inv
    rl2 = rl * rl
    pair_i, pair_j, pair_s = [], [], []
    chunk = max(1, 4_000_000 // max(1, n * len(shifts)))
    svec = shifts @ cell
    for i0 in range(0, n, chunk):
        i1 = min(n, i0 + chunk)
        d = pos[i0:i1, None, None, :] - pos[None, :, None, :] - svec[None, None, :, :]
        r2 = np.einsum("abcd,abcd->abc", d, d)
        ii, jj, ss = np.nonzero(r2 < rl2)
        ii = ii + i0
        keep = ~((ii == jj) & np.all(shifts[ss] == 0, axis=1))
        ii, jj, ss = ii[keep], jj[keep], ss[keep]
        if not full:
            zero = np.all(shifts[ss] == 0, axis=1)
            sel = (zero & (ii < jj)) | (~zero & _lex_positive(shifts[ss]))
            ii, jj, ss = ii[sel], jj[sel], ss[sel]
        pair_i.append(ii)
        pair_j.append(jj)
        pair_s.append(ss)
    pi = np.concatenate(pair_i)
    pj = np.concatenate(pair_j)
    ps = np.concatenate(pair_s)
    order = np.lexsort((ps, pj, pi))
    pi, pj, ps = pi[order], pj[order], ps[order]
    # ghost slots: unique (j, shift) with nonzero shift.  Neighbor position
    # is x_j + shift, so the displacement x_i - x_neighbor matches
    # d = x_i - x_j - shift above.
    zero_shift = int(np.nonzero(np.all(shifts == 0, axis=1))[0][0])
    is_ghost = ps != zero_shift
    key = pj[is_ghost].astype(np.int64) * len(shifts) + ps[is_ghost]
    ukey, inv_key = np.unique(key, return_inverse=True)
    ghost_atom = (ukey // len(shifts)).astype(np.int64)
    ghost_shift = shifts[ukey % len(shifts)]
    nghost = ukey.size
    slot = pj.astype(np.int64).copy()
    slot[is_ghost] = n + inv_key
    x_all = np.concatenate([pos, pos[ghost_atom] + ghost_shift @ cell]) if nghost else pos.copy()
    tag = np.concatenate([np.arange(1, n + 1), ghost_atom + 1]).astype(np.int64)
    types = np.concatenate([struct.types, struct.types[ghost_atom]]).astype(np.int32)
    owner = np.concatenate([np.arange(n), ghost_atom]).astype(np.int64)
    shift = np.concatenate([np.zeros((n, 3)), ghost_shift @ cell]) if nghost else np.zeros((n, 3))
    numneigh = np.bincount(pi, minlength=n).astype(np.int32)
    off = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(numneigh, out=off[1:])
    ilist = np.arange(n, dtype=np.int32)
    return NeighborList(np.ascontiguousarray(x_all), tag, types, n, ilist, numneigh, off,
                        slot.astype(np.int32), full, owner=owner, shift=shift)


# Benchmark workload models (the potential-file half of each configuration).
CONFIG_MODELS = {
    # 256-atom single-element fcc: cutoff 6, 10 Gaussians on [0,5], gtinv
    # order 3 with maxl [8,4], linear model, coefficients N(0,1e-3), seed 2
    "fcc256_order3_linear": dict(elements=["X"], masses=[26.98], cutoff=6.0, model_type=2, maxp=1, maxl=8,
                                 gtinv_order=3, gtinv_maxl=[8, 4], gtinv_sym=[0, 0],
                                 centers=(0.0, 5.0, 10), beta=1.0, sigma=1e-3, seed=2),
    # 256,000-atom binary fcc: cutoff 8, 12 Gaussians on [0,7], gtinv order 4
    # with maxl [12,8,2], quadratic polynomial in the order<=2 invariants
    # (model_type 4, max_p 2), coefficients N(0,1e-4), seed 5
    "binary_order4_quadratic": dict(elements=["Ti", "Al"], masses=[47.87, 26.98], cutoff=8.0, model_type=4,
                                    maxp=2, maxl=12, gtinv_order=4, gtinv_maxl=[12, 8, 2], gtinv_sym=[0, 0, 0],
                                    centers=(0.0, 7.0, 12), beta=1.0, sigma=1e-4, seed=5),
    # the whole gtinv table: invariants of order 2..6 with gtinv_maxl
    # [20, 10, 2, 1, 1] (l up to 20, the order-5 and order-6 families), 20
    # Gaussians, two species; linear (model_type 2, max_p 1) and quadratic in
    # the order<=2 invariants (model_type 4, max_p 2)
    "binary_order6_l20_linear": dict(elements=["Ti", "Al"], masses=[47.87, 26.98], cutoff=6.0, model_type=2,
                                     maxp=1, maxl=20, gtinv_order=6, gtinv_maxl=[20, 10, 2, 1, 1],
                                     gtinv_sym=[0, 0, 0, 0, 0], centers=(0.0, 5.0, 20), beta=1.0, sigma=1e-3,
                                     seed=21),
    "binary_order6_l20_quadratic": dict(elements=["Ti", "Al"], masses=[47.87, 26.98], cutoff=6.0, model_type=4,
                                        maxp=2, maxl=20, gtinv_order=6, gtinv_maxl=[20, 10, 2, 1, 1],
                                        gtinv_sym=[0, 0, 0, 0, 0], centers=(0.0, 5.0, 20), beta=1.0, sigma=1e-4,
                                        seed=22),
    # order 3 with l up to 10: the order-3 general-loop kernel (signatures
    # beyond the unrolled set)
    "binary_order3_l10": dict(elements=["Ti", "Al"], masses=[47.87, 26.98], cutoff=6.0, model_type=2, maxp=1,
                              maxl=12, gtinv_order=3, gtinv_maxl=[12, 10], gtinv_sym=[0, 0], centers=(0.0, 5.0, 10),
                              beta=1.0, sigma=1e-3, seed=23),
}


def write_config_potential(path, name, n_coeff):
    """Write the named benchmark potential with seeded N(0, sigma^2)
    coefficients; n_coeff comes from the model enumeration."""
    c = CONFIG_MODELS[name]
    lo, hi, n = c["centers"]
    params = [(c["beta"], mu) for mu in np.linspace(lo, hi, n)]
    rng = np.random.default_rng(c["seed"])
    coeffs = rng.normal(0.0, c["sigma"], size=n_coeff)
    write_potential(path, c["elements"], c["cutoff"], c["model_type"], c["maxp"], c["maxl"], c["gtinv_order"],
                    c["gtinv_maxl"], c["gtinv_sym"], coeffs, params, c["masses"])
    return path


def config_potential(path, name, count_fn):
    """Write the named potential; count_fn(path_of_probe_file) -> n_coeff."""
    c = CONFIG_MODELS[name]
    lo, hi, n = c["centers"]
    params = [(c["beta"], mu) for mu in np.linspace(lo, hi, n)]
    # a one-coefficient probe file lets the enumeration report the count
    probe = path + ".probe"
    write_potential(probe, c["elements"], c["cutoff"], c["model_type"], c["maxp"], c["maxl"], c["gtinv_order"],
                    c["gtinv_maxl"], c["gtinv_sym"], [0.0], params, c["masses"])
    n_coeff = count_fn(probe)
    os.remove(probe)
    return write_config_potential(path, name, n_coeff)


def elem_of_type(pot_elements, type_elements):
    """Map LAMMPS types (1-based) to potential element indices, as pair_coeff
    does (pair_mlip_gtinv.cpp:573-582)."""
    out = []
    for name in type_elements:
        if name not in pot_elements:
            raise ValueError("element %s not in potential file" % name)
        out.append(pot_elements.index(name))
    return np.array(out, dtype=np.int32)


def reference_data_path(*parts):
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), *parts)


def resident_lists(nl):
    """Convert a full LAMMPS-shaped list into the device-resident form: for
    every owned atom, neighbor indices into the owned atoms plus the image
    shift s with del = x_i - x_j + s (s = x_owner(j) - x_slot(j))."""
    if not nl.full:
        raise ValueError("resident lists need a full neighbor list")
    off = nl.off.astype(np.int64)
    slots = nl.idx.astype(np.int64)
    idx = nl.owner[slots].astype(np.int32)
    shift = -nl.shift[slots]
    return off, idx, np.ascontiguousarray(shift)


def random_structures(nstruct, nmin, nmax, seed, density=0.06, rmin=1.8):
    """Batched/served workload structures from the native generator
    (mlip_random_structures): log-uniform sizes, random triclinic cells at
    `density`, rmin rejection, two species with a random fraction."""
    import ctypes
    from ._lib import library
    lib = library()
    nat = np.zeros(nstruct, dtype=np.int32)
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
    rc = lib.mlip_random_structures(nstruct, nmin, nmax, density, rmin, seed, P(nat, ctypes.c_int), None, None,
                                    None, 0)
    if rc != 0:
        raise ValueError("bad workload parameters")
    tot = int(nat.sum())
    cells = np.zeros((nstruct, 9))
    x = np.zeros((tot, 3))
    t = np.zeros(tot, dtype=np.int32)
    rc = lib.mlip_random_structures(nstruct, nmin, nmax, density, rmin, seed, P(nat, ctypes.c_int),
                                    P(cells, ctypes.c_double), P(x, ctypes.c_double), P(t, ctypes.c_int), tot)
    if rc != 0:
        raise RuntimeError("rejection sampling did not finish")
    out, at = [], 0
    for k in range(nstruct):
        n = int(nat[k])
        out.append(Structure(cells[k].reshape(3, 3), x[at:at + n], t[at:at + n]))
        at += n
    return out
