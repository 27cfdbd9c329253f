"""Device intermediates against the oracle's: a_nlm of every centre from both
a_nlm kernels (register accumulation and the matrix-core chunk products), in
FP64 and mixed precision, on the benchmark model (two species, l <= 12) and a
fitted regression potential.  Bars: FP64 1e-13 of the largest |a_nlm| of
the structure; mixed 1e-5 of it: the mixed mode forms direction cosines and
the Y_lm recursion (up to 12 steps at l <= 12) in FP32, unit roundoff 6e-8,
and sums about 130 such terms per a_nlm, which puts its error near 1e-6 of
the largest a_nlm (the energy and force bars of the mixed mode are tested
on the evaluation results).  Both kernels must also agree with each other
to the same bars."""
import numpy as np
import pytest

import oracle_lib as O
from common import config_potential, ref
from mlip_hip import PairMLIPGtinv
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu


def device_anlm(p, elem_of_atom):
    """a_nlm [atom][s][n][lm (m <= 0)] from the slot-blocked device layout
    (centres sorted by species, each species padded to blocks of 64)."""
    info = p.model_info()
    nf, nt, maxl = info["n_fn"], info["n_type"], info["maxl"]
    nlmh = (maxl + 1) * (maxl + 2) // 2
    buf = p.debug_fetch(0, 1 << 27)
    nblk = buf.size // (nf * nt * nlmh * 2 * 64)
    a = buf.reshape(nblk, nf, nt, nlmh, 2, 64)
    slots = []
    for t in range(nt):
        mem = list(np.nonzero(elem_of_atom == t)[0])
        slots += mem + [-1] * ((-len(mem)) % 64)
    out = np.zeros((len(elem_of_atom), nt, nf, nlmh), dtype=complex)
    for sl, atom in enumerate(slots):
        if atom < 0:
            continue
        v = a[sl // 64, :, :, :, :, sl % 64]          # [n][s][lm][c]
        out[atom] = (v[..., 0] + 1j * v[..., 1]).transpose(1, 0, 2)
    return out


def oracle_anlm(o, s, maxl):
    ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1, want_anlm=True)
    full = ro["anlm"]                                  # [atom][s][n][(l, m) all m]
    idx = [l * l + l + m for l in range(maxl + 1) for m in range(-l, 1)]
    return full[..., idx]


CASES = [("binary_order4_quadratic", lambda: W.fcc(3, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)),
         ("gtinv-820.mlip", lambda: W.read_lammps_data(ref("structure.lammps")))]


@pytest.mark.parametrize("name,make", CASES)
def test_anlm_kernels_match_oracle(name, make):
    pot = ref(name) if name.endswith(".mlip") else config_potential(name)
    s = make()
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, "Ti", "Al")
    info = p.model_info()
    ao = oracle_anlm(O.Oracle(pot), s, info["maxl"])
    scale = np.abs(ao).max()
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    res = {}
    for prec in ("fp64", "mixed"):
        p.set_precision(prec)
        for kind in ("registers", "mfma"):
            p.set_anlm_kernel(kind)
            p.compute(nl, nl.types)
            a = device_anlm(p, s.types - 1)
            res[(prec, kind)] = a
            err = np.abs(a - ao).max() / scale
            print("%s %s %s: max |a - oracle| / max |a| = %.2e" % (name, prec, kind, err), flush=True)
            assert err <= (1e-13 if prec == "fp64" else 1e-5), (name, prec, kind)
        d = np.abs(res[(prec, "mfma")] - res[(prec, "registers")]).max() / scale
        print("%s %s: max |mfma - registers| / max |a| = %.2e" % (name, prec, d), flush=True)
        assert d <= (1e-13 if prec == "fp64" else 1e-5)
    p.set_precision("fp64")
