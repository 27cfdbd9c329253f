"""The polynomial quadratic form Hq = D W_t on the matrix cores (gemm.hip,
FP64 MFMA) against numpy in FP64, species by species, on the benchmark
model (two species, 492-column forms) and on the reference's fitted
regression potentials.  For a model without cubic terms the GEMM's epilogue
writes H = lin + D W instead, which is checked against lin + D W.  Bar:
every entry within 1e-13 of |lin| + sum |D| |W| (the scale of its terms;
FP64 sums of a few hundred terms in another order)."""
import numpy as np
import pytest

from common import config_potential, ref
from mlip_hip import PairMLIPGtinv
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu


def check_forms(p, tag, fused=None):
    lay = p.debug_fetch(6, 8)
    n_pr = int(lay[0])
    if fused is not None:
        assert int(lay[7]) == fused, tag
    lin = p.debug_fetch(7, 2 * n_pr).reshape(2, n_pr) if int(lay[7]) else np.zeros((2, n_pr))
    D = p.debug_fetch(1, 1 << 28)
    Hq = p.debug_fetch(4, 1 << 28)
    forms = p.debug_fetch(5, 1 << 26)
    nslot = 64 * (int(lay[3]) + int(lay[6]))
    D = D[:nslot * n_pr].reshape(nslot, n_pr)
    Hq = Hq[:nslot * n_pr].reshape(nslot, n_pr)
    row0 = 0
    worst = 0.0
    for t in range(2):
        ofs, w, nb = int(lay[1 + 3 * t]), int(lay[2 + 3 * t]), int(lay[3 + 3 * t])
        rows = slice(row0, row0 + 64 * nb)
        row0 += 64 * nb
        if nb == 0 or w == 0:
            continue
        Wt = forms[ofs:ofs + w * w].reshape(w, w)
        assert np.array_equal(Wt, Wt.T)
        ref_ = lin[t, :w] + D[rows, :w] @ Wt
        scale = np.abs(lin[t, :w]) + np.abs(D[rows, :w]) @ np.abs(Wt) + 1e-300
        err = (np.abs(Hq[rows, :w] - ref_) / scale).max()
        worst = max(worst, err)
    print("%s: Hq max error / term scale = %.2e" % (tag, worst), flush=True)
    assert worst <= 1e-13, tag


def test_benchmark_model_forms():
    pot = config_potential("binary_order4_quadratic")
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", pot, "Ti", "Al")
    # 500 atoms: 4 slot blocks per species, a partial last block each
    s = W.fcc(5, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4)
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    p.compute(nl, nl.types)
    check_forms(p, "binary500", fused=1)


@pytest.mark.parametrize("pot", ["gtinv-197.mlip", "gtinv-434.mlip", "gtinv-820.mlip"])
def test_regression_forms(pot):
    p = PairMLIPGtinv()
    p.settings()
    p.coeff("*", "*", ref(pot), "Ti", "Al")
    s = W.read_lammps_data(ref("structure.lammps"))
    nl = W.build_neighbor_list(s, p.init_one(), full=True)
    p.compute(nl, nl.types)
    check_forms(p, pot)
