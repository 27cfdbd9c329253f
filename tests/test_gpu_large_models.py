"""GPU parity on models beyond the reference's fixtures, against the CPU
oracle (itself pinned for them by rotation invariance of every feature and
F = -grad E, tests/test_oracle_large.py):

* the whole gtinv table, gtinv_order 6 with gtinv_maxl [20, 10, 2, 1, 1]
  (order-2 families up to l = 20, order-3 up to l = 10, the order-5 and
  order-6 families), 20 Gaussians, two species, linear (model_type 2) and
  quadratic in the order<=2 invariants (model_type 4, 1.6 M coefficients);
* order 3 with gtinv_maxl [12, 10] (the order-3 general-loop kernel).

Bars: FP64 energy 1e-10 eV/atom, forces 1e-8 of max |F|, virial 1e-8
relative; mixed precision 1e-6 eV/atom and 1e-4 eV/A."""
import numpy as np
import pytest

import oracle_lib as O
from common import config_potential, ref, rel_force_err
from mlip_hip import PairMLIPGtinv
from mlip_hip import workloads as W

pytestmark = pytest.mark.gpu

MODELS = ["binary_order6_l20_linear", "binary_order6_l20_quadratic", "binary_order3_l10"]

_pairs = {}


def gpu_pair(name):
    if name not in _pairs:
        p = PairMLIPGtinv()
        p.settings()
        p.coeff("*", "*", config_potential(name), "Ti", "Al")
        _pairs[name] = p
    return _pairs[name]


def structures():
    return [("TiAl3-32", W.read_lammps_data(ref("structure.lammps"))),
            ("fcc32", W.fcc(2, 4.05, 0.15, 2, type_fraction=0.4, type_seed=3))]


@pytest.mark.parametrize("name", MODELS)
def test_large_model_parity(name):
    o = O.Oracle(config_potential(name))
    p = gpu_pair(name)
    for tag, s in structures():
        ro = o.compute(W.build_neighbor_list(s, o.cutoff), s.types - 1)
        for full in (False, True):
            nl = W.build_neighbor_list(s, p.init_one(), full=full)
            p.set_precision("fp64")
            r = p.compute(nl, nl.types)
            de = abs(r["energy"] - ro["energy"]) / s.natoms
            fe = rel_force_err(r["forces"], ro["forces"])
            ve = np.abs(r["virial"] - ro["virial"]).max() / np.abs(ro["virial"]).max()
            p.set_precision("mixed")
            rm = p.compute(nl, nl.types)
            p.set_precision("fp64")
            dem = abs(rm["energy"] - ro["energy"]) / s.natoms
            fem = np.abs(rm["forces"] - ro["forces"]).max()
            print("%s %s full=%d: FP64 dE/atom=%.2e relF=%.2e relV=%.2e | mixed dE/atom=%.2e dF=%.2e eV/A"
                  % (name, tag, full, de, fe, ve, dem, fem), flush=True)
            assert de <= 1e-10 and fe <= 1e-8 and ve <= 1e-8, (name, tag, full)
            assert dem <= 1e-6 and fem <= 1e-4, (name, tag, full)
