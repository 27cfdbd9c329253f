"""NVE on the CPU oracles against the reference's regression runs
(test/regression/gtinv-*.in and pair-*.in: fix nve, 100 steps from rest,
metal units).

Bar: thermo values within the last printed digit of log.lammps.* (8
significant digits), final positions within 1e-8 A of dump.atom.*."""
import json
import os

import numpy as np
import pytest

import md_oracle as MD
import oracle_lib as O
from common import GOLD, ref
from mlip_hip import workloads as W


def goldens():
    return json.load(open(os.path.join(GOLD, "regression_goldens.json")))


def digits_tol(value):
    """Half a unit in the 8th significant digit, as LAMMPS prints thermo."""
    if value == 0:
        return 1e-12
    return 0.51 * 10.0 ** (np.floor(np.log10(abs(value))) - 7)


def check_thermo(got, want, tag):
    for key, val in want.items():
        assert abs(got[key] - val) <= digits_tol(val), (tag, key, got[key], val)


def potential_masses(path):
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    return [float(t) for t in lines[-1].split("#")[0].split()]


@pytest.mark.parametrize("name", ["gtinv-197", "gtinv-434", "gtinv-820", "pair-17", "pair-60"])
def test_nve_100_steps_matches_log_and_dump(name):
    g = goldens()[name]
    s = W.read_lammps_data(ref(g["structure"]))
    o = O.Oracle(ref(g["potential"]))
    # pair_coeff comes after read_data: the potential file's masses are used
    mass = potential_masses(ref(g["potential"]))
    x, v, th = MD.run_nve(o, s, mass, 0.001, 100, record=(0, 100))
    check_thermo(th[0], {k: g["step0"][k] for k in ("E_pair", "Press")}, name + " step 0")
    check_thermo(th[100], g["step100_dt0.001"], name + " step 100")
    dump = MD.read_dump(ref(g["dump100_dt0.001"]))
    d = MD.periodic_diff(x, dump, s.cell)
    print("%s: max |x - dump| = %.2e A" % (name, np.abs(d).max()))
    assert np.abs(d).max() < 1e-8


def test_nve_stress_run_dt5fs():
    g = goldens()["gtinv-197"]
    s = W.read_lammps_data(ref(g["structure"]))
    o = O.Oracle(ref(g["potential"]))
    mass = potential_masses(ref(g["potential"]))
    _, _, th = MD.run_nve(o, s, mass, 0.005, 100, record=(50, 100))
    check_thermo(th[50], g["step50_dt0.005"], "step 50")
    check_thermo(th[100], g["step100_dt0.005"], "step 100")
