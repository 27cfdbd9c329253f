"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a transposed operand) fails at least one of them:

* the reference's own unit-test known answers (test/test_mlip_features.cpp);
* the feature counts of test/regression/README.md:11-17 (index enumeration);
* the step-0 energy and all six pressure components of the regression logs
  (the CPU pair style itself, 8 significant digits);
* scipy spherical harmonics at general angles (tests/golden/ylm_scipy.json);
* F = -dE/dx by central differences, sum F = 0, rotation/translation
  invariance of E, virial against a finite homogeneous strain.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from mlip_hip import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF = os.path.join(GOLD, "reference")
NKTV2P = 1.6021765e6  # LAMMPS metal units: eV/A^3 -> bar


def goldens():
    return json.load(open(os.path.join(GOLD, "regression_goldens.json")))


def run_case(pot, struct_file, elements, full=False):
    o = O.Oracle(os.path.join(REF, pot))
    s = W.read_lammps_data(os.path.join(REF, struct_file))
    nl = W.build_neighbor_list(s, o.cutoff, full=full)
    em = W.elem_of_type(["Ti", "Al"], elements)
    return o, s, nl, o.compute(nl, em[s.types - 1])


# ---------------------------------------------------------------- unit KATs
def test_cartesian_to_spherical_kat():
    th, ph = O.cart_to_sph(1.0, 1.0, 0.0)  # test_mlip_features.cpp:6-11
    assert th == pytest.approx(math.pi / 2, rel=0, abs=1e-15)
    assert ph == pytest.approx(math.pi / 4, rel=0, abs=1e-15)


def test_ylm_theta0_kat():
    y, _ = O.ylm(2, 0.0, math.pi)  # test_mlip_features.cpp:27-42
    expect = [math.sqrt(1 / (4 * math.pi)), 0, math.sqrt(3 / (4 * math.pi)), 0, 0,
              math.sqrt(5 / (4 * math.pi))]
    np.testing.assert_allclose(y.real, expect, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(y.imag, 0.0, atol=1e-15)


def test_ylm_dtheta_kat():
    _, d = O.ylm(1, math.pi / 2, math.pi)  # test_mlip_features.cpp:44-59
    np.testing.assert_allclose(d.real, [0, 0, -math.sqrt(3 / (4 * math.pi))], atol=1e-15)
    np.testing.assert_allclose(d.imag, 0.0, atol=1e-15)


def test_fn_kat():
    fn, _ = O.fn(4.0, [[math.log(2), 0.0], [math.log(2), 2.0]], 2.0)  # :80-101
    np.testing.assert_allclose(fn, [0.03125, 0.5], rtol=1e-15)


def test_fn_derivative_fd():
    params = [[1.0, 0.0], [0.7, 2.5], [1.3, 5.0]]
    for r in (0.9, 2.2, 5.7):
        fn0, fd = O.fn(6.0, params, r)
        h = 1e-6
        fp, _ = O.fn(6.0, params, r + h)
        fm, _ = O.fn(6.0, params, r - h)
        np.testing.assert_allclose(fd, (fp - fm) / (2 * h), rtol=1e-7, atol=1e-10)


def test_ylm_against_scipy():
    g = json.load(open(os.path.join(GOLD, "ylm_scipy.json")))
    for d, re, im in zip(g["dirs"], g["ylm_re"], g["ylm_im"]):
        th, ph = O.cart_to_sph(*d)
        y, _ = O.ylm(g["lmax"], th, ph)
        np.testing.assert_allclose(y.real, re, atol=2e-14)
        np.testing.assert_allclose(y.imag, im, atol=2e-14)


def test_ylm_dtheta_fd():
    for th, ph in ((0.7, 2.1), (2.3, -1.2), (1.4, 0.3)):
        y, d = O.ylm(8, th, ph)
        h = 1e-6
        yp, _ = O.ylm(8, th + h, ph)
        ym, _ = O.ylm(8, th - h, ph)
        np.testing.assert_allclose(d, (yp - ym) / (2 * h), atol=1e-8)


# ---------------------------------------------------------------- indexing
@pytest.mark.parametrize("name", ["gtinv-197", "gtinv-820", "gtinv-434"])
def test_feature_counts(name):
    g = goldens()
    o = O.Oracle(os.path.join(REF, g[name]["potential"]))
    assert o.info["n_coeff"] == g["feature_counts"][name]


# ---------------------------------------------------------------- goldens
@pytest.mark.parametrize("name", ["gtinv-197", "gtinv-434", "gtinv-820", "gtinv-197-conv64"])
def test_step0_regression_goldens(name):
    g = goldens()[name]
    o, s, nl, r = run_case(g["potential"], g["structure"], g["elements"])
    st = g["step0"]
    assert round(r["energy"], 5) == pytest.approx(st["E_pair"], abs=1.1e-5)
    v = r["virial"] / s.volume * NKTV2P
    assert (v[0] + v[1] + v[2]) / 3 == pytest.approx(st["Press"], abs=2e-4)
    for k, key in enumerate(["Pxx", "Pyy", "Pzz", "Pxy", "Pxz", "Pyz"]):
        if key in st:
            assert v[k] == pytest.approx(st[key], rel=1e-7, abs=1e-4)
    if "Fmax" in st:
        assert np.abs(r["forces"]).max() < 1e-12


def test_triclinic_equals_conventional():
    g = goldens()
    c = g["gtinv-197-conv64"]
    t = g["gtinv-197-tricli64"]
    _, sc, _, rc = run_case(c["potential"], c["structure"], c["elements"])
    _, st, _, rt = run_case(t["potential"], t["structure"], t["elements"])
    pc = rc["virial"][:3].sum() / 3 / sc.volume * NKTV2P
    pt = rt["virial"][:3].sum() / 3 / st.volume * NKTV2P
    assert abs(pc - pt) <= t["press_delta"]
    assert rc["energy"] == pytest.approx(rt["energy"], abs=1e-5)


# ---------------------------------------------------------------- physics
def small_system(seed=3, n=2, a=4.05, sigma=0.12):
    s = W.fcc(n, a, sigma, seed, type_fraction=0.4, type_seed=seed + 1)
    return s


def energy(o, s):
    nl = W.build_neighbor_list(s, o.cutoff)
    return o.compute(nl, s.types - 1)


def test_forces_are_minus_gradient():
    o = O.Oracle(os.path.join(REF, "gtinv-820.mlip"))
    s = small_system()
    r = energy(o, s)
    rng = np.random.default_rng(0)
    # E carries ~1e-12 eV of summation noise; h = 1e-4 keeps both the
    # truncation (h^2) and the noise (1e-12/h) terms near 1e-8 eV/A.
    h = 1e-4
    for i in rng.choice(s.natoms, 4, replace=False):
        for k in range(3):
            sp = W.Structure(s.cell, s.positions.copy(), s.types)
            sp.positions[i, k] += h
            sm = W.Structure(s.cell, s.positions.copy(), s.types)
            sm.positions[i, k] -= h
            fd = -(energy(o, sp)["energy"] - energy(o, sm)["energy"]) / (2 * h)
            assert r["forces"][i, k] == pytest.approx(fd, rel=1e-6, abs=1e-7)


def test_sum_forces_zero_and_invariance():
    o = O.Oracle(os.path.join(REF, "gtinv-197.mlip"))
    s = small_system(seed=5)
    r = energy(o, s)
    np.testing.assert_allclose(r["forces"].sum(0), 0.0, atol=1e-10)
    # rigid rotation of the whole periodic system
    ang = 0.37
    R = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0], [0, 0, 1]])
    R = R @ np.array([[1, 0, 0], [0, math.cos(0.8), -math.sin(0.8)], [0, math.sin(0.8), math.cos(0.8)]])
    sr = W.Structure(s.cell @ R.T, s.positions @ R.T, s.types)
    rr = energy(o, sr)
    assert rr["energy"] == pytest.approx(r["energy"], rel=1e-12)
    np.testing.assert_allclose(rr["forces"], r["forces"] @ R.T, atol=1e-9)
    st = W.Structure(s.cell, s.positions + np.array([0.3, -1.1, 2.2]), s.types)
    assert energy(o, st)["energy"] == pytest.approx(r["energy"], rel=1e-12)


def test_virial_is_strain_derivative():
    o = O.Oracle(os.path.join(REF, "gtinv-434.mlip"))
    s = small_system(seed=7)
    r = energy(o, s)
    h = 1e-5
    # W_ab = -dE/de_ab under x -> (1+e) x for symmetric strain
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for k, (a, b) in enumerate(pairs):
        def e_at(eps):
            F = np.eye(3)
            F[a, b] += eps
            if a != b:
                F[b, a] += eps
            st = W.Structure(s.cell @ F.T, s.positions @ F.T, s.types)
            return energy(o, st)["energy"]
        dE = (e_at(h) - e_at(-h)) / (2 * h)
        scale = 1.0 if a == b else 2.0
        assert -dE / scale == pytest.approx(r["virial"][k], rel=1e-6, abs=1e-6)
