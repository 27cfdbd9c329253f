"""Feature indexing of the HIP engine's model layer against the oracle,
bit-exact: feature families (irreps x type combinations, with their type
masks), quadratic and cubic product lists, and the coefficient count, for
the reference regression potentials and both benchmark models.  Runs on the
CPU (the enumeration needs no device)."""
import numpy as np
import pytest

import oracle_lib as O
from common import config_potential, ref
from mlip_hip import inspect_model

POTS = [ref("gtinv-197.mlip"), ref("gtinv-434.mlip"), ref("gtinv-820.mlip")]


def _pots():
    return POTS + [config_potential("fcc256_order3_linear"), config_potential("binary_order4_quadratic")]


@pytest.mark.parametrize("idx", range(5))
def test_tables_match_oracle(idx):
    pot = _pots()[idx]
    o = O.Oracle(pot)
    info = inspect_model(pot, 0)
    assert info["n_coeff"] == o.info["n_coeff"]
    assert info["n_gtinv"] == o.info["n_gtinv"]
    assert info["n_des"] == o.info["n_des"]
    fam = inspect_model(pot, 1)
    ofam = o.linear_terms()
    # oracle rows: irrep, order, tc[6], mask ; same layout
    np.testing.assert_array_equal(fam, ofam)
    np.testing.assert_array_equal(inspect_model(pot, 2), o.combs(2))
    np.testing.assert_array_equal(inspect_model(pot, 3).reshape(-1, 3), o.combs(3))


def test_benchmark_model_sizes():
    # 256k-atom model: order 4 [12,8,2], two species, quadratic in order<=2
    info = inspect_model(config_potential("binary_order4_quadratic"), 0)
    assert (info["n_irreps"], info["n_gtinv"], info["n_des"], info["n_coeff"]) == (82, 902, 10824, 239184)
    info1 = inspect_model(config_potential("fcc256_order3_linear"), 0)
    assert (info1["n_irreps"], info1["n_des"], info1["n_coeff"]) == (24, 240, 240)


def test_bad_potential_files(tmp_path):
    from mlip_hip import MlipError
    p = tmp_path / "bad.mlip"
    p.write_text("Ti Al\n")
    with pytest.raises(MlipError):
        inspect_model(str(p), 0)
    lines = open(ref("gtinv-434.mlip")).read().splitlines()
    lines[10] = "5461 # number of regression coefficients"
    p.write_text("\n".join(lines) + "\n")
    with pytest.raises(MlipError):
        inspect_model(str(p), 0)
    with pytest.raises(MlipError):
        inspect_model(str(tmp_path / "missing.mlip"), 0)


@pytest.mark.parametrize("name", ["pair-17", "pair-60"])
def test_pair_style_tables_match_oracle(name):
    """Radial-only style: feature count and the polynomial terms in
    coefficient order, bit-exact against the oracle's enumeration."""
    import oracle_lib as O
    from mlip_hip import inspect_model
    pot = ref(name + ".mlip")
    o = O.Oracle(pot)
    info = inspect_model(pot, 0)
    assert info["n_des"] == o.info["n_des"] and info["n_coeff"] == o.info["n_coeff"]
    for order in (2, 3):
        assert np.array_equal(inspect_model(pot, order), o.combs(order))
