"""The LAMMPS adapter compiles against the LAMMPS pair-style interface and
runs the calls of a LAMMPS run on the stand-in classes of
tests/lammps_harness (no GPU here: the engine's error reaches LAMMPS's
error->all).  The stand-in's virial flags are checked against what LAMMPS
does to a style that neither sets no_virial_fdotr_compute nor calls
virial_fdotr_compute: with newton on (vflag = VIRIAL_FDOTR) its global
virial is switched off, which is why the adapter sets the flag."""
import numpy as np
import pytest

import adapter_harness as H
from common import ref
from mlip_hip import workloads as W


def small_list():
    s = W.read_lammps_data(ref("structure.lammps"))
    return W.build_neighbor_list(s, 6.0, full=True)


def test_harness_builds():
    assert H.ensure_built()


def test_fdotr_flags_of_a_plain_style(tmp_path):
    nl = small_list()
    rc, err, out = H.run(tmp_path, "fdotr_probe", ref("gtinv-820.mlip"), ["Ti", "Al"], nl,
                         H.ENERGY_GLOBAL | H.ENERGY_ATOM, H.VIRIAL_FDOTR | H.VIRIAL_ATOM)
    assert rc == 0, err
    assert out["vflag_fdotr"] == 1 and out["vflag_global"] == 0
    assert out["eflag_atom"] and out["vflag_atom"]
    # a style without init_style of its own asks for the default half list
    assert out["request_half"] == 1 and out["request_full"] == 0
    rc, err, out = H.run(tmp_path, "fdotr_probe", ref("gtinv-820.mlip"), ["Ti", "Al"], nl,
                         H.ENERGY_GLOBAL, H.VIRIAL_PAIR)
    assert rc == 0, err
    assert out["vflag_fdotr"] == 0 and out["vflag_global"] == H.VIRIAL_PAIR


def test_adapter_without_device_stops_with_lammps_error(tmp_path):
    if H.hip_device_count() > 0:
        pytest.skip("device present")
    nl = small_list()
    rc, err, out = H.run(tmp_path, "mlip_gtinv/hip", ref("gtinv-820.mlip"), ["Ti", "Al"], nl,
                         H.ENERGY_GLOBAL, H.VIRIAL_FDOTR)
    assert rc == 3 and out is None
    assert err.startswith("ERROR:"), err


def test_adapter_reports_a_missing_potential_file(tmp_path):
    nl = small_list()
    rc, err, out = H.run(tmp_path, "mlip_gtinv/hip", ref("no-such-file.mlip"), ["Ti", "Al"], nl,
                         H.ENERGY_GLOBAL, H.VIRIAL_FDOTR)
    assert rc == 3 and out is None
    assert err.startswith("ERROR:") and "no-such-file.mlip" in err, err
