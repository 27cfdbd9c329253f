"""Shared helpers for the test suite: fixture paths and benchmark-model
potential files written on demand into a temporary directory."""
import os
import tempfile

import numpy as np

from mlip_hip import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = os.path.join(GOLD, "reference")
_TMP = tempfile.mkdtemp(prefix="mlip_tests_")


def ref(name):
    return os.path.join(REF, name)


def config_potential(name):
    """Path of the named benchmark potential (workloads.CONFIG_MODELS)."""
    from mlip_hip import inspect_model
    path = os.path.join(_TMP, name + ".mlip")
    if not os.path.exists(path):
        W.config_potential(path, name, lambda probe: inspect_model(probe, 4))
    return path


def rel_force_err(f, f_ref):
    scale = max(np.abs(f_ref).max(), 1e-300)
    return np.abs(f - f_ref).max() / scale
