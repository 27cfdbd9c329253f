"""The C-ABI library loads without a GPU and exports every function that
include/mlip_hip.h declares (no compute call is made here)."""
import ctypes
import os
import re

from mlip_hip import LIB_PATH

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mlip_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mlip_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(LIB_PATH)
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_no_device_is_an_error_not_a_fallback():
    import pytest
    import torch  # noqa: F401  (plumbing check only)
    if torch.cuda.is_available():
        pytest.skip("device present")
    from mlip_hip import MlipError, PairMLIPGtinv
    from common import ref
    p = PairMLIPGtinv()
    with pytest.raises(MlipError):
        p.coeff("*", "*", ref("gtinv-434.mlip"), "Ti", "Al")
