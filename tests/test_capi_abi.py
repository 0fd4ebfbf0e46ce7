"""The C ABI library loads and exports every entry point include/tdbem.h declares; no compute
calls without a GPU.  The product package never touches the oracle."""
import ctypes
import re
from pathlib import Path

import pytest

from tdbem import api

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "tdbem.h").read_text()
    return sorted(set(re.findall(r"\b(tdbem_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(api.LIB_PATH))
    syms = declared_symbols()
    assert len(syms) >= 14
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_create_without_gpu_fails_loudly():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(api.TdbemError):
        api.Context([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], [[0, 1, 2, 3]], E=1, nu=0.25,
                    rho=1, dt=0.1, numstep=3)


def test_element_create_without_gpu_fails_loudly():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(api.TdbemError):
        api.Element(1, api.KERNEL_STEP, 6, E=1.0, nu=0.25, rho=1.0)


def test_product_does_not_reference_the_oracle():
    pkg = ROOT / "td-bem_amd"
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.h")) + \
            list(pkg.rglob("*.cpp")) + [pkg / "Makefile"]:
        text = f.read_text()
        assert "oracle" not in text.lower(), f
