"""The C ABI library loads and exports every symbol include/gammcor_hip.h
declares (no compute calls, so this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(ROOT, "include", "gammcor_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(gmc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_drop_in_entry_points():
    names = declared()
    for n in ("gmc_twono1", "gmc_ab_cas", "gmc_erpasymm1", "gmc_aceneerpa", "gmc_aceinteg", "gmc_acecorr",
              "gmc_acecorr_dev", "gmc_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from gammcor_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) <= set(_lib.exported_symbols())


def test_version_without_gpu():
    from gammcor_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    assert lib.gmc_version() == 3
