"""The C ABI of libobme_hip.so without a GPU: the library loads, exports
every function include/obme.h declares, its struct layouts match the ctypes
mirrors, and configuration errors come back as status codes with a message
before any device call."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oceanbiome import abi, runtime
from oceanbiome.grid import RectilinearGrid
from oceanbiome.models import LOBSTER, NPZD, CarbonateSystem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "obme.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(obme_[a-z0-9_]+)\s*\(", text)))


def lib():
    if not os.path.exists(runtime.LIB_PATH):
        pytest.fail(f"{runtime.LIB_PATH} is not built (run __graft_entry__.build())")
    return runtime.library()


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in runtime.EXPORTED_SYMBOLS:
        assert n in names, n


def test_library_exports_every_declared_function():
    L = lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match():
    assert runtime.abi_layout_mismatches() == []


def test_version_string():
    v = lib().obme_version().decode()
    assert v and "gfx950" in v


def _create(cfg):
    L = lib()
    h = C.c_void_p()
    st = L.obme_create(C.byref(cfg), C.byref(h))
    return st, h, L.obme_last_error().decode()


def _cfg(bgc, precision=abi.OBME_FP64):
    return bgc.config(precision)


def test_invalid_configs_are_rejected_before_device_work():
    g = RectilinearGrid(size=(4, 4, 8), extent=(1, 1, 8))
    b = NPZD(g)
    cfg = _cfg(b)
    cfg.model = 99
    st, h, msg = _create(cfg)
    assert st == abi.OBME_ERR_INVALID and "model" in msg and not h.value

    cfg = _cfg(b)
    cfg.precision = 7
    st, _, msg = _create(cfg)
    assert st == abi.OBME_ERR_INVALID and "precision" in msg

    cfg = _cfg(b)
    cfg.carbonate_system = 1
    st, _, msg = _create(cfg)
    assert st == abi.OBME_ERR_UNSUPPORTED

    cfg = _cfg(b)
    cfg.sinking.scheme = 7
    st, _, msg = _create(cfg)
    assert st == abi.OBME_ERR_INVALID and "sinking scheme" in msg

    cfg = _cfg(NPZD(g, scale_negatives=True))
    cfg.sinking.scheme = abi.OBME_WENO5
    st, _, msg = _create(cfg)
    assert st == abi.OBME_ERR_UNSUPPORTED and "without negative scaling" in msg

    cfg = _cfg(b)
    cfg.grid.Nz = 0
    st, _, msg = _create(cfg)
    assert st == abi.OBME_ERR_INVALID

    cfg = _cfg(LOBSTER(g, carbonate_system=CarbonateSystem()))
    cfg.sinking.speed[0] = 1e-4   # NO3 does not sink
    st, _, msg = _create(cfg)
    assert st == abi.OBME_ERR_UNSUPPORTED and "drift" in msg


def test_null_arguments():
    L = lib()
    h = C.c_void_p()
    assert L.obme_create(None, C.byref(h)) == abi.OBME_ERR_INVALID
    assert L.obme_destroy(None) == abi.OBME_OK


def test_required_tracer_counts():
    g = RectilinearGrid(size=(2, 2, 4), extent=(1, 1, 4))
    assert NPZD(g).required_biogeochemical_tracers() == ("N", "P", "Z", "T", "D")
    assert LOBSTER(g).required_biogeochemical_tracers() == ("NO₃", "NH₄", "P", "Z", "sPOM", "bPOM", "DOM")
    assert LOBSTER(g, carbonate_system=CarbonateSystem()).required_biogeochemical_tracers()[-2:] == ("DIC", "Alk")
