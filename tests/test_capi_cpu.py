"""CPU-side checks of the boundary: the HIP library loads and exports every
function include/pb_capi.h declares, the Python binding knows all of them,
and the config struct layout matches the header.  No compute call is made
(there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pb_capi.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pb_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("pb_create", "pb_destroy", "pb_set_state", "pb_get_state", "pb_substep", "pb_step",
                 "pb_wavespace_host", "pb_nonlinear_host", "pb_last_error", "pb_timing", "pb_timing_read"):
        assert must in names


def test_library_exports_every_symbol():
    from channel_passive_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_binding_covers_header():
    from channel_passive_amd import capi
    names = set(declared())
    assert names <= set(capi.SIGNATURES)


def test_config_layout_matches_header():
    """offsets of struct pb_config as the header lays it out"""
    from channel_passive_amd.config import PbConfig
    assert PbConfig.nx.offset == 0 and PbConfig.labs_k.offset == 24 and PbConfig.stretch.offset == 32
    assert PbConfig.rk_step.offset == 32 + 8 * 15 and ctypes.sizeof(PbConfig) == 160


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from channel_passive_amd import capi
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(capi, "_lib", None)
    with pytest.raises(capi.PbError):
        capi.load()


def test_fortran_module_binds_every_entry_point():
    """include/poongback_gpu.f90 has an interface with bind(C, name=...) for
    every function of the header, with the same argument count."""
    f90 = open(os.path.join(ROOT, "include", "poongback_gpu.f90")).read()
    bound = dict(re.findall(r"function\s+(pb_\w+)\s*\(([^)]*)\)\s*(?:&\s*)?bind\(C", f90))
    text = open(HEADER).read()
    for name in declared():
        assert name in bound, name
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)\s*;", text)
        n_c = 0 if m.group(1).strip() in ("", "void") else len(m.group(1).split(","))
        n_f = 0 if not bound[name].strip() else len(bound[name].split(","))
        assert n_c == n_f, (name, n_c, n_f)
