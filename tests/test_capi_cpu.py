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
