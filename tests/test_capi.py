"""The C-ABI library loads without a GPU and exports every function the
public header declares (no compute calls)."""
import ctypes
import os
import re

import pytest

from justpic_hip import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "justpic_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(jp_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_core_entry_points():
    names = declared_functions()
    for n in ("jp_advection", "jp_move_particles", "jp_particle2grid", "jp_grid2particle",
              "jp_inject_particles", "jp_init_particles", "jp_ctx_create"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_functions()) == set(_abi.exported_symbols())


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="library not built")
def test_philox_host_helper_matches_published_vector():
    lib = _abi.load_library()
    ctr = (ctypes.c_uint32 * 4)(0, 0, 0, 0)
    key = (ctypes.c_uint32 * 2)(0, 0)
    out = (ctypes.c_uint32 * 4)()
    lib.jp_philox4x32_10(ctr, key, out)
    assert list(out) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]


def test_package_fails_loudly_without_library(tmp_path):
    with pytest.raises(_abi.HipLibraryMissing):
        _abi.load_library(str(tmp_path / "missing.so"))
