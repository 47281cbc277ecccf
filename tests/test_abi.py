"""The C-ABI library loads and exports every entry point include/cc3d_amd.h declares.

CPU-only: loading the library and resolving symbols needs no GPU; no compute
entry point is called here.
"""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cc3d_amd.h")


def declared_functions():
  text = open(HEADER).read()
  text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
  names = re.findall(r"\b(cc3d_amd_[a-z0-9_]+)\s*\(", text)
  return sorted(set(names))


def test_header_declares_entry_points():
  names = declared_functions()
  for required in ("cc3d_amd_connected_components3d", "cc3d_amd_estimate_provisional_label_count",
                   "cc3d_amd_statistics", "cc3d_amd_dust_apply"):
    assert required in names


def test_library_exports_every_declared_symbol():
  from cc3d import _lib
  lib = ctypes.CDLL(_lib.LIB_PATH)
  missing = [n for n in declared_functions() if not hasattr(lib, n)]
  assert not missing, f"undefined exports: {missing}"


def test_version_string():
  from cc3d import _lib
  v = _lib.lib().cc3d_amd_version().decode()
  assert "gfx950" in v
