"""The C-ABI library loads and exports every entry point include/cc3d_amd.h declares.

CPU-only: loading the library and resolving symbols needs no GPU.  Entry
points are called here only with arguments they reject before any device
work (argument checks return a status code and a message).
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


def test_argument_checks_without_device_work():
  from cc3d import _lib
  lib = _lib.lib()
  vp, i64, u64, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int
  buf = ctypes.create_string_buffer(64)  # never read: every call below is rejected first
  p = ctypes.cast(buf, vp)

  f = lib.cc3d_amd_voxel_connectivity_graph
  f.restype, f.argtypes = ci, [vp, ci, i64, i64, i64, ci, vp, vp]
  lib.cc3d_amd_graphs_last_error.restype = ctypes.c_char_p
  assert f(p, 0, 4, 4, 4, 5, p, None) == _lib.ERR_RUNTIME
  assert b"Only 4 and 8 2D" in lib.cc3d_amd_graphs_last_error()
  assert f(p, 0, 4, 4, 2, 8, p, None) == _lib.ERR_RUNTIME
  assert lib.cc3d_amd_graphs_last_error() == b"sz must be 1 for 2D connectivities."
  assert f(p, 0, 0, 4, 4, 26, p, None) == _lib.ERR_INVALID
  assert f(p, 4, 4, 4, 4, 26, p, None) == _lib.ERR_INVALID  # float labels

  lib.cc3d_amd_follow_last_error.restype = ctypes.c_char_p
  k = lib.cc3d_amd_keep_labels
  k.restype, k.argtypes = ci, [vp, ci, i64, vp, i64, u64, vp, vp, ci, vp]
  assert k(p, 4, 8, p, 1, 3, p, p, 2, None) == _lib.ERR_INVALID  # float labels
  assert k(p, 2, 8, p, 1, 3, p, p, 5, None) == _lib.ERR_INVALID  # float output
  assert k(p, 2, 8, p, -1, 3, p, p, 2, None) == _lib.ERR_INVALID
  assert b"keep_labels" in lib.cc3d_amd_follow_last_error()

  d = lib.cc3d_amd_dust_copy
  d.restype = ci
  d.argtypes = [vp, vp, ci, vp, ci, i64, vp, u64, ctypes.c_double, ctypes.c_double, ci, ci, vp, vp, vp]
  assert d(p, p, 2, p, 4, 8, p, 3, 1.0, 0.0, 0, 0, p, p, None) == _lib.ERR_INVALID  # float labels
  assert d(None, p, 2, p, 2, 8, p, 3, 1.0, 0.0, 0, 0, p, p, None) == _lib.ERR_INVALID

  s = lib.cc3d_amd_statistics
  s.restype, s.argtypes = ci, [vp, ci, i64, i64, i64, ci, ci, u64, vp, vp, vp, vp, vp]
  assert s(p, 2, 4, 4, 4, 3, 1, 3, None, p, None, None, None) == _lib.ERR_INVALID  # counts missing
  assert s(p, 2, 4, 4, 4, 3, 1, 3, p, p, p, None, None) == _lib.ERR_INVALID  # centroids need sums
  assert s(p, 4, 4, 4, 4, 3, 1, 3, p, p, None, None, None) == _lib.ERR_INVALID  # float labels


def test_batch_layout_with_every_tile_family():
  """cc3d_amd_batch_workspace_bytes plans a batch on the host alone: volumes
  of all three tile families, modes and periodic flags lay out, and the
  checks reject what a batch does not cover."""
  from cc3d import _lib as L
  lib = L.lib()
  fams = [L.FLAG_TILES_VOXELS, L.FLAG_TILES_TALL, L.FLAG_TILES_ROWS]
  n = 9
  vt = (L.Volume * n)()
  for k in range(n):
    vt[k].inp = 0x1000 * (k + 1)  # never dereferenced: no device work
    vt[k].sx, vt[k].sy, vt[k].sz = 40 + 7 * k, 30 + k, 20 + 3 * k
    vt[k].connectivity = (26, 18, 6)[k % 3] | (fams[(k // 3) % 3] << 8)
  full = lib.cc3d_amd_batch_workspace_bytes(vt, n, L.BATCH_MULTILABEL)
  assert full > 0
  assert lib.cc3d_amd_batch_workspace_bytes(vt, n - 3, L.BATCH_MULTILABEL) < full
  vt[2].connectivity |= L.VOLUME_PERIODIC
  assert lib.cc3d_amd_batch_workspace_bytes(vt, n, L.BATCH_BINARY) > 0
  assert lib.cc3d_amd_batch_workspace_bytes(vt, n, L.BATCH_CONTINUOUS) == 0  # periodic + continuous
  vt[2].connectivity = 8
  vt[2].sz = 1
  multi = lib.cc3d_amd_batch_workspace_bytes(vt, n, L.BATCH_MULTILABEL)
  binary = lib.cc3d_amd_batch_workspace_bytes(vt, n, L.BATCH_BINARY)  # the plane on the 2D tiles, block keys
  cont = lib.cc3d_amd_batch_workspace_bytes(vt, n, L.BATCH_CONTINUOUS)  # the plane on the 2D tiles, its min/max
  assert multi > 0 and binary > 0 and cont > 0
  vt[2].sz = 2  # 2D connectivities need a single plane
  assert lib.cc3d_amd_batch_workspace_bytes(vt, n, L.BATCH_BINARY) == 0
  vt[2].sz = 1
  assert lib.cc3d_amd_batch_workspace_bytes(vt, n, 3) == 0  # no such mode
