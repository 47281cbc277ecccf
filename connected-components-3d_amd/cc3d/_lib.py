"""ctypes binding of libcc3d_amd.so (include/cc3d_amd.h).

The library is the only compute path: when it is missing, or when no GPU is
visible, every labelling call raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libcc3d_amd.so")

OK, ERR_RUNTIME, ERR_INVALID, ERR_HIP, ERR_NOMEM = 0, 1, 2, 3, 4
FLAG_CONTINUOUS, FLAG_COUNT_PROVISIONAL, FLAG_NO_SINGLE_ROW_RULE = 1, 2, 4
FLAG_TILES_ROWS, FLAG_TILES_VOXELS, FLAG_TILES_TALL = 8, 16, 32
VOLUME_PERIODIC = 1 << 16
BATCH_MULTILABEL, BATCH_BINARY, BATCH_CONTINUOUS = 0, 1, 2

# dtype codes of the C ABI
DT_U8, DT_U16, DT_U32, DT_U64, DT_F32, DT_F64 = 0, 1, 2, 3, 4, 5


class Stats(ctypes.Structure):
  _fields_ = [
    ("N", ctypes.c_uint64),
    ("provisional", ctypes.c_uint64),
    ("epl", ctypes.c_uint64),
    ("first_foreground_row", ctypes.c_int64),
    ("last_foreground_row", ctypes.c_int64),
  ]


class Volume(ctypes.Structure):
  _fields_ = [
    ("inp", ctypes.c_void_p),
    ("sx", ctypes.c_int64),
    ("sy", ctypes.c_int64),
    ("sz", ctypes.c_int64),
    ("connectivity", ctypes.c_int64),
  ]


_lib = None


class LibraryMissing(ImportError):
  pass


def lib():
  """Load (once) and return the HIP library; raise if it has not been built."""
  global _lib
  if _lib is not None:
    return _lib
  if not os.path.exists(LIB_PATH):
    raise LibraryMissing(
      f"{LIB_PATH} is missing: build it with `make -C connected-components-3d_amd` "
      "(or __graft_entry__.build()). There is no CPU fallback."
    )
  L = ctypes.CDLL(LIB_PATH)
  vp, i64, u64, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int
  L.cc3d_amd_connected_components3d.restype = ci
  L.cc3d_amd_connected_components3d.argtypes = [
    vp, ci, i64, i64, i64, i64, i64, vp, vp, ci, ctypes.POINTER(u64), ci, ci, vp]
  L.cc3d_amd_estimate_provisional_label_count.restype = ci
  L.cc3d_amd_estimate_provisional_label_count.argtypes = [
    vp, ci, i64, i64, ctypes.POINTER(u64), ctypes.POINTER(i64), ctypes.POINTER(i64), vp]
  L.cc3d_amd_workspace_bytes.restype = ctypes.c_size_t
  L.cc3d_amd_workspace_bytes.argtypes = [i64, i64, i64, i64, ci, ci]
  L.cc3d_amd_label_begin.restype = ci
  L.cc3d_amd_label_begin.argtypes = [
    ctypes.POINTER(vp), vp, ci, i64, i64, i64, i64, vp, ci, ci, ci, vp, ctypes.c_size_t, vp]
  L.cc3d_amd_job_stats.restype = ci
  L.cc3d_amd_job_stats.argtypes = [vp, ctypes.POINTER(Stats)]
  L.cc3d_amd_label_finish.restype = ci
  L.cc3d_amd_label_finish.argtypes = [vp, vp, ci]
  L.cc3d_amd_job_free.restype = None
  L.cc3d_amd_job_free.argtypes = [vp]
  L.cc3d_amd_union_find_length.restype = u64
  L.cc3d_amd_union_find_length.argtypes = [i64, i64, i64, ci, ci]
  L.cc3d_amd_batch_workspace_bytes.restype = ctypes.c_size_t
  L.cc3d_amd_batch_workspace_bytes.argtypes = [ctypes.POINTER(Volume), ci, ci]
  L.cc3d_amd_batch_create.restype = ci
  L.cc3d_amd_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Volume), ci, ci, ci, vp, ctypes.c_size_t, vp]
  L.cc3d_amd_batch_set_delta.restype = ci
  L.cc3d_amd_batch_set_delta.argtypes = [vp, vp]
  L.cc3d_amd_batch_set_outputs.restype = ci
  L.cc3d_amd_batch_set_outputs.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(ci)]
  for name in ("cc3d_amd_batch_begin", "cc3d_amd_batch_finish", "cc3d_amd_batch_graph_capture",
               "cc3d_amd_batch_graph_launch"):
    getattr(L, name).restype = ci
    getattr(L, name).argtypes = [vp]
  L.cc3d_amd_batch_stats.restype = ci
  L.cc3d_amd_batch_stats.argtypes = [vp, ctypes.POINTER(Stats)]
  L.cc3d_amd_batch_free.restype = None
  L.cc3d_amd_batch_free.argtypes = [vp]
  L.cc3d_amd_special_row.restype = ci
  L.cc3d_amd_special_row.argtypes = [
    vp, ci, ci, i64, i64, i64, vp, ci, ctypes.POINTER(u64), ctypes.POINTER(ci), ctypes.POINTER(u64), vp]
  L.cc3d_amd_special_row_last_error.restype = ctypes.c_char_p
  L.cc3d_amd_special_row_last_error.argtypes = []
  L.cc3d_amd_last_error.restype = ctypes.c_char_p
  L.cc3d_amd_last_error.argtypes = []
  L.cc3d_amd_version.restype = ctypes.c_char_p
  L.cc3d_amd_version.argtypes = []
  _lib = L
  return L


def check(rc: int) -> None:
  """Map a status code onto the exception the reference would raise."""
  if rc == OK:
    return
  msg = lib().cc3d_amd_last_error().decode(errors="replace")
  if rc == ERR_RUNTIME:
    raise RuntimeError(msg)
  if rc == ERR_INVALID:
    raise ValueError(msg)
  if rc == ERR_NOMEM:
    raise MemoryError(msg)
  raise RuntimeError(f"HIP error: {msg}")
