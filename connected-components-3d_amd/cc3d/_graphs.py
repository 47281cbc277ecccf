"""voxel_connectivity_graph on the GPU (cc3d/fastcc3d.pyx:1021-1176,
cc3d/cc3d_graphs.hpp:31-236)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from ._core import DimensionError, _require_gpu, _stream, _torch

_UNSIGNED = {8: np.uint64, 4: np.uint32, 2: np.uint16, 1: np.uint8}
_CODE = {1: L.DT_U8, 2: L.DT_U16, 4: L.DT_U32, 8: L.DT_U64}


def _bind():
  lib = L.lib()
  if not hasattr(lib, "_graphs_bound"):
    vp, i64, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    lib.cc3d_amd_voxel_connectivity_graph.restype = ci
    lib.cc3d_amd_voxel_connectivity_graph.argtypes = [vp, ci, i64, i64, i64, ci, vp, vp]
    lib.cc3d_amd_graphs_last_error.restype = ctypes.c_char_p
    lib.cc3d_amd_color_connectivity_graph.restype = ci
    lib.cc3d_amd_color_connectivity_graph.argtypes = [vp, ci, i64, i64, i64, ci, vp, ctypes.POINTER(ctypes.c_uint64), vp]
    lib.cc3d_amd_color_last_error.restype = ctypes.c_char_p
    lib._graphs_bound = True
  return lib


def voxel_connectivity_graph(data: np.ndarray, connectivity: int = 26) -> np.ndarray:
  """Per-voxel bitfield of the directions whose neighbour holds the same label
  (cc3d/fastcc3d.pyx:1021-1176): uint8 for 4/8/6, uint32 for 18/26, a
  direction out of the volume counting as open.  Same bit layout, dtype,
  shape and F order as the reference; the bitfields are computed on the GPU."""
  dims = len(data.shape)
  if dims not in (1, 2, 3):
    raise DimensionError("Only 1D, 2D, and 3D arrays supported. Got: " + str(dims))
  if dims == 2 and connectivity not in (4, 8, 6, 18, 26):
    raise ValueError("Only 4, 8, and 6, 18, 26 connectivities are supported for 2D images. Got: " + str(connectivity))
  elif dims != 2 and connectivity not in (6, 18, 26):
    raise ValueError("Only 6, 18, and 26 connectivities are supported for 3D images. Got: " + str(connectivity))
  out_dtype = np.dtype(np.uint8 if connectivity in (4, 8, 6) else np.uint32)
  if data.size == 0:
    return np.zeros(shape=(0,), dtype=out_dtype)
  data = np.asfortranarray(data)
  while len(data.shape) < 3:
    data = data[..., np.newaxis]
  if not (data.dtype in (np.uint64, np.int64, np.uint32, np.int32, np.uint16, np.int16, np.uint8, np.int8)
          or data.dtype == np.bool_):
    raise TypeError("Type {} not currently supported.".format(data.dtype))
  lab = data.view(_UNSIGNED[data.dtype.itemsize])
  sx, sy, sz = lab.shape
  _require_gpu()
  torch = _torch()
  flat = np.ravel(lab, order="F")
  dev = torch.from_numpy(flat.view(np.uint8)).to("cuda")
  out = torch.empty(flat.size * out_dtype.itemsize, dtype=torch.uint8, device="cuda")
  lib = _bind()
  rc = lib.cc3d_amd_voxel_connectivity_graph(ctypes.c_void_p(dev.data_ptr()), _CODE[lab.itemsize], sx, sy, sz,
                                             int(connectivity), ctypes.c_void_p(out.data_ptr()), _stream())
  if rc != L.OK:
    msg = lib.cc3d_amd_graphs_last_error().decode(errors="replace")
    raise (RuntimeError if rc == L.ERR_RUNTIME else ValueError)(msg)
  graph = out.cpu().numpy().view(out_dtype)
  if dims == 3:
    return graph.reshape((sx, sy, sz), order="F")
  elif dims == 2:
    return graph.reshape((sx, sy), order="F")
  return graph.reshape((sx,), order="F")


def color_connectivity_graph(vcg, connectivity=26, return_N=False):
  """Connected components of a voxel connectivity graph (cc3d/fastcc3d.pyx:941-1018):
  every voxel is labelled, uint32 in F order, numbered by first voxel.  Same
  arguments, checks and results as the reference; labelled on the GPU."""
  vcg = np.asarray(vcg)
  dims = len(vcg.shape)
  if dims not in (2, 3):
    raise DimensionError("Only 2D, and 3D arrays supported. Got: " + str(dims))
  if dims == 2 and connectivity not in [4, 8, 6, 26]:
    raise ValueError(f"Only 4 and 8 connectivity is supported for 2D images. Got: {connectivity}")
  elif dims != 2 and connectivity not in [6, 26]:
    raise ValueError(f"Only 6 and 26 connectivity are supported for 3D images. Got: {connectivity}")
  dtype = vcg.dtype
  if dtype not in [np.uint8, np.uint32]:
    raise ValueError(f"Only uint8 and uint32 are supported. Got: {vcg.dtype}")
  if vcg.size == 0:
    return np.zeros([0] * dims, dtype=np.uint32, order="F")
  while vcg.ndim < 3:
    vcg = vcg[..., np.newaxis]
  vcg = np.asfortranarray(vcg)
  sx, sy, sz = vcg.shape
  if connectivity in [18, 26] and sz > 1 and dtype != np.uint32:
    raise ValueError(f"Only uint32 is supported for 18 and 26 connected. Got: {vcg.dtype}")
  _require_gpu()
  torch = _torch()
  flat = np.ravel(vcg, order="F")
  dev = torch.from_numpy(flat.view(np.uint8)).to("cuda")
  out = torch.empty(flat.size, dtype=torch.int32, device="cuda")
  n = ctypes.c_uint64(0)
  lib = _bind()
  rc = lib.cc3d_amd_color_connectivity_graph(ctypes.c_void_p(dev.data_ptr()), flat.itemsize, sx, sy, sz,
                                             int(connectivity), ctypes.c_void_p(out.data_ptr()), ctypes.byref(n),
                                             _stream())
  if rc != L.OK:
    msg = lib.cc3d_amd_color_last_error().decode(errors="replace")
    raise (RuntimeError if rc == L.ERR_RUNTIME else MemoryError if rc == L.ERR_NOMEM else ValueError)(msg)
  labels = out.cpu().numpy().view(np.uint32).reshape((sx, sy, sz), order="F")
  while labels.ndim > dims:
    labels = labels[..., 0]
  if return_N:
    return (labels, int(n.value))
  return labels
