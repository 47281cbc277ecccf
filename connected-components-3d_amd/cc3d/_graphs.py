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
