"""The labelling surface: cc3d.connected_components on the MI355X library.

Mirrors the binding layer of the reference (cc3d/fastcc3d.pyx:128-642): the
same arguments, validation, dtype views, output-dtype selection and reshapes,
with the labelling itself done by libcc3d_amd.so on the GPU.  Inputs may be
numpy arrays (copied to HBM and back), CPU torch tensors (returned as CPU
tensors, like the reference) or device torch tensors (labelled in place in HBM
and returned on the device).
"""
from __future__ import annotations

import ctypes
import operator
import os
from functools import reduce
from typing import Any, Optional, Sequence, Tuple, Union

import numpy as np

from . import _lib as L


class DimensionError(Exception):
  """The array has the wrong number of dimensions."""
  pass


_NP_TO_CODE = {
  np.dtype(np.uint8): L.DT_U8, np.dtype(np.int8): L.DT_U8, np.dtype(np.bool_): L.DT_U8,
  np.dtype(np.uint16): L.DT_U16, np.dtype(np.int16): L.DT_U16,
  np.dtype(np.uint32): L.DT_U32, np.dtype(np.int32): L.DT_U32,
  np.dtype(np.uint64): L.DT_U64, np.dtype(np.int64): L.DT_U64,
  np.dtype(np.float32): L.DT_F32, np.dtype(np.float64): L.DT_F64,
}
_UNSIGNED_VIEW = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
_OUT_CODE = {np.dtype(np.uint16): L.DT_U16, np.dtype(np.uint32): L.DT_U32, np.dtype(np.uint64): L.DT_U64}

_SUPPORTED = ("Supported: bool, int8, int16, int32, int64, uint8, uint16, uint32, uint64, "
              "float16, float32, float64")


def _torch():
  import torch
  return torch


def _is_torch(x) -> bool:
  return hasattr(x, "cpu") and hasattr(x, "data_ptr")


def _torch_np_dtype(t) -> np.dtype:
  torch = _torch()
  table = {
    torch.bool: np.bool_, torch.uint8: np.uint8, torch.int8: np.int8,
    torch.int16: np.int16, torch.int32: np.int32, torch.int64: np.int64,
    torch.float16: np.float16, torch.float32: np.float32, torch.float64: np.float64,
  }
  for name, npt in (("uint16", np.uint16), ("uint32", np.uint32), ("uint64", np.uint64)):
    if hasattr(torch, name):
      table[getattr(torch, name)] = npt
  if t.dtype not in table:
    raise TypeError(f"Type {t.dtype} is not currently supported. {_SUPPORTED}")
  return np.dtype(table[t.dtype])


def _torch_out_dtype(npdt):
  torch = _torch()
  return {np.dtype(np.uint16): torch.uint16, np.dtype(np.uint32): torch.uint32,
          np.dtype(np.uint64): torch.uint64}[np.dtype(npdt)]


def _strides_contig(shape, strides, itemsize) -> Tuple[bool, bool]:
  """(c_contiguous, f_contiguous) with numpy's rules (size-1 axes ignored)."""
  def ok(order):
    expect = itemsize
    axes = range(len(shape) - 1, -1, -1) if order == "C" else range(len(shape))
    for a in axes:
      if shape[a] == 1:
        continue
      if strides[a] != expect:
        return False
      expect *= shape[a]
    return True
  return ok("C"), ok("F")


class _Vol:
  """A volume resident in HBM, x fastest, plus what the binding needs to know."""

  def __init__(self, data):
    self.is_torch = _is_torch(data)
    self.device_input = False
    if self.is_torch:
      t = data
      self.np_dtype = _torch_np_dtype(t)
      itemsize = self.np_dtype.itemsize
      shape = tuple(t.shape)
      strides = tuple(s * itemsize for s in t.stride())
      c, f = _strides_contig(shape, strides, itemsize)
      self.device_input = t.device.type != "cpu"
      self.torch_device = t.device
      self.t = t
    else:
      a = np.asarray(data)
      self.np_dtype = a.dtype
      shape, c, f = a.shape, a.flags.c_contiguous, a.flags.f_contiguous
      self.a = a
    self.shape = tuple(int(s) for s in shape)
    self.c_contig, self.f_contig = c, f

  @property
  def size(self) -> int:
    return int(np.prod(self.shape, dtype=np.int64)) if len(self.shape) else 1

  def device_buffer(self, order: str):
    """1D device tensor holding the volume in memory order (x fastest)."""
    torch = _torch()
    if self.is_torch:
      t = self.t
      if not (self.c_contig or self.f_contig):
        t = t.contiguous()  # the reference copies to C order (np.copy(order='C'))
        self.c_contig, self.f_contig = True, len([s for s in self.shape if s != 1]) <= 1
      if t.dtype == torch.bool:
        t = t.view(torch.uint8)
      flat = t.as_strided((t.numel(),), (1,))
      if not self.device_input:
        flat = flat.to("cuda")
      return flat
    a = self.a
    if not (a.flags.c_contiguous or a.flags.f_contiguous):
      a = np.copy(a, order=order)
    flat = np.ravel(a, order="K")
    if flat.dtype == np.bool_:
      flat = flat.view(np.uint8)
    elif flat.dtype == np.float16:
      flat = flat.view(np.uint16)
    return torch.from_numpy(flat).to("cuda")


def _even_ceil(n: int) -> int:
  return n << 1 if n & 1 else n


def _delta_bytes(view_dtype: np.dtype, delta) -> bytes:
  """The C++ `T delta` argument (cc3d/fastcc3d.pyx:471-608 passes the Python
  number through Cython's conversion to T)."""
  if view_dtype.kind == "f":
    return np.array([delta], dtype=view_dtype).tobytes()
  d = int(delta)
  bits = view_dtype.itemsize * 8
  name = f"uint{bits}_t"
  if d < 0:
    raise OverflowError(f"can't convert negative value to {name}")
  if d >= (1 << bits):
    raise OverflowError(f"value too large to convert to {name}")
  return int(d).to_bytes(view_dtype.itemsize, "little")


def _stream():
  torch = _torch()
  return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_gpu():
  torch = _torch()
  if not torch.cuda.is_available():
    raise RuntimeError("cc3d (MI355X build) needs a visible GPU; there is no CPU fallback.")


def estimate_provisional_labels(data) -> Tuple[int, int, int]:
  """(epl, first_foreground_row, last_foreground_row) -- cc3d/fastcc3d.pyx:168-242
  on the GPU (cc3d/cc3d.hpp:287-315)."""
  vol = _Vol(data)
  if vol.np_dtype not in _NP_TO_CODE and vol.np_dtype != np.float16:
    raise TypeError(f"Type {vol.np_dtype} is not currently supported. {_SUPPORTED}")
  if vol.np_dtype == np.float16:
    raise TypeError(f"Type {vol.np_dtype} is not currently supported. {_SUPPORTED}")
  if vol.size == 0:
    return (0, -1, -1)
  sx = vol.shape[0] if vol.f_contig else vol.shape[-1]
  _require_gpu()
  order = "F" if vol.f_contig else "C"
  buf = vol.device_buffer(order)
  epl = ctypes.c_uint64(0)
  f = ctypes.c_int64(0)
  l = ctypes.c_int64(0)
  L.check(L.lib().cc3d_amd_estimate_provisional_label_count(
    ctypes.c_void_p(buf.data_ptr()), _NP_TO_CODE[vol.np_dtype], sx, vol.size,
    ctypes.byref(epl), ctypes.byref(f), ctypes.byref(l), _stream()))
  return (int(epl.value), int(f.value), int(l.value))


class _Plan:
  """Everything the binding decides about one volume before labelling it
  (cc3d/fastcc3d.pyx:307-453)."""
  pass


def _prepare(data, max_labels, connectivity, delta, out_dtype, out_file, periodic_boundary, binary_image,
             return_N):
  """Validation, dims, dtype views and binary detection of one volume.

  Returns ("done", result) for an empty volume, else ("plan", _Plan)."""
  vol = _Vol(data)
  dims = len(vol.shape)
  if dims not in (1, 2, 3):
    raise DimensionError("Only 1D, 2D, and 3D arrays supported. Got: " + str(dims))
  if dims == 2 and connectivity not in (4, 8, 6, 18, 26):
    raise ValueError("Only 4, 8, and 6, 18, 26 connectivities are supported for 2D images. Got: " + str(connectivity))
  elif dims != 2 and connectivity not in (6, 18, 26):
    raise ValueError("Only 6, 18, and 26 connectivities are supported for 3D images. Got: " + str(connectivity))
  if periodic_boundary and connectivity not in (4, 8, 6):
    raise ValueError(f"periodic_boundary is not yet implemented for {connectivity}-connectivity.")
  if periodic_boundary and delta != 0:
    raise ValueError("periodic_boundary is not yet implemented continuous data.")

  if vol.size == 0:
    dt = out_dtype if out_dtype is not None else vol.np_dtype
    out = np.zeros(shape=(0,), dtype=dt)
    if vol.is_torch:
      out = _torch().from_numpy(out)
      if vol.device_input:
        out = out.to(vol.torch_device)
    return "done", ((out, 0) if return_N else out)

  order = "F" if vol.f_contig else "C"
  shape = list(vol.shape)
  while len(shape) < 3:
    shape = ([1] + shape) if order == "C" else (shape + [1])
  padded_shape = list(shape)
  np_dtype = vol.np_dtype
  if np_dtype == np.float16:
    if delta == 0:
      np_dtype = np.dtype(np.uint16)
    else:
      raise TypeError("float16 is not supported for continuous images (delta != 0).")
  if np_dtype not in _NP_TO_CODE:
    raise TypeError(f"Type {np_dtype} is not currently supported. {_SUPPORTED}")
  if order == "C":
    shape.reverse()
  sx, sy, sz = shape
  voxels = sx * sy * sz

  binary_image = bool(binary_image) or np_dtype == np.bool_
  binary_flag_given = binary_image

  if np.issubdtype(np_dtype, np.floating):
    delta = float(delta)
    binary_image = binary_image or (delta == np.finfo(np_dtype).max)
  else:
    delta = int(delta)
    binary_image = binary_image or (delta == np.iinfo(np_dtype).max)

  if out_dtype is not None:
    out_dtype = np.dtype(out_dtype)
    if out_dtype not in (np.uint16, np.uint32, np.uint64):
      raise ValueError(
        f"Explicitly defined out_dtype ({out_dtype}) must be one of: "
        f"np.uint16, np.uint32, np.uint64"
      )

  pl = _Plan()
  pl.data, pl.vol, pl.dims, pl.order = data, vol, dims, order
  pl.sx, pl.sy, pl.sz, pl.voxels = sx, sy, sz, voxels
  pl.connectivity, pl.delta, pl.out_dtype, pl.out_file = connectivity, delta, out_dtype, out_file
  pl.periodic_boundary, pl.return_N = bool(periodic_boundary), return_N
  pl.binary_image, pl.binary_flag_given = binary_image, binary_flag_given
  pl.view_dtype = np.dtype(_UNSIGNED_VIEW[np_dtype.itemsize]) if np_dtype.kind != "f" else np_dtype
  pl.code = _NP_TO_CODE[np_dtype]
  pl.ufv = _even_ceil(padded_shape[0]) * _even_ceil(padded_shape[1]) * _even_ceil(padded_shape[2])

  try:
    pl.delta_b = _delta_bytes(pl.view_dtype, delta)
  except OverflowError:
    # The reference converts delta to T only when it calls the C++ core, after
    # choosing the output dtype: report a too-small out_dtype first.
    if binary_flag_given:
      ml = _binary_bound(pl, voxels)
    else:
      ml = min(estimate_provisional_labels(data)[0], voxels)
      if binary_image:
        ml = _binary_bound(pl, ml)
    _choose_out_dtype(pl, ml)
    raise

  flags = 0
  if not binary_image and delta != 0:
    flags |= L.FLAG_CONTINUOUS
  # Binary 2D-8 provisional labels (one per new 2x2 block) can only reach the
  # union-find length through an explicit uint16 output; count them then.
  nblocks = ((sx + 1) >> 1) * ((sy + 1) >> 1)
  pl.q1_count = (binary_image and connectivity == 8 and out_dtype is not None
                 and out_dtype == np.uint16 and nblocks >= 65535)
  if pl.q1_count:
    flags |= L.FLAG_COUNT_PROVISIONAL
  pl.flags = flags
  return "plan", pl


def _binary_bound(pl, ml: int) -> int:
  if pl.connectivity in (4, 6):
    return min(ml, (pl.ufv // 2) + 1)
  # the reference tests `connectivity == (8,18)`, which never holds
  return min(ml, (pl.ufv // 8) + 1)


def _choose_out_dtype(pl, max_labels: int):
  if pl.out_dtype is not None:
    if np.iinfo(pl.out_dtype).max < max_labels:
      raise ValueError(
        f"Explicitly defined out_dtype ({pl.out_dtype}) is too small "
        f"to contain the estimated maximum number of labels ({max_labels})."
      )
    return pl.out_dtype
  if max_labels < np.iinfo(np.uint16).max:
    return np.dtype(np.uint16)
  if max_labels < np.iinfo(np.uint32).max:
    return np.dtype(np.uint32)
  return np.dtype(np.uint64)


def _decide(pl, st):
  """Out dtype, N and the reference's errors from the stage-1 counters
  (cc3d/fastcc3d.pyx:381-434, cc3d/cc3d.hpp:107-111)."""
  if pl.binary_flag_given:
    epl, ffr, lfr = pl.voxels, 0, pl.sy
  else:
    epl, ffr, lfr = int(st.epl), int(st.first_foreground_row), int(st.last_foreground_row)
  max_labels = min(epl, pl.voxels)
  if pl.binary_image:
    max_labels = _binary_bound(pl, max_labels)
  odt = _choose_out_dtype(pl, max_labels)

  special_row = (pl.delta == 0 and ffr == lfr and ffr >= 0)
  pl.special_row = ffr if special_row else None
  N = int(st.N)
  if not special_row:
    uf_len = int(L.lib().cc3d_amd_union_find_length(
      max_labels, pl.voxels, pl.connectivity, int(pl.binary_image), _OUT_CODE[odt]))
    if uf_len == 0:
      N = 0  # the reference returns before labelling (max_labels == 0)
    elif (not (pl.binary_image and pl.connectivity == 8) or pl.q1_count) and int(st.provisional) >= uf_len:
      raise RuntimeError("maximum length exception")
  return odt, N


def connected_components(
  data,
  max_labels: int = -1,
  connectivity: int = 26,
  return_N: bool = False,
  delta: Union[int, float] = 0,
  out_dtype: Optional[Any] = None,
  out_file=None,
  periodic_boundary: bool = False,
  binary_image: bool = False,
):
  """Connected components applied to 1D, 2D or 3D images on the GPU.

  Same arguments and results as cc3d.connected_components
  (cc3d/fastcc3d.pyx:244-626): labels numbered 1..N in order of first
  appearance in memory (x fastest), output dtype uint16/uint32/uint64 chosen
  from the provisional-label estimate unless given.  `max_labels` is accepted
  and ignored, as in the reference.
  """
  kind, pl = _prepare(data, max_labels, connectivity, delta, out_dtype, out_file, periodic_boundary,
                      binary_image, return_N)
  if kind == "done":
    return pl
  return _label_one(pl)


def _label_one(pl):
  _require_gpu()
  from . import _stack
  if _stack.streamable(pl):
    return _stack.label_streamed(pl)
  torch = _torch()
  buf = pl.vol.device_buffer(pl.order)
  sx, sy, sz, connectivity = pl.sx, pl.sy, pl.sz, pl.connectivity
  flags = pl.flags | _tile_flags(pl, buf)
  ws_bytes = int(L.lib().cc3d_amd_workspace_bytes(sx, sy, sz, connectivity, int(pl.binary_image), flags))
  if ws_bytes == 0:
    L.check(L.ERR_RUNTIME if connectivity in (4, 8) and sz != 1 else L.ERR_INVALID)
  ws = torch.empty(ws_bytes, dtype=torch.uint8, device=buf.device)
  stream = _stream()
  job = ctypes.c_void_p(0)
  dbuf = ctypes.create_string_buffer(pl.delta_b, max(len(pl.delta_b), 8))
  L.check(L.lib().cc3d_amd_label_begin(
    ctypes.byref(job), ctypes.c_void_p(buf.data_ptr()), pl.code, sx, sy, sz, connectivity,
    ctypes.cast(dbuf, ctypes.c_void_p), flags, int(pl.periodic_boundary), int(pl.binary_image),
    ctypes.c_void_p(ws.data_ptr()), ws_bytes, stream))
  try:
    st = L.Stats()
    L.check(L.lib().cc3d_amd_job_stats(job, ctypes.byref(st)))
    odt, N = _decide(pl, st)
    out_dev = torch.empty(pl.voxels, dtype=_torch_out_dtype(odt), device=buf.device)
    if pl.special_row is None:
      L.check(L.lib().cc3d_amd_label_finish(job, ctypes.c_void_p(out_dev.data_ptr()), _OUT_CODE[odt]))
  finally:
    L.lib().cc3d_amd_job_free(job)
  if pl.special_row is not None:
    N = _special_row(pl, buf, odt, out_dev)

  return _deliver(out_dev, pl.vol, sx, sy, sz, pl.dims, pl.order, odt, pl.out_file, N, pl.return_N)


def _tile_flags(pl, buf) -> int:
  """Tile family of the local kernel for this volume (cc3d/_tune.py)."""
  from . import _tune
  if pl.flags & L.FLAG_CONTINUOUS:
    # continuous links have no row-tile kernel; tall tiles only when forced
    return _tune._FLAG["tall"] if _tune.FORCE == "tall" else 0
  isz = pl.view_dtype.itemsize
  return _tune.tiles(pl.connectivity, pl.binary_image, isz, pl.voxels, pl.sz > 1,
                     lambda: _tune.sample_density(buf, isz, pl.voxels))


_ROW_FAILURES = {
  1: (ValueError, "cannot convert float NaN to integer"),
  2: (OverflowError, "cannot convert float infinity to integer"),
  3: (OverflowError, "Python int too large to convert to C long"),
}


def _special_row(pl, buf, odt, out_dev) -> int:
  """epl_special_row (cc3d/fastcc3d.pyx:644-679) on the GPU: the binding labels
  a volume whose only foreground is one row (delta == 0) without the C++ core.

  Its `cdef int64_t last_label` compares each voxel with the int64 conversion
  of its run's first value, so float runs break on non-integer values, and the
  conversion raises for NaN, infinities and values outside the int64 range
  (uint64 >= 2^63 included).  Each label is stored into the output before the
  value is converted, so a label past the output dtype's range raises first
  (numpy's out-of-bounds OverflowError).  Writes every voxel of out_dev and
  returns N."""
  n = ctypes.c_uint64(0)
  kind = ctypes.c_int(0)
  rank = ctypes.c_uint64(0)
  check_u64 = int(pl.vol.np_dtype == np.uint64)
  lib = L.lib()
  rc = lib.cc3d_amd_special_row(
    ctypes.c_void_p(buf.data_ptr()), pl.code, check_u64, pl.sx, pl.voxels, pl.special_row,
    ctypes.c_void_p(out_dev.data_ptr()), _OUT_CODE[odt], ctypes.byref(n), ctypes.byref(kind),
    ctypes.byref(rank), _stream())
  if rc != L.OK:
    msg = lib.cc3d_amd_special_row_last_error().decode(errors="replace")
    raise (MemoryError if rc == L.ERR_NOMEM else ValueError if rc == L.ERR_INVALID else RuntimeError)(msg)
  outmax = int(np.iinfo(odt).max)
  if kind.value and int(rank.value) <= outmax:
    exc, msg = _ROW_FAILURES[kind.value]
    raise exc(msg)
  if int(n.value) > outmax:
    raise OverflowError(f"Python integer {outmax + 1} out of bounds for {np.dtype(odt).name}")
  return int(n.value)


def _deliver(out_dev, vol, sx, sy, sz, dims, order, out_dtype, out_file, N, return_N):
  """_final_reshape (cc3d/fastcc3d.pyx:628-642) plus numpy/torch/memmap hand-back."""
  torch = _torch()
  if dims == 3:
    shp = (sz, sy, sx) if order == "C" else (sx, sy, sz)
  elif dims == 2:
    shp = (sy, sx) if order == "C" else (sx, sy)
  else:
    shp = (sx,)
  if vol.is_torch and vol.device_input and out_file is None:
    if order == "C":
      out = out_dev.view(shp)
    else:
      strides = [1]
      for s in shp[:-1]:
        strides.append(strides[-1] * s)
      out = out_dev.as_strided(shp, tuple(strides))
    return (out, N) if return_N else out
  host = out_dev.cpu().numpy()
  if out_file is not None:
    if isinstance(out_file, str):
      with open(out_file, "wb") as f:
        os.ftruncate(f.fileno(), host.size * np.dtype(out_dtype).itemsize)
    mm = np.memmap(out_file, order="F", dtype=out_dtype, shape=(host.size,))
    mm[:] = host
    host = mm
  out = host.reshape(shp, order=order)
  if vol.is_torch:
    out = torch.from_numpy(out)
    if vol.device_input:
      out = out.to(vol.torch_device)
  return (out, N) if return_N else out
