"""Labelling in z-slabs: connected_components_stack (cc3d/__init__.py:353-501)
and volumes of 2^32-1 voxels or more.

Each slab is labelled on the GPU by the two-stage labelling call (labels
1..N_k in the slab's own first-appearance order) and kept as uint32 labels.
Slab k's labels stand for provisional ids offset_k + label; the seam between
consecutive slabs is joined in a 64-bit forest over the ids on the device
(csrc/stack.hip).  Since provisional ids grow with first appearance over the
whole volume and the forest keeps the smaller id as root, the final label of
an id is 1 + the rank of its root: exactly the labels of one call over the
whole volume.  A second pass maps every slab through that table into the
output (an array, a device tensor, or an out_file memmap written slab by
slab).  The GPU holds one slab at a time plus the forest (8 bytes per
provisional id).
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Optional

import numpy as np

from . import _lib as L
from ._core import (_OUT_CODE, _decide, _prepare, _require_gpu, _special_row, _stream, _torch, _torch_out_dtype)

# Volumes of at least this many voxels are labelled in slabs (one labelling
# call takes fewer than 2^32-1 voxels); slabs hold at most _SLAB_VOXELS.
_STREAM_VOXELS = (1 << 32) - 1
_SLAB_VOXELS = 1 << 31


def _bind():
  lib = L.lib()
  if not hasattr(lib, "_stack_bound"):
    vp, i64, u64, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int
    lib.cc3d_amd_stack_seam.restype = ci
    lib.cc3d_amd_stack_seam.argtypes = [vp, vp, vp, vp, ci, i64, i64, u64, u64, ci, ci, vp, vp]
    lib.cc3d_amd_stack_seam_provisional.restype = ci
    lib.cc3d_amd_stack_seam_provisional.argtypes = [vp, vp, ci, i64, i64, ci, ci, vp, vp]
    lib.cc3d_amd_forest_init.restype = ci
    lib.cc3d_amd_forest_init.argtypes = [vp, u64, u64, vp]
    lib.cc3d_amd_forest_labels.restype = ci
    lib.cc3d_amd_forest_labels.argtypes = [vp, u64, vp, ctypes.POINTER(u64), vp]
    lib.cc3d_amd_offset_labels.restype = ci
    lib.cc3d_amd_offset_labels.argtypes = [vp, u64, u64, vp, vp]
    lib.cc3d_amd_map_labels.restype = ci
    lib.cc3d_amd_map_labels.argtypes = [vp, u64, vp, vp, ci, vp]
    lib.cc3d_amd_stack_last_error.restype = ctypes.c_char_p
    lib._stack_bound = True
  return lib


def _check(rc: int):
  if rc != L.OK:
    msg = _bind().cc3d_amd_stack_last_error().decode(errors="replace")
    raise (MemoryError if rc == L.ERR_NOMEM else ValueError if rc == L.ERR_INVALID else RuntimeError)(msg)


def _ptr(t) -> ctypes.c_void_p:
  return ctypes.c_void_p(t.data_ptr())


class _Streamer:
  """Pass 1 over the slabs of one volume (all with the same sx, sy): labels,
  seams, counters; then the table of final labels."""

  def __init__(self, sx: int, sy: int, code: int, connectivity: int, binary: bool, flags: int, delta_b: bytes):
    self.torch = _torch()
    self.sx, self.sy = sx, sy
    self.code, self.connectivity, self.binary, self.flags = code, connectivity, binary, flags
    self.delta_b = delta_b
    self.nid = 0  # provisional ids used (ids are 1..nid)
    self.forest = self.torch.zeros(1 << 16, dtype=self.torch.int64, device="cuda")
    self.prev = None  # (last plane values, last plane labels, offset)
    self.offsets = []
    self.epl = 0
    self.prov = 0
    self.first_row, self.last_row = None, None
    self.seam_prov = self.torch.zeros(1, dtype=self.torch.int64, device="cuda")
    self.z = 0  # planes seen

  def _grow(self, need: int):
    if need <= self.forest.numel():
      return
    cap = self.forest.numel()
    while cap < need:
      cap *= 2
    f = self.torch.zeros(cap, dtype=self.torch.int64, device="cuda")
    f[:self.forest.numel()] = self.forest
    self.forest = f

  def slab(self, dev_bytes, nz: int):
    """Labels one slab of nz planes (dev_bytes: its values on the device, x
    fastest); returns its uint32 labels on the device."""
    torch = self.torch
    lib = L.lib()
    sx, sy = self.sx, self.sy
    from . import _tune
    isz = dev_bytes.numel() // (sx * sy * nz)
    flags = self.flags | _tune.tiles(self.connectivity, self.binary, isz, sx * sy * nz, nz > 1,
                                     lambda: _tune.sample_density(dev_bytes, isz, sx * sy * nz))
    ws_bytes = int(lib.cc3d_amd_workspace_bytes(sx, sy, nz, self.connectivity, int(self.binary), flags))
    if ws_bytes == 0:
      L.check(L.ERR_INVALID)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    job = ctypes.c_void_p(0)
    dbuf = ctypes.create_string_buffer(self.delta_b, max(len(self.delta_b), 8))
    stream = _stream()
    L.check(lib.cc3d_amd_label_begin(
      ctypes.byref(job), _ptr(dev_bytes), self.code, sx, sy, nz, self.connectivity, ctypes.cast(dbuf, ctypes.c_void_p),
      flags | L.FLAG_NO_SINGLE_ROW_RULE, 0, int(self.binary), _ptr(ws), ws_bytes, stream))
    try:
      st = L.Stats()
      L.check(lib.cc3d_amd_job_stats(job, ctypes.byref(st)))
      labels = torch.empty(sx * sy * nz, dtype=torch.int32, device="cuda")
      L.check(lib.cc3d_amd_label_finish(job, _ptr(labels), _OUT_CODE[np.dtype(np.uint32)]))
    finally:
      lib.cc3d_amd_job_free(job)
    del ws
    n = int(st.N)
    self.epl += int(st.epl)
    self.prov += int(st.provisional)
    if int(st.last_foreground_row) >= 0:
      f = int(st.first_foreground_row) + sy * self.z
      l_ = int(st.last_foreground_row) + sy * self.z
      self.first_row = f if self.first_row is None else self.first_row
      self.last_row = l_
    lib_s = _bind()
    off = self.nid
    self._grow(off + n + 1)
    _check(lib_s.cc3d_amd_forest_init(_ptr(self.forest), off + 1, off + n + 1, stream))
    itemsize = dev_bytes.numel() // (sx * sy * nz)
    first_vals = dev_bytes[:sx * sy * itemsize]
    if self.prev is not None:
      pv, pl, poff = self.prev
      _check(lib_s.cc3d_amd_stack_seam(_ptr(pv), _ptr(pl), _ptr(first_vals), _ptr(labels), self.code, sx, sy, poff,
                                       off, self.connectivity, int(self.binary), _ptr(self.forest), stream))
      _check(lib_s.cc3d_amd_stack_seam_provisional(_ptr(pv), _ptr(first_vals), self.code, sx, sy, self.connectivity,
                                                   int(self.binary), _ptr(self.seam_prov), stream))
    last = sx * sy * (nz - 1)
    self.prev = (dev_bytes[last * itemsize:].clone(), labels[last:].clone(), off)
    self.offsets.append(off)
    self.nid = off + n
    self.z += nz
    return labels

  def provisional(self) -> int:
    """Provisional labels of one pass over the whole volume."""
    return self.prov - int(self.seam_prov.item())

  def table(self):
    """(lut, N): the final label of every provisional id (device, uint64)."""
    lut = self.torch.empty(self.nid + 1, dtype=self.torch.int64, device="cuda")
    n = ctypes.c_uint64(0)
    _check(_bind().cc3d_amd_forest_labels(_ptr(self.forest), self.nid, _ptr(lut), ctypes.byref(n), _stream()))
    self.forest = None
    return lut, int(n.value)


def _map_slab(labels_u32, offset: int, lut, out_code: int, out_tensor):
  """out_tensor = final labels of one slab's uint32 labels."""
  torch = _torch()
  lib = _bind()
  ids = torch.empty(labels_u32.numel(), dtype=torch.int64, device="cuda")
  _check(lib.cc3d_amd_offset_labels(_ptr(labels_u32), labels_u32.numel(), offset, _ptr(ids), _stream()))
  _check(lib.cc3d_amd_map_labels(_ptr(ids), ids.numel(), _ptr(lut), _ptr(out_tensor), out_code, _stream()))


def _open_out(out_file, n: int, dtype):
  if isinstance(out_file, str):
    with open(out_file, "wb") as f:
      os.ftruncate(f.fileno(), n * np.dtype(dtype).itemsize)
  return np.memmap(out_file, order="F", dtype=dtype, shape=(n,))


# ---------------------------------------------------------------------------
# connected_components_stack
# ---------------------------------------------------------------------------
def connected_components_stack(stacked_images: Iterable, connectivity: int = 26, return_N: bool = False,
                               binary_image: bool = False, out_file=None):
  """Connected components of a volume given as a sequence of z-slabs of equal
  x and y (cc3d/__init__.py:353-501), labelled one slab at a time on the GPU.

  The reference compresses its result into a crackle array (crackle is not
  available here); this returns the labels as a uint64 array of shape
  (sx, sy, total z) in F order -- or, with out_file, as a memmap written slab
  by slab -- numbered 1..N by first appearance in x-fastest order over the
  whole volume, i.e. exactly connected_components of the concatenated
  volume.  return_N: (labels, N)."""
  if connectivity not in (6, 26):
    raise ValueError(f"Connectivity must be 6 or 26. Got: {connectivity}")
  _require_gpu()
  torch = _torch()
  streamer = None
  parts = []  # (host uint32 labels of one slab, offset)
  shape_xy = None
  code = None
  for image in stacked_images:
    img = np.asfortranarray(np.asarray(image))
    if img.ndim != 3:
      raise ValueError(f"stacked images must be 3D. Got: {img.ndim}")
    if img.size == 0:
      continue
    kind, pl = _prepare(img, -1, connectivity, 0, None, None, False, binary_image, True)
    if shape_xy is None:
      shape_xy = (pl.sx, pl.sy)
      code = pl.code
      streamer = _Streamer(pl.sx, pl.sy, pl.code, connectivity, pl.binary_image, pl.flags, pl.delta_b)
    elif (pl.sx, pl.sy) != shape_xy:
      raise ValueError(f"stacked images must share x and y: {shape_xy} then {(pl.sx, pl.sy)}")
    elif pl.code != code:
      raise ValueError("stacked images must share one dtype")
    flat = np.ravel(img, order="F").view(pl.view_dtype)
    plane = pl.sx * pl.sy
    depth = max(1, _SLAB_VOXELS // plane)
    for z0 in range(0, pl.sz, depth):
      nz = min(depth, pl.sz - z0)
      dev = torch.from_numpy(flat[z0 * plane:(z0 + nz) * plane].view(np.uint8)).to("cuda")
      labels = streamer.slab(dev, nz)
      parts.append((labels.cpu().numpy().view(np.uint32), streamer.offsets[-1]))
      del dev, labels
  if streamer is None:
    out = np.zeros((0, 0, 0), dtype=np.uint64)
    return (out, 0) if return_N else out
  lut, N = streamer.table()
  sx, sy = shape_xy
  total = sum(p.size for p, _ in parts)
  flat_out = _open_out(out_file, total, np.uint64) if out_file is not None else np.empty(total, np.uint64)
  at = 0
  for i, (lab, off) in enumerate(parts):
    dev = torch.from_numpy(lab.view(np.int32)).to("cuda")
    res = torch.empty(lab.size, dtype=torch.int64, device="cuda")
    _map_slab(dev, off, lut, _OUT_CODE[np.dtype(np.uint64)], res)
    flat_out[at:at + lab.size] = res.cpu().numpy().view(np.uint64)
    at += lab.size
    parts[i] = None
  out = flat_out.reshape((sx, sy, total // (sx * sy)), order="F")
  return (out, N) if return_N else out


# ---------------------------------------------------------------------------
# connected_components for volumes of 2^32-1 voxels or more
# ---------------------------------------------------------------------------
def streamable(pl) -> bool:
  return (pl.voxels >= _STREAM_VOXELS and pl.sz > 1 and not pl.periodic_boundary
          and (pl.binary_image or pl.delta == 0) and pl.sx * pl.sy < _STREAM_VOXELS)


class _Counts:
  """The stage-1 counters of the whole volume, shaped like L.Stats."""

  def __init__(self, N, provisional, epl, first, last):
    self.N, self.provisional, self.epl = N, provisional, epl
    self.first_foreground_row = first if first is not None else (1 << 63) - 1
    self.last_foreground_row = last if last is not None else -1


def label_streamed(pl):
  """connected_components for a volume too large for one labelling call:
  the same labels, N, out dtype and errors (cc3d/fastcc3d.pyx:381-434),
  computed slab by slab along the volume's outermost memory axis."""
  from ._core import _deliver
  _require_gpu()
  torch = _torch()
  sx, sy, sz = pl.sx, pl.sy, pl.sz
  plane = sx * sy
  depth = max(1, _SLAB_VOXELS // plane)
  vol = pl.vol
  dev_input = vol.is_torch and vol.device_input
  if dev_input:
    src = pl.data
    if pl.order == "F":
      src = src.permute(*reversed(range(src.ndim)))
    flat_dev = src.contiguous().reshape(-1).view(torch.uint8)
    isz = pl.view_dtype.itemsize
  else:
    arr = pl.data.numpy() if vol.is_torch else np.asarray(pl.data)
    arr = arr if (arr.flags.f_contiguous or arr.flags.c_contiguous) else np.ascontiguousarray(arr)
    flat_host = arr.reshape(-1, order="F" if pl.order == "F" else "C").view(pl.view_dtype)
  streamer = _Streamer(sx, sy, pl.code, pl.connectivity, pl.binary_image, pl.flags, pl.delta_b)
  slabs = []  # (z0, nz, host uint32 labels)
  for z0 in range(0, sz, depth):
    nz = min(depth, sz - z0)
    if dev_input:
      dev = flat_dev[z0 * plane * isz:(z0 + nz) * plane * isz]
    else:
      dev = torch.from_numpy(flat_host[z0 * plane:(z0 + nz) * plane].view(np.uint8)).to("cuda")
    labels = streamer.slab(dev, nz)
    slabs.append((z0, nz, labels.cpu().numpy().view(np.uint32)))
    del dev, labels
  st = _Counts(streamer.nid, streamer.provisional(), streamer.epl, streamer.first_row, streamer.last_row)
  odt, n_dec = _decide(pl, st)
  V = pl.voxels
  out_dev = None
  if dev_input and pl.out_file is None:
    out_dev = torch.empty(V, dtype=_torch_out_dtype(odt), device="cuda")
    target = None
  else:
    target = _open_out(pl.out_file, V, odt) if pl.out_file is not None else np.empty(V, odt)
  N = 0
  if pl.special_row is not None:
    # the single foreground row: the binding's own row labelling on its slab
    z = pl.special_row // sy
    for z0, nz, lab in slabs:
      piece = torch.zeros(nz * plane, dtype=_torch_out_dtype(odt), device="cuda")
      if z0 <= z < z0 + nz:
        sub = _SlabPlan(pl, nz, pl.special_row - sy * z0)
        if dev_input:
          buf = flat_dev[z0 * plane * isz:(z0 + nz) * plane * isz]
        else:
          buf = torch.from_numpy(flat_host[z0 * plane:(z0 + nz) * plane].view(np.uint8)).to("cuda")
        N = _special_row(sub, buf, odt, piece)
      _put(out_dev, target, z0 * plane, piece)
  else:
    lut, N = streamer.table()
    if n_dec == 0:  # max_labels 0: the reference returns before labelling
      lut.zero_()
      N = 0
    for k, (z0, nz, lab) in enumerate(slabs):
      dev = torch.from_numpy(lab.view(np.int32)).to("cuda")
      piece = torch.empty(nz * plane, dtype=_torch_out_dtype(odt), device="cuda")
      _map_slab(dev, streamer.offsets[k], lut, _OUT_CODE[odt], piece)
      _put(out_dev, target, z0 * plane, piece)
  if out_dev is not None:
    return _deliver(out_dev, vol, sx, sy, sz, pl.dims, pl.order, odt, None, N, pl.return_N)
  return _deliver_host(target, vol, sx, sy, sz, pl.dims, pl.order, odt, N, pl.return_N)


class _SlabPlan:
  """The fields _special_row reads, for one slab."""

  def __init__(self, pl, nz, row):
    self.code, self.sx, self.vol = pl.code, pl.sx, pl.vol
    self.voxels = pl.sx * pl.sy * nz
    self.special_row = row


def _put(out_dev, target, at: int, piece):
  if out_dev is not None:
    out_dev[at:at + piece.numel()] = piece
  else:
    target[at:at + piece.numel()] = piece.cpu().numpy().view(target.dtype)


def _deliver_host(flat, vol, sx, sy, sz, dims, order, odt, N, return_N):
  torch = _torch()
  if dims == 3:
    shp = (sz, sy, sx) if order == "C" else (sx, sy, sz)
  elif dims == 2:
    shp = (sy, sx) if order == "C" else (sx, sy)
  else:
    shp = (sx,)
  out = flat.reshape(shp, order=order)
  if vol.is_torch:
    out = torch.from_numpy(np.asarray(out))
    if vol.device_input:
      out = out.to(vol.torch_device)
  return (out, N) if return_N else out
