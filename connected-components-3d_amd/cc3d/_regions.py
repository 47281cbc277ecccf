"""Region adjacency (contacts, region_graph) and run-length rendering (runs,
draw, each) on the GPU.

Drop-ins for cc3d/fastcc3d.pyx:1180-1372 (C++ at cc3d/cc3d_graphs.hpp:238-523):
same arguments, checks, results and errors.  contacts' sums are the float32
sums the reference accumulates, bit for bit (csrc/contacts.hip explains how
they are recovered without its sequential walk); runs are found with a
bitmask scan and a radix sort (csrc/runs.hip), and draw / each paint the runs
with one workgroup per run.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Iterator, List, Tuple

import numpy as np

from . import _lib as L
from ._core import _require_gpu, _stream, _torch

# label dtype codes of cc3d_amd_contacts_run
_CT = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.uint32): 2, np.dtype(np.uint64): 3,
       np.dtype(np.int8): 6, np.dtype(np.int16): 7, np.dtype(np.int32): 8, np.dtype(np.int64): 9,
       np.dtype(np.bool_): 0}
_RUN_TYPES = (np.dtype(np.uint8), np.dtype(np.bool_), np.dtype(np.uint16), np.dtype(np.uint32), np.dtype(np.uint64))
_UNSIGNED = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def _bind():
  lib = L.lib()
  if not hasattr(lib, "_regions_bound"):
    vp, i64, u64, ci, cf = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
    lib.cc3d_amd_contacts_run.restype = ci
    lib.cc3d_amd_contacts_run.argtypes = [vp, ci, i64, i64, i64, ci, ci, cf, cf, cf, ctypes.POINTER(vp),
                                          ctypes.POINTER(u64), vp]
    lib.cc3d_amd_contacts_fetch.restype = ci
    lib.cc3d_amd_contacts_fetch.argtypes = [vp, vp, vp, vp]
    lib.cc3d_amd_contacts_free.restype = None
    lib.cc3d_amd_contacts_free.argtypes = [vp]
    lib.cc3d_amd_contacts_last_error.restype = ctypes.c_char_p
    lib.cc3d_amd_extract_runs.restype = ci
    lib.cc3d_amd_extract_runs.argtypes = [vp, ci, u64, ctypes.POINTER(vp), ctypes.POINTER(u64), vp]
    lib.cc3d_amd_runs_device.restype = ci
    lib.cc3d_amd_runs_device.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]
    lib.cc3d_amd_runs_fetch.restype = ci
    lib.cc3d_amd_runs_fetch.argtypes = [vp, vp, vp, vp]
    lib.cc3d_amd_runs_free.restype = None
    lib.cc3d_amd_runs_free.argtypes = [vp]
    lib.cc3d_amd_draw.restype = ci
    lib.cc3d_amd_draw.argtypes = [vp, ci, u64, u64, vp, vp, u64, vp]
    lib.cc3d_amd_runs_last_error.restype = ctypes.c_char_p
    lib._regions_bound = True
  return lib


def _raise(rc: int, msg: str):
  if rc == L.ERR_RUNTIME:
    raise RuntimeError(msg)
  if rc == L.ERR_NOMEM:
    raise MemoryError(msg)
  if rc == L.ERR_INVALID:
    raise ValueError(msg)
  raise RuntimeError(msg)


def _is_torch(x) -> bool:
  try:
    import torch
  except ImportError:
    return False
  return isinstance(x, torch.Tensor)


def _np_dtype_of(t) -> np.dtype:
  torch = _torch()
  return np.dtype({torch.uint8: np.uint8, torch.int8: np.int8, torch.int16: np.int16, torch.int32: np.int32,
                   torch.int64: np.int64, torch.bool: np.bool_, torch.uint16: np.uint16, torch.uint32: np.uint32,
                   torch.uint64: np.uint64}.get(t.dtype, np.float32))


# ---------------------------------------------------------------------------
# contacts / region_graph
# ---------------------------------------------------------------------------
def contacts(labels, connectivity: int = 26, surface_area: bool = True,
             anisotropy=(1, 1, 1)) -> Dict[Tuple[int, int], float]:
  """The label pairs in contact and their contact area (surface_area) or
  neighbour count (cc3d/fastcc3d.pyx:1198-1252): {(lo, hi): float}.  The
  image is read as 3D with x = axis 0 (fewer axes get trailing ones); faces
  weigh the product of the other two anisotropy weights, edges and corners
  nothing; 4/8 are the 2D neighbourhoods and need a single plane."""
  if _is_torch(labels) and labels.device.type == "cpu":
    labels = labels.numpy()
  dev = _is_torch(labels)
  while len(labels.shape) < 3:
    labels = labels[..., None]
  anisotropy = tuple(anisotropy)
  while len(anisotropy) < 3:
    anisotropy = tuple(anisotropy) + (1,)
  if connectivity not in (4, 8, 6, 18, 26):
    raise ValueError(f"Only (2d) 4, 8, (3d) 6, 18, and 26 connectivities are supported. Got: {connectivity}")
  dtype = _np_dtype_of(labels) if dev else np.asarray(labels).dtype
  if dtype not in _CT:
    raise TypeError("No matching signature found")
  sx, sy, sz = (int(s) for s in labels.shape)
  if sx * sy * sz == 0:
    axis = [i for i, s in enumerate((sx, sy, sz)) if s == 0][0]
    raise IndexError("Out of bounds on buffer access (axis %d)" % axis)
  if connectivity in (4, 8) and sz != 1:
    raise RuntimeError("z thickness must be 1 for 2d region graph extraction.")
  _require_gpu()
  torch = _torch()
  if dev:
    flat = labels.permute(2, 1, 0).contiguous().reshape(-1)
  else:
    host = np.ravel(np.asfortranarray(labels), order="F")
    if host.dtype == np.bool_:
      host = host.view(np.uint8)
    flat = torch.from_numpy(host.view(np.uint8)).to("cuda")
  wx, wy, wz = (float(np.float32(a)) for a in anisotropy)
  lib = _bind()
  res = ctypes.c_void_p()
  n = ctypes.c_uint64(0)
  rc = lib.cc3d_amd_contacts_run(ctypes.c_void_p(flat.data_ptr()), _CT[dtype], sx, sy, sz, int(connectivity),
                                 int(bool(surface_area)), wx, wy, wz, ctypes.byref(res), ctypes.byref(n), _stream())
  if rc != L.OK:
    _raise(rc, lib.cc3d_amd_contacts_last_error().decode(errors="replace"))
  try:
    m = int(n.value)
    lo = np.empty(m, np.uint64)
    hi = np.empty(m, np.uint64)
    val = np.empty(m, np.float32)
    rc = lib.cc3d_amd_contacts_fetch(res, lo.ctypes.data, hi.ctypes.data, val.ctypes.data)
    if rc != L.OK:
      _raise(rc, lib.cc3d_amd_contacts_last_error().decode(errors="replace"))
  finally:
    lib.cc3d_amd_contacts_free(res)
  if dtype.kind == "i":
    lo, hi = lo.view(np.int64), hi.view(np.int64)
  order = np.lexsort((hi, lo))
  lo, hi, val = lo[order].tolist(), hi[order].tolist(), val[order].tolist()
  return {(a, b): v for a, b, v in zip(lo, hi, val)}


def region_graph(labels, connectivity: int = 26) -> set:
  """The set of label pairs in contact (cc3d/fastcc3d.pyx:1180-1196)."""
  return set(contacts(labels, connectivity=connectivity).keys())


# ---------------------------------------------------------------------------
# runs / draw / each
# ---------------------------------------------------------------------------
def _flat_view(arr):
  """The flat array in memory order, as _reshape(arr, (arr.size,)) gives it
  (cc3d/fastcc3d.pyx:129-161): a view for F- or C-contiguous arrays, else a
  C-order copy."""
  if arr.flags.f_contiguous:
    return arr.reshape(-1, order="F") if arr.ndim else arr.reshape(1)
  if arr.flags.c_contiguous:
    return arr.reshape(-1)
  return arr.reshape((arr.size,))


class _Runs:
  """Runs of one labelling, kept on the device (label, start, end sorted by
  label, position within a label) together with their host copy."""

  def __init__(self, flat_dev, itemsize: int, n: int):
    self.lib = _bind()
    self.n = n
    self.handle = ctypes.c_void_p()
    cnt = ctypes.c_uint64(0)
    rc = self.lib.cc3d_amd_extract_runs(ctypes.c_void_p(flat_dev.data_ptr() if n else 0), itemsize, n,
                                        ctypes.byref(self.handle), ctypes.byref(cnt), _stream())
    if rc != L.OK:
      _raise(rc, self.lib.cc3d_amd_runs_last_error().decode(errors="replace"))
    self.count = int(cnt.value)
    self.label = np.empty(self.count, np.uint64)
    self.start = np.empty(self.count, np.uint64)
    self.end = np.empty(self.count, np.uint64)
    rc = self.lib.cc3d_amd_runs_fetch(self.handle, self.label.ctypes.data, self.start.ctypes.data,
                                      self.end.ctypes.data)
    if rc != L.OK:
      self.close()
      _raise(rc, self.lib.cc3d_amd_runs_last_error().decode(errors="replace"))
    lab, st, en = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    self.lib.cc3d_amd_runs_device(self.handle, ctypes.byref(lab), ctypes.byref(st), ctypes.byref(en))
    self.dstart, self.dend = st.value or 0, en.value or 0
    # label groups: [first, last) index range of each label's runs
    if self.count:
      cut = np.flatnonzero(np.diff(self.label)) + 1
      self.bounds = np.concatenate(([0], cut, [self.count]))
    else:
      self.bounds = np.zeros(1, np.int64)

  def groups(self):
    """(label, first, last) per label, label 0 left out except for a one-voxel
    array (cc3d/cc3d_graphs.hpp:477-487)."""
    for g in range(len(self.bounds) - 1):
      b, e = int(self.bounds[g]), int(self.bounds[g + 1])
      key = int(self.label[b])
      if key == 0 and self.n != 1:
        continue
      yield key, b, e

  def as_dict(self) -> Dict[int, List[Tuple[int, int]]]:
    pairs = list(zip(self.start.tolist(), self.end.tolist()))
    return {key: pairs[b:e] for key, b, e in self.groups()}

  def close(self):
    if self.handle:
      self.lib.cc3d_amd_runs_free(self.handle)
      self.handle = ctypes.c_void_p()

  def __del__(self):
    try:
      self.close()
    except Exception:
      pass


def _runs_of(labels):
  """(_Runs, flat memory-order shape info) for numpy arrays and tensors."""
  _require_gpu()
  torch = _torch()
  if _is_torch(labels):
    if labels.device.type == "cpu":
      labels = labels.numpy()
    else:
      dtype = _np_dtype_of(labels)
      if dtype not in _RUN_TYPES:
        raise TypeError("Unsupported type: " + str(dtype))
      flat = labels.reshape(-1)
      if flat.dtype == torch.bool:
        flat = flat.view(torch.uint8)
      return _Runs(flat, dtype.itemsize, int(flat.numel()))
  labels = np.asarray(labels)
  flat = _flat_view(labels)
  if flat.dtype not in _RUN_TYPES:
    raise TypeError("Unsupported type: " + str(flat.dtype))
  host = np.ascontiguousarray(flat).view(_UNSIGNED[flat.dtype.itemsize])
  dev = torch.from_numpy(host.view(np.uint8)).to("cuda") if host.size else None
  return _Runs(dev, flat.dtype.itemsize, int(flat.size))


def runs(labels) -> Dict[int, List[Tuple[int, int]]]:
  """Where each label lies (cc3d/fastcc3d.pyx:1254-1277): {label: [(start,
  end), ...]} over the flat array in memory order, labels ascending, runs in
  position order, label 0 left out (except for a one-element array)."""
  r = _runs_of(labels)
  try:
    return r.as_dict()
  finally:
    r.close()


def _value_of(label, dtype: np.dtype) -> int:
  if dtype == np.bool_:
    return int(label != 0)
  v = int(label)
  info = np.iinfo(dtype)
  if v < info.min or v > info.max:
    raise OverflowError(f"value too large to convert to {dtype.name}_t")
  return v


def _draw_device(dev_flat, itemsize: int, n: int, value: int, dstart: int, dend: int, nruns: int):
  lib = _bind()
  rc = lib.cc3d_amd_draw(ctypes.c_void_p(dev_flat.data_ptr()), itemsize, n, value, ctypes.c_void_p(dstart),
                         ctypes.c_void_p(dend), nruns, _stream())
  if rc != L.OK:
    _raise(rc, lib.cc3d_amd_runs_last_error().decode(errors="replace"))


def draw(label, runs, image):
  """Paints label over the runs of the flat image (cc3d/fastcc3d.pyx:1279-1308):
  a bool image gets label != 0.  Returns the flat view of image (a C-order
  copy for a non-contiguous image).  A run ending past the image or not
  longer than zero raises RuntimeError("Invalid run.") after the runs before
  it are painted."""
  _require_gpu()
  torch = _torch()
  dev_in = _is_torch(image) and image.device.type != "cpu"
  if _is_torch(image) and not dev_in:
    image = image.numpy()
  if dev_in:
    dtype = _np_dtype_of(image)
    flat = image.reshape(-1)
  else:
    flat = _flat_view(image)
    dtype = flat.dtype
  if dtype not in (np.dtype(np.bool_), np.dtype(np.uint8), np.dtype(np.uint16), np.dtype(np.uint32),
                   np.dtype(np.uint64)):
    raise TypeError("Unsupported type: " + str(dtype))
  rl = [(int(s), int(e)) for s, e in runs]
  for s, e in rl:
    if s < 0 or e < 0:
      raise OverflowError("can't convert negative value to size_t")
  value = _value_of(label, dtype)
  n = int(flat.numel()) if dev_in else int(flat.size)
  if not rl:
    return flat
  ra = torch.tensor(np.array(rl, dtype=np.uint64).view(np.int64), device="cuda")  # (k, 2)
  st, en = ra[:, 0].contiguous(), ra[:, 1].contiguous()
  if dev_in:
    target = flat.view(torch.uint8) if flat.dtype == torch.bool else flat
    _draw_device(target, dtype.itemsize, n, value, st.data_ptr(), en.data_ptr(), len(rl))
    return flat
  host = flat.view(_UNSIGNED[dtype.itemsize]) if dtype != np.bool_ else flat.view(np.uint8)
  dev = torch.from_numpy(np.ascontiguousarray(host).view(np.uint8)).to("cuda")
  try:
    _draw_device(dev, dtype.itemsize, n, value, st.data_ptr(), en.data_ptr(), len(rl))
  finally:
    host[...] = dev.cpu().numpy().view(host.dtype)
  return flat


class _ImageIterator:
  """cc3d.each's iterator (cc3d/fastcc3d.pyx:1336-1372): one image per nonzero
  label, painted on the device from the device-resident runs."""

  def __init__(self, labels, binary: bool, in_place: bool):
    self.r = _runs_of(labels)
    self.dev = _is_torch(labels) and labels.device.type != "cpu"
    if _is_torch(labels) and not self.dev:
      labels = labels.numpy()
    self.shape = tuple(labels.shape)
    src = _np_dtype_of(labels) if self.dev else labels.dtype
    self.dtype = np.dtype(np.bool_) if binary else np.dtype(src)
    self.order = "C" if self.dev or not labels.flags.f_contiguous else "F"
    self.in_place = in_place
    self.keys = list(self.r.groups())

  def __len__(self):
    return len(self.keys) - int(any(k == 0 for k, _, _ in self.keys))

  def __iter__(self) -> Iterator[Tuple[int, object]]:
    torch = _torch()
    n = self.r.n
    item = self.dtype.itemsize
    buf = torch.zeros(max(n * item, 1), dtype=torch.uint8, device="cuda")
    host = None
    if self.in_place:
      host = np.zeros(self.shape, dtype=self.dtype, order=self.order)
    for key, b, e in self.keys:
      if key == 0:
        continue
      value = _value_of(key, self.dtype)
      _draw_device(buf, item, n, value, self.r.dstart + 8 * b, self.r.dend + 8 * b, e - b)
      if self.dev:
        img = buf[:n * item].view({1: torch.uint8, 2: torch.uint16, 4: torch.uint32, 8: torch.uint64}[item])
        img = (img.view(torch.bool) if self.dtype == np.bool_ else img).clone().reshape(self.shape)
        yield key, img
      elif self.in_place:
        host.setflags(write=1)
        _flat_view(host).view(np.uint8)[...] = buf[:n * item].cpu().numpy()
        host.setflags(write=0)
        yield key, host
        host.setflags(write=1)
      else:
        img = np.empty(self.shape, dtype=self.dtype, order=self.order)
        _flat_view(img).view(np.uint8)[...] = buf[:n * item].cpu().numpy()
        yield key, img
      # erase this label's runs
      _draw_device(buf, item, n, 0, self.r.dstart + 8 * b, self.r.dend + 8 * b, e - b)
      if self.in_place:
        _flat_view(host).view(np.uint8)[...] = buf[:n * item].cpu().numpy()


def each(labels, binary: bool = False, in_place: bool = False):
  """Iterates (label, image) over the nonzero labels in ascending order
  (cc3d/fastcc3d.pyx:1310-1372).  binary: bool images, else the labels'
  dtype holding the label; in_place: one read-only image reused (valid until
  the next step)."""
  return _ImageIterator(labels, binary, in_place)
