"""Checker-side bindings: the C restatement (oracle.c) and the compiled reference.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
bench's ``cpu_baseline`` leg import this module, and only as the checker.  The
product package never imports it.

Two backends share one Python restatement of the binding layer
(cc3d/fastcc3d.pyx:244-642):

* ``"oracle"`` -- oracle/build/liboracle.so, the C restatement of the
  labelling core (cc3d/cc3d.hpp, cc3d/cc3d_binary.hpp,
  cc3d/cc3d_continuous.hpp).
* ``"ref"``    -- oracle/_ref/libcc3d_ref.so, the unmodified reference headers
  compiled in place (oracle/Makefile).  Present wherever the reference tree
  was available at build time; the library itself travels with the repo.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcc3d_ref.so")

DT_CODE = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.uint32): 2,
           np.dtype(np.uint64): 3, np.dtype(np.float32): 4, np.dtype(np.float64): 5}
MODE_MULTI, MODE_BINARY, MODE_CONT = 0, 1, 2

_oracle = None
_ref = None
# "special_row" or "core": which path the last connected_components call took
last_path = None


class DimensionError(Exception):
  pass


def build(force: bool = False) -> None:
  """Compile the oracle (and, where the reference tree exists, oracle/_ref)."""
  import subprocess
  subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _lib_oracle():
  global _oracle
  if _oracle is None:
    if not os.path.exists(ORACLE_SO):
      build()
    lib = ctypes.CDLL(ORACLE_SO)
    lib.oracle_label.restype = ctypes.c_int
    lib.oracle_label.argtypes = [
      ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
      ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
      ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
    ]
    lib.oracle_statistics.restype = ctypes.c_int
    lib.oracle_statistics.argtypes = [
      ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
      ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_set_mutation.restype = None
    lib.oracle_set_mutation.argtypes = [ctypes.c_int]
    lib.oracle_epl.restype = ctypes.c_uint64
    lib.oracle_epl.argtypes = [
      ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
    ]
    _oracle = lib
  return _oracle


def have_ref() -> bool:
  return os.path.exists(REF_SO)


def _lib_ref():
  global _ref
  if _ref is None:
    lib = ctypes.CDLL(REF_SO)
    lib.ref_connected_components3d.restype = ctypes.c_int
    lib.ref_connected_components3d.argtypes = [
      ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
      ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64,
      ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
      ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
    ]
    lib.ref_estimate_provisional_label_count.restype = ctypes.c_uint64
    lib.ref_estimate_provisional_label_count.argtypes = [
      ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
    ]
    _ref = lib
  return _ref


# ---------------------------------------------------------------------------
# core calls (input already viewed as u8/u16/u32/u64/f32/f64, x fastest)
# ---------------------------------------------------------------------------

def core_label(flat: np.ndarray, sx: int, sy: int, sz: int, connectivity: int,
               mode: int, delta=0, periodic: bool = False) -> Tuple[np.ndarray, int, int]:
  """oracle.c: labels as uint64, N, provisional count P."""
  lib = _lib_oracle()
  flat = np.ascontiguousarray(flat).reshape(-1)
  out = np.zeros(flat.size, dtype=np.uint64)
  N = ctypes.c_uint64(0)
  P = ctypes.c_uint64(0)
  du, df = _delta_pair(flat.dtype, delta)
  rc = lib.oracle_label(flat.ctypes.data, DT_CODE[flat.dtype], sx, sy, sz, connectivity,
                        mode, du, df, int(bool(periodic)), out.ctypes.data,
                        ctypes.byref(N), ctypes.byref(P))
  if rc != 0:
    raise RuntimeError("oracle_label failed")
  return out, int(N.value), int(P.value)


def set_mutation(m: int) -> None:
  """Plant a known mistake in the oracle (0 = none); see oracle.c."""
  _lib_oracle().oracle_set_mutation(int(m))


def core_epl(flat: np.ndarray, sx: int):
  lib = _lib_oracle()
  flat = np.ascontiguousarray(flat).reshape(-1)
  f = ctypes.c_int64(0)
  l = ctypes.c_int64(0)
  e = lib.oracle_epl(flat.ctypes.data, DT_CODE[flat.dtype], sx, flat.size,
                     ctypes.byref(f), ctypes.byref(l))
  return int(e), int(f.value), int(l.value)


def ref_core_label(flat: np.ndarray, sx: int, sy: int, sz: int, max_labels: int,
                   connectivity: int, delta, out_dtype, periodic: bool, binary: bool):
  """The compiled reference: (labels in out_dtype, N) or raises RuntimeError."""
  lib = _lib_ref()
  flat = np.ascontiguousarray(flat).reshape(-1)
  out = np.zeros(flat.size, dtype=out_dtype)
  N = ctypes.c_uint64(0)
  du, df = _delta_pair(flat.dtype, delta)
  err = ctypes.create_string_buffer(256)
  rc = lib.ref_connected_components3d(
    flat.ctypes.data, DT_CODE[flat.dtype], sx, sy, sz, max_labels, connectivity,
    df, du, out.ctypes.data, DT_CODE[np.dtype(out_dtype)], ctypes.byref(N),
    int(bool(periodic)), int(bool(binary)), err, 256)
  if rc != 0:
    raise RuntimeError(err.value.decode())
  return out, int(N.value)


def ref_core_epl(flat: np.ndarray, sx: int):
  lib = _lib_ref()
  flat = np.ascontiguousarray(flat).reshape(-1)
  f = ctypes.c_int64(0)
  l = ctypes.c_int64(0)
  e = lib.ref_estimate_provisional_label_count(
    flat.ctypes.data, DT_CODE[flat.dtype], sx, flat.size, ctypes.byref(f), ctypes.byref(l))
  return int(e), int(f.value), int(l.value)


def _delta_pair(dtype, delta):
  if np.dtype(dtype).kind == "f":
    return 0, float(delta)
  d = int(delta)
  if d < 0:
    raise OverflowError("can't convert negative value to unsigned delta")
  return d & 0xFFFFFFFFFFFFFFFF, float(d)


# ---------------------------------------------------------------------------
# union-find length of each reference first pass (for the P >= L check)
# ---------------------------------------------------------------------------

def uf_length(max_labels: int, voxels: int, connectivity: int, binary: bool,
              out_dtype) -> Optional[int]:
  """DisjointSet length the reference allocates, or None when it returns early.

  cc3d/cc3d.hpp:360-366 (and the 18/6/2D twins), cc3d/cc3d_binary.hpp:47-51,
  :350-352, :527-530, :794-797, :964-967, cc3d/cc3d_continuous.hpp:105-107.
  """
  omax = int(np.iinfo(out_dtype).max)
  if binary and connectivity == 8:
    ml = max(min(max_labels + 1, (voxels >> 1) + 1), 1) + 1
    return min(ml, omax)
  if max_labels == 0:
    return None
  ml = max_labels + 1
  if binary and connectivity == 26:
    ml = min(ml, (voxels >> 2) + 1) + 1
  elif binary and connectivity in (6, 4):
    ml = min(ml, (voxels >> 1) + 1) + 1
  else:
    ml = min(ml, voxels + 1)
    if connectivity == 8:
      ml = max(ml, 1)
  return min(ml, omax)


# ---------------------------------------------------------------------------
# restatement of the binding layer, cc3d/fastcc3d.pyx:244-642
# ---------------------------------------------------------------------------

def _even_ceil(n: int) -> int:
  return n << 1 if n & 1 else n


def _view_unsigned(data: np.ndarray) -> np.ndarray:
  dt = data.dtype
  if dt in (np.uint64, np.int64):
    return data.view(np.uint64)
  if dt in (np.uint32, np.int32):
    return data.view(np.uint32)
  if dt in (np.uint16, np.int16):
    return data.view(np.uint16)
  if dt in (np.uint8, np.int8, np.bool_):
    return data.view(np.uint8)
  if dt in (np.float32, np.float64):
    return data
  raise TypeError(f"Type {dt} is not currently supported.")


def estimate_provisional_labels(data: np.ndarray, backend: str = "oracle"):
  """cc3d/fastcc3d.pyx:168-242."""
  sx = data.shape[0] if data.flags.f_contiguous else data.shape[-1]
  flat = np.ravel(_view_unsigned(np.asarray(data)), order="K")
  if backend == "ref":
    return ref_core_epl(flat, sx)
  return core_epl(flat, sx)


def connected_components(data, max_labels=-1, connectivity=26, return_N=False,
                         delta=0, out_dtype=None, periodic_boundary=False,
                         binary_image=False, backend: str = "oracle"):
  """Checker-side cc3d.connected_components (cc3d/fastcc3d.pyx:244-626)."""
  data = np.asarray(data)
  dims = data.ndim
  if dims not in (1, 2, 3):
    raise DimensionError(f"Only 1D, 2D, and 3D arrays supported. Got: {dims}")
  if dims == 2 and connectivity not in (4, 8, 6, 18, 26):
    raise ValueError(f"Only 4, 8, and 6, 18, 26 connectivities are supported for 2D images. Got: {connectivity}")
  elif dims != 2 and connectivity not in (6, 18, 26):
    raise ValueError(f"Only 6, 18, and 26 connectivities are supported for 3D images. Got: {connectivity}")
  if periodic_boundary and connectivity not in (4, 8, 6):
    raise ValueError(f"periodic_boundary is not yet implemented for {connectivity}-connectivity.")
  if periodic_boundary and delta != 0:
    raise ValueError("periodic_boundary is not yet implemented continuous data.")
  if data.size == 0:
    dt = out_dtype if out_dtype is not None else data.dtype
    out = np.zeros((0,), dtype=dt)
    return (out, 0) if return_N else out

  order = "F" if data.flags.f_contiguous else "C"
  while data.ndim < 3:
    data = data[np.newaxis, ...] if order == "C" else data[..., np.newaxis]
  if not data.flags.c_contiguous and not data.flags.f_contiguous:
    data = np.copy(data, order=order)
  if data.dtype == np.float16:
    if delta == 0:
      data = data.view(np.uint16)
    else:
      raise TypeError("float16 is not supported for continuous images (delta != 0).")

  shape = list(data.shape)
  if order == "C":
    shape.reverse()
  sx, sy, sz = shape
  voxels = sx * sy * sz
  dtype = data.dtype
  binary_image = bool(binary_image) or dtype == np.bool_

  if binary_image:
    epl, ffr, lfr = voxels, 0, sy
  else:
    epl, ffr, lfr = estimate_provisional_labels(data, backend)
  max_labels = min(epl, voxels)

  if np.issubdtype(dtype, np.floating):
    delta = float(delta)
    binary_image = binary_image or (delta == np.finfo(dtype).max)
  else:
    delta = int(delta)
    binary_image = binary_image or (delta == np.iinfo(dtype).max)

  ufv = _even_ceil(data.shape[0]) * _even_ceil(data.shape[1]) * _even_ceil(data.shape[2])
  if binary_image:
    if connectivity in (4, 6):
      max_labels = min(max_labels, (ufv // 2) + 1)
    else:  # the reference's `connectivity == (8,18)` is never true
      max_labels = min(max_labels, (ufv // 8) + 1)

  if out_dtype is not None:
    out_dtype = np.dtype(out_dtype)
    if out_dtype not in (np.uint16, np.uint32, np.uint64):
      raise ValueError(f"Explicitly defined out_dtype ({out_dtype}) must be one of: np.uint16, np.uint32, np.uint64")
    if np.iinfo(out_dtype).max < max_labels:
      raise ValueError(f"Explicitly defined out_dtype ({out_dtype}) is too small to contain the estimated maximum number of labels ({max_labels}).")
  elif max_labels < np.iinfo(np.uint16).max:
    out_dtype = np.dtype(np.uint16)
  elif max_labels < np.iinfo(np.uint32).max:
    out_dtype = np.dtype(np.uint32)
  else:
    out_dtype = np.dtype(np.uint64)

  flat = np.ravel(_view_unsigned(data), order="K")
  mode = MODE_BINARY if binary_image else (MODE_MULTI if delta == 0 else MODE_CONT)

  global last_path
  last_path = "special_row" if (delta == 0 and ffr == lfr and ffr >= 0) else "core"
  if delta == 0 and ffr == lfr and ffr >= 0:
    # The C++ core has no such path (the binding handles it in Python), so
    # both backends use this restatement; tests/golden/special_row_vectors.npz
    # (outputs of the reference's compiled binding) pins it.
    out, N = _special_row(ffr, sx, sy, np.ravel(data, order="K"), out_dtype)
  elif backend == "ref":
    out, N = ref_core_label(flat, sx, sy, sz, max_labels, connectivity, delta,
                            out_dtype, periodic_boundary, binary_image)
  else:
    L = uf_length(max_labels, voxels, connectivity, binary_image, out_dtype)
    if L is None:
      out, N = np.zeros(voxels, dtype=out_dtype), 0
    else:
      if connectivity in (4, 8) and sz != 1:
        raise RuntimeError("sz must be 1 for 2D connectivities.")
      lab, N, P = core_label(flat, sx, sy, sz, connectivity, mode, delta, periodic_boundary)
      if P >= L:
        raise RuntimeError("maximum length exception")
      out = lab.astype(out_dtype)

  if dims == 3:
    out = out.reshape((sz, sy, sx) if order == "C" else (sx, sy, sz), order=order)
  elif dims == 2:
    out = out.reshape((sy, sx) if order == "C" else (sx, sy), order=order)
  else:
    out = out.reshape((sx,), order=order)
  return (out, N) if return_N else out


def _int64_of(val) -> int:
  """Cython's conversion of a numpy scalar to `cdef int64_t`: int() of the
  value (truncation for floats), then the C range check."""
  if isinstance(val, np.floating):
    f = float(val)
    if f != f:
      raise ValueError("cannot convert float NaN to integer")
    if f in (float("inf"), float("-inf")):
      raise OverflowError("cannot convert float infinity to integer")
  i = int(val)
  if i >= (1 << 63) or i < -(1 << 63):
    raise OverflowError("Python int too large to convert to C long")
  return i


def _special_row(row: int, sx: int, sy: int, flat: np.ndarray, out_dtype):
  """cc3d/fastcc3d.pyx:644-679: the single-foreground-row shortcut.

  `flat` holds the binding's own `data` in memory order (signed and float
  values as they are; float16 already viewed as uint16).  `last_label` is a C
  int64: each new run stores the int64 conversion of its first value
  (_int64_of, which raises like the reference), and a voxel continues the run
  when it compares equal to that integer.  Each label is stored before the
  conversion, so a label outside out_dtype raises numpy's OverflowError
  first."""
  out = np.zeros(flat.size, dtype=out_dtype)
  outmax = int(np.iinfo(out_dtype).max)
  start = row * sx
  N = 0
  last = 0
  for i in range(sx):
    val = flat[start + i]
    if val == 0:
      last = 0
      continue
    elif val == last:
      out[start + i] = N
      continue
    else:
      N += 1
      if N > outmax:
        raise OverflowError(f"Python integer {N} out of bounds for {np.dtype(out_dtype).name}")
      out[start + i] = N
      last = _int64_of(val)
  return out, N


# ---------------------------------------------------------------------------
# voxel connectivity graph
# ---------------------------------------------------------------------------

# (dx, dy, dz) of each bit, as published by the reference: the docstring at
# cc3d/fastcc3d.pyx:1030-1064 and the expected values of automated_test.py:878-950.
VCG_DIRS_3D = [
  (+1, 0, 0), (-1, 0, 0), (0, +1, 0), (0, -1, 0), (0, 0, +1), (0, 0, -1),
  (+1, +1, 0), (-1, +1, 0), (+1, -1, 0), (-1, -1, 0),
  (+1, 0, +1), (-1, 0, +1), (0, +1, +1), (0, -1, +1), (+1, 0, -1), (-1, 0, -1), (0, +1, -1), (0, -1, -1),
  (+1, +1, +1), (-1, +1, +1), (+1, -1, +1), (-1, -1, +1), (+1, +1, -1), (-1, +1, -1), (+1, -1, -1), (-1, -1, -1),
]
VCG_DIRS_2D = [(+1, 0), (-1, 0), (0, +1), (0, -1), (+1, +1), (-1, +1), (+1, -1), (-1, -1)]


def _vcg_bits(lab: np.ndarray, dirs) -> np.ndarray:
  """Per voxel: bit d cleared where the neighbour at dirs[d] exists and
  differs (cc3d/cc3d_graphs.hpp:31-182: start from all ones, clear both ends
  of every differing pair)."""
  g = np.zeros(lab.shape, dtype=np.uint64)
  for d, off in enumerate(dirs):
    keep = np.ones(lab.shape, dtype=bool)
    src = tuple(slice(max(0, -o), n - max(0, o)) for o, n in zip(off, lab.shape))
    dst = tuple(slice(max(0, o), n - max(0, -o)) for o, n in zip(off, lab.shape))
    keep[src] = lab[src] == lab[dst]  # voxel p at src, its neighbour p + off at dst
    g |= keep.astype(np.uint64) << np.uint64(d)
  return g


def voxel_connectivity_graph(data, connectivity=26, backend: str = "oracle"):
  """Checker-side cc3d.voxel_connectivity_graph (cc3d/fastcc3d.pyx:1021-1176,
  cc3d/cc3d_graphs.hpp:192-236)."""
  data = np.asarray(data)
  dims = data.ndim
  if dims not in (1, 2, 3):
    raise DimensionError("Only 1D, 2D, and 3D arrays supported. Got: " + str(dims))
  if dims == 2 and connectivity not in (4, 8, 6, 18, 26):
    raise ValueError("Only 4, 8, and 6, 18, 26 connectivities are supported for 2D images. Got: " + str(connectivity))
  elif dims != 2 and connectivity not in (6, 18, 26):
    raise ValueError("Only 6, 18, and 26 connectivities are supported for 3D images. Got: " + str(connectivity))
  out_dtype = np.uint8 if connectivity in (4, 8, 6) else np.uint32
  if data.size == 0:
    return np.zeros(shape=(0,), dtype=out_dtype)
  data = np.asfortranarray(data)
  while data.ndim < 3:
    data = data[..., np.newaxis]
  if data.dtype in (np.uint64, np.int64):
    lab = data.view(np.uint64)
  elif data.dtype in (np.uint32, np.int32):
    lab = data.view(np.uint32)
  elif data.dtype in (np.uint16, np.int16):
    lab = data.view(np.uint16)
  elif data.dtype in (np.uint8, np.int8, bool):
    lab = data.view(np.uint8)
  else:
    raise TypeError("Type {} not currently supported.".format(data.dtype))
  sx, sy, sz = lab.shape
  if backend == "ref":
    lib = _lib_ref()
    lib.ref_voxel_connectivity_graph.restype = ctypes.c_int
    lib.ref_voxel_connectivity_graph.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                                 ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int]
    flat = np.ravel(lab, order="F")
    out = np.zeros(flat.size, dtype=out_dtype)
    rc = lib.ref_voxel_connectivity_graph(flat.ctypes.data, DT_CODE[flat.dtype], sx, sy, sz, connectivity,
                                          out.ctypes.data, out.itemsize)
    if rc != 0:
      raise RuntimeError("reference voxel_connectivity_graph failed")
    graph = out.reshape((sx, sy, sz), order="F")
  else:
    if connectivity in (4, 8):
      if sz != 1:
        raise RuntimeError("sz must be 1 for 2D connectivities.")
      g = _vcg_bits(lab[:, :, 0], VCG_DIRS_2D)[:, :, np.newaxis]
      mask = 0xFF if connectivity == 8 else 0x0F
    else:
      g = _vcg_bits(lab, VCG_DIRS_3D)
      mask = {26: 0x3FFFFFF, 18: 0x3FFFF, 6: 0x3F}[connectivity]
    graph = np.asfortranarray((g & np.uint64(mask)).astype(out_dtype))
  if dims == 3:
    return graph.reshape((sx, sy, sz), order="F")
  elif dims == 2:
    return graph.reshape((sx, sy), order="F")
  return graph.reshape((sx,), order="F")


# ---------------------------------------------------------------------------
# statistics
# ---------------------------------------------------------------------------

def statistics_raw(labels: np.ndarray, N: Optional[int] = None):
  """cc3d/fastcc3d.pyx:681-938 over oracle.c's oracle_statistics: (counts
  uint32[N+1], bbox uint32[N+1, 2*ndim] as min,max per numpy axis (min
  0xFFFFFFFF for absent labels), centroids float64[N+1, ndim] (NaN for absent
  labels)).  labels: a 2D or 3D unsigned (or bool) array, C or F contiguous."""
  lab = labels
  if lab.dtype == np.bool_:
    lab = lab.view(np.uint8)
  if lab.dtype.kind == "i":
    lab = lab.view(f"u{lab.dtype.itemsize}")
  f_order = bool(lab.flags.f_contiguous)
  if not (lab.flags.c_contiguous or f_order):
    lab = np.ascontiguousarray(lab)
  nd = lab.ndim
  s = list(lab.shape) + [1] * (3 - nd)
  if N is None:
    N = int(lab.max())
  counts = np.zeros(N + 1, dtype=np.uint32)
  bbox = np.zeros((N + 1) * 6, dtype=np.uint32)
  sums = np.zeros((N + 1) * 3, dtype=np.uint64)
  rc = _lib_oracle().oracle_statistics(lab.ctypes.data, DT_CODE[lab.dtype], s[0], s[1], s[2], int(f_order), N,
                                       counts.ctypes.data, bbox.ctypes.data, sums.ctypes.data)
  if rc:
    raise ValueError("label above N")
  bbox = bbox.reshape(N + 1, 6)[:, :2 * nd]
  sums = sums.reshape(N + 1, 3)[:, :nd].astype(np.float64)
  with np.errstate(invalid="ignore", divide="ignore"):
    cen = sums / counts[:, None].astype(np.float64)
  cen[counts == 0] = np.nan
  return counts, bbox, cen


# ---------------------------------------------------------------------------
# color_connectivity_graph
# ---------------------------------------------------------------------------

def color_connectivity_graph(vcg, connectivity=26, return_N=False, backend: str = "oracle"):
  """Checker-side cc3d.color_connectivity_graph (cc3d/fastcc3d.pyx:941-1018):
  the binding restated in Python over oracle.c's oracle_color_vcg, or over
  the reference's own color_connectivity_graph_N (oracle/_ref)."""
  vcg = np.asarray(vcg)
  dims = vcg.ndim
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
  flat = np.ravel(vcg, order="F")
  out = np.zeros(flat.size, dtype=np.uint32)
  if backend == "ref":
    lib = _lib_ref()
    lib.ref_color_connectivity_graph.restype = ctypes.c_int
    lib.ref_color_connectivity_graph.argtypes = [
      ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
      ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, ctypes.c_int]
    n = ctypes.c_uint64(0)
    err = ctypes.create_string_buffer(256)
    rc = lib.ref_color_connectivity_graph(flat.ctypes.data, flat.itemsize, sx, sy, sz, int(connectivity),
                                          out.ctypes.data, ctypes.byref(n), err, 256)
    if rc != 0:
      raise RuntimeError(err.value.decode())
    N = int(n.value)
  else:
    if connectivity not in (26, 6, 4, 8):
      raise RuntimeError("Only 4, 8, 6 and 26 connectivities are supported.")
    if sz > 1 and connectivity not in (6, 26):
      raise RuntimeError("Only 6 and 26 connectivity is supported in 3D.")
    if sz == 1:
      variant = 3 if connectivity in (6, 4) else 2
    else:
      variant = 1 if connectivity == 6 else 0
    lib = _lib_oracle()
    lib.oracle_color_vcg.restype = ctypes.c_uint64
    lib.oracle_color_vcg.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int, ctypes.c_void_p]
    N = int(lib.oracle_color_vcg(flat.ctypes.data, flat.itemsize, sx, sy, sz, variant, out.ctypes.data))
  out = out.reshape((sx, sy, sz), order="F")
  while out.ndim > dims:
    out = out[..., 0]
  return (out, N) if return_N else out


# ---------------------------------------------------------------------------
# contacts / region_graph (cc3d/fastcc3d.pyx:1180-1252) and runs / draw /
# each (cc3d/fastcc3d.pyx:1254-1372)
# ---------------------------------------------------------------------------
CT_CODE = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.uint32): 2, np.dtype(np.uint64): 3,
           np.dtype(np.int8): 6, np.dtype(np.int16): 7, np.dtype(np.int32): 8, np.dtype(np.int64): 9,
           np.dtype(bool): 0}


def contacts(labels, connectivity=26, surface_area=True, anisotropy=(1, 1, 1), backend: str = "oracle"):
  """cc3d.contacts restated (cc3d/fastcc3d.pyx:1198-1252): labels padded to 3
  axes, anisotropy to 3 weights, connectivity checked, F order; the C
  restatement (oracle_contacts) or the compiled reference computes the sums.
  Returns {(lo, hi): float}."""
  labels = np.asarray(labels)
  while len(labels.shape) < 3:
    labels = labels[..., np.newaxis]
  anisotropy = tuple(anisotropy)
  while len(anisotropy) < 3:
    anisotropy = tuple(anisotropy) + (1,)
  if connectivity not in (4, 8, 6, 18, 26):
    raise ValueError(f"Only (2d) 4, 8, (3d) 6, 18, and 26 connectivities are supported. Got: {connectivity}")
  if labels.dtype not in CT_CODE:
    raise TypeError("No matching signature found")
  labels = np.asfortranarray(labels)
  sx, sy, sz = labels.shape
  if labels.size == 0:
    raise IndexError("Out of bounds on buffer access (axis %d)" % [i for i, s in enumerate(labels.shape) if s == 0][0])
  code = CT_CODE[labels.dtype]
  signed = labels.dtype.kind == "i"
  flat = np.ravel(labels, order="F")
  if labels.dtype == bool:
    flat = flat.view(np.uint8)
  wx, wy, wz = (float(np.float32(a)) for a in anisotropy)
  cap = int(min(max(1024, 13 * flat.size), 1 << 26))
  lo = np.zeros(cap, np.uint64)
  hi = np.zeros(cap, np.uint64)
  val = np.zeros(cap, np.float32)
  if backend == "ref":
    lib = _lib_ref()
    lib.ref_contacts.restype = ctypes.c_int64
    lib.ref_contacts.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                 ctypes.c_char_p, ctypes.c_int]
    err = ctypes.create_string_buffer(256)
    data = np.ascontiguousarray(flat)
    n = lib.ref_contacts(data.ctypes.data, code, sx, sy, sz, int(connectivity), int(bool(surface_area)), wx, wy, wz,
                         lo.ctypes.data, hi.ctypes.data, val.ctypes.data, cap, err, 256)
    if n == -2:
      raise RuntimeError(err.value.decode())
  else:
    if connectivity in (4, 8) and sz != 1:
      raise RuntimeError("z thickness must be 1 for 2d region graph extraction.")
    lib = _lib_oracle()
    lib.oracle_contacts.restype = ctypes.c_int64
    lib.oracle_contacts.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    wide = flat.astype(np.int64).view(np.uint64) if signed else flat.astype(np.uint64)
    n = lib.oracle_contacts(wide.ctypes.data, int(signed), sx, sy, sz, int(connectivity), int(bool(surface_area)),
                            wx, wy, wz, lo.ctypes.data, hi.ctypes.data, val.ctypes.data, cap)
  if n < 0:
    raise MemoryError("pair capacity exceeded")
  lo, hi, val = lo[:n], hi[:n], val[:n]
  if signed:
    lo, hi = lo.view(np.int64), hi.view(np.int64)
  return {(int(a), int(b)): float(v) for a, b, v in zip(lo.tolist(), hi.tolist(), val.tolist())}


def region_graph(labels, connectivity=26, backend: str = "oracle"):
  """cc3d.region_graph (cc3d/fastcc3d.pyx:1180-1196): the keys of contacts."""
  return set(contacts(labels, connectivity=connectivity, backend=backend).keys())


def _reshape_flat(arr: np.ndarray) -> np.ndarray:
  """_reshape(arr, (arr.size,)) (cc3d/fastcc3d.pyx:129-161): a view in memory
  order for F- or C-contiguous arrays, else a C-order copy."""
  if arr.flags.f_contiguous:
    return np.ravel(arr, order="F") if arr.ndim else arr.reshape(arr.size)
  if arr.flags.c_contiguous:
    return np.ravel(arr, order="C")
  return arr.reshape((arr.size,))


def runs(labels, backend: str = "oracle"):
  """cc3d.fastcc3d.runs restated (cc3d/fastcc3d.pyx:1254-1277,
  cc3d/cc3d_graphs.hpp:470-501): {label: [(start, end), ...]} over the
  flat array in memory order."""
  labels = np.asarray(labels)
  flat = _reshape_flat(labels)
  if flat.dtype not in (np.uint8, np.bool_, np.uint16, np.uint32, np.uint64):
    raise TypeError("Unsupported type: " + str(flat.dtype))
  if flat.dtype == np.bool_:
    flat = flat.view(np.uint8)
  n = flat.size
  lab = np.zeros(max(n, 1), np.uint64)
  st = np.zeros(max(n, 1), np.uint64)
  en = np.zeros(max(n, 1), np.uint64)
  if backend == "ref":
    lib = _lib_ref()
    lib.ref_runs.restype = ctypes.c_int64
    lib.ref_runs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_int64]
    data = np.ascontiguousarray(flat)
    m = lib.ref_runs(data.ctypes.data if n else None, CT_CODE[data.dtype], n, lab.ctypes.data, st.ctypes.data,
                     en.ctypes.data, max(n, 1))
    out = {}
    for a, s, e in zip(lab[:m].tolist(), st[:m].tolist(), en[:m].tolist()):
      out.setdefault(a, []).append((s, e))
    return out
  lib = _lib_oracle()
  lib.oracle_runs.restype = ctypes.c_int64
  lib.oracle_runs.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
  wide = flat.astype(np.uint64)
  m = lib.oracle_runs(wide.ctypes.data, n, lab.ctypes.data, st.ctypes.data, en.ctypes.data)
  out = {}
  for a, s, e in zip(lab[:m].tolist(), st[:m].tolist(), en[:m].tolist()):
    out.setdefault(a, []).append((s, e))
  return {k: out[k] for k in sorted(out)}


def draw(label, rns, image):
  """cc3d.fastcc3d.draw restated (cc3d/fastcc3d.pyx:1279-1308,
  cc3d/cc3d_graphs.hpp:503-523): paints label (label != 0 for a bool image)
  over each run of the flat view of image, checking each run as it comes
  ("Invalid run.": end past the image, or start >= end); returns the flat view."""
  flat = _reshape_flat(image)
  if flat.dtype not in (np.bool_, np.uint8, np.uint16, np.uint32, np.uint64):
    raise TypeError("Unsupported type: " + str(flat.dtype))
  val = (label != 0) if flat.dtype == np.bool_ else label
  for s, e in rns:
    if e > flat.size or s >= e:
      raise RuntimeError("Invalid run.")
    flat[s:e] = val
  return flat


def each(labels, binary=False, in_place=False):
  """cc3d.each restated (cc3d/fastcc3d.pyx:1310-1372): (label, image) per
  nonzero label in label order."""
  labels = np.asarray(labels)
  all_runs = runs(labels)
  order = "F" if labels.flags.f_contiguous else "C"
  dtype = bool if binary else labels.dtype
  for key, rns in all_runs.items():
    if key == 0:
      continue
    img = np.zeros(labels.shape, dtype=dtype, order=order)
    draw(key, rns, img)
    yield key, img
