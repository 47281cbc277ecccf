"""statistics, dust and largest_k on the GPU (cc3d/fastcc3d.pyx:681-938,
cc3d/__init__.py:71-155, :199-279)."""
from __future__ import annotations

import ctypes
from typing import Union

import numpy as np

from . import _lib as L
from ._core import _NP_TO_CODE, _Vol, _is_torch, _require_gpu, _stream, _torch, connected_components


def _bind():
  lib = L.lib()
  if not hasattr(lib, "_follow_bound"):
    vp, i64, u64, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int
    lib.cc3d_amd_label_range.restype = ci
    lib.cc3d_amd_label_range.argtypes = [vp, ci, i64, ci, ctypes.POINTER(u64), ctypes.POINTER(i64),
                                         ctypes.POINTER(i64), vp]
    lib.cc3d_amd_statistics.restype = ci
    lib.cc3d_amd_statistics.argtypes = [vp, ci, i64, i64, i64, ci, ci, u64, vp, vp, vp, vp, vp]
    lib.cc3d_amd_dust_apply.restype = ci
    lib.cc3d_amd_dust_apply.argtypes = [vp, ci, vp, ci, i64, vp, u64, ctypes.c_double, ctypes.c_double,
                                        ci, ci, vp, vp, vp]
    lib.cc3d_amd_dust_copy.restype = ci
    lib.cc3d_amd_dust_copy.argtypes = [vp, vp, ci, vp, ci, i64, vp, u64, ctypes.c_double, ctypes.c_double,
                                       ci, ci, vp, vp, vp]
    lib.cc3d_amd_keep_labels.restype = ci
    lib.cc3d_amd_keep_labels.argtypes = [vp, ci, i64, vp, i64, u64, vp, vp, ci, vp]
    lib.cc3d_amd_follow_last_error.restype = ctypes.c_char_p
    lib._follow_bound = True
  return lib


def _fcheck(rc):
  if rc != L.OK:
    msg = _bind().cc3d_amd_follow_last_error().decode(errors="replace")
    if rc == L.ERR_INVALID:
      raise ValueError(msg)
    raise RuntimeError(msg)


def _device_stats(buf, code, s0, s1, s2, ndim, f_order, N, bbox=True, centroids=True):
  """counts (uint32), bbox (uint32 min/max pairs), centroids (float64) as device
  tensors; bbox / centroids are None when not asked for (not computed)."""
  torch = _torch()
  dev = buf.device
  nd = 2 if ndim == 2 else 3
  counts = torch.empty(N + 1, dtype=torch.int32, device=dev)
  bb = torch.empty((N + 1) * 2 * nd, dtype=torch.int32, device=dev) if bbox else None
  cen = torch.empty((N + 1) * nd, dtype=torch.float64, device=dev) if centroids else None
  sums = torch.empty((N + 1) * nd, dtype=torch.int64, device=dev) if centroids else None
  ptr = lambda t: ctypes.c_void_p(None if t is None else t.data_ptr())
  _fcheck(_bind().cc3d_amd_statistics(
    ctypes.c_void_p(buf.data_ptr()), code, s0, s1, s2, nd, int(f_order), N,
    ptr(counts), ptr(bb), ptr(cen), ptr(sums), _stream()))
  return counts, bb, cen


def _label_volume(out_labels):
  """The checks of cc3d/fastcc3d.pyx:716-748 on a label volume, which is then
  resident in HBM.  None for an empty volume, else (buf, code, sx, sy, s2,
  ndim, f_order, N) with N the largest label."""
  vol = _Vol(out_labels)
  shape = list(vol.shape)
  order = "F" if vol.f_contig else "C"
  while len(shape) < 2:
    shape = shape + [1]
  ndim = len(shape)
  sx, sy = shape[0], shape[1]
  sz = shape[2] if ndim > 2 else 1
  voxels = sx * sy * sz
  if voxels == 0:
    return None
  npdt = vol.np_dtype
  if npdt == np.bool_:
    npdt = np.dtype(np.uint8)
  if npdt.kind not in "ui":
    raise TypeError(f"statistics needs integer labels, got {vol.np_dtype}")
  _require_gpu()
  buf = vol.device_buffer(order)
  code = _NP_TO_CODE[npdt]
  lib = _bind()
  umax = ctypes.c_uint64(0)
  smin = ctypes.c_int64(0)
  smax = ctypes.c_int64(0)
  signed = npdt.kind == "i"
  _fcheck(lib.cc3d_amd_label_range(ctypes.c_void_p(buf.data_ptr()), code, voxels, int(signed),
                                   ctypes.byref(umax), ctypes.byref(smin), ctypes.byref(smax), _stream()))
  N = int(smax.value) if signed else int(umax.value)
  if N < 0:
    raise OverflowError("can't convert negative value to uint64_t")
  if N > voxels:
    raise ValueError(
      f"Statistics can only be computed on volumes containing labels with values lower than the number of voxels. Max: {N}"
    )
  if signed and int(smin.value) < 0:
    raise ValueError(
      f"Statistics can only be computed on volumes containing labels with values >= 0. Min: {int(smin.value)}"
    )
  # memory layout: the 2D/3D array in its own order; axes as numpy sees them
  return buf, code, sx, sy, sz, ndim, int(order == "F"), N


def statistics(out_labels, no_slice_conversion: bool = False) -> dict:
  """Voxel counts, bounding boxes and centroids per label (cc3d/fastcc3d.pyx:681-938).

  Counts are uint32, bounding boxes uint16 (uint32 when a side exceeds 65535)
  as [min0, max0, min1, max1(, min2, max2)] along the numpy axes, centroids
  float64 (NaN for absent labels).  Computed on the GPU; returned as numpy.
  """
  lv = _label_volume(out_labels)
  if lv is None:
    return {"voxel_counts": None, "bounding_boxes": None, "centroids": None}
  buf, code, sx, sy, sz, ndim, f_order, N = lv
  voxels = sx * sy * sz
  counts, bbox, cen = _device_stats(buf, code, sx, sy, sz, ndim, f_order, N)
  nd = 2 if ndim == 2 else 3
  counts = counts.cpu().numpy().view(np.uint32)
  bb = bbox.cpu().numpy().view(np.uint32)
  big = max(sx, sy, sz) > np.iinfo(np.uint16).max
  bdt = np.uint32 if big else np.uint16
  bb = np.minimum(bb, np.iinfo(bdt).max).astype(bdt).reshape((N + 1, 2 * nd))
  cen = cen.cpu().numpy().reshape((N + 1, nd))
  output = {"voxel_counts": counts, "bounding_boxes": bb, "centroids": cen}
  if no_slice_conversion:
    return output
  slices = []
  for row in bb:
    if all(int(row[2 * a]) < voxels for a in range(nd)):
      slices.append(tuple(slice(int(row[2 * a]), int(row[2 * a + 1]) + 1) for a in range(nd)))
    else:
      slices.append(None)
  output["bounding_boxes"] = slices
  return output


def _view_as_unsigned(img):
  """cc3d/__init__.py:281-293."""
  if np.issubdtype(img.dtype, np.unsignedinteger) or img.dtype == bool:
    return img
  if img.dtype in (np.int8, np.int16, np.int32, np.int64):
    return img.view(f"u{img.dtype.itemsize}")
  return img


def dust(img, threshold: Union[int, float, tuple, list], connectivity: int = 26, in_place: bool = False,
         binary_image: bool = False, precomputed_ccl: bool = False, invert: bool = False,
         return_N: bool = False):
  """Remove connected components smaller than threshold (cc3d/__init__.py:71-155).

  Labelling, voxel counts and the masking pass all run on the GPU; numpy
  inputs are copied to HBM once and the result is copied back.
  """
  torch = _torch()
  if _is_torch(img):
    raise TypeError("dust takes numpy arrays (as the reference does)")
  orig_dtype = img.dtype
  img = _view_as_unsigned(img)
  if not in_place:
    img = np.copy(img)
  order = "F" if img.flags.f_contiguous else "C"
  _require_gpu()

  if precomputed_ccl:
    # the reference's statistics() checks on the (unsigned view of the)
    # labels: N = max label must not exceed the voxel count
    # (cc3d/fastcc3d.pyx:731-737); the labels are then resident in HBM
    lv = _label_volume(img) if img.size else None
    N = lv[-1] if lv is not None else 0
    img_dev = lv[0] if lv is not None else _Vol(img).device_buffer(order)
    cc_dev = img_dev
    cc_dtype = np.dtype(img.dtype) if img.dtype != np.bool_ else np.dtype(np.uint8)
  else:
    img_dev = _Vol(img).device_buffer(order)
    # img_dev holds img packed in `order` (a non-contiguous view is copied)
    tview = torch.as_strided(img_dev, img.shape, _packed_strides(img.shape, order))
    if img.dtype == np.bool_:
      tview = tview.view(torch.bool)  # keeps the reference's bool => binary rule
    cc_t, N = connected_components(tview, connectivity=connectivity, return_N=True, binary_image=bool(binary_image))
    cc_dev = cc_t.as_strided((cc_t.numel(),), (1,))
    cc_dtype = np.dtype({torch.uint16: np.uint16, torch.uint32: np.uint32, torch.uint64: np.uint64}[cc_t.dtype])

  shape = list(img.shape)
  while len(shape) < 2:
    shape = shape + [1]
  ndim = len(shape)
  sz = shape[2] if ndim > 2 else 1
  counts, _, _ = _device_stats(cc_dev, _NP_TO_CODE[cc_dtype], shape[0], shape[1], sz, ndim,
                               int(order == "F"), N, bbox=False, centroids=False)

  if isinstance(threshold, (tuple, list)):
    lo, hi, ranged = float(threshold[0]), float(threshold[1]), 1
  else:
    lo, hi, ranged = float(threshold), 0.0, 0
  erase = torch.empty(N + 1, dtype=torch.uint8, device=img_dev.device)
  nmask = torch.zeros(1, dtype=torch.int64, device=img_dev.device)
  img_code = _NP_TO_CODE[np.dtype(img.dtype) if img.dtype != np.bool_ else np.dtype(np.uint8)]
  _fcheck(_bind().cc3d_amd_dust_apply(
    ctypes.c_void_p(img_dev.data_ptr()), img_code, ctypes.c_void_p(cc_dev.data_ptr()),
    _NP_TO_CODE[cc_dtype], img.size,
    ctypes.c_void_p(counts.data_ptr()), N, lo, hi, ranged, int(bool(invert)),
    ctypes.c_void_p(erase.data_ptr()), ctypes.c_void_p(nmask.data_ptr()), _stream()))
  n_to_mask = int(nmask.item())
  dust_N = n_to_mask if invert else N - n_to_mask

  if n_to_mask == 0:
    # nothing masked: the reference returns the (unsigned view of the) image
    if invert:
      img = np.zeros(img.shape, dtype=img.dtype, order="F")
    return (img, dust_N) if return_N else img

  host = img_dev.cpu().numpy().view(img.dtype if img.dtype != np.bool_ else np.uint8)
  if img.dtype == np.bool_:
    host = host.view(np.bool_)
  # write through the caller's array (a strided view included), as the
  # reference's img[mask] = 0 does
  img[...] = host.reshape(img.shape, order=order)
  img = img.view(orig_dtype)
  return (img, dust_N) if return_N else img


def _packed_strides(shape, order: str):
  """Element strides of a dense array of `shape` in memory order `order`."""
  strides = [0] * len(shape)
  acc = 1
  axes = range(len(shape) - 1, -1, -1) if order == "C" else range(len(shape))
  for a in axes:
    strides[a] = acc
    acc *= shape[a]
  return tuple(strides)


def largest_k(img, k: int, connectivity: int = 26, delta: Union[int, float] = 0, return_N: bool = False,
              binary_image: bool = False, precomputed_ccl: bool = False):
  """The k largest connected components (cc3d/__init__.py:199-279), numbered
  1..k in ascending size order as the reference does without fastremap.

  Labelling, voxel counts and the renumbering pass run on the GPU.  The
  choice of the k labels is the reference's own numpy code (argpartition and
  a stable sort by count) on the same uint32 counts, so ties break the same
  way.  k == 1 returns the boolean mask of the largest component.
  """
  assert k >= 0
  torch = _torch()
  if _is_torch(img):
    raise TypeError("largest_k takes numpy arrays (as the reference does)")
  order = "C" if img.flags.c_contiguous else "F"
  if k == 0:
    return np.zeros(img.shape, dtype=np.uint16, order=order)
  if precomputed_ccl:
    cc_labels = np.copy(img, order="F")
    N = np.max(cc_labels)
    if N <= k:
      return (cc_labels, N) if return_N else cc_labels
    cc_shape, cc_dtype, cc_c = cc_labels.shape, cc_labels.dtype, cc_labels.flags.c_contiguous
    lv = _label_volume(cc_labels)
  elif img.size and (img.flags.c_contiguous or img.flags.f_contiguous) and img.dtype != np.float16:
    # the labels stay in HBM: one copy of the image in, one of the result out
    _require_gpu()
    io = "F" if img.flags.f_contiguous else "C"
    img_dev = _Vol(img).device_buffer(io)
    tview = torch.as_strided(img_dev, img.shape, _packed_strides(img.shape, io))
    if img.dtype == np.bool_:
      tview = tview.view(torch.bool)  # keeps the reference's bool => binary rule
    cc_t, N = connected_components(tview, connectivity=connectivity, return_N=True, delta=delta,
                                   binary_image=bool(binary_image))
    if N <= k:
      cc_labels = cc_t.cpu().numpy()  # the input's shape and memory order
      return (cc_labels, N) if return_N else cc_labels
    cc_shape, cc_c = img.shape, img.flags.c_contiguous
    cc_dtype = np.dtype({torch.uint16: np.uint16, torch.uint32: np.uint32, torch.uint64: np.uint64}[cc_t.dtype])
    lv = _label_volume(cc_t)
  else:
    cc_labels, N = connected_components(img, connectivity=connectivity, return_N=True, delta=delta,
                                        binary_image=bool(binary_image))
    if N <= k:
      return (cc_labels, N) if return_N else cc_labels
    cc_shape, cc_dtype, cc_c = cc_labels.shape, cc_labels.dtype, cc_labels.flags.c_contiguous
    lv = _label_volume(cc_labels)

  buf, code, sx, sy, sz, ndim, f_order, NL = lv
  counts, _, _ = _device_stats(buf, code, sx, sy, sz, ndim, f_order, NL, bbox=False, centroids=False)
  cts = counts.cpu().numpy().view(np.uint32)
  if k == 1:
    keep = [int(np.argmax(cts[1:])) + 1]
    out_np = np.dtype(np.bool_)
  else:
    preserve = np.argpartition(cts[1:], len(cts) - k - 1)[-k:]
    preserve += 1
    keep = list(map(int, sorted(preserve, key=lambda label: cts[label])))
    out_np = np.dtype(cc_dtype)
  voxels = sx * sy * sz
  keep_t = torch.tensor(keep, dtype=torch.int64).to(buf.device)
  lut = torch.empty(NL + 1, dtype=torch.int32, device=buf.device)
  out_t = torch.empty(voxels * out_np.itemsize, dtype=torch.uint8, device=buf.device)
  _fcheck(_bind().cc3d_amd_keep_labels(
    ctypes.c_void_p(buf.data_ptr()), code, voxels, ctypes.c_void_p(keep_t.data_ptr()), len(keep), NL,
    ctypes.c_void_p(lut.data_ptr()), ctypes.c_void_p(out_t.data_ptr()), _NP_TO_CODE[out_np], _stream()))
  host = out_t.cpu().numpy()
  cc_out = np.zeros(cc_shape, dtype=out_np, order="C" if cc_c else "F")
  flat = np.ravel(cc_out, order="K")
  flat[:] = host.view(flat.dtype)
  if return_N:
    return cc_out, (1 if k == 1 else len(keep))
  return cc_out
