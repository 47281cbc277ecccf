"""GPU tests of the entry points a reference binding calls directly.

* cc3d_amd_connected_components3d -- the replacement of the reference core
  cc3d::connected_components3d<T,U> (cc3d/cc3d_continuous.hpp:394-455, bound at
  cc3d/fastcc3d.pyx:67-74) -- is called through ctypes on the inputs of every
  committed reference vector, with explicit max_labels (0, small values that
  overflow the union-find, the estimate, more than the voxels), every output
  dtype and the vectors' periodic / binary / delta arguments.  Its labels, N
  and status are compared with the reference core compiled from its own
  sources (oracle/_ref) called with the same arguments.
* the binding's single-foreground-row shortcut (cc3d/fastcc3d.pyx:644-679)
  through cc3d.connected_components, against the reference binding's own
  results (tests/golden/special_row_vectors.npz).
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

_VIEW = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
_CODE = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.uint32): 2, np.dtype(np.uint64): 3,
         np.dtype(np.float32): 4, np.dtype(np.float64): 5}
_OUT = [(np.uint16, 1), (np.uint32, 2), (np.uint64, 3)]


def _core_args(arr, kw):
  """What the reference binding hands its C++ core (cc3d/fastcc3d.pyx:336-613):
  the 3D volume in memory order as its unsigned (or float) view, sx, sy, sz."""
  data = arr
  order = "F" if data.flags.f_contiguous else "C"
  while data.ndim < 3:
    data = data[np.newaxis, ...] if order == "C" else data[..., np.newaxis]
  if data.dtype == np.float16:
    data = data.view(np.uint16)
  shape = list(data.shape)
  if order == "C":
    shape.reverse()
  flat = np.ravel(data, order="K")
  if flat.dtype.kind != "f":
    flat = flat.view(_VIEW[flat.dtype.itemsize])
  binary = bool(kw.get("binary_image", False)) or arr.dtype == np.bool_
  return flat, shape[0], shape[1], shape[2], binary


def _delta_host(dtype, delta):
  if np.dtype(dtype).kind == "f":
    return np.array([delta], dtype=dtype)
  return np.array([int(delta)], dtype=dtype)


def test_core_entry_point_matches_reference_core(gpu):
  import torch
  if not oracle.have_ref():
    pytest.skip("oracle/_ref not built")
  import make_golden
  from cc3d import _lib as L
  lib = L.lib()
  rng = np.random.default_rng(4242)
  cases = make_golden.load()
  ncalls = nerr = 0
  for i, (arr, kw, _out, _N, _err) in enumerate(cases):
    if arr.size == 0:
      continue
    conn = kw.get("connectivity", 26)
    periodic = bool(kw.get("periodic_boundary", False))
    delta = kw.get("delta", 0)
    if periodic and delta:
      continue  # the binding rejects this before calling the core
    flat, sx, sy, sz, binary = _core_args(arr, kw)
    if flat.dtype.kind == "f":
      delta = float(delta)
    elif int(delta) < 0:
      continue
    voxels = sx * sy * sz
    epl = oracle.ref_core_epl(flat, sx)[0]
    dev = torch.from_numpy(np.ascontiguousarray(flat)).to(gpu)
    dh = _delta_host(flat.dtype, delta)
    picks = [0, 1, 3, max(epl, 1), voxels + 5]
    for ml in [picks[int(rng.integers(0, len(picks)))], max(epl, 1)]:
      odt, ocode = _OUT[int(rng.integers(0, 3))]
      tag = (f"vector {i} shape={arr.shape} dtype={arr.dtype} conn={conn} periodic={periodic} binary={binary} "
             f"delta={delta} max_labels={ml} out={np.dtype(odt).name}")
      print(tag, flush=True)
      try:
        exp, expN = oracle.ref_core_label(flat, sx, sy, sz, ml, conn, delta, odt, periodic, binary)
        exp_err = None
      except RuntimeError as e:
        exp, expN, exp_err = None, None, str(e)
      out = torch.empty(voxels, dtype={np.uint16: torch.uint16, np.uint32: torch.uint32,
                                       np.uint64: torch.uint64}[odt], device=gpu)
      N = ctypes.c_uint64(0)
      rc = lib.cc3d_amd_connected_components3d(
        ctypes.c_void_p(dev.data_ptr()), _CODE[flat.dtype], sx, sy, sz, ml, conn, dh.ctypes.data,
        ctypes.c_void_p(out.data_ptr()), ocode, ctypes.byref(N), int(periodic), int(binary),
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
      ncalls += 1
      if exp_err is not None:
        nerr += 1
        assert rc == L.ERR_RUNTIME, f"{tag}: reference raised {exp_err!r}, entry point returned {rc}"
        msg = lib.cc3d_amd_last_error().decode()
        assert msg == exp_err, f"{tag}: {msg!r} vs {exp_err!r}"
        continue
      assert rc == L.OK, f"{tag}: status {rc}: {lib.cc3d_amd_last_error().decode()}"
      torch.cuda.synchronize()
      got = out.cpu().numpy()
      assert int(N.value) == expN, f"{tag}: N {N.value} vs {expN}"
      assert np.array_equal(got, exp), f"{tag}: labels differ"
  assert ncalls > 1000 and nerr > 20, (ncalls, nerr)
  # the dispatcher's own errors, checked before any labelling (also with max_labels == 0)
  vol = np.zeros((4, 3, 2), dtype=np.uint32)
  vol[1, 1, 1] = 7
  flat = vol.ravel(order="F")
  dev = torch.from_numpy(flat).to(gpu)
  for conn, periodic, delta in ((8, False, 0), (4, True, 0), (10, False, 0), (26, True, 3), (6, True, 1)):
    for ml in (0, 5):
      try:
        oracle.ref_core_label(flat, 4, 3, 2, ml, conn, delta, np.uint32, periodic, False)
        exp_err = None
      except RuntimeError as e:
        exp_err = str(e)
      out = torch.empty(flat.size, dtype=torch.uint32, device=gpu)
      N = ctypes.c_uint64(0)
      dh = np.array([delta], dtype=np.uint32)
      rc = lib.cc3d_amd_connected_components3d(
        ctypes.c_void_p(dev.data_ptr()), 2, 4, 3, 2, ml, conn, dh.ctypes.data, ctypes.c_void_p(out.data_ptr()), 2,
        ctypes.byref(N), int(periodic), 0, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
      tag = f"conn={conn} periodic={periodic} delta={delta} max_labels={ml}"
      if exp_err is None:
        assert rc == L.OK, tag
      else:
        assert rc == L.ERR_RUNTIME and lib.cc3d_amd_last_error().decode() == exp_err, tag


def test_special_row_matches_reference_binding(gpu):
  import torch
  import cc3d
  import special_row_cases as sr
  cases = sr.load()
  for i, (arr, kw, out, N, et, em) in enumerate(cases):
    print(f"special row {i} shape={arr.shape} dtype={arr.dtype} kw={kw}", flush=True)
    sr.check(cc3d.connected_components, arr, kw, out, N, et, em, f"special row {i}")
  # device-resident input: the labels stay in HBM and equal the same vectors
  ndev = 0
  for i, (arr, kw, out, N, et, em) in enumerate(cases):
    if arr.dtype == np.float16 or et is not None or not arr.flags.c_contiguous:
      continue
    t = torch.from_numpy(arr).to(gpu)
    got, gN = cc3d.connected_components(t, return_N=True, **kw)
    assert got.device.type == "cuda"
    assert gN == N and np.array_equal(got.cpu().numpy(), out), f"special row {i} (device tensor)"
    ndev += 1
  assert ndev >= 10


@pytest.mark.parametrize("order", ["C", "F"])
def test_out_file_and_cpu_tensor(gpu, tmp_path, order):
  """out_file (cc3d/fastcc3d.pyx:436-443: a file sized to the output and
  memory-mapped, returned in the input's shape) and CPU torch tensors in and
  out (:306-309, :620-622), against the oracle."""
  import torch
  import cc3d
  import synth
  vol = synth.voronoi((37, 29, 23), 30, seed=5)
  vol = np.asfortranarray(vol) if order == "F" else np.ascontiguousarray(vol)
  exp, eN = oracle.connected_components(vol, connectivity=26, return_N=True)
  path = str(tmp_path / f"labels_{order}.bin")
  got, N = cc3d.connected_components(vol, connectivity=26, return_N=True, out_file=path)
  assert isinstance(got, np.memmap) and N == eN
  assert got.dtype == exp.dtype and got.shape == exp.shape and np.array_equal(got, exp)
  assert got.flags.f_contiguous == exp.flags.f_contiguous
  del got
  assert os.path.getsize(path) == exp.size * exp.dtype.itemsize
  back = np.fromfile(path, dtype=exp.dtype)
  assert np.array_equal(back, np.ravel(exp, order="K"))
  # an existing file with other contents: the reference reopens it with "wb"
  # (emptied, then sized); every voxel is written here
  with open(path, "wb") as f:
    f.write(b"\xff" * (exp.size * exp.dtype.itemsize))
  got = cc3d.connected_components(vol, connectivity=26, out_file=path)
  assert np.array_equal(got, exp)
  # CPU torch tensors come back as CPU torch tensors with the same labels
  t = torch.from_numpy(vol)
  gt, tN = cc3d.connected_components(t, connectivity=26, return_N=True)
  assert isinstance(gt, torch.Tensor) and gt.device.type == "cpu" and tN == eN
  assert np.array_equal(gt.numpy(), exp)
