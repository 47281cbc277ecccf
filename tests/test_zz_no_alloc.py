"""The no-allocation entry point cc3d_amd_connected_components3d_noalloc
against cc3d_amd_connected_components3d, through the C ABI.

Each case makes two calls in a row, the second on a smaller input, so the
second runs over scratch the first one left behind.  The bar is the entry
point's own: the same labels, N and status, bit for bit.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_CODE = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.uint32): 2, np.dtype(np.uint64): 3,
         np.dtype(np.float32): 4, np.dtype(np.float64): 5}


def _lib():
  from cc3d import _lib as L
  lib = L.lib()
  f = lib.cc3d_amd_connected_components3d_noalloc
  f.restype = ctypes.c_int
  f.argtypes = lib.cc3d_amd_connected_components3d.argtypes
  return L, lib


def _call(fn, vol, conn, max_labels, delta, periodic, binary, gpu):
  import torch
  sx, sy, sz = (list(vol.shape) + [1, 1])[:3]
  flat = np.ravel(vol, order="F")
  dev = torch.from_numpy(flat).to(gpu)
  out = torch.full((flat.size,), 7, dtype=torch.uint32, device=gpu)  # stale contents: every voxel is written
  N = ctypes.c_uint64(0)
  d = np.array([delta], dtype=flat.dtype)
  rc = fn(ctypes.c_void_p(dev.data_ptr()), _CODE[flat.dtype], sx, sy, sz, max_labels, conn, d.ctypes.data,
          ctypes.c_void_p(out.data_ptr()), 2, ctypes.byref(N), int(periodic), int(binary),
          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
  torch.cuda.synchronize()
  return rc, int(N.value), out.cpu().numpy()


def _cases():
  import synth
  rng = np.random.default_rng(77)
  big = synth.voronoi((160, 150, 140), 900, seed=3)
  small = synth.voronoi((70, 41, 33), 60, seed=4)
  b2d = (rng.random((300, 257)) < 0.45).astype(np.uint8)
  b2d_small = (rng.random((61, 40)) < 0.45).astype(np.uint8)
  tones = synth.tones((96, 80, 40), 50, seed=5)
  tones_small = synth.tones((33, 20, 17), 9, seed=6)
  # (volume, connectivity, max_labels, delta, periodic, binary)
  return [
    ((big, 26, big.size, 0, False, False), (small, 26, small.size, 0, False, False)),
    ((big, 6, big.size, 0, True, False), (small, 18, small.size, 0, False, False)),
    ((b2d, 8, b2d.size, 0, False, True), (b2d_small, 4, b2d_small.size, 0, True, True)),
    ((tones, 26, tones.size, np.float32(20), False, False),
     (tones_small, 6, tones_small.size, np.float32(20), False, False)),
    ((big, 26, 3, 0, False, False), (small, 26, 3, 0, False, False)),  # "maximum length exception"
  ]


def test_noalloc_matches_entry_point(gpu):
  L, lib = _lib()
  for k, pair in enumerate(_cases()):
    for n, (vol, conn, ml, delta, periodic, binary) in enumerate(pair):
      tag = f"case {k} call {n}: shape={vol.shape} dtype={vol.dtype} conn={conn} max_labels={ml}"
      print(tag, flush=True)
      exp = _call(lib.cc3d_amd_connected_components3d, vol, conn, ml, delta, periodic, binary, gpu)
      got = _call(lib.cc3d_amd_connected_components3d_noalloc, vol, conn, ml, delta, periodic, binary, gpu)
      assert got[0] == exp[0], f"{tag}: status {got[0]} vs {exp[0]}"
      if exp[0] != L.OK:
        assert lib.cc3d_amd_last_error().decode() == "maximum length exception", tag
        continue
      assert got[1] == exp[1], f"{tag}: N {got[1]} vs {exp[1]}"
      assert np.array_equal(got[2], exp[2]), f"{tag}: labels differ"


def test_noalloc_rejects_volumes_past_its_scratch(gpu):
  import torch
  L, lib = _lib()
  dev = torch.zeros(1, dtype=torch.uint32, device=gpu)
  out = torch.zeros(1, dtype=torch.uint32, device=gpu)
  N = ctypes.c_uint64(0)
  d = np.zeros(1, dtype=np.uint32)
  # geometry only (sizes are checked before any pointer is used)
  rc = lib.cc3d_amd_connected_components3d_noalloc(
    ctypes.c_void_p(dev.data_ptr()), 2, 1024, 1024, 512, 10, 26, d.ctypes.data, ctypes.c_void_p(out.data_ptr()), 2,
    ctypes.byref(N), 0, 0, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
  assert rc == L.ERR_INVALID
  assert "bytes of scratch" in lib.cc3d_amd_last_error().decode()
