"""Deterministic synthetic volumes (Voronoi labels, Bernoulli masks, tone images).

Thin ctypes wrapper over synth.c, compiled to libcc3d_synth.so next to it.
Arrays are returned x fastest (numpy Fortran order, shape (sx, sy, sz)).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "libcc3d_synth.so")
_lib = None


def build() -> None:
  subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fPIC", "-shared", "-pthread", "-o", SO + ".tmp",
                  os.path.join(HERE, "synth.c"), "-lm"], check=True)
  os.replace(SO + ".tmp", SO)


def _l():
  global _lib
  if _lib is None:
    if not os.path.exists(SO):
      build()
    L = ctypes.CDLL(SO)
    i64, u64, d, vp, ci = ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p, ctypes.c_int
    L.synth_voronoi.argtypes = [u64, i64, i64, i64, i64, u64, u64, d, d, vp, ci, ci]
    L.synth_voronoi.restype = ci
    L.synth_bernoulli.argtypes = [u64, i64, d, vp]
    L.synth_bernoulli.restype = ci
    L.synth_tones.argtypes = [u64, i64, i64, i64, i64, vp, ci, ctypes.c_float, vp, ci]
    L.synth_tones.restype = ci
    L.synth_voronoi_regions.argtypes = [u64, i64, i64, i64, i64, vp, ci]
    L.synth_voronoi_regions.restype = ci
    _lib = L
  return _lib


def _threads() -> int:
  try:
    return max(1, min(16, len(os.sched_getaffinity(0))))
  except AttributeError:
    return 4


def voronoi(shape, nseeds, seed=0, id_range=(1, 2**31), p_membrane=0.8, p_salt=0.02,
            dtype=np.uint32) -> np.ndarray:
  """Voronoi label volume; id_range=(lo, hi) with hi=None for full 64-bit nonzero IDs."""
  sx, sy, sz = (list(shape) + [1, 1])[:3]
  out = np.zeros((sx, sy, sz), dtype=dtype, order="F")
  lo, hi = id_range
  rc = _l().synth_voronoi(seed, sx, sy, sz, nseeds, lo, 0 if hi is None else hi, p_membrane, p_salt,
                          out.ctypes.data, np.dtype(dtype).itemsize, _threads())
  if rc:
    raise RuntimeError("synth_voronoi failed")
  return out.reshape(shape, order="F")


def bernoulli(shape, p, seed=0) -> np.ndarray:
  n = int(np.prod(shape))
  out = np.zeros(shape, dtype=np.uint8, order="F")
  _l().synth_bernoulli(seed, n, p, out.ctypes.data)
  return out


def tones(shape, nseeds, tone_values=(0.0, 100.0, 200.0), amp=5.0, seed=0) -> np.ndarray:
  sx, sy, sz = (list(shape) + [1, 1])[:3]
  out = np.zeros((sx, sy, sz), dtype=np.float32, order="F")
  tv = np.asarray(tone_values, dtype=np.float32)
  rc = _l().synth_tones(seed, sx, sy, sz, nseeds, tv.ctypes.data, len(tv), amp, out.ctypes.data, _threads())
  if rc:
    raise RuntimeError("synth_tones failed")
  return out.reshape(shape, order="F")
