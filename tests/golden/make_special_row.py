#!/usr/bin/env python3
"""Writes tests/golden/special_row_vectors.npz: the reference binding's own
results for volumes whose only foreground is one row.

That shortcut (epl_special_row, cc3d/fastcc3d.pyx:644-679, taken at :469-470)
is Cython code in the reference's binding, outside the C++ core that
oracle/_ref compiles, so only the binding itself can answer it.  This script
compiles the reference's fastcc3d.pyx with Cython and g++ OUTSIDE the
repository (default /tmp/cc3d_ref_binding; nothing it builds is kept here),
imports it, and records inputs, keyword arguments and results -- labels and N,
or the exception type and message.  The cases cover float16/32/64 and 64-bit
integer rows: integer-valued runs, repeated non-integer values, runs broken
and joined by truncation, NaN, +-0, infinities, values at and beyond +-2^63,
uint64 >= 2^63, rows that overflow a uint16 output, in 1D, 2D and 3D, both
memory orders.

Run in this container (it needs /root/reference):
    python tests/golden/make_special_row.py
"""
import json
import os
import subprocess
import sys
import sysconfig

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("CC3D_REFERENCE", "/root/reference")
BUILD = os.environ.get("CC3D_REF_BINDING_BUILD", "/tmp/cc3d_ref_binding")
OUT = os.path.join(HERE, "special_row_vectors.npz")


def build_binding():
  pyx = os.path.join(REF, "cc3d", "fastcc3d.pyx")
  ext = sysconfig.get_config_var("EXT_SUFFIX")
  so = os.path.join(BUILD, "fastcc3d" + ext)
  if os.path.exists(so):
    return
  os.makedirs(BUILD, exist_ok=True)
  cpp = os.path.join(BUILD, "fastcc3d.cpp")
  subprocess.run([sys.executable, "-m", "cython", "-3", "--cplus", pyx, "-o", cpp], check=True)
  subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", "-I" + os.path.join(REF, "cc3d"),
                  "-I" + sysconfig.get_paths()["include"], "-I" + np.get_include(), cpp, "-o", so], check=True)


def rows():
  """(values, dtype) pairs of one foreground row."""
  f = []
  for dt in ("float32", "float64"):
    for vals in ([2.5, 2.5], [2.5, 2.0], [2.5, 2.7, 2.9], [-2.5, -2.0, -2.0, -2.5], [-0.0, 0.5, 0.0, -0.3, 0.3],
                 [1, 1, 2, 2.0, 0, 2], [3.0, 3.7, 3.0, 3.2, 3.0], [0.9, 0.1, 0.0, 0.9], [7.0, 7.0, 7.0],
                 [np.nan], [1, np.nan], [2.5, 2.0, np.nan, 1.0], [np.inf], [-np.inf, 1], [1, np.inf, np.nan],
                 [1e30], [2.0 ** 63], [-(2.0 ** 63)], [-(2.0 ** 63), -(2.0 ** 63)], [-(2.0 ** 64)],
                 [2.0 ** 62, 2.0 ** 62 + 0.0], [2.0 ** 63 - 2.0 ** 40], [5.5, 1e30], [0.0, -0.0, 1.5, 1.0]):
      f.append((vals, dt))
  rng = np.random.default_rng(1234)
  for k in range(12):
    dt = "float32" if k % 2 else "float64"
    v = rng.choice([0.0, 0.5, 1.0, 1.5, 2.0, -1.5, -1.0], size=int(rng.integers(5, 60)))
    f.append((v.tolist(), dt))
  for vals in ([2 ** 63 + 5], [2 ** 63 - 1, 2 ** 63 - 1], [5, 2 ** 63, 5], [2 ** 64 - 1], [7, 7, 0, 7, 2 ** 63 + 1],
               [1, 2, 2, 3]):
    f.append((vals, "uint64"))
  for vals in ([-5, -5, 3, 3, 0, -(2 ** 63)], [2 ** 63 - 1, -1, -1], [-(2 ** 63), 2 ** 63 - 1]):
    f.append((vals, "int64"))
  for vals in ([-3, -3, 0, 2], [5, 5, 6]):
    f.append((vals, "int8"))
    f.append((vals, "int32"))
  f.append(([1.5, 1.5, 0.0, 2.0, np.nan], "float16"))
  f.append(([-0.0, 0.0, 1.0, 1.0], "float16"))
  return f


def place(vals, dt, layout, rng):
  """Put the row into a volume of the given layout; everything else is 0."""
  vals = np.array(vals, dtype=dt) if dt not in ("uint64", "int64") else np.array(vals, dtype=np.dtype(dt))
  n = vals.size
  if layout == "1d":
    return vals.copy()
  if layout in ("2dF", "2dC"):
    sy = int(rng.integers(1, 5))
    y = int(rng.integers(0, sy))
    a = np.zeros((n, sy), dtype=dt, order="F")
    a[:, y] = vals
    if layout == "2dC":
      a = np.ascontiguousarray(a.T)
    return a
  sy, sz = int(rng.integers(1, 4)), int(rng.integers(1, 4))
  y, z = int(rng.integers(0, sy)), int(rng.integers(0, sz))
  a = np.zeros((n, sy, sz), dtype=dt, order="F")
  a[:, y, z] = vals
  if layout == "3dC":
    a = np.ascontiguousarray(a.transpose(2, 1, 0))
  return a


def main():
  build_binding()
  sys.path.insert(0, BUILD)
  import fastcc3d  # the reference's binding, built above

  rng = np.random.default_rng(20261203)
  arrays = {}
  meta = []
  cases = []
  layouts = ["1d", "2dF", "2dC", "3dF", "3dC"]
  for j, (vals, dt) in enumerate(rows()):
    layout = layouts[j % len(layouts)]
    kw = {"connectivity": 26 if layout[0] != "2" else int(rng.choice([4, 8, 26]))}
    if kw["connectivity"] in (4, 8, 6) and rng.random() < 0.3:
      kw["periodic_boundary"] = True
    cases.append((place(vals, dt, layout, rng), kw))
  # rows that outgrow a uint16 output: 70,000 starts at an estimate of 1
  cases.append((np.full(70000, 2.5, np.float32), {}))
  cases.append((np.full(70000, 2.5, np.float32), {"out_dtype": "uint32"}))
  cases.append((np.concatenate([np.full(65535, 2.5, np.float32), [np.nan]]).astype(np.float32), {}))
  cases.append((np.concatenate([np.full(65534, 2.5, np.float32), [np.nan]]).astype(np.float32), {}))
  cases.append((np.concatenate([np.full(65534, 2.5, np.float64), [2.0 ** 63]]), {}))

  for i, (arr, kw) in enumerate(cases):
    call = dict(kw)
    if "out_dtype" in call:
      call["out_dtype"] = np.dtype(call["out_dtype"])
    rec = {"kwargs": kw, "dtype": str(arr.dtype), "order": "F" if arr.flags.f_contiguous else "C"}
    try:
      out, N = fastcc3d.connected_components(arr, return_N=True, **call)
      rec["N"] = int(N)
      rec["out_dtype"] = str(out.dtype)
      arrays[f"out_{i}"] = out
    except Exception as e:  # the reference's exception is the expected result
      rec["exc_type"] = type(e).__name__
      rec["exc_msg"] = str(e)
    arrays[f"in_{i}"] = arr
    meta.append(rec)
  arrays["meta"] = np.array(json.dumps(meta))
  np.savez_compressed(OUT, **arrays)
  nexc = sum(1 for m in meta if "exc_type" in m)
  print(f"wrote {OUT}: {len(meta)} cases, {nexc} raising")


if __name__ == "__main__":
  main()
