#!/usr/bin/env python3
"""Writes tests/golden/ref_vectors.npz: input/output vectors of the reference.

Outputs come from the reference implementation compiled from its own sources
(oracle/_ref/libcc3d_ref.so, built by oracle/Makefile) called through the
binding restatement in oracle/oracle.py.  Cases that take the binding's
single-foreground-row shortcut (cc3d/fastcc3d.pyx:469-470, :644-679), which
is binding code and not part of the compiled core, are answered by the
reference's binding itself, compiled outside the repository by
make_special_row.build_binding.  Inputs are small seeded random
volumes that exercise every path of the labelling core: multilabel, binary and
continuous at every connectivity, periodic boundaries, both memory orders,
all input dtypes, explicit output dtypes, and the quirk cases (binary 2D-8
block numbering and stale labels on odd sides, binary 2D-4 row 0, continuous
2D-8 range shortcut, union-find overflow).

Run in this container, where oracle/_ref and /root/reference exist:
    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)
import oracle  # noqa: E402

DTYPES = ["uint8", "uint16", "int32", "uint32", "uint64", "int64", "float32", "float64", "int8", "int16",
          "bool", "float16"]


def case(rng, i):
  nd = int(rng.choice([1, 2, 3]))
  if nd == 1:
    shape = (int(rng.integers(1, 40)),)
    conn = int(rng.choice([6, 18, 26]))
  elif nd == 2:
    shape = tuple(int(s) for s in rng.integers(1, 23, size=2))
    conn = int(rng.choice([4, 8, 6, 18, 26]))
  else:
    shape = tuple(int(s) for s in rng.integers(1, 11, size=3))
    conn = int(rng.choice([6, 18, 26]))
  dt = str(rng.choice(DTYPES))
  nvals = int(rng.integers(2, 6))
  arr = rng.integers(0, nvals, size=shape)
  if dt == "bool":
    arr = arr > 0
  else:
    arr = arr.astype(dt)
    if dt.startswith("float") and rng.random() < 0.2:
      arr = arr + rng.random(size=shape).astype(dt)  # non-integer values
  order = "F" if rng.random() < 0.5 else "C"
  arr = np.asfortranarray(arr) if order == "F" else np.ascontiguousarray(arr)
  kind = str(rng.choice(["multi", "binary", "cont"]))
  kw = {"connectivity": conn}
  if kind == "binary":
    kw["binary_image"] = True
  if kind == "cont" and dt not in ("bool", "float16"):
    kw["delta"] = int(rng.integers(1, 3)) if not dt.startswith("float") else float(rng.choice([0.5, 1.0, 2.5]))
  if kind != "cont" and conn in (4, 6, 8) and rng.random() < 0.35:
    kw["periodic_boundary"] = True
  if rng.random() < 0.3:
    kw["out_dtype"] = str(rng.choice(["uint16", "uint32", "uint64"]))
  return arr, order, kw


def run_ref(arr, kw, binding=None):
  """The reference's result: its C++ core compiled into oracle/_ref behind the
  binding restatement, or -- for the single-foreground-row shortcut, which
  lives in the binding rather than the core -- the reference's compiled
  binding itself (make_special_row.build_binding)."""
  kw2 = dict(kw)
  if "out_dtype" in kw2:
    kw2["out_dtype"] = np.dtype(kw2["out_dtype"])
  try:
    out, N = oracle.connected_components(arr, return_N=True, backend="ref", **kw2)
    res = (out, int(N), "")
  except Exception as e:
    res = (None, -1, type(e).__name__)
  if oracle.last_path == "special_row":
    if binding is None:
      raise SystemExit("single-row case needs the reference binding (tests/golden/make_special_row.py)")
    try:
      out, N = binding.connected_components(arr, return_N=True, **kw2)
      res = (out, int(N), "")
    except Exception as e:
      res = (None, -1, type(e).__name__)
  return res


def main():
  if not oracle.have_ref():
    raise SystemExit("oracle/_ref/libcc3d_ref.so is missing: run `make -C oracle ref` where the reference exists")
  import make_special_row
  make_special_row.build_binding()
  sys.path.insert(0, make_special_row.BUILD)
  import fastcc3d as binding  # the reference's binding, built outside the repository
  rng = np.random.default_rng(20261202)
  blobs = {}
  meta = []
  cases = [case(rng, i) for i in range(600)]
  # quirk cases: binary 2D-8 odd sides (stale labels), binary 2D-4 row 0,
  # continuous 2D-8 range shortcut, union-find overflow
  extra = [
    (np.array([[0, 0, 0], [0, 0, 1], [0, 0, 1]], dtype=np.uint8), "C", {"connectivity": 8, "binary_image": True}),
    (np.array([1, 0, 1, 0, 1, 0, 1, 0], dtype=bool), "C", {"connectivity": 26}),
    (np.array([[3, 3, 5, 5, 7], [1, 1, 1, 1, 1]], dtype=np.uint8), "C", {"connectivity": 4, "binary_image": True}),
    (np.array([[10, 50, 90], [90, 10, 50]], dtype=np.uint8), "C", {"connectivity": 8, "delta": 45}),
  ]
  for k in range(40):
    shp = (int(rng.integers(3, 30)) | 1, int(rng.integers(3, 30)))
    a = (rng.random(shp) < 0.4).astype(np.uint8)
    extra.append((np.asfortranarray(a) if k % 2 else a, "F" if k % 2 else "C",
                  {"connectivity": 8, "binary_image": True}))
  for k in range(20):
    a = rng.integers(0, 4, size=(int(rng.integers(2, 20)), int(rng.integers(2, 20)))).astype(np.uint8)
    extra.append((a, "C", {"connectivity": 4, "binary_image": True}))
  for k in range(20):
    a = rng.integers(0, 200, size=(int(rng.integers(2, 16)), int(rng.integers(2, 16)))).astype(np.uint8)
    extra.append((a, "C", {"connectivity": 8, "delta": int(rng.integers(20, 120))}))
  cases += extra
  for i, (arr, order, kw) in enumerate(cases):
    out, N, err = run_ref(arr, kw, binding)
    blobs[f"in{i}"] = arr if arr.dtype != np.float16 else arr.view(np.uint16)
    if out is not None:
      blobs[f"out{i}"] = out
    meta.append({"i": i, "order": order, "dtype": str(arr.dtype), "shape": list(arr.shape), "kw": kw,
                 "N": N, "err": err})
  blobs["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
  np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **blobs)
  print(f"wrote {len(cases)} cases")


def load():
  """(arr, kw, expected_out or None, N, err) for every committed case."""
  z = np.load(os.path.join(HERE, "ref_vectors.npz"), allow_pickle=False)
  meta = json.loads(bytes(z["meta"]).decode())
  res = []
  for m in meta:
    i = m["i"]
    arr = z[f"in{i}"]
    if m["dtype"] == "float16":
      arr = arr.view(np.float16)
    arr = np.asfortranarray(arr) if m["order"] == "F" else np.ascontiguousarray(arr)
    kw = dict(m["kw"])
    if "out_dtype" in kw:
      kw["out_dtype"] = np.dtype(kw["out_dtype"])
    out = z[f"out{i}"] if f"out{i}" in z.files else None
    res.append((arr, kw, out, m["N"], m["err"]))
  return res


if __name__ == "__main__":
  main()
