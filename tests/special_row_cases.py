"""Loader of tests/golden/special_row_vectors.npz (written by
tests/golden/make_special_row.py from the reference's compiled binding)."""
import json
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "special_row_vectors.npz")


def load():
  """[(arr, kwargs, expected_out or None, N, exc_type, exc_msg)]"""
  z = np.load(PATH, allow_pickle=False)
  meta = json.loads(str(z["meta"]))
  res = []
  for i, m in enumerate(meta):
    arr = z[f"in_{i}"]
    arr = np.asfortranarray(arr) if m["order"] == "F" else np.ascontiguousarray(arr)
    kw = dict(m["kwargs"])
    if "out_dtype" in kw:
      kw["out_dtype"] = np.dtype(kw["out_dtype"])
    out = z[f"out_{i}"] if f"out_{i}" in z.files else None
    res.append((arr, kw, out, m.get("N"), m.get("exc_type"), m.get("exc_msg")))
  return res


def check(fn, arr, kw, out, N, exc_type, exc_msg, tag):
  try:
    got, gN = fn(arr, return_N=True, **kw)
  except Exception as e:
    assert exc_type is not None, f"{tag}: raised {e!r}; the reference returns N={N}"
    assert type(e).__name__ == exc_type and str(e) == exc_msg, f"{tag}: {e!r} vs {exc_type}({exc_msg})"
    return
  assert exc_type is None, f"{tag}: returned N={gN}; the reference raises {exc_type}({exc_msg})"
  got = np.asarray(got)
  assert gN == N, f"{tag}: N {gN} vs {N}"
  assert got.dtype == out.dtype and got.shape == out.shape, f"{tag}: {got.dtype}{got.shape} vs {out.dtype}{out.shape}"
  assert np.array_equal(got, out), f"{tag}: labels differ"
