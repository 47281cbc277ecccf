"""Pins the oracle (CPU restatement) before it is trusted as the GPU checker.

Checks, all on the CPU:
  * the committed reference vectors (tests/golden/ref_vectors.npz, produced by
    the reference compiled from its own sources);
  * the compiled reference itself on fresh seeded inputs, when oracle/_ref is
    present;
  * scipy.ndimage.label on binary volumes (the reference's own cross-check,
    automated_test.py:499-554);
  * brute-force breadth-first labelling on tiny multilabel volumes;
  * estimate_provisional_labels against a direct numpy count;
  * planted mistakes in the oracle are caught by the golden vectors.
"""
import os
import sys
from collections import deque

import numpy as np
import pytest

import oracle

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_golden  # noqa: E402


def _compare(arr, kw, exp_out, expN, err):
  try:
    out, N = oracle.connected_components(arr, return_N=True, backend="oracle", **kw)
  except Exception as e:
    assert err and type(e).__name__ == err, f"oracle raised {e!r}, reference {err or 'ok'}"
    return
  assert not err, f"oracle ok, reference raised {err}"
  assert N == expN
  assert out.dtype == exp_out.dtype and out.shape == exp_out.shape
  assert np.array_equal(out, exp_out)


def test_oracle_matches_reference_vectors():
  for arr, kw, out, N, err in make_golden.load():
    _compare(arr, kw, out, N, err)


@pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built")
def test_oracle_matches_compiled_reference_random():
  rng = np.random.default_rng(99)
  for i in range(300):
    arr, order, kw = make_golden.case(rng, i)
    kw2 = dict(kw)
    if "out_dtype" in kw2:
      kw2["out_dtype"] = np.dtype(kw2["out_dtype"])
    try:
      ref = oracle.connected_components(arr, return_N=True, backend="ref", **kw2)
      _compare(arr, kw2, ref[0], ref[1], "")
    except RuntimeError:
      with pytest.raises(RuntimeError):
        oracle.connected_components(arr, return_N=True, backend="oracle", **kw2)


@pytest.mark.parametrize("conn", [26, 18, 6])
def test_binary_3d_matches_scipy(conn):
  import scipy.ndimage
  rng = np.random.default_rng(conn)
  labels = rng.integers(0, 2, (40, 37, 33)).astype(bool)
  if conn == 26:
    st = np.ones((3, 3, 3))
  elif conn == 18:
    st = np.array([[[0, 1, 0], [1, 1, 1], [0, 1, 0]], [[1, 1, 1]] * 3, [[0, 1, 0], [1, 1, 1], [0, 1, 0]]])
  else:
    st = None
  ours, N = oracle.connected_components(labels, connectivity=conn, return_N=True)
  sp, Ns = scipy.ndimage.label(labels, structure=st)
  assert N == Ns
  assert np.array_equal(ours, sp)


@pytest.mark.parametrize("conn", [4, 8])
def test_binary_2d_partition_matches_scipy(conn):
  """2D binary partitions equal scipy's (8-conn numbering follows 2x2 blocks)."""
  import scipy.ndimage
  rng = np.random.default_rng(conn + 100)
  img = rng.integers(0, 2, (61, 47)).astype(bool)
  st = np.ones((3, 3)) if conn == 8 else None
  ours, N = oracle.connected_components(img, connectivity=conn, return_N=True)
  sp, Ns = scipy.ndimage.label(img, structure=st)
  assert N == Ns
  pairs = set(zip(ours[img].tolist(), sp[img].tolist()))
  assert len(pairs) == N  # a bijection between the two labelings


def _bfs_label(vol, conn):
  """Textbook BFS labelling, numbered by first appearance in x-fastest order."""
  sx, sy, sz = vol.shape
  offs = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
          if (dx, dy, dz) != (0, 0, 0)]
  m = {26: 3, 18: 2, 6: 1}[conn]
  offs = [o for o in offs if sum(map(abs, o)) <= m]
  lab = np.zeros(vol.shape, dtype=np.int64)
  n = 0
  for z in range(sz):
    for y in range(sy):
      for x in range(sx):
        if vol[x, y, z] == 0 or lab[x, y, z]:
          continue
        n += 1
        lab[x, y, z] = n
        q = deque([(x, y, z)])
        while q:
          a = q.popleft()
          for o in offs:
            b = (a[0] + o[0], a[1] + o[1], a[2] + o[2])
            if 0 <= b[0] < sx and 0 <= b[1] < sy and 0 <= b[2] < sz and not lab[b] and vol[b] == vol[a]:
              lab[b] = n
              q.append(b)
  return lab, n


@pytest.mark.parametrize("conn", [26, 18, 6])
def test_multilabel_matches_bfs(conn):
  rng = np.random.default_rng(7 + conn)
  for _ in range(6):
    vol = np.asfortranarray(rng.integers(0, 4, (9, 8, 7)).astype(np.uint16))
    ours, N = oracle.connected_components(vol, connectivity=conn, return_N=True, out_dtype=np.uint32)
    exp, n = _bfs_label(vol, conn)
    assert N == n
    assert np.array_equal(ours.astype(np.int64), exp)


def test_epl_matches_numpy_count():
  rng = np.random.default_rng(3)
  for _ in range(10):
    vol = np.asfortranarray(rng.integers(0, 3, (17, 9, 5)).astype(np.uint32))
    epl, f, l = oracle.estimate_provisional_labels(vol)
    flat = np.ravel(vol, order="F").reshape(-1, 17)
    starts = (flat != 0) & np.concatenate([np.ones((flat.shape[0], 1), bool), flat[:, 1:] != flat[:, :-1]], 1)
    assert epl == int(starts.sum())
    rows = np.nonzero(starts.any(1))[0]
    assert (f, l) == ((int(rows[0]), int(rows[-1])) if len(rows) else (-1, -1))


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(mutation):
  """Each planted oracle mistake makes at least one golden vector disagree."""
  cases = make_golden.load()
  oracle.set_mutation(mutation)
  try:
    caught = False
    for arr, kw, out, N, err in cases:
      try:
        _compare(arr, kw, out, N, err)
      except AssertionError:
        caught = True
        break
    assert caught, f"mutation {mutation} went unnoticed"
  finally:
    oracle.set_mutation(0)


# ---- known answers transcribed from the reference's own test suite ---------

def test_known_2d_cross():
  img = np.ones((17, 17), dtype=np.uint8)
  img[:, 8] = 0
  img[8, :] = 0
  gt = np.zeros((17, 17), dtype=np.uint32)
  gt[:8, :8], gt[:8, 9:], gt[9:, :8], gt[9:, 9:] = 1, 2, 3, 4
  for conn in (4, 8, 6, 18, 26):
    assert np.array_equal(oracle.connected_components(img, connectivity=conn), gt)
  img[9:, 9:] = 2
  assert np.array_equal(oracle.connected_components(img, connectivity=26), gt)


def test_known_diagonals():
  inp = np.array([
    [0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
    [1, 0, 0, 1, 0, 0, 1, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0],
    [0, 1, 0, 0, 1, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0],
    [1, 0, 1, 0, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0, 0],
    [0, 0, 1, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0],
    [0, 1, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0],
  ], dtype=np.uint32)
  gt8 = np.array([
    [0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
    [2, 0, 0, 1, 0, 0, 3, 0, 4, 4, 0, 0, 0, 0, 0, 0, 0],
    [0, 2, 0, 0, 1, 0, 0, 0, 4, 4, 0, 0, 0, 0, 0, 0, 0],
    [2, 0, 2, 0, 0, 0, 4, 4, 0, 0, 4, 4, 0, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 4, 4, 0, 0, 4, 4, 0, 0, 0, 0, 0],
    [0, 0, 5, 0, 6, 0, 0, 0, 0, 4, 0, 0, 0, 0, 0, 0, 0],
    [0, 5, 0, 0, 0, 6, 0, 0, 4, 0, 0, 0, 0, 0, 0, 0, 0],
  ], dtype=np.uint32)
  for conn in (8, 18, 26):
    assert np.array_equal(oracle.connected_components(inp, connectivity=conn), gt8)
  gt4 = np.array([
    [0, 0, 1, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
    [3, 0, 0, 4, 0, 0, 5, 0, 6, 6, 0, 0, 0, 0, 0, 0, 0],
    [0, 7, 0, 0, 8, 0, 0, 0, 6, 6, 0, 0, 0, 0, 0, 0, 0],
    [9, 0, 10, 0, 0, 0, 11, 11, 0, 0, 12, 12, 0, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 11, 11, 0, 0, 12, 12, 0, 0, 0, 0, 0],
    [0, 0, 13, 0, 14, 0, 0, 0, 0, 15, 0, 0, 0, 0, 0, 0, 0],
    [0, 16, 0, 0, 0, 17, 0, 0, 18, 0, 0, 0, 0, 0, 0, 0, 0],
  ], dtype=np.uint32)
  for conn in (4, 6):
    assert np.array_equal(oracle.connected_components(inp, connectivity=conn), gt4)


def test_known_cross_with_intruder():
  img = np.ones((5, 5), dtype=np.uint8)
  img[:, 2] = 0
  img[2, :] = 0
  img[3:, 3:] = 2
  img[3, 3] = 1
  gt = np.array([[1, 1, 0, 2, 2], [1, 1, 0, 2, 2], [0, 0, 0, 0, 0], [3, 3, 0, 4, 5], [3, 3, 0, 5, 5]])
  for conn in (4, 6, 8, 18, 26):
    assert np.array_equal(oracle.connected_components(img, connectivity=conn), gt)


def test_known_periodic_counts():
  lab = np.zeros((10, 10), dtype=np.uint32, order="F")
  lab[:, 0] = 1
  lab[:, -1] = 1
  for conn in (4, 8):
    assert oracle.connected_components(lab, connectivity=conn, return_N=True)[1] == 2
    assert oracle.connected_components(lab, connectivity=conn, return_N=True, periodic_boundary=True)[1] == 1
  lab[0, 0], lab[-1, 0], lab[0, -1], lab[-1, -1] = 2, 3, 3, 2
  assert oracle.connected_components(lab, connectivity=8, return_N=True, periodic_boundary=True)[1] == 3
  v = np.zeros((10, 10, 10), dtype=np.uint32, order="F")
  v[1, :, :] = 1
  v[-2, :, :] = 1
  assert oracle.connected_components(v, connectivity=6, return_N=True, periodic_boundary=True)[1] == 2
  v[:] = 0
  v[0, :, :] = 1
  v[-1, :, :] = 1
  assert oracle.connected_components(v, connectivity=6, return_N=True, periodic_boundary=True)[1] == 1


def test_known_out_dtypes():
  assert oracle.connected_components(np.zeros((64, 64, 64), np.uint8)).dtype == np.uint16
  with pytest.raises(ValueError):
    oracle.connected_components(np.arange(41 ** 3, dtype=np.uint32) + 1, out_dtype=np.uint16)
  with pytest.raises(ValueError):
    oracle.connected_components(np.ones((4, 4, 4), np.uint8), out_dtype=np.uint8)


def test_vcg_reference_known_answers():
  """automated_test.py:878-950: the expected bitfields of the reference's own
  voxel-graph tests (2D 3x3 at 4/8, 3D 3x3x3 at 26, centre voxel zeroed)."""
  labels = np.ones((3, 3), dtype=np.uint8)
  assert np.all(oracle.voxel_connectivity_graph(labels, connectivity=4))
  labels[1, 1] = 0
  g4 = oracle.voxel_connectivity_graph(labels, connectivity=4)
  gt4 = np.array([[0x0f, 0b00001011, 0x0f], [0b00001110, 0x00, 0b00001101], [0x0f, 0b00000111, 0x0f]],
                 dtype=np.uint8)
  assert g4.dtype == np.uint8 and np.all(gt4.T == g4)
  g8 = oracle.voxel_connectivity_graph(labels, connectivity=8)
  gt8 = np.array([[0b11101111, 0b11111011, 0b11011111], [0b11111110, 0x00, 0b11111101],
                  [0b10111111, 0b11110111, 0b01111111]], dtype=np.uint8)
  assert np.all(gt8.T == g8)
  E = {
    (-1, -1, -1): 0b01111111111111111111111111, (+1, -1, -1): 0b10111111111111111111111111,
    (-1, +1, -1): 0b11011111111111111111111111, (+1, +1, -1): 0b11101111111111111111111111,
    (-1, -1, +1): 0b11110111111111111111111111, (+1, -1, +1): 0b11111011111111111111111111,
    (-1, +1, +1): 0b11111101111111111111111111, (+1, +1, +1): 0b11111110111111111111111111,
    (0, -1, -1): 0b11111111011111111111111111, (0, +1, -1): 0b11111111101111111111111111,
    (-1, 0, -1): 0b11111111110111111111111111, (+1, 0, -1): 0b11111111111011111111111111,
    (0, -1, +1): 0b11111111111101111111111111, (0, +1, +1): 0b11111111111110111111111111,
    (-1, 0, +1): 0b11111111111111011111111111, (+1, 0, +1): 0b11111111111111101111111111,
    (-1, -1, 0): 0b11111111111111110111111111, (+1, -1, 0): 0b11111111111111111011111111,
    (-1, +1, 0): 0b11111111111111111101111111, (+1, +1, 0): 0b11111111111111111110111111,
    (0, 0, -1): 0b11111111111111111111011111, (0, 0, +1): 0b11111111111111111111101111,
    (0, -1, 0): 0b11111111111111111111110111, (0, +1, 0): 0b11111111111111111111111011,
    (-1, 0, 0): 0b11111111111111111111111101, (+1, 0, 0): 0b11111111111111111111111110,
  }
  labels = np.ones((3, 3, 3), dtype=np.uint8)
  g = oracle.voxel_connectivity_graph(labels, connectivity=26)
  assert g.dtype == np.uint32 and np.all(g == 0x3FFFFFF)
  labels[1, 1, 1] = 0
  g = oracle.voxel_connectivity_graph(labels, connectivity=26)
  # voxel (x, y, z) sees the zeroed centre in direction (1 - x, 1 - y, 1 - z)
  for x in range(3):
    for y in range(3):
      for z in range(3):
        want = 0 if (x, y, z) == (1, 1, 1) else E[(1 - x, 1 - y, 1 - z)]
        assert int(g[x, y, z]) == want, (x, y, z)


def test_vcg_oracle_matches_reference_build():
  """The restatement against the reference's extract_voxel_connectivity_graph
  compiled from its own sources, on random labels of every supported kind."""
  if not oracle.have_ref():
    pytest.skip("reference build absent")
  rng = np.random.default_rng(1021)
  cases = []
  for shape in [(17, 9, 5), (1, 7, 3), (6, 1, 1), (33, 20)]:
    for conn in ([4, 8, 6, 18, 26] if len(shape) == 2 else [6, 18, 26]):
      for dt in (np.uint8, np.int16, np.uint32, np.int64, bool):
        cases.append((shape, conn, dt))
  cases.append(((40,), 26, np.uint16))
  for shape, conn, dt in cases:
    a = rng.integers(0, 3, size=shape)
    a = (a > 0) if dt == bool else a.astype(dt)
    for arr in (a, np.asfortranarray(a)):
      exp = oracle.voxel_connectivity_graph(arr, connectivity=conn, backend="ref")
      got = oracle.voxel_connectivity_graph(arr, connectivity=conn)
      assert got.dtype == exp.dtype and got.shape == exp.shape and np.array_equal(got, exp), (shape, conn, dt)


def test_oracle_special_row_matches_reference_binding():
  """The single-foreground-row shortcut (cc3d/fastcc3d.pyx:644-679) against
  the reference's compiled binding: float truncation of the run value, NaN /
  infinity / >= 2^63 conversion errors, uint16 overflow."""
  import special_row_cases as sr
  cases = sr.load()
  assert len(cases) >= 70 and sum(1 for c in cases if c[4]) >= 20
  for i, (arr, kw, out, N, et, em) in enumerate(cases):
    sr.check(lambda a, **k: oracle.connected_components(a, backend="oracle", **k), arr, kw, out, N, et, em,
             f"special row {i}")


def test_special_row_vectors_catch_planted_mistake(monkeypatch):
  """Labelling the row as runs of equal values (no int64 conversion) is
  caught by the vectors."""
  import special_row_cases as sr
  monkeypatch.setattr(oracle, "_int64_of", lambda v: v)
  bad = 0
  for arr, kw, out, N, et, em in sr.load():
    try:
      sr.check(lambda a, **k: oracle.connected_components(a, backend="oracle", **k), arr, kw, out, N, et, em, "")
    except AssertionError:
      bad += 1
  assert bad >= 20


def _np_statistics(lab):
  """numpy closed forms: bincount counts, per-label min/max of coordinates,
  coordinate sums / counts."""
  N = int(lab.max())
  nd = lab.ndim
  counts = np.bincount(lab.ravel().astype(np.int64), minlength=N + 1).astype(np.uint32)
  idx = np.nonzero(np.ones(lab.shape, bool))
  flat = lab.ravel()
  bb = np.zeros((N + 1, 2 * nd), dtype=np.int64)
  cen = np.full((N + 1, nd), np.nan)
  for a in range(nd):
    mn = np.full(N + 1, 0xFFFFFFFF, np.int64)
    mx = np.zeros(N + 1, np.int64)
    np.minimum.at(mn, flat, idx[a])
    np.maximum.at(mx, flat, idx[a])
    bb[:, 2 * a], bb[:, 2 * a + 1] = mn, mx
    s = np.zeros(N + 1, np.int64)
    np.add.at(s, flat, idx[a])
    with np.errstate(invalid="ignore", divide="ignore"):
      cen[:, a] = np.where(counts == 0, np.nan, s / counts)
  return counts, bb, cen


@pytest.mark.parametrize("order", ["C", "F"])
@pytest.mark.parametrize("shape,dtype", [((40, 33, 21), np.uint32), ((70, 45), np.uint16), ((17, 1, 9), np.uint8),
                                         ((9, 300), np.uint64)])
def test_oracle_statistics_matches_numpy(order, shape, dtype):
  rng = np.random.default_rng(sum(shape))
  lab = rng.integers(0, 30, size=shape).astype(dtype)
  lab[lab == 7] = 0  # an absent label
  lab = np.asfortranarray(lab) if order == "F" else np.ascontiguousarray(lab)
  c, b, cen = oracle.statistics_raw(lab)
  c2, b2, cen2 = _np_statistics(lab)
  assert np.array_equal(c, c2) and np.array_equal(b.astype(np.int64), b2)
  assert np.array_equal(np.isnan(cen), np.isnan(cen2)) and np.array_equal(np.nan_to_num(cen), np.nan_to_num(cen2))


def test_oracle_statistics_reference_known_answer():
  """cc3d's own statistics test values (automated_test.py:962-1006)."""
  for order in ("C", "F"):
    labels = np.zeros((123, 128, 125), dtype=np.uint8, order=order)
    labels[10:20, 10:20, 10:20] = 1
    labels[40:50, 40:50, 40:51] = 2
    c, b, cen = oracle.statistics_raw(labels)
    assert c[1] == 1000 and c[2] == 10 * 10 * 11
    assert np.allclose(cen[1], [14.5, 14.5, 14.5]) and np.allclose(cen[2], [44.5, 44.5, 45])
    assert list(b[0]) == [0, 122, 0, 127, 0, 124]
    assert list(b[1]) == [10, 19, 10, 19, 10, 19] and list(b[2]) == [40, 49, 40, 49, 40, 50]


def test_oracle_color_graph_known_answers():
  """color_connectivity_graph: the reference's own cases (automated_test.py:1436-1568)."""
  import ccg_known_answers as ka
  ka.check_all(lambda v, **k: oracle.color_connectivity_graph(v, backend="oracle", **k))


@pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built")
def test_oracle_color_graph_matches_compiled_reference():
  import ccg_known_answers as ka
  rng = np.random.default_rng(5)
  for vcg, conn in ka.random_cases(rng, 250):
    a, na = oracle.color_connectivity_graph(vcg, connectivity=conn, return_N=True, backend="ref")
    b, nb = oracle.color_connectivity_graph(vcg, connectivity=conn, return_N=True, backend="oracle")
    assert na == nb and a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b), (vcg.shape, conn)
  ka.check_all(lambda v, **k: oracle.color_connectivity_graph(v, backend="ref", **k))


def test_color_graph_checks_catch_planted_mistake(monkeypatch):
  """Reading q's bits instead of p's is caught: on a directed graph (bits set
  on one side only) the partition changes."""
  import ccg_known_answers as ka
  vcg = np.zeros((4, 3), dtype=np.uint8)
  vcg[1, 0] = 0b10  # voxel (1, 0) joins (0, 0); (0, 0) has no +x bit
  a, n = oracle.color_connectivity_graph(vcg, connectivity=4, return_N=True)
  assert n == 11 and a[0, 0] == a[1, 0]
  flipped = np.zeros((4, 3), dtype=np.uint8)
  flipped[0, 0] = 0b01  # only the +x side set: no join
  b, m = oracle.color_connectivity_graph(flipped, connectivity=4, return_N=True)
  assert m == 12


def test_oracle_contacts_known_answers():
  """contacts / region_graph: the reference's own cases (automated_test.py:689-877)."""
  import region_known_answers as ka
  ka.check_contacts(lambda v, **k: oracle.contacts(v, **k), lambda v, **k: oracle.region_graph(v, **k))
  if oracle.have_ref():
    ka.check_contacts(lambda v, **k: oracle.contacts(v, backend="ref", **k),
                      lambda v, **k: oracle.region_graph(v, backend="ref", **k))


@pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built")
def test_oracle_contacts_match_compiled_reference():
  """The restatement against extract_region_graph compiled from its header,
  float sums compared bit for bit (non-dyadic weights included)."""
  import region_known_answers as ka
  rng = np.random.default_rng(17)
  for lab, conn, sa, an in ka.random_label_volumes(rng, 200):
    a = oracle.contacts(lab, connectivity=conn, surface_area=sa, anisotropy=an)
    b = oracle.contacts(lab, connectivity=conn, surface_area=sa, anisotropy=an, backend="ref")
    assert a == b, (lab.shape, lab.dtype, conn, sa, an)


def test_contacts_checks_catch_planted_mistakes():
  """What the checks above must catch: ignoring the reference's doubled
  corner offsets on the x borders changes neighbour counts; summing the
  areas in another order changes non-dyadic sums."""
  lab = np.zeros((3, 3, 3), dtype=np.uint32)
  lab[0, 1, 1] = 1
  lab[0, 0, 0] = 2  # (0, -1, -1) from (0, 1, 1): the edge, also named by a corner offset at x = 0
  res = oracle.contacts(lab, connectivity=26, surface_area=False)
  assert res[(1, 2)] == 2  # counted twice, as the reference does
  res = oracle.contacts(lab, connectivity=18, surface_area=False)
  assert res[(1, 2)] == 1
  # order matters for non-dyadic sums: 0.1f added 10 times is not 10 * 0.1f
  acc = np.float32(0)
  for _ in range(10):
    acc = np.float32(acc + np.float32(0.1))
  assert acc != np.float32(10 * np.float32(0.1))


@pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built")
def test_oracle_runs_match_compiled_reference():
  rng = np.random.default_rng(3)
  for i in range(200):
    n = int(rng.integers(0, 60))
    dt = [np.uint8, np.uint16, np.uint32, np.uint64, bool][i % 5]
    lab = rng.integers(0, 3, size=n).astype(dt)
    a, b = oracle.runs(lab), oracle.runs(lab, backend="ref")
    assert a == b and list(a) == list(b)
  assert oracle.runs(np.zeros(1, np.uint8)) == {0: [(0, 1)]}  # the one-voxel special case


def test_oracle_each_known_answer():
  """cc3d.each (automated_test.py:660-676) on the restatement."""
  import region_known_answers as ka

  def each(labels, binary=False, in_place=False):
    class It:
      def __len__(self):
        r = oracle.runs(labels)
        return len(r) - int(0 in r)

      def __iter__(self):
        return oracle.each(labels, binary=binary, in_place=in_place)
    return It()
  ka.check_each(each, side=24)


def test_tile_family_rule():
  """The tuner's rule: the default family (rows at 6, voxels otherwise) unless
  another is at least MARGIN faster; ties keep the default."""
  from cc3d import _tune
  m = _tune.MARGIN
  assert _tune.decide(26, {"voxels": 1.0, "tall": 1.0 - m / 2, "rows": 2.0}) == "voxels"
  assert _tune.decide(26, {"voxels": 1.0, "tall": 1.0 - 2 * m, "rows": 2.0}) == "tall"
  assert _tune.decide(26, {"voxels": 1.0, "tall": 0.9, "rows": 0.8}) == "rows"
  assert _tune.decide(6, {"voxels": 0.9, "tall": 0.95, "rows": 1.0}) == "voxels"
  assert _tune.decide(6, {"voxels": 1.0, "tall": 1.0, "rows": 1.0}) == "rows"
  assert _tune.decide(18, {"voxels": 1.0, "rows": 1.5}) == "voxels"


def test_tile_table_lookup():
  """The tile-family table: every key present with both timings, choices
  following its rule, and the dispatcher's lookup (density sampled only
  when the choice depends on it)."""
  import json
  import torch
  from cc3d import _lib as L
  from cc3d import _tune
  with open(_tune.TABLE_PATH) as f:
    doc = json.load(f)
  ent = doc["entries"]
  for conn in (26, 18, 6):
    for mode, sizes in (("multi", (1, 2, 4, 8)), ("binary", (1,))):
      for isz in sizes:
        for cls in ("s", "m", "l"):
          row = ent[f"{conn}/{mode}/{isz}/{cls}"]
          for e in row.values():
            ms = {f: e[f"ms_{f}"] for f in _tune.FAMILIES if f"ms_{f}" in e}
            assert {"voxels", "rows"} <= set(ms)
            assert e["choice"] == _tune.decide(conn, ms)
  calls = []

  def density():
    calls.append(1)
    return 0.1
  flags = {"rows": L.FLAG_TILES_ROWS, "voxels": L.FLAG_TILES_VOXELS, "tall": L.FLAG_TILES_TALL}
  for k in ent:
    conn, mode, isz, cls = k.split("/")
    voxels = {"s": 64 ** 3, "m": 160 ** 3, "l": 400 ** 3}[cls]
    calls.clear()
    flag = _tune.tiles(int(conn), mode == "binary", int(isz), voxels, True, density)
    row = ent[k]
    fam, loss = _tune.unsampled_choice(row)
    sampled = len({v["choice"] for v in row.values()}) > 1 and loss > _tune.sample_ms()
    if sampled:
      exp = row[_tune.density_class(0.1)]["choice"]
    elif len({v["choice"] for v in row.values()}) == 1:
      exp = next(iter(row.values()))["choice"]
    else:
      exp = fam
    assert flag == flags[exp], k
    assert bool(calls) == sampled, k
  # the unsampled family minimises the worst loss against the per-density choices
  row = {"sparse": {"choice": "voxels", "ms_voxels": 1.0, "ms_tall": 1.01, "ms_rows": 2.0},
         "dense": {"choice": "tall", "ms_voxels": 1.2, "ms_tall": 1.0, "ms_rows": 2.0}}
  fam, loss = _tune.unsampled_choice(row)
  assert fam == "tall" and abs(loss - 0.01) < 1e-9
  assert _tune.tiles(26, False, 4, 100, False, density) == 0  # a single plane: no choice
  g = torch.Generator().manual_seed(0)
  buf = (torch.rand(100000, generator=g) < 0.3).to(torch.int32) * 7
  assert abs(_tune.sample_density(buf.view(torch.uint8), 4, 100000) - 0.3) < 0.04
