"""statistics and dust on the GPU against CPU restatements.

statistics is checked against numpy (bincount for counts, per-label min/max of
np.nonzero coordinates, sums / counts for centroids -- exact integer sums, so
the float64 division must match bit for bit).  dust is checked against a
restatement of cc3d/__init__.py:71-155 driven by the oracle's labels.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _np_stats(labels):
  lab = labels
  while lab.ndim < 2:
    lab = lab[..., None]
  N = int(lab.max())
  nd = lab.ndim
  counts = np.bincount(lab.ravel().astype(np.int64), minlength=N + 1).astype(np.uint32)
  idx = np.nonzero(np.ones(lab.shape, bool))
  flat = lab.ravel()  # C order, matching idx
  bb = np.zeros((N + 1, 2 * nd), dtype=np.int64)
  cen = np.full((N + 1, nd), np.nan)
  for a in range(nd):
    mn = np.full(N + 1, np.iinfo(np.uint16).max if max(lab.shape) <= 65535 else np.iinfo(np.uint32).max, np.int64)
    mx = np.zeros(N + 1, np.int64)
    np.minimum.at(mn, flat, idx[a])
    np.maximum.at(mx, flat, idx[a])
    bb[:, 2 * a], bb[:, 2 * a + 1] = mn, mx
    s = np.zeros(N + 1, np.int64)
    np.add.at(s, flat, idx[a])
    with np.errstate(invalid="ignore", divide="ignore"):
      c = s.astype(np.float64) / counts.astype(np.float64)
    cen[:, a] = np.where(counts == 0, np.nan, c)
  return counts, bb, cen


@pytest.mark.parametrize("order", ["C", "F"])
@pytest.mark.parametrize("shape", [(40, 33, 21), (70, 45), (200,)])
def test_statistics_match_numpy(gpu, order, shape):
  rng = np.random.default_rng(len(shape))
  lab = rng.integers(0, 30, size=shape).astype(np.uint32)
  lab[lab == 7] = 0  # an absent label
  _check_stats(np.asfortranarray(lab) if order == "F" else np.ascontiguousarray(lab))


def _check_stats(lab):
  import cc3d
  st = cc3d.statistics(lab, no_slice_conversion=True)
  counts, bb, cen = _np_stats(lab)
  assert np.array_equal(st["voxel_counts"], counts)
  assert np.array_equal(st["bounding_boxes"].astype(np.int64), bb)
  assert np.array_equal(np.isnan(st["centroids"]), np.isnan(cen))
  assert np.array_equal(np.nan_to_num(st["centroids"]), np.nan_to_num(cen))


@pytest.mark.parametrize("shape,order,dtype", [
  ((1, 50, 40), "F", np.uint16),    # unit fastest axis: the next axis is contiguous
  ((40, 50, 1), "C", np.uint32),
  ((1, 1, 300), "F", np.uint8),
  ((130, 17, 9), "F", np.uint64),   # partial bricks on every axis
  ((9, 17, 130), "C", np.uint32),
  ((5000, 3), "F", np.uint32),      # few long rows
  ((3, 70000), "C", np.uint16),
])
def test_statistics_brick_edges(gpu, shape, order, dtype):
  print("statistics brick case", shape, order, np.dtype(dtype).name, flush=True)
  rng = np.random.default_rng(11)
  lab = rng.integers(0, 12, size=shape).astype(dtype)
  lab = np.asfortranarray(lab) if order == "F" else np.ascontiguousarray(lab)
  _check_stats(lab)


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
def test_statistics_table_overflow(gpu, dtype):
  """Nearly every voxel its own label: most runs find no free slot in the
  per-brick table and go straight to global memory."""
  rng = np.random.default_rng(12)
  lab = rng.integers(0, 64 * 64 * 16, size=(64, 64, 16)).astype(dtype)  # labels < voxels
  lab[:, :, 3] = 0
  _check_stats(np.asfortranarray(lab))


def test_statistics_many_bricks_per_workgroup(gpu):
  """A segmented volume with more bricks than resident workgroups, so each
  workgroup walks several bricks and carries label 0's totals across them."""
  import synth
  lab = synth.voronoi((256, 256, 160), 3000, seed=13, id_range=(1, 400000))
  _check_stats(lab)
  _check_stats(np.ascontiguousarray(lab[:, :, :97]))


def test_statistics_reference_known_answer(gpu):
  """cc3d's own statistics test values (automated_test.py:962-1006)."""
  import cc3d
  for order in ("C", "F"):
    labels = np.zeros((123, 128, 125), dtype=np.uint8, order=order)
    labels[10:20, 10:20, 10:20] = 1
    labels[40:50, 40:50, 40:51] = 2
    st = cc3d.statistics(labels)
    assert st["voxel_counts"][1] == 1000
    assert st["voxel_counts"][2] == 10 * 10 * 11
    assert np.allclose(st["centroids"][1, :], [14.5, 14.5, 14.5])
    assert np.allclose(st["centroids"][2, :], [44.5, 44.5, 45])
    assert st["bounding_boxes"][0] == (slice(0, 123), slice(0, 128), slice(0, 125))
    assert st["bounding_boxes"][1] == (slice(10, 20), slice(10, 20), slice(10, 20))
    assert st["bounding_boxes"][2] == (slice(40, 50), slice(40, 50), slice(40, 51))
  st = cc3d.statistics(np.zeros((0, 1, 1), dtype=np.uint8))
  assert st == {"voxel_counts": None, "bounding_boxes": None, "centroids": None}


def _dust_restated(img, threshold, connectivity, invert):
  """cc3d/__init__.py:71-155 with the oracle's labels."""
  orig = img.dtype
  img = img.view(f"u{img.dtype.itemsize}") if img.dtype.kind == "i" else img
  img = np.copy(img)
  cc, N = oracle.connected_components(img, connectivity=connectivity, return_N=True)
  sizes = np.bincount(cc.ravel().astype(np.int64), minlength=N + 1)
  if isinstance(threshold, (tuple, list)):
    to_mask = [i for i in range(1, N + 1) if not (threshold[0] <= sizes[i] < threshold[1])]
  else:
    to_mask = [i for i in range(1, N + 1) if sizes[i] < threshold]
  dust_N = len(to_mask) if invert else N - len(to_mask)
  if not to_mask:
    if invert:
      img = np.zeros(img.shape, dtype=img.dtype, order="F")
    return img, dust_N
  mask = np.isin(cc, to_mask, invert=invert)
  img[mask] = 0
  return img.view(orig), dust_N


@pytest.mark.parametrize("invert", [False, True])
@pytest.mark.parametrize("threshold", [20, (25, 491)])
@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint32])
def test_dust_matches_restatement(gpu, invert, threshold, dtype):
  import cc3d
  rng = np.random.default_rng(5)
  img = rng.integers(0, 5, size=(60, 50, 40)).astype(dtype)
  got, gN = cc3d.dust(img, threshold=threshold, connectivity=26, invert=invert, return_N=True)
  exp, eN = _dust_restated(img, threshold, 26, invert)
  assert gN == eN
  assert got.dtype == exp.dtype
  assert np.array_equal(got, exp)


def test_dust_reference_known_answer(gpu):
  """automated_test.py:1160-1200 (dust static)."""
  import cc3d
  for conn in (6, 18, 26):
    labels = np.zeros((100, 100, 10), dtype=np.uint8, order="F")
    labels[:5, :5, :1] = 1
    labels[20:40, 20:40, :] = 2
    rec, N = cc3d.dust(labels, threshold=26, connectivity=conn, return_N=True)
    assert list(np.unique(rec)) == [0, 2] and N == 1
    rec, N = cc3d.dust(labels, threshold=26, connectivity=conn, invert=True, return_N=True)
    assert list(np.unique(rec)) == [0, 1] and N == 1


def _dust_full_restated(img, threshold, connectivity=26, in_place=False, precomputed_ccl=False, invert=False):
  """cc3d/__init__.py:101-155 step by step, with the oracle's labels and the
  checks of the reference's statistics() (cc3d/fastcc3d.pyx:716-748)."""
  orig = img.dtype
  img = img.view(f"u{img.dtype.itemsize}") if img.dtype.kind == "i" else img
  if not in_place:
    img = np.copy(img)
  if precomputed_ccl:
    cc, N = img, int(np.max(img))
  else:
    cc, N = oracle.connected_components(img, connectivity=connectivity, return_N=True)
  if N > cc.size:
    raise ValueError(
      f"Statistics can only be computed on volumes containing labels with values lower than the number of voxels. Max: {N}")
  sizes = np.bincount(cc.ravel().astype(np.int64), minlength=N + 1)
  if isinstance(threshold, (tuple, list)):
    to_mask = [i for i in range(1, N + 1) if not (threshold[0] <= sizes[i] < threshold[1])]
  else:
    to_mask = [i for i in range(1, N + 1) if sizes[i] < threshold]
  dust_N = len(to_mask) if invert else N - len(to_mask)
  if not to_mask:
    if invert:
      img = np.zeros(img.shape, dtype=img.dtype, order="F")
    return img, dust_N
  mask = np.isin(cc, to_mask, invert=invert)
  img[mask] = 0
  return img.view(orig), dust_N


@pytest.mark.parametrize("precomputed", [False, True])
@pytest.mark.parametrize("invert", [False, True])
def test_dust_in_place_strided_view(gpu, precomputed, invert):
  """in_place=True on a non-contiguous view writes through to the caller's
  array, as the reference's img[mask] = 0 does."""
  import cc3d
  rng = np.random.default_rng(17 + 2 * precomputed + invert)
  base = rng.integers(0, 4, size=(50, 40, 30)).astype(np.uint16)
  if precomputed:
    lab, _ = oracle.connected_components(base, connectivity=26, return_N=True, out_dtype=np.uint32)
    base = lab.astype(np.int32)
  a = base.copy()
  b = base.copy()
  va, vb = a[:, ::2, 1:], b[:, ::2, 1:]
  assert not (va.flags.c_contiguous or va.flags.f_contiguous)
  got, gN = cc3d.dust(va, threshold=6, connectivity=26, in_place=True, precomputed_ccl=precomputed,
                      invert=invert, return_N=True)
  exp, eN = _dust_full_restated(vb, 6, 26, in_place=True, precomputed_ccl=precomputed, invert=invert)
  assert gN == eN
  assert got.dtype == exp.dtype and np.array_equal(got, exp)
  assert np.array_equal(a, b), "the caller's array differs from the reference's in-place result"
  # not in place: the caller's array is untouched
  c = base.copy()
  got2 = cc3d.dust(c[:, ::2, 1:], threshold=6, connectivity=26, precomputed_ccl=precomputed, invert=invert)
  assert np.array_equal(c, base) and np.array_equal(got2, exp)


def test_dust_precomputed_label_checks(gpu):
  """precomputed_ccl runs the reference statistics() check: a label above the
  voxel count (a negative signed label viewed as unsigned included) raises
  ValueError with the reference's message."""
  import cc3d
  img = np.zeros((4, 4, 4), dtype=np.uint32)
  img[0, 0, 0] = 1000
  with pytest.raises(ValueError) as e1:
    cc3d.dust(img, threshold=3, precomputed_ccl=True)
  with pytest.raises(ValueError) as e2:
    _dust_full_restated(img, 3, precomputed_ccl=True)
  assert str(e1.value) == str(e2.value)
  neg = np.zeros((4, 4, 4), dtype=np.int16)
  neg[1, 2, 3] = -2
  with pytest.raises(ValueError) as e1:
    cc3d.dust(neg, threshold=3, precomputed_ccl=True)
  with pytest.raises(ValueError) as e2:
    _dust_full_restated(neg, 3, precomputed_ccl=True)
  assert str(e1.value) == str(e2.value)


def _largest_k_restated(img, k, connectivity=26, delta=0, return_N=False, binary_image=False,
                        precomputed_ccl=False):
  """cc3d/__init__.py:199-279 without fastremap (absent here, as in a plain
  install), with the oracle's labels and numpy's counts."""
  order = "C" if img.flags.c_contiguous else "F"
  if k == 0:
    return np.zeros(img.shape, dtype=np.uint16, order=order)
  if precomputed_ccl:
    cc = np.copy(img, order="F")
    N = np.max(cc)
  else:
    cc, N = oracle.connected_components(img, connectivity=connectivity, return_N=True, delta=delta,
                                        binary_image=bool(binary_image))
  if N <= k:
    return (cc, N) if return_N else cc
  cts = np.bincount(cc.ravel().astype(np.int64), minlength=int(N) + 1).astype(np.uint32)
  if k == 1:
    out = cc == (np.argmax(cts[1:]) + 1)
    return (out, 1) if return_N else out
  preserve = np.argpartition(cts[1:], len(cts) - k - 1)[-k:] + 1
  keep = list(map(int, sorted(preserve, key=lambda label: cts[label])))
  out = np.zeros(cc.shape, dtype=cc.dtype, order="C" if cc.flags.c_contiguous else "F")
  for i, label in enumerate(keep):
    out[cc == label] = i + 1
  return (out, len(keep)) if return_N else out


def _same_array(a, b):
  return (a.dtype == b.dtype and a.shape == b.shape and a.flags.c_contiguous == b.flags.c_contiguous
          and a.flags.f_contiguous == b.flags.f_contiguous and np.array_equal(a, b))


@pytest.mark.parametrize("k", [0, 1, 2, 5, 50])
@pytest.mark.parametrize("order,connectivity", [("C", 26), ("F", 6)])
def test_largest_k_matches_restatement(gpu, k, order, connectivity):
  """The reference's own case (automated_test.py:1286-1321: uint8 noise with 5
  values, 100^3), checked exactly, ties included."""
  import cc3d
  rng = np.random.default_rng(1286)
  img = rng.integers(0, 5, size=(100, 100, 100)).astype(np.uint8)
  img = np.asfortranarray(img) if order == "F" else np.ascontiguousarray(img)
  got = cc3d.largest_k(img, k=k, connectivity=connectivity, return_N=True)
  exp = _largest_k_restated(img, k, connectivity=connectivity, return_N=True)
  if k == 0:
    assert _same_array(got, exp)
    return
  assert _same_array(got[0], exp[0]) and int(got[1]) == int(exp[1])


def test_largest_k_variants(gpu):
  """precomputed labels (signed, C order; also N <= k), binary and continuous
  inputs, a 2D image, and inputs labelled through the host path (a strided
  view, float16)."""
  import cc3d
  import synth
  vor = synth.voronoi((96, 80, 40), 60, seed=7, id_range=(1, 1000))
  cc = oracle.connected_components(vor, connectivity=26).astype(np.int32)
  cc = np.ascontiguousarray(cc)
  cases = [
    (cc, dict(k=7, precomputed_ccl=True, return_N=True)),
    (cc, dict(k=1, precomputed_ccl=True)),
    (cc, dict(k=10 ** 6, precomputed_ccl=True, return_N=True)),
    (vor > 500, dict(k=3, binary_image=True, connectivity=18)),
    (synth.tones((64, 64, 32), 40, seed=3), dict(k=4, delta=20, connectivity=26)),
    (np.random.default_rng(9).integers(0, 3, size=(300, 200)).astype(np.uint16), dict(k=6, connectivity=8)),
    (np.random.default_rng(10).integers(0, 4, size=(40, 60, 30)).astype(np.int16)[:, ::2, :], dict(k=5)),
    (np.random.default_rng(11).integers(0, 4, size=(50, 40, 20)).astype(np.float16), dict(k=3)),
  ]
  for i, (img, kw) in enumerate(cases):
    print("largest_k case", i, img.shape, img.dtype, kw, flush=True)
    got = cc3d.largest_k(img, **kw)
    exp = _largest_k_restated(img, **kw)
    if kw.get("return_N"):
      assert _same_array(got[0], exp[0]) and int(got[1]) == int(exp[1]), i
    else:
      assert _same_array(got, exp), i


def test_vcg_reference_known_answers(gpu):
  """automated_test.py:878-950 through the product."""
  import cc3d
  labels = np.ones((3, 3), dtype=np.uint8)
  assert np.all(cc3d.voxel_connectivity_graph(labels, connectivity=8))
  labels[1, 1] = 0
  g = cc3d.voxel_connectivity_graph(labels, connectivity=4)
  gt = np.array([[0x0f, 0b00001011, 0x0f], [0b00001110, 0x00, 0b00001101], [0x0f, 0b00000111, 0x0f]],
                dtype=np.uint8)
  assert g.dtype == np.uint8 and np.all(gt.T == g)
  labels = np.ones((3, 3, 3), dtype=np.uint8)
  labels[1, 1, 1] = 0
  g = cc3d.voxel_connectivity_graph(labels, connectivity=26)
  assert g.dtype == np.uint32 and np.array_equal(g, oracle.voxel_connectivity_graph(labels, connectivity=26))
  assert int(g[0, 0, 0]) == 0b11111110111111111111111111  # the centre is at (+1, +1, +1)


@pytest.mark.parametrize("shape,conns", [((70, 45, 33), (6, 18, 26)), ((300, 257), (4, 8, 6, 26)),
                                         ((1000,), (26,)), ((1, 40, 9), (18,))])
@pytest.mark.parametrize("dt", [np.uint8, np.int16, np.uint32, np.uint64, bool])
def test_vcg_matches_oracle_and_reference(gpu, shape, conns, dt):
  import cc3d
  rng = np.random.default_rng(sum(shape))
  a = rng.integers(0, 3, size=shape)
  a = (a > 0) if dt == bool else a.astype(dt)
  for arr in (a, np.asfortranarray(a)):
    for conn in conns:
      print("vcg case", shape, np.dtype(dt).name, conn, arr.flags.f_contiguous, flush=True)
      got = cc3d.voxel_connectivity_graph(arr, connectivity=conn)
      exp = oracle.voxel_connectivity_graph(arr, connectivity=conn)
      assert got.dtype == exp.dtype and got.shape == exp.shape and got.flags.f_contiguous
      assert np.array_equal(got, exp)
      if oracle.have_ref():
        assert np.array_equal(got, oracle.voxel_connectivity_graph(arr, connectivity=conn, backend="ref"))


def test_vcg_errors(gpu):
  import cc3d
  with pytest.raises(ValueError):
    cc3d.voxel_connectivity_graph(np.zeros((4, 4, 4), np.uint8), connectivity=8)
  with pytest.raises(TypeError):
    cc3d.voxel_connectivity_graph(np.zeros((4, 4), np.float32), connectivity=8)
  with pytest.raises(cc3d.DimensionError):
    cc3d.voxel_connectivity_graph(np.zeros((2, 2, 2, 2), np.uint8))
  assert cc3d.voxel_connectivity_graph(np.zeros((0, 3), np.uint8), connectivity=4).shape == (0,)


def test_color_graph_reference_known_answers(gpu):
  """automated_test.py:1436-1568 through the product."""
  import cc3d
  import ccg_known_answers as ka
  ka.check_all(cc3d.color_connectivity_graph)


def test_color_graph_matches_oracle(gpu):
  import cc3d
  import ccg_known_answers as ka
  rng = np.random.default_rng(11)
  for i, (vcg, conn) in enumerate(ka.random_cases(rng, 120)):
    print("colour case", i, vcg.shape, vcg.dtype, conn, flush=True)
    got, gn = cc3d.color_connectivity_graph(vcg, connectivity=conn, return_N=True)
    exp, en = oracle.color_connectivity_graph(vcg, connectivity=conn, return_N=True)
    assert got.dtype == exp.dtype and got.shape == exp.shape and got.flags.f_contiguous
    assert gn == en and np.array_equal(got, exp), i
    if oracle.have_ref() and i % 4 == 0:
      assert np.array_equal(got, oracle.color_connectivity_graph(vcg, connectivity=conn, backend="ref"))


@pytest.mark.parametrize("conn", [6, 26])
def test_color_graph_full_size_round_trip(gpu, conn):
  """A 256^3 label volume with no background: the colouring of its own
  connectivity graph is its connected-components labelling (both number by
  first voxel in F order), and it equals the oracle's."""
  import cc3d
  import synth
  lab = np.asfortranarray(synth.voronoi((256, 256, 256), 3000, seed=9).astype(np.uint32) + 1)
  lab[::7, ::5, ::3] = 5  # broken-up regions
  vcg = cc3d.voxel_connectivity_graph(lab, connectivity=conn)
  got, gn = cc3d.color_connectivity_graph(vcg, connectivity=conn, return_N=True)
  cc, cn = cc3d.connected_components(lab, connectivity=conn, return_N=True)
  assert gn == cn and np.array_equal(got, cc)
  exp, en = oracle.color_connectivity_graph(vcg, connectivity=conn, return_N=True)
  assert en == gn and np.array_equal(got, exp)


def test_color_graph_errors(gpu):
  import cc3d
  with pytest.raises(cc3d.DimensionError):
    cc3d.color_connectivity_graph(np.zeros((2, 2, 2, 2), np.uint32))
  with pytest.raises(ValueError):
    cc3d.color_connectivity_graph(np.zeros((4, 4), np.uint8), connectivity=18)
  with pytest.raises(ValueError):
    cc3d.color_connectivity_graph(np.zeros((4, 4, 4), np.uint32), connectivity=8)
  with pytest.raises(ValueError):
    cc3d.color_connectivity_graph(np.zeros((4, 4), np.uint16), connectivity=4)
