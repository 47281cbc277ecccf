"""Labelling in z-slabs: connected_components_stack (cc3d/__init__.py:353-501)
and connected_components on volumes streamed through the GPU slab by slab,
against the oracle on the whole volume."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _whole(vol, conn, binary):
  exp, n = oracle.connected_components(np.asfortranarray(vol), connectivity=conn, binary_image=binary,
                                       return_N=True)
  return exp.astype(np.uint64), n


@pytest.mark.parametrize("binary_image", [False, True])
@pytest.mark.parametrize("connectivity", [6, 26])
def test_stack_reference_cases(gpu, binary_image, connectivity):
  """automated_test.py:1602-1641: a stack of ones cut by one empty plane, and
  a random image cut into four slabs."""
  import cc3d
  stack = [np.ones([100, 100, 100], dtype=np.uint32) for _ in range(4)]
  stack += [np.zeros([100, 100, 1], dtype=np.uint32)]
  stack += [np.ones([100, 100, 100], dtype=np.uint32) for _ in range(2)]
  arr = cc3d.connected_components_stack(stack, connectivity=connectivity, binary_image=binary_image)
  ans = np.ones([100, 100, 601], dtype=np.uint32)
  ans[:, :, 400] = 0
  ans[:, :, 401:] = 2
  assert np.all(np.unique(arr[:]) == [0, 1, 2])
  assert np.all(arr[:] == ans)
  rng = np.random.default_rng(0)
  image = rng.integers(0, 100, size=[100, 100, 11], dtype=np.uint8)
  exp, n = _whole(image, connectivity, binary_image)
  stack = [image[:, :, :2], image[:, :, 2:6], image[:, :, 6:9], image[:, :, 9:11]]
  res, N = cc3d.connected_components_stack(stack, connectivity=connectivity, binary_image=binary_image,
                                           return_N=True)
  assert N == n and res.dtype == np.uint64 and np.array_equal(res, exp)


@pytest.mark.parametrize("connectivity", [6, 26])
def test_stack_random_slabs(gpu, connectivity, tmp_path):
  import cc3d
  import synth
  rng = np.random.default_rng(connectivity)
  cases = [synth.voronoi((90, 70, 64), 120, seed=3).astype(np.uint32),
           (rng.random((64, 50, 40)) < 0.35).astype(np.uint8),
           rng.integers(0, 3, size=(33, 17, 29)).astype(np.int16),
           synth.tones((60, 40, 30), 8, seed=2).astype(np.float32)]
  for i, vol in enumerate(cases):
    for binary in (False, True):
      cuts = sorted(set(int(c) for c in rng.integers(1, vol.shape[2], size=int(rng.integers(2, 6)))))
      bounds = [0] + cuts + [vol.shape[2]]
      parts = [vol[:, :, a:b] for a, b in zip(bounds[:-1], bounds[1:])]
      if i % 2:
        parts = [np.ascontiguousarray(p) for p in parts]  # C-order slabs
      print("stack case", i, vol.shape, vol.dtype, binary, bounds, flush=True)
      exp, n = _whole(vol, connectivity, binary)
      path = str(tmp_path / f"stack_{i}_{binary}.bin")
      res, N = cc3d.connected_components_stack(iter(parts), connectivity=connectivity, binary_image=binary,
                                               return_N=True, out_file=path)
      assert isinstance(res, np.memmap) and N == n and np.array_equal(res, exp), (i, binary)
  with pytest.raises(ValueError):
    cc3d.connected_components_stack([np.ones((4, 4, 4), np.uint8)], connectivity=18)


_FORCED = [
  # (volume maker, kwargs)
  (lambda s: s.voronoi((80, 60, 50), 200, seed=1).astype(np.uint32), dict(connectivity=26)),
  (lambda s: s.voronoi((80, 60, 50), 200, seed=1).astype(np.uint64), dict(connectivity=18)),
  (lambda s: s.voronoi((80, 60, 50), 200, seed=1).astype(np.uint16), dict(connectivity=6)),
  (lambda s: s.bernoulli((70, 50, 40), 0.3, seed=4), dict(connectivity=26, binary_image=True)),
  (lambda s: s.bernoulli((70, 50, 40), 0.5, seed=5), dict(connectivity=6, binary_image=True)),
  (lambda s: s.tones((64, 48, 36), 10, seed=6).astype(np.float64), dict(connectivity=26)),
  (lambda s: np.asfortranarray(s.voronoi((50, 40, 30), 60, seed=7).astype(np.int32)), dict(connectivity=26)),
  (lambda s: s.bernoulli((66, 44, 33), 0.45, seed=8), dict(connectivity=26, out_dtype=np.uint16)),
]


@pytest.mark.parametrize("case", range(len(_FORCED)))
def test_connected_components_streamed(gpu, monkeypatch, case):
  """connected_components routed through the slab path (forced by lowering
  its size threshold): labels, N, dtype and errors as the oracle gives them
  for the whole volume."""
  import cc3d
  import synth
  make, kw = _FORCED[case]
  vol = make(synth)
  plane = vol.shape[0] * vol.shape[1] if vol.flags.f_contiguous else vol.shape[-1] * vol.shape[-2]
  monkeypatch.setattr(cc3d._stack, "_STREAM_VOXELS", 1)
  monkeypatch.setattr(cc3d._stack, "_SLAB_VOXELS", plane * 7)  # slabs of 7 planes
  print("streamed case", case, vol.shape, vol.dtype, kw, flush=True)
  try:
    exp, en = oracle.connected_components(vol, return_N=True, **kw)
    exp_err = None
  except Exception as e:  # noqa: BLE001
    exp_err = e
  if exp_err is not None:
    with pytest.raises(type(exp_err), match=str(exp_err)):
      cc3d.connected_components(vol, return_N=True, **kw)
    return
  got, gn = cc3d.connected_components(vol, return_N=True, **kw)
  assert gn == en and got.dtype == exp.dtype and got.shape == exp.shape and np.array_equal(got, exp)


def test_streamed_single_row_and_empty(gpu, monkeypatch):
  import cc3d
  monkeypatch.setattr(cc3d._stack, "_STREAM_VOXELS", 1)
  monkeypatch.setattr(cc3d._stack, "_SLAB_VOXELS", 20 * 10 * 3)
  vol = np.zeros((20, 10, 12), dtype=np.float32)
  vol[:, 4, 9] = [0, 1.5, 1.5, 2, 2, 2, 0, 3, 3, 0, 1, 1, 1, 0, 0, 7, 7.5, 7.5, 0, 1]  # the one foreground row
  exp, en = oracle.connected_components(vol, return_N=True)
  got, gn = cc3d.connected_components(vol, return_N=True)
  assert gn == en and np.array_equal(got, exp) and got.dtype == exp.dtype
  vol = np.zeros((20, 10, 12), dtype=np.uint8)
  exp, en = oracle.connected_components(vol, return_N=True)
  got, gn = cc3d.connected_components(vol, return_N=True)
  assert gn == en == 0 and np.array_equal(got, exp)


def test_more_than_2_32_voxels(gpu):
  """A 2048 x 2048 x 1025 uint8 volume (4.3 G voxels, past the 2^32-1 limit of
  one labelling call), labelled in slabs of 512 planes: four bands along x,
  cut into z-segments by empty planes at z = 300 and 700, so the components
  of the middle and last segments cross the slab seams at z = 512 and 1024.
  N and the labels are checked against the closed form on the planes around
  each seam and cut, and on a sample of voxels."""
  import cc3d
  sx, sy, sz = 2048, 2048, 1025
  vol = np.empty((sx, sy, sz), dtype=np.uint8, order="F")
  band = (np.arange(sx) // 512 + 1).astype(np.uint8)
  vol[...] = band[:, None, None]
  cuts = [300, 700]  # empty planes: segments [0,300) [301,700) [701,1025)
  for c in cuts:
    vol[:, :, c] = 0
  print("volume built", flush=True)
  labels, N = cc3d.connected_components(vol, connectivity=26, return_N=True)
  assert N == 12 and labels.dtype == np.uint32
  # first appearance: segment s, band b -> 4 * s + b
  for z in (0, 299, 301, 511, 512, 513, 699, 701, 1023, 1024):
    seg = sum(z > c for c in cuts)
    assert np.array_equal(labels[:, 0, z], (4 * seg + band).astype(np.uint32)), z
    assert np.all(labels[:, :, z] == labels[:, :1, z])
  for c in cuts:
    assert not labels[:, :, c].any()
  rng = np.random.default_rng(1)
  idx = rng.integers(0, [sx, sy, sz], size=(100000, 3))
  got = labels[idx[:, 0], idx[:, 1], idx[:, 2]]
  seg = (idx[:, 2] > cuts[0]).astype(np.int64) + (idx[:, 2] > cuts[1])
  exp = np.where(np.isin(idx[:, 2], cuts), 0, 4 * seg + band[idx[:, 0]])
  assert np.array_equal(got, exp.astype(np.uint32))
