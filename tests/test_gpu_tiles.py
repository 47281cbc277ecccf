"""All tile families of the local kernel (voxel and tall voxel tiles,
csrc/local.hpp; row tiles, csrc/rows.hpp) give the same labels on every
connectivity and mode, alone and mixed in one batch; the autotuned choice
changes nothing but speed."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _vols():
  import synth
  rng = np.random.default_rng(8)
  return [synth.voronoi((70, 50, 40), 90, seed=1).astype(np.uint32),
          synth.voronoi((64, 33, 17), 30, seed=2).astype(np.uint64),
          (rng.random((50, 40, 30)) < 0.3).astype(np.uint8),
          rng.integers(0, 4, size=(37, 29, 23)).astype(np.uint16)]


@pytest.mark.parametrize("family", ["voxels", "tall", "rows"])
@pytest.mark.parametrize("connectivity", [6, 18, 26])
def test_families_match_oracle(gpu, monkeypatch, family, connectivity):
  import cc3d
  from cc3d import _tune
  monkeypatch.setattr(_tune, "FORCE", family)
  for i, vol in enumerate(_vols()):
    binary = vol.dtype == np.uint8
    got, n = cc3d.connected_components(vol, connectivity=connectivity, binary_image=binary, return_N=True)
    exp, en = oracle.connected_components(vol, connectivity=connectivity, binary_image=binary, return_N=True)
    assert n == en and np.array_equal(got, exp), (family, connectivity, i)


def test_batch_mixes_families(gpu, monkeypatch):
  """Volumes of one batch labelled by different tile families (the tuner
  forced per volume by size) give the per-volume results."""
  import cc3d
  from cc3d import _tune
  real = _tune.tiles

  def alternate(connectivity, binary, itemsize, voxels, is3d, density):
    if not is3d:
      return 0
    return _tune._FLAG[_tune.FAMILIES[(voxels // 7) % 3]]
  monkeypatch.setattr(_tune, "tiles", alternate)
  import synth
  vols = [synth.voronoi((40 + 7 * k, 30 + 3 * k, 20 + k), 20 + k, seed=k).astype(np.uint32) for k in range(8)]
  conns = [26, 6, 18, 26, 6, 18, 6, 26]
  res = cc3d.connected_components_batch(vols, connectivity=conns, return_N=True)
  fams = {(int(np.prod(v.shape)) // 7) % 3 for v in vols}
  assert fams == {0, 1, 2}  # every family present
  for v, c, (lab, n) in zip(vols, conns, res):
    exp, en = oracle.connected_components(v, connectivity=c, return_N=True)
    assert n == en and np.array_equal(lab, exp)
  monkeypatch.setattr(_tune, "tiles", real)


@pytest.mark.parametrize("family", ["voxels", "tall"])
def test_voxel_families_continuous_and_periodic(gpu, monkeypatch, family):
  """Continuous links and periodic boundaries on both voxel-tile families."""
  import cc3d
  from cc3d import _tune
  import synth
  monkeypatch.setattr(_tune, "FORCE", family)
  v = synth.tones((70, 61, 45), 40, seed=3)
  for conn in (26, 18, 6):
    got, n = cc3d.connected_components(v, connectivity=conn, delta=20, return_N=True)
    exp, en = oracle.connected_components(v, connectivity=conn, delta=20, return_N=True)
    assert n == en and np.array_equal(got, exp), ("continuous", conn)
  rng = np.random.default_rng(4)
  p = rng.integers(0, 3, size=(45, 38, 29)).astype(np.uint16)
  got, n = cc3d.connected_components(p, connectivity=6, periodic_boundary=True, return_N=True)
  exp, en = oracle.connected_components(p, connectivity=6, periodic_boundary=True, return_N=True)
  assert n == en and np.array_equal(got, exp), "periodic"
