"""Batched labelling on the GPU: every result equals the single-volume path.

connected_components_batch has no reference counterpart; its contract is that
result i is connected_components(volumes[i], ...) bit for bit (labels, N,
dtype, exceptions), and the single-volume path is itself pinned against the
reference (test_gpu_parity.py).  The batch of Voronoi volumes mirrors the
batched workload (uneven sides, connectivity 6 or 26 per volume) at a small
size; graph replays are checked against eager runs, with eager work between
replays.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
  (la, na), (lb, nb) = a, b
  la = la.cpu().numpy() if hasattr(la, "cpu") else la
  lb = lb.cpu().numpy() if hasattr(lb, "cpu") else lb
  return na == nb and la.dtype == lb.dtype and la.shape == lb.shape and np.array_equal(la, lb)


def _mixed_volumes(rng, n):
  import synth
  vols, conns = [], []
  for i in range(n):
    kind = i % 4
    if kind == 0:
      shape = tuple(int(s) for s in rng.choice([17, 40, 64, 96], size=3))
      v = synth.voronoi(shape, max(2, int(np.prod(shape)) // 8000), seed=100 + i)
    elif kind == 1:
      shape = tuple(int(s) for s in rng.integers(1, 50, size=3))
      v = rng.integers(0, 4, size=shape).astype(np.uint32)
    elif kind == 2:
      shape = tuple(int(s) for s in rng.integers(1, 120, size=2))
      v = rng.integers(0, 3, size=shape).astype(np.uint32)
    else:
      shape = (int(rng.integers(1, 300)),)
      v = rng.integers(0, 3, size=shape).astype(np.uint32)
    if rng.random() < 0.5:
      v = np.asfortranarray(v)
    vols.append(v)
    conns.append(int(rng.choice([6, 18, 26])) if v.ndim != 2 else int(rng.choice([4, 8, 6, 18, 26])))
  return vols, conns


def test_batch_matches_single(gpu):
  import cc3d
  rng = np.random.default_rng(7)
  vols, conns = _mixed_volumes(rng, 24)
  got = cc3d.connected_components_batch(vols, connectivity=conns, return_N=True)
  for i, (v, c) in enumerate(zip(vols, conns)):
    exp = cc3d.connected_components(v, connectivity=c, return_N=True)
    assert _same(got[i], exp), f"volume {i} shape={v.shape} conn={c}"


@pytest.mark.parametrize("conn", [6, 26])
def test_batch_binary_matches_single(gpu, conn):
  import cc3d
  import synth
  vols = [synth.bernoulli((int(a), int(b), int(c)), p, seed=s)
          for s, (a, b, c, p) in enumerate([(64, 40, 33, 0.3), (96, 96, 20, 0.5), (31, 70, 45, 0.1)])]
  got = cc3d.connected_components_batch(vols, connectivity=conn, binary_image=True, return_N=True,
                                        out_dtype=np.uint32)
  for i, v in enumerate(vols):
    exp = cc3d.connected_components(v, connectivity=conn, binary_image=True, return_N=True, out_dtype=np.uint32)
    assert _same(got[i], exp), f"volume {i}"


def test_batch_unbatchable_volumes_fall_back_to_single(gpu):
  import cc3d
  rng = np.random.default_rng(3)
  vols = [rng.integers(0, 3, size=(30, 40)).astype(np.uint8), rng.integers(0, 3, size=(20, 21, 22)).astype(np.uint8)]
  got = cc3d.connected_components_batch(vols, connectivity=[8, 6], binary_image=True, return_N=True)
  for v, c, g in zip(vols, [8, 6], got):
    assert _same(g, cc3d.connected_components(v, connectivity=c, binary_image=True, return_N=True))
  got = cc3d.connected_components_batch(vols, connectivity=[4, 6], periodic_boundary=True, return_N=True)
  for v, c, g in zip(vols, [4, 6], got):
    assert _same(g, cc3d.connected_components(v, connectivity=c, periodic_boundary=True, return_N=True))


def test_batch_raises_like_single(gpu):
  import cc3d
  big = np.arange(1, 70001, dtype=np.uint32).reshape((70, 1000))
  small = np.ones((10, 10, 10), dtype=np.uint32)
  with pytest.raises(ValueError) as e_single:
    cc3d.connected_components(big, out_dtype=np.uint16)
  with pytest.raises(ValueError) as e_batch:
    cc3d.connected_components_batch([small, big], out_dtype=np.uint16)
  assert str(e_batch.value) == str(e_single.value)


def test_voronoi_batch_uneven(gpu):
  """The batched workload's recipe at a reduced size: sides from {64, 96, 128},
  one seed per 8,000 voxels, membranes, connectivity 6 or 26 per volume."""
  import cc3d
  import synth
  rng = np.random.default_rng(96)
  vols, conns = [], []
  for i in range(16):
    shape = tuple(int(s) for s in rng.choice([64, 96, 128], size=3))
    vols.append(np.asfortranarray(synth.voronoi(shape, max(1, int(np.prod(shape)) // 8000), seed=1000 + i)))
    conns.append(int(rng.choice([6, 26])))
  got = cc3d.connected_components_batch(vols, connectivity=conns, out_dtype=np.uint32, return_N=True)
  for i, (v, c) in enumerate(zip(vols, conns)):
    assert _same(got[i], cc3d.connected_components(v, connectivity=c, out_dtype=np.uint32, return_N=True)), i


def test_label_batch_graph_replays(gpu):
  """Graph replays equal eager runs, with eager GPU work between replays, and
  follow the current input contents."""
  import torch
  import cc3d
  import synth
  shapes = [(64, 64, 64), (96, 64, 32), (40, 130, 17)]
  vols = [torch.from_numpy(np.ravel(synth.voronoi(s, 40, seed=i), order="F")).cuda().as_strided(
          s, (1, s[0], s[0] * s[1])) for i, s in enumerate(shapes)]
  lb = cc3d.LabelBatch(vols, connectivity=[26, 6, 18], out_dtype=np.uint32)
  eager = [(o.clone(), N) for o, N in lb.run()]
  for rep in range(4):
    junk = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work between replays
    torch.cuda.synchronize()
    got = lb.replay()
    for k, (o, N) in enumerate(got):
      assert N == eager[k][1] and torch.equal(o, eager[k][0]), f"replay {rep} volume {k}"
    del junk
  # new contents in the same input buffers
  new = synth.voronoi(shapes[0], 300, seed=77)
  vols[0].copy_(torch.from_numpy(np.asfortranarray(new)).cuda())
  torch.cuda.synchronize()
  got = lb.replay()
  exp = cc3d.connected_components(new, connectivity=26, out_dtype=np.uint32, return_N=True)
  assert got[0][1] == exp[1]
  assert np.array_equal(got[0][0].cpu().numpy(), exp[0])
  lb.close()
