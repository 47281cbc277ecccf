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


def _tones(rng, shape, dtype):
  """Smooth ramps plus noise: continuous-valued volumes whose delta links
  both join and split regions."""
  idx = np.indices(shape).sum(axis=0).astype(np.float64)
  v = np.floor(idx * 3.0 + rng.normal(0, 4, size=shape)) % 97
  v[rng.random(shape) < 0.1] = 0
  return v.astype(dtype)


@pytest.mark.parametrize("dtype,delta", [(np.float32, 4.5), (np.uint8, 3), (np.uint16, 7)])
def test_batch_continuous_matches_single(gpu, dtype, delta):
  import cc3d
  rng = np.random.default_rng(11)
  vols, conns = [], []
  for shape, conn in [((40, 33, 17), 26), ((64, 64, 20), 18), ((31, 50, 9), 6), ((70, 90), 4), ((55, 43), 6),
                      ((80, 61), 26), ((33, 47), 18), ((66, 20), 8), ((1, 1, 5), 26), ((129, 3, 2), 26)]:
    v = _tones(rng, shape, dtype)
    vols.append(np.asfortranarray(v) if rng.random() < 0.5 else v)
    conns.append(conn)
  got = cc3d.connected_components_batch(vols, connectivity=conns, delta=delta, return_N=True)
  for i, (v, c) in enumerate(zip(vols, conns)):
    exp = cc3d.connected_components(v, connectivity=c, delta=delta, return_N=True)
    assert _same(got[i], exp), f"volume {i} shape={v.shape} conn={c}"


@pytest.mark.parametrize("binary", [False, True])
def test_batch_periodic_matches_single(gpu, binary):
  import cc3d
  rng = np.random.default_rng(12)
  vols, conns = [], []
  for shape, conn in [((20, 21, 22), 6), ((64, 40, 33), 6), ((30, 40), 4), ((50, 37), 8), ((33, 28), 6),
                      ((1, 17, 9), 6), ((45, 1, 1), 6)]:
    vols.append(rng.integers(0, 3, size=shape).astype(np.uint16))
    conns.append(conn)
  row = np.zeros((12, 9), dtype=np.uint16)  # one foreground row: the single-row shortcut ignores periodicity
  row[4, :] = [1, 1, 0, 2, 2, 0, 0, 1, 1]
  vols.append(row)
  conns.append(4)
  got = cc3d.connected_components_batch(vols, connectivity=conns, periodic_boundary=True, binary_image=binary,
                                        return_N=True)
  for i, (v, c) in enumerate(zip(vols, conns)):
    exp = cc3d.connected_components(v, connectivity=c, periodic_boundary=True, binary_image=binary, return_N=True)
    assert _same(got[i], exp), f"volume {i} shape={v.shape} conn={c}"


def test_batch_binary_planes_matches_single(gpu):
  import cc3d
  import synth
  vols = [synth.bernoulli((70, 90, 1), 0.45, seed=1)[:, :, 0], synth.bernoulli((130, 33, 1), 0.3, seed=2)[:, :, 0],
          synth.bernoulli((40, 40, 1), 0.6, seed=3)[:, :, 0], synth.bernoulli((20, 30, 40), 0.4, seed=4)]
  conns = [26, 18, 6, 26]
  got = cc3d.connected_components_batch(vols, connectivity=conns, binary_image=True, return_N=True)
  for i, (v, c) in enumerate(zip(vols, conns)):
    exp = cc3d.connected_components(v, connectivity=c, binary_image=True, return_N=True)
    assert _same(got[i], exp), f"volume {i} shape={v.shape} conn={c}"


def _job_stats(v, sx, sy, sz, conn, code, mode, delta_b, periodic):
  """Stage-1 counters of the single-volume entry for one device volume."""
  import ctypes
  from cc3d import _lib as L
  lib = L.lib()
  flags = L.FLAG_CONTINUOUS if mode == L.BATCH_CONTINUOUS else 0
  dbuf = ctypes.create_string_buffer(delta_b, 8)
  job = ctypes.c_void_p(0)
  L.check(lib.cc3d_amd_label_begin(ctypes.byref(job), ctypes.c_void_p(v.data_ptr()), code, sx, sy, sz, conn,
                                   ctypes.cast(dbuf, ctypes.c_void_p), flags, int(periodic),
                                   int(mode == L.BATCH_BINARY), None, 0, None))
  st = L.Stats()
  try:
    L.check(lib.cc3d_amd_job_stats(job, ctypes.byref(st)))
  finally:
    lib.cc3d_amd_job_free(job)
  return st


@pytest.mark.parametrize("mode,periodic", [(0, True), (1, True), (1, False), (2, False), (0, False)])
def test_batch_stats_match_single_entry(gpu, mode, periodic):
  """The batch's stage-1 counters (N, provisional labels, epl, foreground
  rows) equal the single-volume entry's for every volume, through the C ABI."""
  import ctypes
  import torch
  from cc3d import _lib as L
  rng = np.random.default_rng(20 + mode)
  # device volumes stored x fastest: torch shape (sz, sy, sx) of a C-order array
  shapes = [((30, 41, 17), 6), ((64, 33, 1), 4), ((50, 60, 1), 8), ((20, 20, 20), 6), ((97, 1, 1), 6),
            ((51, 37, 1), 8), ((70, 1, 1), 8)]
  if not periodic:
    shapes += [((40, 30, 20), 26), ((40, 30, 20), 18), ((77, 45, 1), 26)]
  dt = np.float32 if mode == L.BATCH_CONTINUOUS else np.uint8
  code = L.DT_F32 if dt == np.float32 else L.DT_U8
  delta_b = np.array([3.0], dtype=np.float32).tobytes() if mode == L.BATCH_CONTINUOUS else b""
  vols = []
  for (sx, sy, sz), _ in shapes:
    a = _tones(rng, (sz, sy, sx), dt) if mode == L.BATCH_CONTINUOUS else rng.integers(0, 3, size=(sz, sy, sx)).astype(dt)
    vols.append(torch.from_numpy(a).cuda())
  vt = (L.Volume * len(vols))()
  for k, (((sx, sy, sz), conn), v) in enumerate(zip(shapes, vols)):
    vt[k].inp = v.data_ptr()
    vt[k].sx, vt[k].sy, vt[k].sz = sx, sy, sz
    vt[k].connectivity = conn | (L.VOLUME_PERIODIC if periodic else 0)
  lib = L.lib()
  torch.cuda.synchronize()
  b = ctypes.c_void_p(0)
  L.check(lib.cc3d_amd_batch_create(ctypes.byref(b), vt, len(vols), code, mode, None, 0, None))
  try:
    if mode == L.BATCH_CONTINUOUS:
      dbuf = ctypes.create_string_buffer(delta_b, 8)
      L.check(lib.cc3d_amd_batch_set_delta(b, ctypes.cast(dbuf, ctypes.c_void_p)))
    L.check(lib.cc3d_amd_batch_begin(b))
    st = (L.Stats * len(vols))()
    L.check(lib.cc3d_amd_batch_stats(b, st))
  finally:
    lib.cc3d_amd_batch_free(b)
  for k, (((sx, sy, sz), conn), v) in enumerate(zip(shapes, vols)):
    one = _job_stats(v, sx, sy, sz, conn, code, mode, delta_b.ljust(8, b"\0"), periodic)
    got = (st[k].N, st[k].provisional, st[k].epl, st[k].first_foreground_row, st[k].last_foreground_row)
    exp = (one.N, one.provisional, one.epl, one.first_foreground_row, one.last_foreground_row)
    assert got == exp, f"volume {k} shape={(sx, sy, sz)} conn={conn}: batch {got} single {exp}"


def test_batch_continuous_needs_delta(gpu):
  import ctypes
  import torch
  from cc3d import _lib as L
  v = torch.zeros(8 * 8 * 8, dtype=torch.float32, device="cuda")
  vt = (L.Volume * 1)()
  vt[0].inp = v.data_ptr()
  vt[0].sx = vt[0].sy = vt[0].sz = 8
  vt[0].connectivity = 26
  lib = L.lib()
  b = ctypes.c_void_p(0)
  L.check(lib.cc3d_amd_batch_create(ctypes.byref(b), vt, 1, L.DT_F32, L.BATCH_CONTINUOUS, None, 0, None))
  try:
    assert lib.cc3d_amd_batch_begin(b) == L.ERR_INVALID
    assert b"delta" in lib.cc3d_amd_last_error()
  finally:
    lib.cc3d_amd_batch_free(b)
  vt[0].connectivity = 26 | L.VOLUME_PERIODIC
  assert lib.cc3d_amd_batch_workspace_bytes(vt, 1, L.BATCH_CONTINUOUS) == 0


def test_label_batch_continuous_graph_replays(gpu):
  import torch
  import cc3d
  rng = np.random.default_rng(5)
  arrs = [_tones(rng, s, np.float32) for s in [(48, 40, 30), (64, 64, 1), (20, 70, 33)]]
  vols = [torch.from_numpy(a).cuda() for a in arrs]
  lb = cc3d.LabelBatch(vols, connectivity=[26, 6, 18], out_dtype=np.uint32, delta=2.5)
  eager = [(o.clone(), N) for o, N in lb.run()]
  for k, (a, c) in enumerate(zip(arrs, [26, 6, 18])):
    exp = cc3d.connected_components(a, connectivity=c, delta=2.5, out_dtype=np.uint32, return_N=True)
    assert eager[k][1] == exp[1] and np.array_equal(eager[k][0].cpu().numpy(), exp[0]), k
  for rep in range(3):
    junk = torch.randn(1 << 20, device="cuda").sum()
    torch.cuda.synchronize()
    got = lb.replay()
    for k, (o, N) in enumerate(got):
      assert N == eager[k][1] and torch.equal(o, eager[k][0]), f"replay {rep} volume {k}"
    del junk
  lb.close()


@pytest.mark.parametrize("binary,conn", [(True, 8), (True, 4), (False, 8)])
def test_batch_2d_images_match_single(gpu, binary, conn):
  """2D images with 2D connectivities: binary 2D-8 (block numbering and, for
  odd sides, the reference's stale background labels), binary 2D-4, and
  continuous 2D-8 (global min/max rule) batched on the 2D tiles."""
  import cc3d
  import synth
  rng = np.random.default_rng(40 + conn + binary)
  vols = []
  for k, shape in enumerate([(130, 77), (64, 64), (33, 201), (1, 57), (90, 1), (255, 129), (8, 8)]):
    if binary:
      v = synth.bernoulli(shape + (1,), 0.45, seed=k)[:, :, 0]
    else:
      v = _tones(rng, shape, np.float32)
    vols.append(np.asfortranarray(v) if k % 2 else v)
  kw = dict(binary_image=True) if binary else dict(delta=6.5)
  got = cc3d.connected_components_batch(vols, connectivity=conn, return_N=True, **kw)
  for i, v in enumerate(vols):
    exp = cc3d.connected_components(v, connectivity=conn, return_N=True, **kw)
    assert _same(got[i], exp), f"volume {i} shape={v.shape}"
  if binary:  # periodic planes on the 2D tiles as well
    got = cc3d.connected_components_batch(vols, connectivity=conn, return_N=True, periodic_boundary=True, **kw)
    for i, v in enumerate(vols):
      exp = cc3d.connected_components(v, connectivity=conn, return_N=True, periodic_boundary=True, **kw)
      assert _same(got[i], exp), f"periodic volume {i} shape={v.shape}"
