"""GPU parity: cc3d (HIP library) against the oracle and the compiled reference.

Every case runs the product through its public Python surface (which calls the
C ABI of libcc3d_amd.so) and the checker through oracle.connected_components
(the same binding-layer restatement over oracle.c, or over the compiled
reference in oracle/_ref when that library is present).  The bar is bit-exact:
same labels, same N, same output dtype, same exception type.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _run(fn, arr, kw):
  try:
    out = fn(arr, return_N=True, **kw)
    return ("ok", out[0], out[1])
  except Exception as e:  # error parity is part of the contract
    return ("err", type(e).__name__, str(e))


def _same(a, b):
  if a[0] != b[0]:
    return False
  if a[0] == "err":
    return a[1] == b[1]
  return a[2] == b[2] and a[1].dtype == b[1].dtype and a[1].shape == b[1].shape and np.array_equal(a[1], b[1])


def _check(cc3d, arr, kw, tag=""):
  print(f"case {tag} shape={arr.shape} dtype={arr.dtype} F={arr.flags.f_contiguous} kw={kw}", flush=True)
  got = _run(cc3d.connected_components, arr, kw)
  exp = _run(lambda a, **k: oracle.connected_components(a, backend="oracle", **k), arr, kw)
  assert _same(got, exp), f"{tag}: product {got[0]} vs oracle {exp[0]}: {got[2] if got[0]=='ok' else got} / {exp[2] if exp[0]=='ok' else exp}"
  if oracle.have_ref():
    ref = _run(lambda a, **k: oracle.connected_components(a, backend="ref", **k), arr, kw)
    assert _same(got, ref), f"{tag}: product vs reference build differ"


def _rand_case(rng, big=False):
  nd = int(rng.choice([2, 3]))
  if nd == 2:
    hi = 200 if big else 70
    shape = tuple(int(s) for s in rng.integers(1, hi, size=2))
    conn = int(rng.choice([4, 8, 6, 18, 26]))
  else:
    hi = 60 if big else 24
    shape = tuple(int(s) for s in rng.integers(1, hi, size=3))
    conn = int(rng.choice([6, 18, 26]))
  dt = rng.choice([np.uint8, np.uint16, np.int32, np.uint32, np.uint64, np.int64, np.float32, np.float64,
                   np.int8, np.int16, bool, np.float16])
  kind = str(rng.choice(["multi", "binary", "cont"]))
  nvals = int(rng.integers(2, 6))
  arr = rng.integers(0, nvals, size=shape)
  arr = (arr > 0) if dt == bool else arr.astype(dt)
  if rng.random() < 0.5:
    arr = np.asfortranarray(arr)
  kw = dict(connectivity=conn)
  if kind == "binary":
    kw["binary_image"] = True
  if kind == "cont" and dt not in (bool, np.float16):
    kw["delta"] = int(rng.integers(1, 3))
  if kind != "cont" and conn in (4, 6, 8) and rng.random() < 0.3:
    kw["periodic_boundary"] = True
  if rng.random() < 0.3:
    kw["out_dtype"] = [np.uint16, np.uint32, np.uint64][int(rng.integers(0, 3))]
  return arr, kw


def test_reference_vectors(gpu):
  """Every committed reference vector (tests/golden/ref_vectors.npz: outputs of the
  reference compiled from its own sources) through the product, bit for bit."""
  import os
  import sys
  import cc3d
  sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
  import make_golden
  cases = make_golden.load()
  assert len(cases) > 600
  for i, (arr, kw, out, N, err) in enumerate(cases):
    print(f"vector {i} shape={arr.shape} dtype={arr.dtype} kw={kw}", flush=True)
    got = _run(cc3d.connected_components, arr, kw)
    if err:
      assert got[0] == "err" and got[1] == err, f"vector {i}: reference raised {err}, product {got[:2]}"
      continue
    assert got[0] == "ok", f"vector {i}: product raised {got[1]}: {got[2]}"
    assert got[2] == N, f"vector {i}: N {got[2]} vs {N}"
    assert got[1].dtype == out.dtype and got[1].shape == out.shape, f"vector {i}: dtype/shape"
    assert np.array_equal(got[1], out), f"vector {i}: labels differ"


def test_random_small(gpu):
  import cc3d
  rng = np.random.default_rng(12345)
  for i in range(400):
    arr, kw = _rand_case(rng)
    _check(cc3d, arr, kw, f"small#{i}")


def test_random_tile_crossing(gpu):
  """Shapes that span several 32x8x8 (3D) / 64x32 (2D) tiles with partial edges."""
  import cc3d
  rng = np.random.default_rng(777)
  for i in range(80):
    arr, kw = _rand_case(rng, big=True)
    _check(cc3d, arr, kw, f"big#{i}")


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64, np.float32])
def test_queue_overflow_6conn(gpu, dtype):
  """Noise with three labels: most voxels start a run, so a tile queues more
  voxels than the work queue holds and the rest are unioned after it."""
  import cc3d
  rng = np.random.default_rng(66)
  arr = rng.integers(0, 3, size=(96, 40, 24)).astype(dtype)
  _check(cc3d, np.asfortranarray(arr), dict(connectivity=6), f"overflow6 {np.dtype(dtype).name}")


@pytest.mark.parametrize("conn", [26, 18, 6])
@pytest.mark.parametrize("order", ["C", "F"])
def test_voronoi_multilabel(gpu, conn, order):
  import cc3d
  import synth
  vol = synth.voronoi((97, 83, 71), 400, seed=5)
  if order == "C":
    vol = np.ascontiguousarray(vol)
  _check(cc3d, vol, dict(connectivity=conn, out_dtype=np.uint32), "voronoi")


@pytest.mark.parametrize("p", [0.1, 0.3, 0.5])
@pytest.mark.parametrize("conn", [6, 26])
def test_bernoulli_binary(gpu, p, conn):
  import cc3d
  import synth
  vol = synth.bernoulli((130, 70, 50), p, seed=int(p * 10) + conn)
  _check(cc3d, vol, dict(connectivity=conn, binary_image=True, out_dtype=np.uint32), "bernoulli")


def test_bernoulli_2d8(gpu):
  import cc3d
  import synth
  for shape in [(512, 512), (333, 517), (1, 129), (129, 1)]:
    vol = synth.bernoulli(shape, 0.5, seed=9)
    _check(cc3d, vol, dict(connectivity=8, binary_image=True, out_dtype=np.uint32), "bern2d8")


def test_tones_continuous(gpu):
  import cc3d
  import synth
  vol = synth.tones((90, 70, 40), 150, seed=4)
  _check(cc3d, vol, dict(connectivity=26, delta=20, out_dtype=np.uint32), "tones")


def test_config0_voronoi_256(gpu):
  """256^3 uint32 Voronoi, 1500 seeds, membranes 0.8, salt 2 %, 26-conn, uint32, return_N."""
  import cc3d
  import synth
  vol = synth.voronoi((256, 256, 256), 1500, seed=2024)
  _check(cc3d, vol, dict(connectivity=26, out_dtype=np.uint32), "config0")


def test_device_tensor_in_out(gpu):
  import torch
  import cc3d
  import synth
  vol = synth.voronoi((64, 48, 40), 100, seed=8)
  exp, expN = oracle.connected_components(vol, connectivity=18, return_N=True)
  flat = torch.from_numpy(np.ravel(vol, order="F")).cuda()
  t = flat.as_strided(vol.shape, (1, vol.shape[0], vol.shape[0] * vol.shape[1]))
  out, N = cc3d.connected_components(t, connectivity=18, return_N=True)
  assert out.device.type == "cuda"
  assert N == expN
  assert np.array_equal(out.cpu().numpy(), exp)
  tc = torch.from_numpy(np.ascontiguousarray(vol)).cuda()
  out2 = cc3d.connected_components(tc, connectivity=18)
  exp2 = oracle.connected_components(np.ascontiguousarray(vol), connectivity=18)
  assert np.array_equal(out2.cpu().numpy(), exp2)


@pytest.mark.parametrize("connectivity", [6, 18, 26])
def test_single_row_special_case(gpu, connectivity):
  """automated_test.py:453-483 with a fixed row: one foreground row is labelled
  as runs of equal values, out dtype uint16."""
  import cc3d
  sx, sy, sz = 256, 257, 252
  y, z = 131, 77
  img = np.zeros((sx, sy, sz), dtype=np.uint8, order="F")
  img[:, y, z] = 6
  out = cc3d.connected_components(img, connectivity=connectivity)
  epl, start, end = cc3d.estimate_provisional_labels(img)
  assert (epl, start, end) == (1, y + sy * z, y + sy * z)
  assert out.dtype == np.uint16
  gt = np.zeros(img.shape, dtype=np.uint8, order="F")
  gt[:, y, z] = 1
  assert np.array_equal(out, gt)
  img[:100, y, z] = 3
  gt[100:, y, z] = 2
  epl, start, end = cc3d.estimate_provisional_labels(img)
  out = cc3d.connected_components(img, connectivity=connectivity)
  assert epl == 2 and start == end
  assert np.array_equal(out, gt)


def test_max_labels_is_ignored(gpu):
  """automated_test.py:490-497: max_labels does not change the result."""
  import cc3d
  img = np.arange(0, 64 ** 3).astype(np.uint64).reshape((64, 64, 64))
  real = cc3d.connected_components(img, max_labels=64 * 64 * 64)
  assert np.array_equal(real, cc3d.connected_components(img, max_labels=0))
  assert np.array_equal(real, cc3d.connected_components(img, max_labels=-50))
  assert np.array_equal(real.ravel(), np.arange(64 ** 3))


def test_512_cube_all_distinct(gpu):
  """automated_test.py:486-488 (every voxel its own component) with the result
  checked: voxel i of a C-order arange gets label i."""
  import cc3d
  img = np.arange(0, 512 ** 3, dtype=np.uint64).reshape((512, 512, 512))
  out, N = cc3d.connected_components(img, return_N=True)
  assert N == 512 ** 3 - 1
  assert out.dtype == np.uint32
  assert np.array_equal(out.ravel(), np.arange(512 ** 3, dtype=np.uint32))


@pytest.mark.parametrize("shape,seed", [((128, 256, 384), 30001), ((192, 288, 128), 30030), ((64, 384, 384), 30080)])
def test_voronoi_6conn_large_repeatable(gpu, shape, seed):
  """Regression: large 6-connected Voronoi volumes (deep global forests) must
  match the reference on every one of several runs, as a numpy input and as
  a device tensor."""
  import torch
  import cc3d
  import synth
  vol = synth.voronoi(shape, max(1, int(np.prod(shape)) // 8000), seed=seed)
  backend = "ref" if oracle.have_ref() else "oracle"
  exp, expN = oracle.connected_components(vol, connectivity=6, out_dtype=np.uint32, return_N=True, backend=backend)
  dvol = torch.from_numpy(np.ravel(vol, order="F")).cuda().as_strided(shape, (1, shape[0], shape[0] * shape[1]))
  for run in range(3):
    print(f"run {run} shape={shape} seed={seed}", flush=True)
    got, N = cc3d.connected_components(vol, connectivity=6, out_dtype=np.uint32, return_N=True)
    assert N == expN and np.array_equal(got, exp), f"numpy input, run {run}"
    got, N = cc3d.connected_components(dvol, connectivity=6, out_dtype=np.uint32, return_N=True)
    assert N == expN and np.array_equal(got.cpu().numpy(), exp), f"device input, run {run}"


@pytest.mark.parametrize("shape", [(5000, 5000, 3), (5504, 5504, 3)])
def test_large_plane_staging(gpu, shape):
  """uint64 volumes whose tile box spans about 1.8 GB (32-bit buffer offsets
  near their limit) and about 2.2 GB (past it: the 64-bit address path)."""
  import cc3d
  import synth
  vol = synth.voronoi(shape, 3000, seed=55, id_range=(1, None), dtype=np.uint64)
  for conn in (26, 6):
    _check(cc3d, vol, dict(connectivity=conn, out_dtype=np.uint32), f"largeplane{shape}")
