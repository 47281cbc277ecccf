"""Parity at the full sizes of the benchmark workloads.

Each case labels the workload's own input on the GPU and compares labels, N
and output dtype bit for bit with the reference compiled from its own
sources (oracle/_ref), or with the oracle where that build is absent:

  * 512^3 uint8 Bernoulli p in {0.1, 0.3, 0.5}, binary_image, 6 and 26;
  * 16384^2 uint8 Bernoulli p = 0.5, binary_image, 8-connected (the 2x2
    block numbering of the binary 2D path);
  * 512^3 uint64 Voronoi, 40,000 seeds, 64-bit IDs, 6, 18 and 26;
  * 512^3 float32 three-tone image, 5,000 seeds, continuous delta = 20, 26.

  * 512^3 uint32 Voronoi, 12,000 seeds, 26-connected: the benchmark's
    headline input;
  * the batched workload: 96 uint32 Voronoi volumes with sides drawn from
    {64, 96, ..., 384} and connectivity from {6, 26}, labelled as one batch
    (eager launches and a replayed HIP graph), each volume against the
    reference;
  * the served workload's large request: a 1024x1024x512 uint32 Voronoi
    volume with 200,000 seeds, 6-connected, served in a mix with small
    requests; labels against the reference, voxel counts and bounding boxes
    against the oracle's statistics, the dusted copy and dust_N against the
    dust rule applied to those counts.

(The 256^3 uint32 Voronoi case is in test_gpu_parity.py.)
"""
import concurrent.futures
import os
import functools

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _cmp(arr, tag, **kw):
  import cc3d
  backend = "ref" if oracle.have_ref() else "oracle"
  print(f"config {tag} shape={arr.shape} dtype={arr.dtype} kw={kw} checker={backend}", flush=True)
  got, N = cc3d.connected_components(arr, return_N=True, **kw)
  exp, eN = oracle.connected_components(arr, return_N=True, backend=backend, **kw)
  assert N == eN, f"{tag}: N {N} vs {eN}"
  assert got.dtype == exp.dtype and got.shape == exp.shape, tag
  assert np.array_equal(got, exp), f"{tag}: labels differ"


@functools.lru_cache(maxsize=1)
def _voronoi_u64():
  import synth
  return synth.voronoi((512, 512, 512), 40000, seed=64, id_range=(1, None), dtype=np.uint64)


@pytest.mark.parametrize("p", [0.1, 0.3, 0.5])
@pytest.mark.parametrize("conn", [6, 26])
def test_binary_bernoulli_512(gpu, p, conn):
  import synth
  vol = synth.bernoulli((512, 512, 512), p, seed=int(p * 10) * 100 + conn)
  _cmp(vol, f"binary p={p} conn={conn}", connectivity=conn, binary_image=True, out_dtype=np.uint32)


def test_binary_bernoulli_2d_16384(gpu):
  import synth
  vol = synth.bernoulli((16384, 16384), 0.5, seed=16384)
  _cmp(vol, "binary 2D p=0.5", connectivity=8, binary_image=True, out_dtype=np.uint32)


@pytest.mark.parametrize("conn", [6, 18, 26])
def test_uint64_voronoi_512(gpu, conn):
  _cmp(_voronoi_u64(), f"uint64 voronoi conn={conn}", connectivity=conn, out_dtype=np.uint32)


def test_float32_tones_512(gpu):
  import synth
  vol = synth.tones((512, 512, 512), 5000, seed=32)
  _cmp(vol, "float32 tones delta=20", connectivity=26, delta=20, out_dtype=np.uint32)


def test_uint32_voronoi_512_headline(gpu):
  import synth
  vol = synth.voronoi((512, 512, 512), 12000, seed=512)
  _cmp(vol, "uint32 voronoi 512^3 headline", connectivity=26, out_dtype=np.uint32)


def _ref_labels(vols, conns):
  """The reference (or the oracle) on every volume, one host thread each."""
  backend = "ref" if oracle.have_ref() else "oracle"
  workers = max(1, min(16, len(os.sched_getaffinity(0))))
  with concurrent.futures.ThreadPoolExecutor(workers) as ex:
    return list(ex.map(lambda vc: oracle.connected_components(vc[0], connectivity=vc[1], return_N=True,
                                                              out_dtype=np.uint32, backend=backend),
                       zip(vols, conns)))


def test_batched_96_volumes(gpu):
  """The batched workload at its full size (the recipe of bench.py's batched
  leg), eager and as a replayed graph, each volume against the reference."""
  import torch
  import cc3d
  import synth
  rng = np.random.default_rng(96)
  sides = np.arange(64, 385, 32)
  vols, conns = [], []
  for i in range(96):
    shape = tuple(int(s) for s in rng.choice(sides, size=3))
    vols.append(synth.voronoi(shape, max(1, int(np.prod(shape)) // 8000), seed=30000 + i))
    conns.append(int(rng.choice([6, 26])))
  print(f"batched workload: {len(vols)} volumes, {sum(v.size for v in vols)} voxels", flush=True)
  exp = _ref_labels(vols, conns)
  print("batched workload: reference done", flush=True)
  dvols = [torch.from_numpy(np.ravel(v, order="F")).to(gpu).as_strided(v.shape, (1, v.shape[0], v.shape[0] * v.shape[1]))
           for v in vols]
  got = cc3d.connected_components_batch(dvols, connectivity=conns, out_dtype=np.uint32, return_N=True)
  for k, ((g, n), (e, en)) in enumerate(zip(got, exp)):
    assert n == en, f"volume {k} {vols[k].shape} conn={conns[k]}: N {n} vs {en}"
    assert np.array_equal(g.cpu().numpy(), e), f"volume {k} {vols[k].shape} conn={conns[k]}: labels differ"
  print("batched workload: eager batch matches", flush=True)
  lb = cc3d.LabelBatch(dvols, connectivity=conns, out_dtype=np.uint32)
  lb.capture()
  for rep in range(2):
    res = lb.replay()
    torch.cuda.synchronize()
    for k, ((g, n), (e, en)) in enumerate(zip(res, exp)):
      assert n == en and np.array_equal(g.cpu().numpy(), e), f"graph replay {rep} volume {k}"
  lb.close()


def test_served_large_request(gpu):
  """The served workload's large request (1024x1024x512, 200,000 seeds,
  6-connected) in a mix with small 26-connected requests."""
  import torch
  import cc3d
  import synth
  rng = np.random.default_rng(41)
  small = []
  for i in range(6):
    shape = tuple(int(s) for s in rng.integers(128, 257, size=3))
    small.append(synth.voronoi(shape, max(1, int(np.prod(shape)) // 8000), seed=40000 + i))
  large = synth.voronoi((1024, 1024, 512), 200000, seed=41000)
  vols = small[:3] + [large] + small[3:]
  conns = [26, 26, 26, 6, 26, 26, 26]
  print("served: volumes generated", flush=True)
  exp = _ref_labels(vols, conns)
  print("served: reference done", flush=True)
  srv = cc3d.LabelServer(workers=4)
  try:
    reqs = [srv.submit(torch.from_numpy(np.ravel(v, order="F")).to(gpu).as_strided(
              v.shape, (1, v.shape[0], v.shape[0] * v.shape[1])), c) for v, c in zip(vols, conns)]
    for r in reqs:
      assert r.done.wait(300), "request did not finish"
  finally:
    srv.close()
  for k, (r, v, (e, en)) in enumerate(zip(reqs, vols, exp)):
    assert r.error is None, f"request {k}: {r.error!r}"
    res = r.result
    tag = f"request {k} {v.shape} conn={conns[k]}"
    lab = res["labels"].cpu().numpy()
    assert res["N"] == en, f"{tag}: N {res['N']} vs {en}"
    assert np.array_equal(lab, e), f"{tag}: labels differ"
    counts, bbox, _ = oracle.statistics_raw(e, en)
    assert np.array_equal(res["voxel_counts"].cpu().numpy().view(np.uint32), counts), f"{tag}: counts"
    assert np.array_equal(res["bounding_boxes"].cpu().numpy().view(np.uint32), bbox), f"{tag}: bounding boxes"
    small_lab = counts < 100
    small_lab[0] = False
    dusted = np.where(small_lab[e], 0, v)
    assert np.array_equal(res["dusted"].cpu().numpy(), dusted), f"{tag}: dusted copy"
    assert res["dust_N"] == en - int(small_lab[1:].sum()), f"{tag}: dust_N"
    print(f"served: {tag} ok (N={en})", flush=True)
