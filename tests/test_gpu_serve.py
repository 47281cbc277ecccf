"""The labelling service on the GPU.

Every request's result (labels, N, voxel counts, bounding boxes, dusted
image, dust N) must equal the library calls on that volume alone
(connected_components, statistics, dust) bit for bit, whether the request
ran alone or in a mix with others (batched, overlapping on the device).
Admission must keep the reserved footprint within the HBM budget.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(v):
  import torch
  sx, sy, sz = v.shape
  return torch.from_numpy(np.ravel(v, order="F")).cuda().as_strided(v.shape, (1, sx, sx * sy))


def _requests(n=10, seed=4):
  import synth
  rng = np.random.default_rng(seed)
  vols, conns = [], []
  for i in range(n):
    shape = tuple(int(s) for s in rng.integers(48, 100, size=3))
    vols.append(synth.voronoi(shape, max(1, int(np.prod(shape)) // 8000), seed=500 + i))
    conns.append(26 if i % 5 else 6)
  return vols, conns


def _host(res):
  return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in res.items()}


def _equal(a, b):
  for k in ("labels", "voxel_counts", "bounding_boxes", "dusted"):
    if not np.array_equal(a[k], b[k]):
      return k
  for k in ("N", "dust_N"):
    if a[k] != b[k]:
      return k
  return None


def test_serve_matches_library(gpu):
  import cc3d
  vols, conns = _requests(6)
  srv = cc3d.LabelServer(workers=2)
  reqs = [srv.submit(_dev(v), c) for v, c in zip(vols, conns)]
  for r in reqs:
    r.done.wait(120)
    assert r.error is None, r.error
  srv.close()
  for v, c, r in zip(vols, conns, reqs):
    res = _host(r.result)
    lab, N = cc3d.connected_components(v, connectivity=c, out_dtype=np.uint32, return_N=True)
    assert res["N"] == N and np.array_equal(res["labels"], lab)
    st = cc3d.statistics(lab, no_slice_conversion=True)
    assert np.array_equal(res["voxel_counts"].view(np.uint32), st["voxel_counts"])
    assert np.array_equal(res["bounding_boxes"].view(np.uint32).astype(np.int64),
                          st["bounding_boxes"].astype(np.int64))
    dusted, dN = cc3d.dust(v, threshold=100, connectivity=c, return_N=True)
    assert res["dust_N"] == dN
    assert np.array_equal(res["dusted"], dusted)


def test_serve_same_alone_and_in_mix(gpu):
  """Each result is identical (exact; no tolerance) alone and in a mix."""
  import cc3d
  vols, conns = _requests(12, seed=9)
  alone = []
  srv = cc3d.LabelServer(workers=1, max_batch=1)
  for v, c in zip(vols, conns):
    r = srv.submit(_dev(v), c)
    r.done.wait(120)
    assert r.error is None, r.error
    alone.append(_host(r.result))
  srv.close()
  srv = cc3d.LabelServer(workers=4, max_batch=8)
  reqs = [srv.submit(_dev(v), c) for v, c in zip(vols, conns)]
  for r in reqs:
    r.done.wait(120)
    assert r.error is None, r.error
  srv.close()
  assert max(r.batched_with for r in reqs) > 1  # the mix really batched requests
  for i, r in enumerate(reqs):
    k = _equal(_host(r.result), alone[i])
    assert k is None, f"request {i}: {k} differs between alone and mix"


def test_admission_respects_budget(gpu):
  import cc3d
  from cc3d._serve import request_footprint
  vols, conns = _requests(5, seed=2)
  dv = [_dev(v) for v in vols]
  fp = [request_footprint(d, c) for d, c in zip(dv, conns)]
  srv = cc3d.LabelServer(hbm_budget=max(fp), workers=3, max_batch=4)
  reqs = [srv.submit(d, c) for d, c in zip(dv, conns)]
  for r in reqs:
    r.done.wait(120)
    assert r.error is None, r.error
  srv.close()
  # at most one request's admission footprint fits at a time; once a request
  # has swapped its workspace for its (smaller) per-label outputs, the next
  # may start, but what is reserved never exceeds the budget
  assert srv.peak_reserved <= max(fp)
  spans = sorted((r.t_start, r.t_done) for r in reqs)
  assert all(b0 >= a0 for (a0, _), (b0, _) in zip(spans, spans[1:]))
  too_big = cc3d.LabelServer(hbm_budget=min(fp) - 1, workers=1)
  r = too_big.submit(dv[0], conns[0])
  assert isinstance(r.error, MemoryError)
  too_big.close()
