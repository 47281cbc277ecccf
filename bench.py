#!/usr/bin/env python3
"""Headline benchmark: 512^3 uint32 26-connected multilabel labelling on one MI355X.

Workload (the labelling path's headline configuration): a 512x512x512 uint32
Voronoi volume with 12,000 uniform seeds (the seed density of the 256^3,
1,500-seed case), seed IDs uniform in [1, 2^31), membranes (a voxel with a
different-region 26-neighbour is zeroed with probability 0.8) and 2 % salt
(voxels overwritten by the ID of a random seed); x fastest; 26-connectivity,
out_dtype uint32, return_N.  Generated deterministically (synth/synth.c).

One step = one cc3d.connected_components call on the volume already resident
in HBM (a device tensor), returning a device uint32 label volume and N.

Prints one JSON line.  Also reports:
  roofline      the dominant kernel's algorithmic bytes / its mean duration,
                measured with HIP events on the launch stream during the timed
                steps; traffic from the committed rocprofv3 counter summary
  kernels       every kernel's mean duration, from one extra untimed pass with
                events around each launch
  path_roofline the whole call: 8 bytes/voxel (4 in + 4 out) / step time,
                beside the published peak and the bandwidth a plain device
                copy of the same volume achieves on this GPU (copy_gbs)
  cpu_baseline  the reference implementation (oracle/_ref, compiled from the
                reference sources) timed on one host core on the same volume,
                whose output is also compared bit-for-bit with the GPU's
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "connected-components-3d_amd"), os.path.join(ROOT, "oracle"), ROOT):
  if p not in sys.path:
    sys.path.insert(0, p)

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "Gvoxels/s 512³ 26-conn CCL (bit-exact); batched/served vs plain; p99 latency"
SHAPE = (512, 512, 512)
NSEEDS = 12000
SEED = 512


def _progress(msg: str) -> None:
  print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _profile_rows(lib):
  names = ctypes.create_string_buffer(64 * 64)
  ms = (ctypes.c_double * 64)()
  cnt = (ctypes.c_int * 64)()
  n = lib.cc3d_amd_profile_collect(names, 64, ms, cnt, 64)
  rows = {}
  for i in range(n):
    nm = names.raw[i * 64:(i + 1) * 64].split(b"\0", 1)[0].decode()
    rows[nm] = (ms[i], cnt[i])
  return rows


def _copy_gbs(src) -> float:
  """Achievable HBM bandwidth: a device-to-device copy of the tensor src
  (read + write bytes / time), the best of 10 timed copies."""
  import torch
  dst = torch.empty_like(src)
  dst.copy_(src)
  best = float("inf")
  for _ in range(10):
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    dst.copy_(src)
    b.record()
    b.synchronize()
    best = min(best, a.elapsed_time(b) * 1e-3)
  del dst
  return 2.0 * src.numel() * src.element_size() / best / 1e9


def _traffic_from_profiles(kernel: str):
  """HBM bytes per launch of `kernel` from the committed counter summary, if any."""
  path = os.path.join(ROOT, "profiles", "traffic.json")
  try:
    with open(path) as f:
      d = json.load(f)
    return d.get(kernel, {}).get("bytes_per_launch")
  except Exception:
    return None


def batched_leg(torch, cc3d, synth, dev, reps=3, cpu=True):
  """96 uint32 Voronoi volumes with uneven sides: plain per-volume calls vs one
  batch (two-stage launches, and a captured HIP graph), plus the reference on
  the host cores over the same batch."""
  rng = np.random.default_rng(96)
  sides = np.arange(64, 385, 32)
  vols, conns = [], []
  t0 = time.time()
  for i in range(96):
    shape = tuple(int(s) for s in rng.choice(sides, size=3))
    vols.append(synth.voronoi(shape, max(1, int(np.prod(shape)) // 8000), seed=30000 + i))
    conns.append(int(rng.choice([6, 26])))
  gen_s = time.time() - t0
  V = int(sum(v.size for v in vols))
  _progress(f"batched: generated {len(vols)} volumes, {V} voxels in {gen_s:.1f} s")
  dvols = [torch.from_numpy(np.ravel(v, order="F")).to(dev).as_strided(v.shape, (1, v.shape[0], v.shape[0] * v.shape[1]))
           for v in vols]
  torch.cuda.synchronize()

  def plain():
    return [cc3d.connected_components(d, connectivity=c, out_dtype=np.uint32, return_N=True)
            for d, c in zip(dvols, conns)]

  ref = plain()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(reps):
    plain()
  torch.cuda.synchronize()
  plain_s = (time.perf_counter() - t0) / reps
  _progress(f"batched: plain per-volume {plain_s * 1e3:.1f} ms")

  lb = cc3d.LabelBatch(dvols, connectivity=conns, out_dtype=np.uint32)
  lb.run(deliver=False)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(reps):
    lb.run(deliver=False)
  torch.cuda.synchronize()
  batch_s = (time.perf_counter() - t0) / reps

  lb.capture()
  lb.replay(deliver=False)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(reps):
    Ns = lb.replay(deliver=False)
  torch.cuda.synchronize()
  graph_s = (time.perf_counter() - t0) / reps
  _progress(f"batched: batch {batch_s * 1e3:.1f} ms, graph {graph_s * 1e3:.1f} ms")
  # bit-exact check of the batch outputs against the plain path (both in memory order)
  exact = all(int(N) == int(r[1]) for N, r in zip(Ns, ref))
  exact = exact and all(torch.equal(o, r[0].as_strided((r[0].numel(),), (1,))) for o, r in zip(lb.outs, ref))
  lb.close()

  cpu_res = None
  if cpu:
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    if oracle.have_ref():
      try:
        ncores = len(os.sched_getaffinity(0))
      except AttributeError:
        ncores = os.cpu_count() or 1
      threads = max(1, min(16, ncores))

      def one(i):
        return oracle.connected_components(vols[i], connectivity=conns[i], out_dtype=np.uint32, return_N=True,
                                           backend="ref")[1]

      t0 = time.perf_counter()
      with ThreadPoolExecutor(max_workers=threads) as ex:
        cpuN = list(ex.map(one, range(len(vols))))
      cpu_s = time.perf_counter() - t0
      _progress(f"batched: reference on {threads} host threads {cpu_s:.2f} s")
      cpu_res = {"value": V / cpu_s / 1e9, "unit": "Gvoxels/s", "cores": threads, "kind": "reference",
                 "sample": f"the whole batch, one volume per host thread ({cpu_s:.2f} s; the reference releases "
                           "the GIL, so threads run like one process per core)",
                 "N_equal_gpu": cpuN == [int(r[1]) for r in ref]}
  return {
    "workload": "96 uint32 Voronoi volumes, sides iid from {64,96,...,384}, 1 seed / 8000 voxels, membranes 0.8, "
                "salt 2%, connectivity 6 or 26 per volume, out uint32",
    "volumes": len(vols), "voxels": V,
    "plain_gvox_s": V / plain_s / 1e9, "plain_ms": plain_s * 1e3,
    "batched_gvox_s": V / batch_s / 1e9, "batched_ms": batch_s * 1e3,
    "graph_gvox_s": V / graph_s / 1e9, "graph_ms": graph_s * 1e3,
    "speedup_vs_plain": plain_s / min(batch_s, graph_s),
    "bit_exact_vs_plain": bool(exact),
    "cpu_baseline": cpu_res,
    "generate_s": gen_s,
  }


def _dev_volume(torch, v, dev):
  sx, sy, sz = v.shape
  return torch.from_numpy(np.ravel(v, order="F")).to(dev).as_strided(v.shape, (1, sx, sx * sy))


def served_leg(torch, cc3d, synth, dev, nreq=1000, loads=(0.5, 0.7, 0.85, 0.95, 1.0, 1.1)):
  """Open-loop Poisson arrivals on the labelling service: 90 % small volumes
  (sides U[128, 256], 26-connected), 10 % 1024x1024x512 (200,000 seeds,
  6-connected) -- exactly nreq / 10 large ones per level, in random order;
  each request also gets device-side voxel counts, bounding boxes and dust
  below 100 voxels.  A closed-loop burst gives a first capacity estimate;
  the open-loop levels are fractions of it, and the capacity reported is the
  highest offered rate the service sustains (achieved within 5 % of offered).
  The plain form serves the same requests one at a time without batching.
  Every 100th request of each level (and its first large one) is checked
  against the library called directly: labels, N, counts, boxes, dusted
  image."""
  rng = np.random.default_rng(4)
  t0 = time.time()
  small = []
  for i in range(24):
    shape = tuple(int(s) for s in rng.integers(128, 257, size=3))
    small.append(_dev_volume(torch, synth.voronoi(shape, max(1, int(np.prod(shape)) // 8000), seed=40000 + i), dev))
  _progress("served: small volumes generated")
  large = _dev_volume(torch, synth.voronoi((1024, 1024, 512), 200000, seed=41000), dev)
  gen_s = time.time() - t0
  _progress(f"served: volumes generated in {gen_s:.1f} s")
  torch.cuda.synchronize()

  def draw(n, seed):
    r = np.random.default_rng(seed)
    nl = n // 10
    kinds = np.array([1] * nl + [0] * (n - nl))
    r.shuffle(kinds)
    return [(large, 6) if k else (small[int(r.integers(len(small)))], 26) for k in kinds]

  def burst(srv, reqs):
    t0 = time.perf_counter()
    rs = [srv.submit(v, c, t_arrive=t0, keep=False) for v, c in reqs]
    for r in rs:
      if not r.done.wait(300):
        raise TimeoutError("a served request did not finish within 300 s")
      if r.error is not None:
        raise r.error
      r.result = None  # the caller is done with it
    return rs, time.perf_counter() - t0

  vox = lambda reqs: sum(int(v.numel()) for v, _ in reqs)
  srv = cc3d.LabelServer(workers=4, max_batch=16)
  try:
    return _served_runs(srv, torch, cc3d, draw, burst, vox, nreq, loads, gen_s)
  finally:
    srv.close()


def _served_check(torch, cc3d, vol, conn, res) -> bool:
  """One served result against the library called directly (on a host copy
  of the request's volume)."""
  vh = vol.cpu().numpy()
  lab, N = cc3d.connected_components(vh, connectivity=conn, out_dtype=np.uint32, return_N=True)
  if res["N"] != N or not np.array_equal(res["labels"].cpu().numpy(), lab):
    return False
  st = cc3d.statistics(lab, no_slice_conversion=True)
  if not np.array_equal(res["voxel_counts"].cpu().numpy().view(np.uint32), st["voxel_counts"]):
    return False
  if not np.array_equal(res["bounding_boxes"].cpu().numpy().view(np.uint32).astype(np.int64),
                        st["bounding_boxes"].astype(np.int64)):
    return False
  dusted, dN = cc3d.dust(vh, threshold=100, connectivity=conn, return_N=True)
  return res["dust_N"] == dN and np.array_equal(res["dusted"].cpu().numpy(), dusted)


def _served_runs(srv, torch, cc3d, draw, burst, vox, nreq, loads, gen_s):
  burst(srv, draw(40, 1))  # warm-up
  _progress("served: warm-up done")
  reqs = draw(nreq, 2)
  rs, cap_s = burst(srv, reqs)
  cap = nreq / cap_s
  _progress(f"served: closed-loop {cap:.1f} requests/s")
  plain = cc3d.LabelServer(workers=1, max_batch=1)
  burst(plain, draw(20, 1))
  _, plain_s = burst(plain, reqs)
  plain.close()
  _progress(f"served: plain {nreq / plain_s:.1f} requests/s")
  levels = []
  checked, exact = 0, True
  for f in loads:
    lam = f * cap
    r = np.random.default_rng(int(f * 1000))
    arrivals = np.cumsum(r.exponential(1.0 / lam, size=nreq))
    reqs = draw(nreq, 3 + int(f * 100))
    first_large = next(i for i, (_, c) in enumerate(reqs) if c == 6)
    sample = {i for i in range(0, nreq, 100)} | {first_large}
    base = time.perf_counter() + 0.01
    rs = []
    for i, (a, (v, c)) in enumerate(zip(arrivals, reqs)):
      now = time.perf_counter()
      if base + a > now:
        time.sleep(base + a - now)
      rs.append(srv.submit(v, c, t_arrive=base + a, keep=i in sample))
    for q in rs:
      if not q.done.wait(300):
        raise TimeoutError("a served request did not finish within 300 s")
      if q.error is not None:
        raise q.error
    end = max(q.t_done for q in rs)
    for i in sorted(sample):
      v, c = reqs[i]
      exact = _served_check(torch, cc3d, v, c, rs[i].result) and exact
      checked += 1
    for q in rs:
      q.result = None
    lat = np.array([q.latency for q in rs]) * 1e3
    big = np.array([c == 6 for _, c in reqs])
    achieved = nreq / (end - base)
    realized = nreq / float(arrivals[-1])  # the arrival rate this draw of the process offered
    _progress(f"served: load {f} offered {lam:.1f} (drawn {realized:.1f}) achieved {achieved:.1f} requests/s")
    levels.append({"load": f, "offered_rps": lam, "offered_drawn_rps": realized, "achieved_rps": achieved,
                   "sustained": bool(achieved >= 0.95 * realized),
                   "gvox_s": vox(reqs) / (end - base) / 1e9, "large_requests": int(big.sum()),
                   "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                   "p50_small_ms": float(np.percentile(lat[~big], 50)),
                   "p99_small_ms": float(np.percentile(lat[~big], 99)),
                   "p50_large_ms": float(np.percentile(lat[big], 50)),
                   "p99_large_ms": float(np.percentile(lat[big], 99)),
                   "mean_batch": float(np.mean([q.batched_with for q in rs]))})
  ok = [lv for lv in levels if lv["sustained"]]
  cap_open = max(lv["offered_rps"] for lv in ok) if ok else None
  mean_vox = vox(draw(nreq, 2)) / nreq
  return {
    "workload": "Poisson arrivals; 90% uint32 Voronoi sides U[128,256] 26-conn, 10% 1024x1024x512 200k seeds "
                "6-conn; each + device counts, bboxes, dust<100; inputs resident in HBM",
    "requests_per_level": nreq,
    "capacity_rps": cap_open,
    "capacity_gvox_s": cap_open * mean_vox / 1e9 if cap_open else None,
    "capacity_rule": "highest open-loop offered rate with achieved >= 95% of the drawn arrival rate",
    "closed_loop_rps": cap, "plain_rps": nreq / plain_s,
    "served_vs_plain": (cap_open / (nreq / plain_s)) if cap_open else None,
    "sampled_results_checked": checked, "sampled_results_bit_exact": bool(exact),
    "peak_reserved_gb": srv.peak_reserved / 1e9, "budget_gb": srv.budget / 1e9,
    "levels": levels, "generate_s": gen_s,
  }


def configs_leg(torch, cc3d, synth, dev, steps=5):
  """The other workloads at full size: GPU throughput with the input resident
  in HBM, the reference on one host core, and a bit-exact check of the two."""
  import oracle
  cases = [("256^3 uint32 Voronoi 1500 seeds 26-conn", lambda: synth.voronoi((256, 256, 256), 1500, seed=2024),
            dict(connectivity=26, out_dtype=np.uint32))]
  for p in (0.1, 0.3, 0.5):
    for conn in (6, 26):
      cases.append((f"512^3 uint8 Bernoulli p={p} binary {conn}-conn",
                    lambda p=p, conn=conn: synth.bernoulli((512, 512, 512), p, seed=int(p * 10) * 100 + conn),
                    dict(connectivity=conn, binary_image=True, out_dtype=np.uint32)))
  cases.append(("16384^2 uint8 Bernoulli p=0.5 binary 8-conn",
                lambda: synth.bernoulli((16384, 16384), 0.5, seed=16384),
                dict(connectivity=8, binary_image=True, out_dtype=np.uint32)))
  u64 = {}
  for conn in (6, 18, 26):
    def gen_u64():
      if "v" not in u64:
        u64["v"] = synth.voronoi((512, 512, 512), 40000, seed=64, id_range=(1, None), dtype=np.uint64)
      return u64["v"]
    cases.append((f"512^3 uint64 Voronoi 40k seeds {conn}-conn", gen_u64, dict(connectivity=conn, out_dtype=np.uint32)))
  cases.append(("512^3 float32 tones, continuous delta=20 26-conn", lambda: synth.tones((512, 512, 512), 5000, seed=32),
                dict(connectivity=26, delta=20, out_dtype=np.uint32)))
  res = []
  for name, gen, kw in cases:
    vol = gen()
    V = vol.size
    if vol.ndim == 3:
      d = torch.from_numpy(np.ravel(vol, order="F")).to(dev).as_strided(vol.shape, (1, vol.shape[0], vol.shape[0] * vol.shape[1]))
    else:
      d = torch.from_numpy(np.ravel(vol, order="F")).to(dev).as_strided(vol.shape, (1, vol.shape[0]))
    out, N = cc3d.connected_components(d, return_N=True, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
      out, N = cc3d.connected_components(d, return_N=True, **kw)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t0) / steps
    path_gbs = (vol.itemsize + 4) * V / gpu_s / 1e9
    row = {"workload": name, "voxels": V, "gvox_s": V / gpu_s / 1e9, "ms": gpu_s * 1e3, "components": int(N),
           "path_gbs": path_gbs, "roofline_frac": path_gbs / HBM_PEAK_GBS,
           "algorithmic_bytes_per_voxel": vol.itemsize + 4}
    if oracle.have_ref():
      t0 = time.perf_counter()
      ref, rN = oracle.connected_components(vol, return_N=True, backend="ref", **kw)
      cpu_s = time.perf_counter() - t0
      row["cpu_reference_gvox_s"] = V / cpu_s / 1e9
      row["bit_exact_vs_reference"] = bool(rN == N and np.array_equal(out.cpu().numpy(), ref))
    _progress(f"configs: {name}: {row['gvox_s']:.1f} Gvox/s")
    res.append(row)
    del d, out
  return res


def main() -> int:
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=20)
  ap.add_argument("--warmup", type=int, default=3)
  ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
  ap.add_argument("--no-batched", action="store_true", help="skip the batched leg")
  ap.add_argument("--no-served", action="store_true", help="skip the served leg")
  ap.add_argument("--no-configs", action="store_true",
                  help="skip timing the other workloads at full size against the reference")
  ap.add_argument("--calib", action="store_true",
                  help="also run one 2 GiB device copy (counter calibration for profiles)")
  args = ap.parse_args()
  if args.gpus != 1:
    print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
          "(volumes are independent: each GPU would serve its own queue of volumes)", file=sys.stderr)
    return 2

  import torch
  import cc3d
  import synth
  from cc3d import _lib

  assert torch.cuda.is_available(), "needs a GPU"
  dev = torch.device("cuda:0")
  t0 = time.time()
  vol = synth.voronoi(SHAPE, NSEEDS, seed=SEED)
  gen_s = time.time() - t0
  V = vol.size
  _progress(f"headline: volume generated in {gen_s:.1f} s")
  flat = torch.from_numpy(np.ravel(vol, order="F")).to(dev)
  sx, sy, sz = SHAPE
  dvol = flat.as_strided(SHAPE, (1, sx, sx * sy))
  torch.cuda.synchronize()

  if args.calib:
    # one 2 GiB device copy of a written buffer, after a 1 GiB write that
    # evicts it from the caches: HBM bytes the counters should report
    a = torch.ones(512 << 20, dtype=torch.int32, device=dev)
    e = torch.zeros(256 << 20, dtype=torch.int32, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b, e

  def step():
    return cc3d.connected_components(dvol, connectivity=26, out_dtype=np.uint32, return_N=True)

  for _ in range(args.warmup):
    out, N = step()
  torch.cuda.synchronize()

  lib = _lib.lib()
  lib.cc3d_amd_profile_enable.argtypes = [ctypes.c_int]
  lib.cc3d_amd_profile_collect.restype = ctypes.c_int
  lib.cc3d_amd_profile_collect.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_int), ctypes.c_int]
  # per-kernel breakdown from one untimed profiled pass of the same steps
  _profile_rows(lib)  # drop anything recorded before
  lib.cc3d_amd_profile_enable(1)
  for _ in range(min(args.steps, 5)):
    step()
  torch.cuda.synchronize()
  lib.cc3d_amd_profile_enable(0)
  rows_all = _profile_rows(lib)
  # timed steps; HIP events on the launch stream around the dominant kernel only
  lib.cc3d_amd_profile_enable(2)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    out, N = step()
  torch.cuda.synchronize()
  elapsed = time.perf_counter() - t0
  lib.cc3d_amd_profile_enable(0)
  rows = dict(rows_all)
  rows.update(_profile_rows(lib))  # k_local: measured over the timed steps
  _progress(f"headline: {args.steps} steps in {elapsed * 1e3:.1f} ms")

  ms_per_step = elapsed * 1e3 / args.steps
  gvox = V * args.steps / elapsed / 1e9

  # dominant kernel and its algorithmic bytes per launch
  per_kernel = {k: {"ms_per_launch": v[0] / max(v[1], 1), "launches": v[1]} for k, v in rows.items()}
  algo = {  # bytes each kernel must move per voxel at minimum
    "k_local": 4.0,      # read the uint32 input once
    "k_finalize": 4.0,   # write the uint32 output once
  }
  dom = max(per_kernel, key=lambda k: per_kernel[k]["ms_per_launch"]) if per_kernel else "k_local"
  dom_ms = per_kernel.get(dom, {}).get("ms_per_launch", float("nan"))
  dom_bytes = algo.get(dom, 0.0) * V
  achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms and dom_bytes else None
  roofline = {
    "bound": "hbm",
    "kernel": dom,
    "achieved": achieved,
    "peak": HBM_PEAK_GBS,
    "unit": "GB/s",
    "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
    "traffic": _traffic_from_profiles(dom),
    "algorithmic_bytes_per_launch": dom_bytes,
  }
  path_gbs = 8.0 * V / (ms_per_step * 1e-3) / 1e9
  kernel_sum_ms = sum(v["ms_per_launch"] for v in per_kernel.values())
  copy_gbs = _copy_gbs(flat)
  path_roofline = {
    "achieved": path_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": path_gbs / HBM_PEAK_GBS,
    "algorithmic_bytes_per_voxel": 8, "kernel_ms_sum": kernel_sum_ms,
    "kernel_frac": (8.0 * V / (kernel_sum_ms * 1e-3) / 1e9) / HBM_PEAK_GBS if kernel_sum_ms else None,
    "copy_gbs": copy_gbs, "frac_of_copy": path_gbs / copy_gbs if copy_gbs else None,
  }

  cpu = None
  exact = None
  if not args.no_cpu:
    import oracle
    if oracle.have_ref():
      t0 = time.perf_counter()
      ref_out, refN = oracle.connected_components(vol, connectivity=26, out_dtype=np.uint32, return_N=True,
                                                  backend="ref")
      cpu_s = time.perf_counter() - t0
      _progress(f"headline: reference on one host core {cpu_s:.2f} s")
      got = out.cpu().numpy()
      exact = bool(refN == N and np.array_equal(got, ref_out))
      cpu = {"value": V / cpu_s / 1e9, "unit": "Gvoxels/s", "cores": 1, "kind": "reference",
             "sample": f"one full 512^3 volume of the workload ({cpu_s:.2f} s single-threaded, "
                       "estimate_provisional_labels + connected_components3d)"}

  batched = None
  if not args.no_batched:
    batched = batched_leg(torch, cc3d, synth, dev, cpu=not args.no_cpu)

  configs = None
  if not args.no_configs:
    torch.cuda.empty_cache()
    configs = configs_leg(torch, cc3d, synth, dev)

  served = None
  if not args.no_served:
    torch.cuda.empty_cache()
    served = served_leg(torch, cc3d, synth, dev)

  line = {
    "metric": METRIC,
    "value": gvox,
    "unit": "Gvoxels/s",
    "n_gpus": 1,
    "steps": args.steps,
    "warmup": args.warmup,
    "ms_per_step": ms_per_step,
    "higher_is_better": True,
    "scaling": "weak",
    "vs_baseline": None,
    "dtype": "u32",
    "data": "synthetic",
    "config": {"workload": "voronoi-512^3-uint32-26conn-multilabel", "shape": list(SHAPE), "seeds": NSEEDS,
               "membrane_p": 0.8, "salt_p": 0.02, "out_dtype": "uint32", "return_N": True,
               "components": int(N), "input_resident_in_hbm": True},
    "roofline": roofline,
    "path_roofline": path_roofline,
    "kernels": per_kernel,
    "cpu_baseline": cpu,
    "bit_exact_vs_reference": exact,
    "batched": batched,
    "served": served,
    "configs": configs,
    "generate_s": gen_s,
  }
  print(json.dumps(line))
  return 0


if __name__ == "__main__":
  sys.exit(main())
