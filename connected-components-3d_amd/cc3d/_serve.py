"""A labelling service: admission by HBM footprint, continuous batching.

No counterpart in the reference (a library call per volume).  A request is a
device-resident volume plus its connectivity.  The service labels it
(connected_components, out_dtype uint32), then computes on the device its
voxel counts and bounding boxes (cc3d.statistics) and removes components
below `dust_threshold` voxels from a copy of the image (cc3d.dust with the
same connectivity, reusing the labels instead of labelling twice).

Scheduling:
  * admission: a request enters service only when its HBM footprint fits the
    free part of the budget (first come, first served; the head of the queue
    is never overtaken by a later request);
  * continuous batching: each worker (one HIP stream each, default priority)
    takes the head request and, when it is small, every other waiting small
    request that fits, up to `max_batch`, and labels them with one batched
    launch (cc3d.connected_components_batch); requests that arrive while
    others run are admitted as memory frees up, and their work overlaps on
    the device with the other workers';
  * every result is identical to labelling the request alone.

Footprint of a request: at admission, the labelling workspace + uint32 labels
+ the dusted copy.  Once the labelling has returned N, the workspace is given
back and the per-label outputs (voxel counts and bounding boxes, 28 bytes for
each of N + 1 labels) are reserved in its place; when they need more than the
workspace gave back, the worker waits for the difference, unless no other
request is in service (then it proceeds, so the service cannot stall).  The
rest is released at completion.  Results belong to the caller after
completion; the budget should leave room for the results callers keep.
"""
from __future__ import annotations

import collections
import ctypes
import threading
import time
from typing import List, Optional

import numpy as np

from . import _lib as L
from ._batch import connected_components_batch
from ._core import _NP_TO_CODE, _torch, connected_components
from ._follow import _bind, _device_stats, _fcheck

_STAT_BYTES_PER_LABEL = 4 + 4 * 6  # counts, bbox


class Request:
  """One labelling request; `done` is set when `result` (or `error`) is ready."""

  def __init__(self, volume, connectivity: int, t_arrive: float):
    self.volume = volume
    self.connectivity = int(connectivity)
    self.t_arrive = t_arrive
    self.t_start: Optional[float] = None
    self.t_done: Optional[float] = None
    self.result: Optional[dict] = None
    self.error: Optional[BaseException] = None
    self.done = threading.Event()
    self.footprint = 0
    self.batched_with = 1
    self.ready = None
    self.keep = True

  @property
  def latency(self) -> float:
    return self.t_done - self.t_arrive


def request_footprint(volume, connectivity: int) -> int:
  """HBM bytes reserved for one request at admission (see module docstring)."""
  return _workspace(volume, connectivity) + _outputs(volume)


def _workspace(volume, connectivity: int) -> int:
  sx, sy, sz = _dims(volume)
  return int(L.lib().cc3d_amd_workspace_bytes(sx, sy, sz, connectivity, 0, 0))


def _outputs(volume) -> int:
  sx, sy, sz = _dims(volume)
  V = sx * sy * sz
  return 4 * V + volume.element_size() * V  # labels + dusted copy


def _dims(volume):
  shape = list(volume.shape) + [1] * (3 - volume.dim())
  return int(shape[0]), int(shape[1]), int(shape[2])


def _followons(vol, labels, N: int, dust_threshold: int):
  """Counts, bounding boxes (device) and the dusted image (device copy)."""
  torch = _torch()
  sx, sy, sz = _dims(vol)
  V = sx * sy * sz
  lab = labels.as_strided((V,), (1,))
  counts, bbox, _ = _device_stats(lab, L.DT_U32, sx, sy, sz, 3, 1, N, centroids=False)
  img = torch.empty_like(vol, memory_format=torch.preserve_format)  # the dusted copy, written in one pass
  flat = img.as_strided((V,), (1,))
  src = vol.detach().as_strided((V,), (1,))
  erase = torch.empty(N + 1, dtype=torch.uint8, device=vol.device)
  nmask = torch.zeros(1, dtype=torch.int64, device=vol.device)
  code = _NP_TO_CODE[np.dtype({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[vol.element_size()])]
  stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
  _fcheck(_bind().cc3d_amd_dust_copy(
    ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(flat.data_ptr()), code, ctypes.c_void_p(lab.data_ptr()),
    L.DT_U32, V, ctypes.c_void_p(counts.data_ptr()), N, float(dust_threshold), 0.0, 0, 0,
    ctypes.c_void_p(erase.data_ptr()), ctypes.c_void_p(nmask.data_ptr()), stream))
  return {"labels": labels, "N": N, "voxel_counts": counts, "bounding_boxes": bbox.view(N + 1, 6),
          "dusted": img, "_nmask": nmask}


class LabelServer:
  """Serves labelling requests on one GPU (see module docstring)."""

  def __init__(self, hbm_budget: Optional[int] = None, workers: int = 4, max_batch: int = 16,
               small_voxels: int = 1 << 24, dust_threshold: int = 100):
    torch = _torch()
    if hbm_budget is None:
      free, _ = torch.cuda.mem_get_info()
      hbm_budget = int(free * 0.5)
    self.budget = int(hbm_budget)
    self.free = self.budget
    self.max_batch = int(max_batch)
    self.small_voxels = int(small_voxels)
    self.dust_threshold = int(dust_threshold)
    self.cv = threading.Condition()
    self.waiting = collections.deque()
    self.active = 0  # requests in service
    self.stats_waiters = 0  # of those, requests waiting for their per-label outputs' memory
    self.peak_reserved = 0  # the most HBM reserved at once
    self.closing = False
    self.streams = [torch.cuda.Stream() for _ in range(workers)]
    self.threads = [threading.Thread(target=self._worker, args=(s,), daemon=True) for s in self.streams]
    for th in self.threads:
      th.start()

  def submit(self, volume, connectivity: int = 26, t_arrive: Optional[float] = None, keep: bool = True) -> Request:
    """volume: an integer device tensor of shape (sx, sy, sz) with x fastest
    (strides 1, sx, sx*sy).  t_arrive defaults to now.  keep=False drops the
    result tensors at completion (only N and dust_N are kept)."""
    sx, sy, sz = _dims(volume)
    if volume.device.type == "cpu" or tuple(volume.stride()) != (1, sx, sx * sy)[:volume.dim()]:
      raise ValueError("requests take device tensors with x fastest (strides 1, sx, sx*sy)")
    r = Request(volume, connectivity, time.perf_counter() if t_arrive is None else t_arrive)
    r.ready = _torch().cuda.Event()
    r.ready.record()  # the volume's producers on the caller's stream come first
    r.keep = bool(keep)
    r.footprint = request_footprint(volume, connectivity)
    if r.footprint > self.budget:
      r.error = MemoryError(f"request needs {r.footprint} bytes of HBM; the budget is {self.budget}")
      r.t_done = time.perf_counter()
      r.done.set()
      return r
    with self.cv:
      self.waiting.append(r)
      self.cv.notify_all()
    return r

  def close(self):
    with self.cv:
      self.closing = True
      self.cv.notify_all()
    for th in self.threads:
      th.join()

  def _note(self):
    self.peak_reserved = max(self.peak_reserved, self.budget - self.free)

  def _small(self, r: Request) -> bool:
    sx, sy, sz = _dims(r.volume)
    return sx * sy * sz <= self.small_voxels

  def _take(self) -> Optional[List[Request]]:
    with self.cv:
      while True:
        if self.waiting and self.waiting[0].footprint <= self.free:
          break
        if self.closing and not self.waiting:
          return None
        self.cv.wait()
      head = self.waiting.popleft()
      self.free -= head.footprint
      self._note()
      group = [head]
      self.active += 1
      if self._small(head):
        keep = collections.deque()
        while self.waiting and len(group) < self.max_batch:
          r = self.waiting.popleft()
          if self._small(r) and r.footprint <= self.free and r.volume.dtype == head.volume.dtype:
            self.free -= r.footprint
            self._note()
            group.append(r)
            self.active += 1
          else:
            keep.append(r)
            break  # first come, first served: stop at the first request that does not join
        while keep:
          self.waiting.appendleft(keep.pop())
      return group

  def _swap_for_stats(self, group: List[Request], ns: List[int]):
    """The labelling workspace of each request, given back, for its
    per-label outputs (see module docstring).  A group waits for memory only
    while some request in service is not waiting too, so the service always
    progresses."""
    with self.cv:
      self.stats_waiters += len(group)
      self.cv.notify_all()
      for r, n in zip(group, ns):
        ws = _workspace(r.volume, r.connectivity)
        stats = _STAT_BYTES_PER_LABEL * (n + 1)
        self.free += ws
        r.footprint -= ws
        while stats > self.free and self.active > self.stats_waiters:
          self.cv.wait()
        self.free -= stats
        self._note()
        r.footprint += stats
      self.stats_waiters -= len(group)
      self.cv.notify_all()

  def _worker(self, stream):
    torch = _torch()
    while True:
      group = self._take()
      if group is None:
        return
      t0 = time.perf_counter()
      for r in group:
        r.t_start = t0
        r.batched_with = len(group)
      try:
        with torch.cuda.stream(stream):
          for r in group:
            stream.wait_event(r.ready)
          if len(group) == 1:
            lab = [connected_components(group[0].volume, connectivity=group[0].connectivity,
                                        out_dtype=np.uint32, return_N=True)]
          else:
            lab = connected_components_batch([r.volume for r in group],
                                             connectivity=[r.connectivity for r in group],
                                             out_dtype=np.uint32, return_N=True)
          self._swap_for_stats(group, [int(N) for _, N in lab])
          res = [_followons(r.volume, labels, int(N), self.dust_threshold) for r, (labels, N) in zip(group, lab)]
          stream.synchronize()
        for r, out in zip(group, res):
          nm = int(out.pop("_nmask").item())
          out["dust_N"] = out["N"] - nm
          r.result = out if r.keep else {"N": out["N"], "dust_N": out["dust_N"]}
        del res, lab
      except BaseException as e:  # reported on the request
        for r in group:
          r.error = e
      t1 = time.perf_counter()
      with self.cv:
        for r in group:
          r.t_done = t1
          self.free += r.footprint
          self.active -= 1
        self.cv.notify_all()
      for r in group:
        r.done.set()
