"""Batched labelling: many volumes through one set of GPU launches.

No counterpart in the reference, which labels one volume per call.  Result i
of `connected_components_batch(volumes, ...)` is exactly
`connected_components(volumes[i], ...)`: the same labels, N, output dtype and
exceptions.  Volumes share one launch of every kernel (cc3d_amd_batch_* in
include/cc3d_amd.h), one batch per input dtype and mode (multilabel, binary,
continuous), whatever their connectivity, and with or without a periodic
boundary.  The rest -- out_file outputs, and binary 2D-8 images whose
provisional blocks must be counted for an explicit uint16 output -- go
through the single-volume path.

`LabelBatch` keeps a planned batch of device-resident volumes with fixed
output buffers, for repeated labelling; it can capture the whole launch
sequence in a HIP graph and replay it.
"""
from __future__ import annotations

import ctypes
from typing import Any, List, Optional, Sequence

import numpy as np

from . import _lib as L
from ._core import (_OUT_CODE, _decide, _deliver, _label_one, _prepare, _require_gpu, _special_row, _stream,
                    _torch, _torch_out_dtype)


def _mode(pl) -> int:
  if pl.binary_image:
    return L.BATCH_BINARY
  return L.BATCH_CONTINUOUS if pl.flags & L.FLAG_CONTINUOUS else L.BATCH_MULTILABEL


def _batchable(pl) -> bool:
  # FLAG_COUNT_PROVISIONAL (binary 2D-8 into an explicit uint16 output past
  # 65,535 blocks) has no batch counterpart
  if pl.out_file is not None or (pl.flags & ~L.FLAG_CONTINUOUS):
    return False
  if pl.connectivity in (26, 18, 6):
    return True
  if pl.connectivity in (8, 4):
    return pl.sz == 1
  return False


def _group_key(pl):
  """Plans of one batch share input dtype, mode and (continuous) delta."""
  mode = _mode(pl)
  return pl.code, mode, (pl.delta_b if mode == L.BATCH_CONTINUOUS else b"")


def _volume_table(plans, bufs):
  from ._core import _tile_flags
  vt = (L.Volume * len(plans))()
  for k, (pl, buf) in enumerate(zip(plans, bufs)):
    vt[k].inp = buf.data_ptr()
    vt[k].sx, vt[k].sy, vt[k].sz = pl.sx, pl.sy, pl.sz
    vt[k].connectivity = (pl.connectivity | (_tile_flags(pl, buf) << 8)
                          | (L.VOLUME_PERIODIC if pl.periodic_boundary else 0))
  return vt


def _create(plans, bufs, ws, ws_bytes, stream):
  """cc3d_amd_batch_create (+ delta for a continuous batch) -> handle."""
  lib = L.lib()
  code, mode, delta_b = _group_key(plans[0])
  vt = _volume_table(plans, bufs)
  b = ctypes.c_void_p(0)
  L.check(lib.cc3d_amd_batch_create(ctypes.byref(b), vt, len(plans), code, mode, ctypes.c_void_p(ws.data_ptr()),
                                    ws_bytes, stream))
  if mode == L.BATCH_CONTINUOUS:
    dbuf = ctypes.create_string_buffer(delta_b, max(len(delta_b), 8))
    rc = lib.cc3d_amd_batch_set_delta(b, ctypes.cast(dbuf, ctypes.c_void_p))
    if rc != L.OK:
      lib.cc3d_amd_batch_free(b)
      L.check(rc)
  return b


def _workspace_bytes(plans, bufs) -> int:
  ws_bytes = int(L.lib().cc3d_amd_batch_workspace_bytes(_volume_table(plans, bufs), len(plans),
                                                         _group_key(plans[0])[1]))
  if ws_bytes == 0:
    L.check(L.ERR_INVALID)
  return ws_bytes


def _label_group(items, results, errors):
  """Labels plans sharing input dtype, mode and delta in one batch."""
  torch = _torch()
  plans = [pl for _, pl in items]
  bufs = [pl.vol.device_buffer(pl.order) for pl in plans]
  n = len(plans)
  lib = L.lib()
  ws_bytes = _workspace_bytes(plans, bufs)
  ws = torch.empty(ws_bytes, dtype=torch.uint8, device=bufs[0].device)
  b = _create(plans, bufs, ws, ws_bytes, _stream())
  try:
    L.check(lib.cc3d_amd_batch_begin(b))
    st = (L.Stats * n)()
    L.check(lib.cc3d_amd_batch_stats(b, st))
    odts, Ns = [], []
    for k, (i, pl) in enumerate(items):
      try:
        odt, N = _decide(pl, st[k])
      except Exception as e:  # the reference raises for this volume
        errors[i] = e
        odt, N = np.dtype(np.uint32), 0
      odts.append(odt)
      Ns.append(N)
    outs = [torch.empty(pl.voxels, dtype=_torch_out_dtype(odt), device=bufs[0].device)
            for pl, odt in zip(plans, odts)]
    ptrs = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
    codes = (ctypes.c_int * n)(*[_OUT_CODE[o] for o in odts])
    L.check(lib.cc3d_amd_batch_set_outputs(b, ptrs, codes))
    L.check(lib.cc3d_amd_batch_finish(b))
    for k, (i, pl) in enumerate(items):
      if i in errors:
        continue
      if pl.special_row is not None:
        try:
          Ns[k] = _special_row(pl, bufs[k], odts[k], outs[k])
        except Exception as e:  # the reference raises for this volume
          errors[i] = e
          continue
      results[i] = _deliver(outs[k], pl.vol, pl.sx, pl.sy, pl.sz, pl.dims, pl.order, odts[k], None, Ns[k],
                            pl.return_N)
  finally:
    lib.cc3d_amd_batch_free(b)


def connected_components_batch(
  volumes: Sequence[Any],
  connectivity=26,
  return_N: bool = False,
  out_dtype: Optional[Any] = None,
  binary_image: bool = False,
  max_labels: int = -1,
  delta=0,
  periodic_boundary: bool = False,
) -> List[Any]:
  """connected_components over a list of volumes; `connectivity` may be one
  value or one per volume.  Returns the list of results (each `labels` or
  `(labels, N)`).  When volumes would raise, the exception of the first such
  volume is raised after the batch has run."""
  n = len(volumes)
  conns = list(connectivity) if isinstance(connectivity, (list, tuple)) else [connectivity] * n
  if len(conns) != n:
    raise ValueError("connectivity must be one value or one per volume")
  results: List[Any] = [None] * n
  errors = {}
  plans = []
  for i, v in enumerate(volumes):
    try:
      kind, pl = _prepare(v, max_labels, conns[i], delta, out_dtype, None, periodic_boundary, binary_image,
                          return_N)
    except Exception as e:
      errors[i] = e
      continue
    if kind == "done":
      results[i] = pl
    else:
      plans.append((i, pl))
  if plans:
    _require_gpu()
  groups = {}
  for i, pl in plans:
    if _batchable(pl):
      groups.setdefault(_group_key(pl), []).append((i, pl))
    else:
      try:
        results[i] = _label_one(pl)
      except Exception as e:
        errors[i] = e
  for items in groups.values():
    _label_group(items, results, errors)
  if errors:
    raise errors[min(errors)]
  return results


class LabelBatch:
  """A planned batch of device-resident volumes with fixed output buffers.

  volumes: device torch tensors (any layout connected_components accepts).
  All must be batchable and share dtype and mode (binary_image, delta);
  out_dtype is fixed (uint16/uint32/uint64) so the whole sequence can run as
  one HIP graph.
  `run()` labels through the two-stage launches; `capture()` records begin +
  finish in a graph and `replay()` relaunches it.  Both return
  [(labels, N)], raising like connected_components when a volume's estimate
  does not fit out_dtype or its provisional labels overflow.
  """

  def __init__(self, volumes, connectivity=26, out_dtype=np.uint32, binary_image=False, stream=None, delta=0,
               periodic_boundary=False):
    torch = _torch()
    _require_gpu()
    n = len(volumes)
    conns = list(connectivity) if isinstance(connectivity, (list, tuple)) else [connectivity] * n
    self.plans = []
    for v, c in zip(volumes, conns):
      kind, pl = _prepare(v, -1, c, delta, out_dtype, None, periodic_boundary, binary_image, True)
      if kind != "plan" or not _batchable(pl):
        raise ValueError("LabelBatch volumes must be non-empty and batchable")
      self.plans.append(pl)
    if len({_group_key(pl) for pl in self.plans}) != 1:
      raise ValueError("LabelBatch volumes must share dtype and mode")
    self.odt = np.dtype(out_dtype)
    self.stream = stream if stream is not None else torch.cuda.Stream()
    self.bufs = [pl.vol.device_buffer(pl.order) for pl in self.plans]
    dev = self.bufs[0].device
    self.n = n
    lib = L.lib()
    ws_bytes = _workspace_bytes(self.plans, self.bufs)
    self.ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    self.outs = [torch.empty(pl.voxels, dtype=_torch_out_dtype(self.odt), device=dev) for pl in self.plans]
    self.b = ctypes.c_void_p(0)
    torch.cuda.current_stream().synchronize()  # inputs and buffers are ready before the batch stream uses them
    self.b = _create(self.plans, self.bufs, self.ws, ws_bytes, ctypes.c_void_p(self.stream.cuda_stream))
    ptrs = (ctypes.c_void_p * n)(*[o.data_ptr() for o in self.outs])
    codes = (ctypes.c_int * n)(*([_OUT_CODE[self.odt]] * n))
    L.check(lib.cc3d_amd_batch_set_outputs(self.b, ptrs, codes))
    self.captured = False

  def _results(self, deliver=True):
    st = (L.Stats * self.n)()
    L.check(L.lib().cc3d_amd_batch_stats(self.b, st))
    res = []
    for k, pl in enumerate(self.plans):
      odt, N = _decide(pl, st[k])
      if pl.special_row is not None:
        N = _special_row(pl, self.bufs[k], odt, self.outs[k])
      if deliver:
        res.append(_deliver(self.outs[k], pl.vol, pl.sx, pl.sy, pl.sz, pl.dims, pl.order, odt, None, N, True))
      else:
        res.append(N)
    return res

  def run(self, deliver=True):
    lib = L.lib()
    L.check(lib.cc3d_amd_batch_begin(self.b))
    L.check(lib.cc3d_amd_batch_finish(self.b))
    return self._results(deliver)

  def capture(self):
    L.check(L.lib().cc3d_amd_batch_graph_capture(self.b))
    self.captured = True

  def replay(self, deliver=True):
    if not self.captured:
      self.capture()
    L.check(L.lib().cc3d_amd_batch_graph_launch(self.b))
    return self._results(deliver)

  def close(self):
    if self.b:
      L.lib().cc3d_amd_batch_free(self.b)
      self.b = ctypes.c_void_p(0)

  def __del__(self):
    try:
      self.close()
    except Exception:
      pass
