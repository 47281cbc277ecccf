"""Tile-family choice for the local labelling kernel, from a measured table.

The local kernel comes in three tile families with identical results: voxel
tiles (32x8x8, one union-find node per voxel, 256 threads; csrc/local.hpp),
tall voxel tiles (32x16x8, 512 threads: fewer tile borders and a smaller halo
per voxel for a larger LDS footprint) and row tiles (64x8x8, x-runs as nodes;
csrc/rows.hpp).  Which is fastest depends on the connectivity, the voxel
type, the volume's size and how much of it is foreground.  `tile_table.json`
(next to this file) holds, per key

    connectivity / mode / voxel bytes / size class / density class

the family chosen and every family's timing.  `python -m cc3d._tune`
measures every key on this GPU with synthetic volumes (Voronoi labellings,
optionally thinned by a Bernoulli mask, and Bernoulli binary images), and the
cost of one density sample, and rewrites the table.  At call time the key's
size class comes from the shape.  The density class is sampled from the
device volume (4096 evenly spaced voxels, one small reduction and a host
synchronisation) only when the table's choice for that key depends on it and
the best single family for the key would lose more than one sample costs on
some density class; otherwise that family is used unsampled.  Keys missing
from the table keep the default family (rows for 6, voxels otherwise).
"""
from __future__ import annotations

import json
import os
import time
import zlib
from typing import Callable, Optional

import numpy as np

from . import _lib as L

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(_HERE, "tile_table.json")

SIZE_CLASSES = ((1 << 19, "s"), (1 << 23, "m"), (None, "l"))  # voxels: <= 512 Ki, <= 8 Mi, more
DENSITY_CLASSES = ((0.3, "sparse"), (0.6, "mid"), (None, "dense"))  # foreground fraction
FAMILIES = ("voxels", "tall", "rows")
_FLAG = {"rows": L.FLAG_TILES_ROWS, "voxels": L.FLAG_TILES_VOXELS, "tall": L.FLAG_TILES_TALL}

_table = None
_doc = None
# measured cost of one density sample (ms), when the table does not hold it
SAMPLE_MS_DEFAULT = 0.03
# the tuner's own override: "rows" / "voxels" / "tall" forces a family for every call
FORCE: Optional[str] = None


def size_class(voxels: int) -> str:
  for lim, name in SIZE_CLASSES:
    if lim is None or voxels <= lim:
      return name
  return SIZE_CLASSES[-1][1]


def density_class(frac: float) -> str:
  for lim, name in DENSITY_CLASSES:
    if lim is None or frac < lim:
      return name
  return DENSITY_CLASSES[-1][1]


def table() -> dict:
  global _table, _doc
  if _table is None:
    try:
      with open(TABLE_PATH) as f:
        _doc = json.load(f)
    except (OSError, ValueError):
      _doc = {}
    _table = _doc.get("entries", {})
  return _table


def sample_ms() -> float:
  table()
  return float(_doc.get("sample_ms", SAMPLE_MS_DEFAULT))


def unsampled_choice(row: dict):
  """(family, loss): the family whose largest loss against the per-density
  choices of a table row is smallest, and that loss in ms."""
  fams = [f for f in FAMILIES if all(f"ms_{f}" in e for e in row.values())]
  chosen = {d: e[f"ms_{e['choice']}"] for d, e in row.items()}
  loss = {f: max(e[f"ms_{f}"] - chosen[d] for d, e in row.items()) for f in fams}
  f = min(fams, key=lambda x: (loss[x], x != row[next(iter(row))]["choice"]))
  return f, loss[f]


def key(connectivity: int, binary: bool, itemsize: int, voxels: int) -> str:
  return f"{int(connectivity)}/{'binary' if binary else 'multi'}/{int(itemsize)}/{size_class(voxels)}"


def sample_density(buf, itemsize: int, voxels: int) -> float:
  """Fraction of nonzero voxels among up to 4096 evenly spaced ones of a
  device buffer (raw bytes, itemsize bytes per voxel)."""
  torch = _torch_mod()
  n = min(4096, voxels)
  step = max(1, voxels // n)
  words = {1: torch.uint8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[itemsize]
  flat = buf.view(torch.uint8).view(words) if buf.dtype != words else buf
  pick = flat.as_strided((n,), (step,))
  return float((pick != 0).sum().item()) / n


def _torch_mod():
  import torch
  return torch


def tiles(connectivity: int, binary: bool, itemsize: int, voxels: int, is3d: bool,
          density: Callable[[], float]) -> int:
  """CC3D_AMD_TILES_* flag bits for one volume (0: the default family)."""
  if FORCE is not None:
    return _FLAG[FORCE]
  if not is3d:
    return 0
  row = table().get(key(connectivity, binary, itemsize, voxels))
  if not row:
    return 0
  choices = {v["choice"] for v in row.values()}
  if len(choices) == 1:
    return _FLAG[choices.pop()]
  fam, loss = unsampled_choice(row)
  if loss <= sample_ms():
    return _FLAG[fam]
  d = row.get(density_class(density()))
  return _FLAG[d["choice"]] if d else 0


# ---------------------------------------------------------------------------
# the measurement that writes the table
# ---------------------------------------------------------------------------
_SHAPES = {"s": (80, 80, 80), "m": (192, 192, 192), "l": (384, 384, 384)}
_DENSITY_P = {"sparse": 0.22, "mid": 0.5, "dense": None}  # Bernoulli mask / foreground probability


def _volume(synth, mode: str, itemsize: int, shape, density: str, seed: int) -> np.ndarray:
  if mode == "binary":
    p = {"sparse": 0.15, "mid": 0.45, "dense": 0.75}[density]
    return synth.bernoulli(shape, p, seed=seed)
  nseeds = max(1, int(np.prod(shape)) // 8000)
  dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[itemsize]
  vol = synth.voronoi(shape, nseeds, seed=seed, id_range=(1, None), dtype=dt) if itemsize == 8 else \
    synth.voronoi(shape, nseeds, seed=seed).astype(dt)
  keep = _DENSITY_P[density]
  if keep is not None:
    vol = vol * synth.bernoulli(shape, keep, seed=seed + 1).astype(dt)
  return vol


# Another family replaces the default one only when it is at least this much
# faster (measured differences below it are within run-to-run noise).
MARGIN = 0.02


def default_family(connectivity: int) -> str:
  return "rows" if connectivity == 6 else "voxels"


def decide(connectivity: int, ms: dict) -> str:
  """The fastest family of {family: ms} if it beats the default by MARGIN."""
  default = default_family(connectivity)
  best = min(ms, key=lambda f: (ms[f], f != default))
  return best if ms[best] < (1.0 - MARGIN) * ms[default] else default


def measure(reps: int = 5, log=print) -> dict:
  """Times both tile families on every key; returns the table."""
  import torch
  import synth
  from ._core import connected_components
  global FORCE
  entries = {}
  started = time.time()
  for conn in (26, 18, 6):
    for mode, sizes in (("multi", (1, 2, 4, 8)), ("binary", (1,))):
      for itemsize in sizes:
        for cls, shape in _SHAPES.items():
          k = f"{conn}/{mode}/{itemsize}/{cls}"
          entries[k] = {}
          for dens in ("sparse", "mid", "dense"):
            vol = _volume(synth, mode, itemsize, shape, dens, seed=zlib.crc32(f"{k}/{dens}".encode()) % 100000)
            d = torch.from_numpy(np.ravel(vol, order="F")).cuda().as_strided(
              shape, (1, shape[0], shape[0] * shape[1]))
            frac = float(np.count_nonzero(vol)) / vol.size
            ms = {}
            for fam in FAMILIES:
              FORCE = fam
              try:
                for _ in range(2):
                  connected_components(d, connectivity=conn, out_dtype=np.uint32, binary_image=mode == "binary")
                torch.cuda.synchronize()
                times = []
                for _ in range(reps):
                  a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                  a.record()
                  connected_components(d, connectivity=conn, out_dtype=np.uint32, binary_image=mode == "binary")
                  b.record()
                  torch.cuda.synchronize()
                  times.append(a.elapsed_time(b))
                ms[fam] = float(np.median(times))
              finally:
                FORCE = None
            choice = decide(conn, ms)
            ent = {"choice": choice, "foreground": round(frac, 3)}
            ent.update({f"ms_{f}": round(ms[f], 4) for f in FAMILIES})
            entries[k][density_class(frac)] = ent
            log(f"[tune {time.time() - started:6.1f}s] {k} {dens} fg={frac:.2f} "
                + " ".join(f"{f} {ms[f]:.3f} ms" for f in FAMILIES) + f" -> {choice}")
            del d
  return entries


def measure_sample(reps: int = 5) -> float:
  """Median ms of one density sample of a 192^3 uint32 device volume."""
  import torch
  buf = torch.zeros(192 ** 3, dtype=torch.int32, device="cuda")
  times = []
  for _ in range(2 + 4 * reps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sample_density(buf, 4, buf.numel())
    times.append((time.perf_counter() - t0) * 1e3)
  return float(np.median(times[2:]))


def main(argv=None) -> int:
  import argparse
  import sys
  ap = argparse.ArgumentParser(description="measure the tile families and write the tile table")
  ap.add_argument("--out", default=TABLE_PATH)
  ap.add_argument("--reps", type=int, default=5)
  args = ap.parse_args(argv)
  here = os.path.dirname(_HERE)
  if here not in sys.path:
    sys.path.insert(0, here)
  entries = measure(args.reps, log=lambda m: print(m, flush=True))
  import torch
  smp = measure_sample(args.reps)
  print(f"[tune] one density sample: {smp:.4f} ms", flush=True)
  doc = {"device": torch.cuda.get_device_name(0), "reps": args.reps,
         "note": "median ms of one connected_components call (device input, uint32 out) per tile family; "
                 "sample_ms: median ms of one density sample (sample_density with its host synchronisation)",
         "sample_ms": round(smp, 4), "entries": entries}
  with open(args.out, "w") as f:
    json.dump(doc, f, indent=1, sort_keys=True)
  print("wrote", args.out)
  return 0


if __name__ == "__main__":
  # run as the package's module, so FORCE is the flag the dispatcher reads
  import importlib
  import sys
  sys.path.insert(0, os.path.dirname(_HERE))
  raise SystemExit(importlib.import_module("cc3d._tune").main())
