"""Parity at the full sizes of the benchmark workloads.

Each case labels the workload's own input on the GPU and compares labels, N
and output dtype bit for bit with the reference compiled from its own
sources (oracle/_ref), or with the oracle where that build is absent:

  * 512^3 uint8 Bernoulli p in {0.1, 0.3, 0.5}, binary_image, 6 and 26;
  * 16384^2 uint8 Bernoulli p = 0.5, binary_image, 8-connected (the 2x2
    block numbering of the binary 2D path);
  * 512^3 uint64 Voronoi, 40,000 seeds, 64-bit IDs, 6, 18 and 26;
  * 512^3 float32 three-tone image, 5,000 seeds, continuous delta = 20, 26.

(The 256^3 and 512^3 uint32 Voronoi cases are in test_gpu_parity.py and in
the benchmark's own check.)
"""
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
