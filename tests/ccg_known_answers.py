"""The reference's own color_connectivity_graph cases (automated_test.py:1436-1568),
as functions of the implementation under test: fn(vcg, connectivity=...)."""
import numpy as np


def check_4(fn, dtype, connectivity):
  vcg = np.array([[]], dtype=dtype)
  assert fn(vcg, connectivity=connectivity).size == 0
  vcg = np.zeros([2, 1], dtype=dtype)
  vcg[:] = 0x0f
  assert np.all(fn(vcg, connectivity=connectivity) == np.array([1, 1], dtype=np.uint32))
  vcg = np.zeros([2, 1], dtype=dtype)
  vcg[:] = 0b1100
  assert np.all(fn(vcg, connectivity=connectivity) == np.array([[1], [2]], dtype=np.uint32))
  vcg = np.zeros([1, 2], dtype=dtype)
  vcg[:] = 0b1100
  assert np.all(fn(vcg, connectivity=connectivity) == np.array([1, 1], dtype=np.uint32))
  vcg[:] = 0b0011
  assert np.all(fn(vcg, connectivity=connectivity) == np.array([1, 2], dtype=np.uint32))
  vcg = np.zeros([2, 2], dtype=dtype)
  assert np.all(fn(vcg, connectivity=connectivity) == np.array([[1, 3], [2, 4]], dtype=np.uint32))
  vcg[:] = 0b1111
  vcg[0, 0] = 0b1110
  vcg[1, 0] = 0b1101
  assert np.all(fn(vcg, connectivity=connectivity) == np.array([[1, 1], [1, 1]], dtype=np.uint32))
  vcg[:] = 0b1111
  vcg[0, 0] = 0b1110
  vcg[1, 0] = 0b1101
  vcg[1, 1] = 0b0111
  assert np.all(fn(vcg, connectivity=connectivity) == np.array([[1, 1], [2, 1]], dtype=np.uint32))
  vcg[:] = 0b1111
  vcg[0, 0] = 0b1010
  vcg[1, 0] = 0b1101
  vcg[0, 1] = 0b0111
  vcg[1, 1] = 0b0111
  assert np.all(fn(vcg, connectivity=connectivity) == np.array([[1, 3], [2, 3]], dtype=np.uint32))


def check_8(fn):
  vcg = np.zeros([2, 2], dtype=np.uint8)
  vcg[0, 0] = 0b10000
  vcg[1, 1] = 0b10000000
  vcg[1, 0] = 0b100000
  vcg[0, 1] = 0b1000000
  assert np.all(fn(vcg, connectivity=8) == np.array([[1, 2], [2, 1]], dtype=np.uint32))
  vcg = np.zeros([2, 2], dtype=np.uint32)
  vcg[0, 0] = 0b1000000
  vcg[1, 1] = 0b1000000000
  vcg[1, 0] = 0b10000000
  vcg[0, 1] = 0b100000000
  assert np.all(fn(vcg, connectivity=8) == np.array([[1, 2], [2, 1]], dtype=np.uint32))


def _octants():
  out = np.zeros((10, 10, 10), dtype=np.uint32, order="F")
  out[:5, :5, :5] = 1
  out[5:, :5, :5] = 2
  out[:5, 5:, :5] = 3
  out[5:, 5:, :5] = 4
  out[:5, :5, 5:] = 5
  out[5:, :5, 5:] = 6
  out[:5, 5:, 5:] = 7
  out[5:, 5:, 5:] = 8
  return out


def _cut(vcg):
  vcg[:, :, 4] = vcg[:, :, 4] & 0b101111
  vcg[:, :, 5] = vcg[:, :, 5] & 0b011111
  vcg[:, 4, :] = vcg[:, 4, :] & 0b111011
  vcg[:, 5, :] = vcg[:, 5, :] & 0b110111
  vcg[4, :, :] = vcg[4, :, :] & 0b111110
  vcg[5, :, :] = vcg[5, :, :] & 0b111101


def check_6(fn, dtype):
  vcg = np.zeros([10, 10, 10], dtype=dtype, order="F")
  vcg[:, :, :] = np.iinfo(dtype).max
  _cut(vcg)
  assert np.all(fn(vcg, connectivity=6) == _octants())


def check_26(fn):
  raised = False
  try:
    fn(np.zeros([2, 1, 2], dtype=np.uint8, order="F"), connectivity=26)
  except ValueError:
    raised = True
  assert raised
  vcg = np.zeros([10, 10, 10], dtype=np.uint32, order="F")
  vcg[:, :, :] = 0b11111111111111111111111111
  assert np.all(fn(vcg, connectivity=26) == 1)
  _cut(vcg)
  assert np.all(fn(vcg, connectivity=26) == _octants())
  for x in (4, 5):
    for y in (4, 5):
      for z in (4, 5):
        vcg[x, y, z] = 0b11111111111111111111000000
  assert np.all(fn(vcg, connectivity=26) == 1)


def check_all(fn):
  for dtype in (np.uint8, np.uint32):
    for conn in (4, 8):
      check_4(fn, dtype, conn)
    check_6(fn, dtype)
  check_8(fn)
  check_26(fn)


def random_cases(rng, n):
  """Random graphs (random bits, and graphs extracted from random labels)
  in 2D and 3D at every connectivity and dtype the binding accepts."""
  out = []
  for i in range(n):
    nd = int(rng.choice([2, 3]))
    if nd == 2:
      shape = tuple(int(s) for s in rng.integers(1, 90, 2))
      conn = int(rng.choice([4, 8, 6, 26]))
    else:
      shape = tuple(int(s) for s in rng.integers(1, 40, 3))
      conn = int(rng.choice([6, 26]))
    dt = np.uint32 if (nd == 3 and conn == 26 and shape[2] > 1) else (np.uint8 if rng.random() < 0.5 else np.uint32)
    hi = 2 ** 26 if dt == np.uint32 else 256
    if rng.random() < 0.5:
      vcg = rng.integers(0, hi, size=shape).astype(dt)
      # dense links: most bits set, as in graphs of labelled images
      vcg = vcg | (rng.integers(0, hi, size=shape).astype(dt) & rng.integers(0, hi, size=shape).astype(dt))
    else:
      lab = rng.integers(0, 3, size=shape).astype(np.uint32)
      vcg = _vcg_of(lab, dt, nd, conn)
    if rng.random() < 0.5:
      vcg = np.asfortranarray(vcg)
    out.append((vcg, conn))
  return out


def _vcg_of(lab, dt, nd, conn):
  """Bits set toward equal-label in-volume neighbours (the graph's layout),
  every direction out of the volume set."""
  if nd == 2 and dt == np.uint8:
    dirs = [(+1, 0, 0), (-1, 0, 0), (0, +1, 0), (0, -1, 0), (+1, +1, 0), (-1, +1, 0), (+1, -1, 0), (-1, -1, 0)]
  else:
    dirs = [(+1, 0, 0), (-1, 0, 0), (0, +1, 0), (0, -1, 0), (0, 0, +1), (0, 0, -1), (+1, +1, 0), (-1, +1, 0),
            (+1, -1, 0), (-1, -1, 0), (+1, 0, +1), (-1, 0, +1), (0, +1, +1), (0, -1, +1), (+1, 0, -1), (-1, 0, -1),
            (0, +1, -1), (0, -1, -1), (+1, +1, +1), (-1, +1, +1), (+1, -1, +1), (-1, -1, +1), (+1, +1, -1),
            (-1, +1, -1), (+1, -1, -1), (-1, -1, -1)]
    if dt == np.uint8:
      dirs = dirs[:6]
  L = lab if lab.ndim == 3 else lab[..., None]
  g = np.zeros(L.shape, dtype=np.uint64)
  sx, sy, sz = L.shape
  for b, (dx, dy, dz) in enumerate(dirs):
    sh = np.roll(L, (-dx, -dy, -dz), axis=(0, 1, 2))
    ok = np.ones(L.shape, bool)
    xs, ys, zs = np.arange(sx)[:, None, None], np.arange(sy)[None, :, None], np.arange(sz)[None, None, :]
    inb = (xs + dx >= 0) & (xs + dx < sx) & (ys + dy >= 0) & (ys + dy < sy) & (zs + dz >= 0) & (zs + dz < sz)
    ok = np.where(inb, sh == L, True)
    g |= ok.astype(np.uint64) << np.uint64(b)
  g = g.astype(dt)
  return g if lab.ndim == 3 else g[..., 0]
