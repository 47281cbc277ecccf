"""The reference's own cases for contacts / region_graph and each
(automated_test.py:666-877), as functions of the implementation under test."""
import numpy as np


def check_contacts(contacts, region_graph, connected_components=None):
  labels = np.zeros((10, 10, 10), dtype=np.uint32)
  labels[0, 0, 0] = 1
  labels[1, 0, 0] = 2
  labels[0, 1, 0] = 3
  labels[0, 0, 1] = 4
  labels[1, 1, 1] = 5
  res = contacts(labels, surface_area=True)
  assert res[(1, 2)] == 1 and res[(1, 3)] == 1 and res[(1, 4)] == 1 and res[(1, 5)] == 0
  assert (1, 9) not in res
  res = contacts(labels, surface_area=True, anisotropy=(2, 2, 2))
  assert res[(1, 2)] == 4 and res[(1, 3)] == 4 and res[(1, 4)] == 4 and res[(1, 5)] == 0
  res = contacts(labels, surface_area=True, anisotropy=(1.5, 10, 10))
  assert res[(1, 2)] == 100 and res[(1, 3)] == 15 and res[(1, 4)] == 15 and res[(1, 5)] == 0
  res = contacts(labels, surface_area=False)
  assert res[(1, 2)] == 1 and res[(1, 3)] == 1 and res[(1, 4)] == 1 and res[(1, 5)] == 1
  if connected_components is not None:
    data = np.ones((3, 3, 3))
    data[1, 1, 1] = 2
    labels_out = connected_components(data, connectivity=6)
    assert contacts(labels_out, connectivity=6, surface_area=True)[(1, 2)] == 6
  data = np.ones((3, 3), dtype=np.uint32)
  data[1, 1] = 2
  res = contacts(data, connectivity=8, surface_area=True, anisotropy=(1.5, 2.3))
  assert np.isclose(res[(1, 2)], (1.5 * 2 + 2.3 * 2))
  res = contacts(data, connectivity=4, surface_area=True, anisotropy=(1.5, 2.3))
  assert np.isclose(res[(1, 2)], (1.5 * 2 + 2.3 * 2))

  # 26
  labels = np.zeros((10, 10, 10), dtype=np.uint32)
  labels[5, 5, 5] = 1
  labels[6, 6, 6] = 2
  labels[4, 4, 6] = 3
  labels[4, 6, 6] = 4
  labels[6, 4, 6] = 5
  labels[4, 4, 4] = 6
  labels[4, 6, 4] = 7
  labels[6, 4, 4] = 8
  labels[6, 6, 4] = 9
  labels[1, :, :] = 10
  assert region_graph(labels, connectivity=26) == set([(1, 2), (1, 3), (1, 4), (1, 5), (1, 6), (1, 7), (1, 8), (1, 9)])
  assert region_graph(labels, connectivity=18) == set()
  assert region_graph(labels, connectivity=6) == set()

  # 18
  labels = np.zeros((10, 10, 10), dtype=np.uint32)
  labels[5, 5, 5] = 1
  labels[5, 6, 6] = 2
  labels[5, 4, 6] = 3
  labels[6, 5, 6] = 4
  labels[4, 5, 6] = 5
  labels[5, 4, 4] = 6
  labels[5, 6, 4] = 7
  labels[6, 5, 4] = 8
  labels[4, 5, 4] = 9
  labels[1, :, :] = 10
  e18 = set([(1, 2), (1, 3), (1, 4), (1, 5), (1, 6), (1, 7), (1, 8), (1, 9),
             (2, 4), (2, 5), (3, 4), (3, 5), (6, 8), (6, 9), (7, 8), (7, 9)])
  assert set(contacts(labels, connectivity=26).keys()) == e18
  assert set(contacts(labels, connectivity=18).keys()) == e18
  assert set(contacts(labels, connectivity=6).keys()) == set()

  # 6
  labels = np.zeros((10, 10, 10), dtype=np.uint32)
  labels[5, 5, 5] = 1
  labels[5, 5, 6] = 2
  labels[5, 5, 4] = 3
  labels[4, 5, 5] = 4
  labels[6, 5, 5] = 5
  labels[5, 4, 5] = 6
  labels[5, 6, 5] = 7
  labels[1, :, :] = 10
  e6 = set([(1, 2), (1, 3), (1, 4), (1, 5), (1, 6), (1, 7), (2, 4), (2, 5), (2, 6), (2, 7),
            (3, 4), (3, 5), (3, 6), (3, 7), (4, 6), (4, 7), (5, 6), (5, 7)])
  assert set(contacts(labels, connectivity=26).keys()) == e6
  assert set(contacts(labels, connectivity=18).keys()) == e6
  assert set(contacts(labels, connectivity=6).keys()) == set([(1, 2), (1, 3), (1, 4), (1, 5), (1, 6), (1, 7)])

  # 2D
  labels = np.zeros((10, 10), dtype=np.uint32)
  labels[5, 5] = 1
  labels[3, 2] = 2
  labels[3, 3] = 3
  labels[7, 5] = 4
  labels[8, 5] = 5
  labels[7, 2] = 6
  labels[8, 2] = 7
  labels[9, 9] = 8
  labels[8, 8] = 9
  labels[6, 8] = 10
  labels[5, 9] = 11
  labels[1, :] = 12
  assert set(contacts(labels, connectivity=4).keys()) == set([(2, 3), (4, 5), (6, 7)])
  assert set(contacts(labels, connectivity=8).keys()) == set([(2, 3), (4, 5), (6, 7), (8, 9), (10, 11)])


def check_each(each, dtypes=(np.uint8, np.uint16, np.uint32, np.uint64), side=128, seed=0):
  rng = np.random.default_rng(seed)
  for dtype in dtypes:
    for order in ("C", "F"):
      for in_place in (True, False):
        for dims in (1, 2, 3):
          shape = [side] * dims
          labels = rng.integers(0, 3, shape).astype(dtype)
          labels = np.asfortranarray(labels) if order == "F" else np.ascontiguousarray(labels)
          seen = []
          for label, img in each(labels, binary=False, in_place=in_place):
            assert np.all(img == (label * (labels == label)))
            seen.append(label)
          assert seen == sorted(int(v) for v in np.unique(labels) if v != 0)
          for label, img in each(labels, binary=True, in_place=in_place):
            assert img.dtype == np.bool_
            assert np.all(img == (labels == label))
          assert len(each(labels)) == len(seen)


def random_label_volumes(rng, n):
  """(labels, connectivity, surface_area, anisotropy) cases over every label
  dtype, 1-3 axes, C and F order, dense and sparse labels."""
  out = []
  types = [np.uint8, np.uint16, np.uint32, np.uint64, np.int8, np.int16, np.int32, np.int64, bool]
  anis = [(1, 1, 1), (2, 2, 2), (4, 4, 40), (1.5, 10, 10), (1.5, 2.3, 0.7), (0.1, 0.2, 0.3), (3,), (), (2.5, 0.5)]
  for i in range(n):
    nd = int(rng.integers(1, 4))
    shape = tuple(int(s) for s in rng.integers(1, 24 if nd == 3 else 70, nd))
    dt = types[int(rng.integers(0, len(types)))]
    if dt is bool:
      lab = rng.random(shape) < 0.5
    else:
      lo = -4 if np.dtype(dt).kind == "i" else 0
      k = int(rng.choice([3, 6, 40]))
      lab = rng.integers(lo, k, size=shape).astype(dt)
      if rng.random() < 0.3 and np.dtype(dt).itemsize == 8:
        lab = (lab.astype(np.int64) * (1 << 40)).astype(dt)  # labels beyond 32 bits
    if rng.random() < 0.4:
      lab = np.asfortranarray(lab)
    conns = [6, 18, 26] + ([4, 8] if nd <= 2 else [])
    conn = int(conns[int(rng.integers(0, len(conns)))])
    out.append((lab, conn, bool(rng.random() < 0.6), anis[int(rng.integers(0, len(anis)))]))
  return out
