"""contacts / region_graph and runs / draw / each on the GPU
(cc3d/fastcc3d.pyx:1180-1372, cc3d/cc3d_graphs.hpp:238-523) against the
oracle restatement and the reference compiled from its header (contact sums
compared bit for bit), and the reference's own known answers."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_contacts_reference_known_answers(gpu):
  import cc3d
  import region_known_answers as ka
  ka.check_contacts(cc3d.contacts, cc3d.region_graph, cc3d.connected_components)


def test_contacts_match_oracle(gpu):
  import cc3d
  import region_known_answers as ka
  rng = np.random.default_rng(23)
  for i, (lab, conn, sa, an) in enumerate(ka.random_label_volumes(rng, 160)):
    print("contacts case", i, lab.shape, lab.dtype, conn, sa, an, flush=True)
    got = cc3d.contacts(lab, connectivity=conn, surface_area=sa, anisotropy=an)
    exp = oracle.contacts(lab, connectivity=conn, surface_area=sa, anisotropy=an)
    assert got == exp, i
    if oracle.have_ref() and i % 4 == 0:
      assert got == oracle.contacts(lab, connectivity=conn, surface_area=sa, anisotropy=an, backend="ref")
    assert cc3d.region_graph(lab, connectivity=conn) == set(
      oracle.contacts(lab, connectivity=conn, anisotropy=an).keys())


@pytest.mark.parametrize("anisotropy,surface_area", [((4, 4, 40), True), ((1.5, 2.3, 0.7), True),
                                                     ((1, 1, 1), False)])
def test_contacts_full_size(gpu, anisotropy, surface_area):
  """A 256^3 Voronoi labelling: exact sums (integer weights), the replayed
  raster-order sums (non-dyadic weights) and neighbour counts."""
  import cc3d
  import synth
  lab = synth.voronoi((256, 256, 256), 3000, seed=13)
  got = cc3d.contacts(lab, connectivity=26, surface_area=surface_area, anisotropy=anisotropy)
  exp = oracle.contacts(lab, connectivity=26, surface_area=surface_area, anisotropy=anisotropy)
  assert len(got) > 1000 and got == exp


def test_contacts_device_tensor_and_errors(gpu):
  import torch
  import cc3d
  rng = np.random.default_rng(2)
  lab = rng.integers(0, 9, size=(30, 20, 10)).astype(np.int32)
  t = torch.from_numpy(lab).cuda()
  assert cc3d.contacts(t, connectivity=18) == oracle.contacts(lab, connectivity=18)
  with pytest.raises(ValueError):
    cc3d.contacts(lab, connectivity=10)
  with pytest.raises(RuntimeError, match="z thickness"):
    cc3d.contacts(lab, connectivity=8)
  with pytest.raises(TypeError):
    cc3d.contacts(lab.astype(np.float32))
  with pytest.raises(IndexError):
    cc3d.contacts(np.zeros((0, 4), np.uint32))


def test_runs_match_oracle_and_reference(gpu):
  import cc3d
  rng = np.random.default_rng(5)
  cases = [np.zeros(0, np.uint8), np.zeros(1, np.uint8), np.array([7], np.uint32), np.zeros((3, 4), np.uint16)]
  for i in range(60):
    nd = int(rng.integers(1, 4))
    shape = tuple(int(s) for s in rng.integers(1, 30, nd))
    dt = [np.uint8, np.uint16, np.uint32, np.uint64, bool][i % 5]
    lab = rng.integers(0, 4, size=shape).astype(dt)
    if i % 3 == 1:
      lab = np.asfortranarray(lab)
    if i % 7 == 2 and nd >= 2:
      lab = lab[:, ::2]  # not contiguous: read in C order
    cases.append(lab)
  big = (rng.random((1 << 22,)) < 0.3).astype(np.uint64) * rng.integers(1, 1 << 40, size=1 << 22).astype(np.uint64)
  cases.append(big)
  for i, lab in enumerate(cases):
    print("runs case", i, lab.shape, lab.dtype, flush=True)
    got = cc3d.fastcc3d.runs(lab)
    exp = oracle.runs(lab)
    assert got == exp and list(got) == list(exp), i
    if oracle.have_ref() and lab.size < 100000:
      assert got == oracle.runs(lab, backend="ref")
  with pytest.raises(TypeError):
    cc3d.fastcc3d.runs(np.zeros(4, np.int32))


def test_draw_matches_restatement(gpu):
  import cc3d
  rng = np.random.default_rng(9)
  for dt in (np.uint8, np.uint16, np.uint32, np.uint64, bool):
    for order in ("C", "F"):
      img = np.zeros((13, 11, 7), dtype=dt, order=order)
      exp = img.copy(order="K")
      rns = [(0, 5), (9, 40), (100, 101), (500, 1001)]
      got = cc3d.fastcc3d.draw(3, rns, img)
      oracle.draw(3, rns, exp)
      assert np.array_equal(img, exp) and got.ndim == 1 and np.shares_memory(got, img)
      # an invalid run: the runs before it are painted, then RuntimeError
      img2 = np.zeros_like(img)
      exp2 = np.zeros_like(img)
      bad = [(1, 4), (20, 10), (30, 35)]
      with pytest.raises(RuntimeError, match="Invalid run."):
        cc3d.fastcc3d.draw(2, bad, img2)
      with pytest.raises(RuntimeError, match="Invalid run."):
        oracle.draw(2, bad, exp2)
      assert np.array_equal(img2, exp2)
      with pytest.raises(RuntimeError, match="Invalid run."):
        cc3d.fastcc3d.draw(2, [(0, img.size + 1)], np.zeros_like(img))
  # a non-contiguous image: the returned C-order copy is painted, the image is not
  base = np.zeros((8, 8), np.uint32)
  view = base[::2, ::2]  # no flat view exists: reshape copies
  out = cc3d.fastcc3d.draw(5, [(0, 3)], view)
  assert out[:3].tolist() == [5, 5, 5] and base.sum() == 0


def test_each_reference_known_answer(gpu):
  """automated_test.py:660-676 through the product."""
  import cc3d
  import region_known_answers as ka
  ka.check_each(cc3d.each, side=128)


def test_each_device_tensor(gpu):
  import torch
  import cc3d
  rng = np.random.default_rng(4)
  lab = rng.integers(0, 5, size=(40, 30, 20)).astype(np.uint16)
  t = torch.from_numpy(lab).cuda()
  n = 0
  for label, img in cc3d.each(t, binary=False):
    assert img.device.type == "cuda"
    assert np.array_equal(img.cpu().numpy(), label * (lab == label))
    n += 1
  assert n == 4
