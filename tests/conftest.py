import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "connected-components-3d_amd"), os.path.join(ROOT, "oracle"), ROOT):
  if p not in sys.path:
    sys.path.insert(0, p)


def pytest_configure(config):
  config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def have_gpu() -> bool:
  try:
    import torch
    return torch.cuda.is_available()
  except Exception:
    return False


@pytest.fixture(scope="session")
def gpu():
  if not have_gpu():
    pytest.skip("no GPU visible")
  import torch
  return torch.device("cuda:0")
