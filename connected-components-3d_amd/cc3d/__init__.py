"""cc3d for MI355X: connected-components labelling on the GPU.

Drop-in for the labelling path of connected-components-3d 4.0.0
(cc3d/__init__.py:6-17 re-exports; cc3d/fastcc3d.pyx:244-626): the same call
signatures and results, computed by hand-written HIP kernels for gfx950
(libcc3d_amd.so, C ABI in include/cc3d_amd.h).
"""
from ._batch import LabelBatch, connected_components_batch
from ._core import DimensionError, connected_components, estimate_provisional_labels
from ._follow import dust, largest_k, statistics
from ._graphs import color_connectivity_graph, voxel_connectivity_graph
from ._regions import contacts, draw, each, region_graph, runs
from ._stack import connected_components_stack
from . import fastcc3d
from ._serve import LabelServer, Request

__version__ = "4.0.0+mi355x.0.1"

__all__ = [
  "connected_components",
  "connected_components_stack",
  "estimate_provisional_labels",
  "statistics",
  "dust",
  "largest_k",
  "voxel_connectivity_graph",
  "color_connectivity_graph",
  "contacts",
  "region_graph",
  "each",
  "DimensionError",
  "connected_components_batch",
  "LabelBatch",
  "LabelServer",
]
