"""Module-level names of the reference's extension module (cc3d/fastcc3d.pyx),
for callers that import them from there, e.g. ``cc3d.fastcc3d.runs`` and
``cc3d.fastcc3d.draw``."""
from ._core import DimensionError, connected_components, estimate_provisional_labels
from ._follow import statistics
from ._graphs import color_connectivity_graph, voxel_connectivity_graph
from ._regions import contacts, draw, each, region_graph, runs

__all__ = [
  "DimensionError",
  "connected_components",
  "estimate_provisional_labels",
  "statistics",
  "voxel_connectivity_graph",
  "color_connectivity_graph",
  "contacts",
  "region_graph",
  "runs",
  "draw",
  "each",
]
