"""MI355X-native (gfx950 HIP) per-timestep path of the multicellular
phase-field / active-gel organoid simulator.

``host`` binds the C-ABI library ``csrc/liborganoid_hip.so`` (include/organoid.h);
``params`` and ``configs`` hold the constants and the seeded workloads.
"""
from .params import OrgParams, grid_units, template_defaults  # noqa: F401
