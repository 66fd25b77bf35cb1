"""floodsim_amd: FloodSim's 2D shallow-water step on MI355X (gfx950).

The product is the C-ABI library libfloodsim_hip.so (include/floodsim_hip.h), its N-API addon
for Node (csrc/napi_addon.c, js/solver-hip.js) and this package's ctypes binding (solver.py).
"""
from .solver import GPUSolver, Pool, FloodSimError, lib  # noqa: F401
