"""CPU checks of the boundary: the C-ABI library loads and exports every symbol include/*.h declares
(no compute call: there is no GPU here), the package fails loudly without its library, and the
fs_params layout matches the header."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            src = open(os.path.join(ROOT, "include", f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"\b(fs_[a-z_0-9]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_header():
    from floodsim_amd import solver
    assert os.path.exists(solver.LIB_PATH), "run __graft_entry__.build() first"
    L = ctypes.CDLL(solver.LIB_PATH)
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(solver.EXPORTS) <= set(names)


def test_version_string():
    from floodsim_amd import solver
    L = solver.lib()
    assert b"gfx950" in L.fs_version()


def test_params_layout():
    from floodsim_amd.solver import fs_params
    # int32 nx, ny; 6 doubles; int32 secondOrder, limiter, bc[4], lts, rows, untiled, hazard,
    # implicitFriction, stepFlux
    assert ctypes.sizeof(fs_params) == 8 + 6 * 8 + 12 * 4
    assert fs_params.untiled.offset == 8 + 6 * 8 + 8 * 4
    assert fs_params.hazard.offset == 8 + 6 * 8 + 9 * 4
    assert fs_params.stepFlux.offset == 8 + 6 * 8 + 11 * 4


def test_missing_library_raises(tmp_path):
    from floodsim_amd import solver
    with pytest.raises(solver.FloodSimError):
        saved = solver._LIB
        solver._LIB = None
        try:
            solver.lib(str(tmp_path / "nope.so"))
        finally:
            solver._LIB = saved
