"""The clearance map behind the kernel's skipping of the DOM search
(ppc_hip_clearance_map): every entry must be a lower bound of the horizontal
distance from any point of its square to the nearest string cylinder, checked
against brute force on random points of the 86-string and 7-string geometries,
tight to a square's diagonal below its 100 m cap; and it must be useful (most of
the footprint well clear of the strings)."""
import ctypes

import numpy as np
import pytest


class Line(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("max", ctypes.c_int32), ("x", ctypes.c_float), ("y", ctypes.c_float),
                ("r", ctypes.c_float), ("h", ctypes.c_float), ("d", ctypes.c_float), ("dl", ctypes.c_float),
                ("dh", ctypes.c_float)]


# squares of 1 m; the 200-string detector's footprint needs 2 m squares to stay under the
# 4M entries the upload allows
@pytest.mark.parametrize("name,step", [("c1", 1.0), ("c2", 1.0), ("cbig", 2.0)])
def test_clearance_map_is_a_lower_bound(name, step):
    from ppc_amd import api, configs

    det = api.Detector(configs.ensure(name), {"WFLA": None, "FWID": None})
    lib = api.hip()
    lib.ppc_hip_clearance_map.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_size_t] + \
        [ctypes.POINTER(ctypes.c_int32)] * 2 + [ctypes.POINTER(ctypes.c_float)] * 2
    nx, ny, x0, y0 = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float(), ctypes.c_float()
    assert lib.ppc_hip_clearance_map(det.tables, step, None, 0, nx, ny, x0, y0) == 0
    cm = np.zeros(nx.value * ny.value, dtype=np.float32)
    assert lib.ppc_hip_clearance_map(det.tables, step, cm.ctypes.data_as(ctypes.c_void_p), cm.size, nx, ny, x0,
                                     y0) == 0
    cm = cm.reshape(ny.value, nx.value)
    lines = np.frombuffer(det.field("sc"), dtype=np.dtype([(f, "<i4" if f in ("n", "max") else "<f4")
                                                           for f, _ in Line._fields_]))
    hr = float(det.scalar("hr"))
    rr = np.maximum(lines["r"], hr)
    rng = np.random.default_rng(1)
    px = x0.value + rng.uniform(0, nx.value * step, 200000)
    py = y0.value + rng.uniform(0, ny.value * step, 200000)
    d = np.min(np.hypot(px[:, None] - lines["x"][None], py[:, None] - lines["y"][None]) - rr[None], axis=1)
    ix = np.floor((px - x0.value) / step).astype(int)
    iy = np.floor((py - y0.value) / step).astype(int)
    ok = (ix >= 0) & (iy >= 0) & (ix < nx.value) & (iy < ny.value)
    bound = cm[iy[ok], ix[ok]]
    assert np.all(bound <= d[ok] + 1e-4)
    # and tight to a diagonal, below the 100 m cap of the entries
    assert np.all(np.minimum(d[ok], 100.0) - bound <= 2 * step * 0.7072 + 0.05 + 1e-3)
    assert np.mean(bound > 5.0) > 0.8
