"""The oracle's guard against non-finite optical depths (the reference's GPU build,
gpu/pro.cu:727-729), forced from the host on the CPU: with a NaN in one layer's
absorption, a photon that reaches the layer ends there and is counted; one that never
reaches it is unchanged (same draws, same hits as without the NaN).  The oracle here is its
build that keeps NaN semantics (oracle/liboracle_nan.so; the kernel is built the same way):
under --fast-math alone a NaN may be folded away by a comparison."""
import ctypes

import numpy as np

import oracle_util as U


def test_nan_layer_ends_photons_that_reach_it(c1_dir):
    from ppc_amd import api

    det = api.Detector(c1_dir, {"WFLA": None, "FWID": None})
    src = (2.0, 1.0, 3.0)
    hmin, rdh = float(det.scalar("hmin")), float(det.scalar("rdh"))
    J = int(round((src[2] - hmin) * rdh)) + 3
    T0 = U.product_tables(det)
    T1 = U.product_tables(det)
    ab = np.frombuffer(bytes(T1.field("abs_")), np.float32).reshape(U.WNUM, U.MAXLYS).copy()
    ab[:, J] = np.nan
    ctypes.memmove(T1.field("abs_"), ab.tobytes(), ab.nbytes)
    rm = det.array("rm", np.uint32)
    lib = U.oracle(nan=True)
    b = U.point_bunch(*src, num=1)
    H0, H1 = (U.Hit * 64)(), (U.Hit * 64)()
    rng = np.random.default_rng(4)
    ended = unchanged = 0
    n = 4000
    for k in range(n):
        st = (int(rng.integers(1, 2 ** 32)), int(rng.integers(0, 2 ** 30)), int(rm[k]))
        c0, c1 = U.Counters(), U.Counters()
        s0, s1 = U.Rng(*st), U.Rng(*st)
        m0 = lib.orc_photon(T0.ptr, ctypes.byref(b), 1, 0, ctypes.byref(s0), H0, 64, ctypes.byref(c0))
        m1 = lib.orc_photon(T1.ptr, ctypes.byref(b), 1, 0, ctypes.byref(s1), H1, 64, ctypes.byref(c1))
        assert c0.nonfinite == 0 and c1.nonfinite <= 1
        if c1.nonfinite:
            ended += 1
            assert c1.steps <= c0.steps  # ended no later than without the NaN
        else:
            unchanged += (s0.x, s0.c) == (s1.x, s1.c) and m0 == m1 and c0.steps == c1.steps
    assert ended > 200
    assert unchanged == n - ended
