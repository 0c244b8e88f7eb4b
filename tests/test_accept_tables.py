"""The tables the device hit post-processing takes (xppc_c_accept_tables), built
by the front-end from the same detector description print() uses, checked on
the CPU against the detector's own fields: one sensor per type with its
angular-sensitivity polynomial, the wavelength bin edges, and per DOM the
radius and axis ratio of the geometry."""
import ctypes

import numpy as np

from test_gpu_accept import AcceptTables

WNUM = 512


def _tables(icedir, env):
    from ppc_amd import api

    det = api.Detector(icedir, env)
    return det, AcceptTables.from_address(det.lib.xppc_c_accept_tables())


def test_accept_tables_follow_the_detector(c1_dir, c2_dir):
    for icedir in (c1_dir, c2_dir):
        det, T = _tables(icedir, {"WFLA": None, "FWID": None})
        gsize = int(det.scalar("gsize", np.int32))
        assert T.gsize == gsize and T.nsensor >= 1 and T.nacc >= 1
        oms = det.array("oms", np.float32).reshape(-1, 5)
        doms = [T.doms[i] for i in range(gsize)]
        assert np.allclose([d.R for d in doms], oms[:, 0]) and np.allclose([d.F for d in doms], oms[:, 1])
        assert all(-1 <= d.sensor < T.nsensor for d in doms)
        assert all(0 <= d.acc < T.nacc for d in doms if d.sensor >= 0)
        assert sum(d.sensor >= 0 for d in doms) > 0.9 * gsize
        wvs = np.ctypeslib.as_array(T.wvs, shape=(2 * WNUM,)).reshape(WNUM, 2)
        ref = det.array("wvs", np.float32).reshape(WNUM, 3)
        assert np.array_equal(wvs[:, 0], ref[:, 1]) and np.array_equal(wvs[:, 1], ref[:, 2])
        rat = np.ctypeslib.as_array(T.rat, shape=(T.nacc * WNUM,)).reshape(T.nacc, WNUM)
        assert np.all(np.isfinite(rat)) and np.all(rat >= 0)  # relative to the spectrum average (may exceed 1)
        for k in range(T.nsensor):
            S = T.sensors[k]
            assert (S.def_ == 1 and 1 <= S.ns <= 16) or (S.def_ == 0 and S.npmt >= 1)


def test_accept_tables_of_multi_pmt_modules():
    """the nextgen ice (om.conf: a 12-PMT and a 2-PMT module on strings 1 and 2): their DOMs
    point to sensors with those PMT counts and the modules' radii and axis ratios; the other
    DOMs keep the one-PMT angular-sensitivity sensor"""
    from ppc_amd import configs

    det, T = _tables(configs.ensure("c1ng"), {"WFLA": "405", "FWID": "10"})
    gsize = int(det.scalar("gsize", np.int32))
    strs = det.array("str", np.int32)
    doms = [T.doms[i] for i in range(gsize)]
    for s, npmt, R, Rz in ((1, 12, 0.16, 0.205), (2, 2, 0.15, 0.27)):
        sel = [d for d, k in zip(doms, strs) if k == s]
        assert len(sel) == 60
        S = T.sensors[sel[0].sensor]
        assert S.def_ == 0 and S.npmt == npmt
        assert all(d.sensor == sel[0].sensor for d in sel)
        assert np.allclose([d.R for d in sel], R) and np.allclose([d.F for d in sel], Rz / R, rtol=1e-6)
    other = [d for d, k in zip(doms, strs) if k > 2]
    assert all(T.sensors[d.sensor].def_ == 1 for d in other)
