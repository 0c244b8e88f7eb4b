"""The tables the device hit post-processing takes (xppc_c_accept_tables), built
by the front-end from the same detector description print() uses, checked on
the CPU against the detector's own fields: one sensor per type with its
angular-sensitivity polynomial, the wavelength bin edges, and per DOM the
radius and axis ratio of the geometry."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_util as U

from test_gpu_accept import AcceptTables

WNUM = 512
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


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


_REF = r'''
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[2])
import oracle_util as U
from test_gpu_accept import AcceptTables
from ppc_amd import api
R = U.ref_tables(sys.argv[1])
ref = U.reference()
ref.ref_dom_accept.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int]
D = api.Detector(sys.argv[1])
T = AcceptTables.from_address(D.lib.xppc_c_accept_tables())
W = 512
rat = np.ctypeslib.as_array(T.rat, shape=(T.nacc * W,)).reshape(T.nacc, W) if T.nacc else None
out, rr = (ctypes.c_float * 10)(), (ctypes.c_float * W)()
bad = used = tilted = cabled = 0
for i in range(T.gsize):
    m = ref.ref_dom_accept(i, out, rr, W)
    omt, typ, rde, hv, azi = out[0], out[1], out[2], out[3], out[4]
    tilt, rmax, crde = np.array(out[5:8]), out[8], out[9]
    d = T.doms[i]
    want = hv > 0 and crde > 0 and m == W
    ok = (d.sensor >= 0) == want and abs(d.azi - azi) <= 1e-6 * max(1.0, abs(azi))
    ok = ok and np.allclose(list(d.tilt), tilt, rtol=0, atol=1e-6)
    if want and d.sensor >= 0:
        used += 1
        ok = ok and abs(d.rde - rde / rmax) <= 1e-6 * (rde / rmax)
        ok = ok and np.allclose(rat[d.acc], np.array(rr[:W]), rtol=1e-6, atol=0)
    tilted += int(abs(tilt[0]) > 1e-6 or abs(tilt[1]) > 1e-6)  # off the vertical
    cabled += int(azi >= 0)
    if not ok:
        bad += 1
        if bad < 5:
            print("MISMATCH dom", i, list(out), d.sensor, d.acc, d.rde, d.azi, list(d.tilt))
print("DOMS", T.gsize, "USED", used, "TILTED", tilted, "CABLED", cabled, "BAD", bad)
'''


@pytest.mark.ref
@pytest.mark.parametrize("cfg", ["c1", "c2x", "c2xb", "c1ng"])
def test_accept_tables_match_reference(cfg):
    """per DOM, the post-processing tables against the reference's own (gpu/ini.cxx:929-1014,
    1425-1440): whether it records (high voltage and an acceptance class), its efficiency over
    its class maximum (eff-f2k), cable azimuth (dx.dat), axis (cx.dat), and its class's
    acceptance per wavelength bin (wv.rde for the high-QE DOMs, om.wv_<type> for nextgen)"""
    if not U.ref_available():
        pytest.skip("oracle/_ref not built (no reference tree)")
    from ppc_amd import configs

    env = dict(os.environ)
    for k in ("WFLA", "FWID"):
        env.pop(k, None)
    env.update({"DTMN": "1", "PYTHONPATH": ROOT + os.pathsep + HERE})
    r = subprocess.run([sys.executable, "-c", _REF, configs.ensure(cfg), HERE], env=env, capture_output=True,
                       text=True, timeout=300, cwd=HERE)
    assert r.returncode == 0, r.stderr[-3000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("DOMS")][-1].split()
    print(" ".join(line))
    assert line[-1] == "0", r.stdout[-3000:]
    assert int(line[3]) > 0.5 * int(line[1])
    if cfg == "c2xb":
        assert int(line[5]) == 9 * 60  # the DOMs of cx.dat
    if cfg.startswith("c2x"):
        assert int(line[7]) == int(line[1])  # every DOM has a cable position
