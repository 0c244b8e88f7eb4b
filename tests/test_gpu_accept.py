"""Hit post-processing on the device (ppc_hip_accept, PPC_ACCEPT=device) against
the reference's host post-processing (print(), gpu/f2k.cxx:247-363).

The device draws its random numbers from a counter-based generator, the host
from the reference's stream, so the two agree statistically:
  * per raw hit, the accepted photon's impact point and direction are a
    deterministic function of the hit and must match the host formulas;
  * the number of accepted hits must match the sum over hits of the acceptance
    probability (relative DOM efficiency x angular sensitivity, cable shadow)
    within 5 sigma of its binomial spread;
  * through the library pipeline (flasher on the 7-string ice, deterministic
    seeds, so both runs see the same raw hits), host and device acceptance give
    per-DOM counts that agree (two-sample chi2 p > 0.01) and the same
    wavelength distribution (KS p > 0.01).
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WNUM = 512


class Sensor(ctypes.Structure):
    _fields_ = [("def_", ctypes.c_int32), ("ns", ctypes.c_int32), ("npmt", ctypes.c_int32), ("pmt0", ctypes.c_int32),
                ("mas", ctypes.c_float), ("beta", ctypes.c_float), ("fx", ctypes.c_float), ("Rr", ctypes.c_float),
                ("Rz", ctypes.c_float), ("cable", ctypes.c_float), ("s", ctypes.c_float * 16)]


class DomAccept(ctypes.Structure):
    _fields_ = [("sensor", ctypes.c_int32), ("acc", ctypes.c_int32), ("rde", ctypes.c_float), ("azi", ctypes.c_float),
                ("tilt", ctypes.c_float * 3), ("R", ctypes.c_float), ("F", ctypes.c_float), ("pad", ctypes.c_float)]


class AcceptTables(ctypes.Structure):
    _fields_ = [("nsensor", ctypes.c_int32), ("sensors", ctypes.POINTER(Sensor)), ("npmt", ctypes.c_int32),
                ("pmt_dirs", ctypes.POINTER(ctypes.c_float)), ("nacc", ctypes.c_int32),
                ("rat", ctypes.POINTER(ctypes.c_float)), ("gsize", ctypes.c_int32),
                ("doms", ctypes.POINTER(DomAccept)), ("wvs", ctypes.POINTER(ctypes.c_float))]


def test_device_acceptance_matches_host_formulas(c1_dir):
    from ppc_amd import api

    det = api.Detector(c1_dir, {"WFLA": "405", "FWID": "10", "DTMN": "1"})
    T = AcceptTables.from_address(det.lib.xppc_c_accept_tables())
    assert T.nsensor >= 1 and T.gsize == 420
    dev = api.Device(0, deterministic=True)
    dev.upload(det)
    dev.upload_accept(det)
    b = det.flasher_bunch(1, 30, num=40_000_000)
    print("device acceptance: c1 flasher, %d photons" % b.num, flush=True)
    dev.run([b])
    h = dev.hits()
    a = dev.accept(seed=12345)
    # the same launch again: its hits stored in another order, post-processed identically
    dev.seed(0)
    dev.run([b])
    h2 = dev.hits()
    a2 = dev.accept(seed=12345)
    dev.close()
    key = lambda x: np.lexsort((x["dph"], x["dth"], x["pph"], x["pth"], x["z"], x["n"], x["t"], x["i"]))
    o1, o2 = key(h), key(h2)
    assert np.array_equal(h[o1], h2[o2]) and np.array_equal(a[o1], a2[o2])
    assert len(a) == len(h) > 20000
    ok = a["pmt"] >= 0
    # geometry of the accepted photons (gpu/f2k.cxx:279-283, 344-351)
    dom = np.array([(T.doms[i].R, T.doms[i].F) for i in range(T.gsize)], dtype=np.float32)
    R, F = dom[h["i"], 0], dom[h["i"], 1]
    f32 = np.float32
    nx = np.sin(h["pth"]) * np.cos(h["pph"])
    ny = np.sin(h["pth"]) * np.sin(h["pph"])
    nz = np.cos(h["pth"])
    dts = np.where(F < 0, f32(1), np.sin(h["dth"]))
    rx, ry, rz = dts * np.cos(h["dph"]), dts * np.sin(h["dph"]), F * np.cos(h["dth"])
    assert np.allclose(a["n"][ok], np.stack([nx, ny, nz], 1)[ok], atol=2e-6)
    assert np.allclose(a["r"][ok], -R[ok, None] * np.stack([rx, ry, rz], 1)[ok], atol=2e-6)
    # wavelengths inside their bins
    wvs = np.ctypeslib.as_array(T.wvs, shape=(WNUM * 2,)).reshape(WNUM, 2)
    lo, hi = wvs[h["z"], 0], wvs[h["z"], 1]
    assert np.all((a["wv"][ok] >= lo[ok]) & (a["wv"][ok] <= hi[ok]))
    # acceptance probability per hit: efficiency x angular sensitivity (one-PMT sensors here)
    rat = np.ctypeslib.as_array(T.rat, shape=(T.nacc * WNUM,)).reshape(T.nacc, WNUM)
    p = np.zeros(len(h))
    for k, hit in enumerate(h):
        d = T.doms[int(hit["i"])]
        if d.sensor < 0:
            continue
        S = T.sensors[d.sensor]
        assert S.def_ == 1 and d.azi < 0  # this ice: one PMT, no cable
        pe = min(1.0, rat[d.acc, hit["z"]] * d.rde)
        x = nx[k] * d.tilt[0] + ny[k] * d.tilt[1] + nz[k] * d.tilt[2]
        poly = sum(S.s[j] * x ** j for j in range(S.ns))
        pa = min(1.0, max(0.0, poly / S.mas)) if S.mas > 0 else float(rx[k] * d.tilt[0] + ry[k] * d.tilt[1]
                                                                       + rz[k] * d.tilt[2] > S.s[0])
        p[k] = pe * pa
    want, sd = p.sum(), np.sqrt((p * (1 - p)).sum())
    print("accepted %d of %d, expected %.1f +- %.1f" % (ok.sum(), len(h), want, sd), flush=True)
    assert abs(ok.sum() - want) < 5 * sd + 1


_LIB_RUN = r"""
import json, os, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from ppc_amd import api
det = api.Detector(sys.argv[2], {"WFLA": "405", "FWID": "10", "DTMN": "1"})
det.start(0)
det.lib.xppc_c_flset(1, 30)
det.lib.xppc_c_sett(0.0, 0.0, 1.0, 1, 0, 0, 0.0)
det.lib.xppc_c_flone(int(float(sys.argv[3])))
det.lib.xppc_c_eout()
h = det.hitz()
p, raw, ms = det.totals()
det.lib.xppc_c_efin()
det.stop()
print(json.dumps({"str": h["str"].tolist(), "dom": h["dom"].tolist(), "wv": h["wv"].tolist(), "raw": raw,
                  "photons": p}))
"""


def test_library_device_acceptance_matches_host(c1_dir):
    out = {}
    for mode in ("host", "device"):  # one process each: the host draws advance the shared seed
        env = dict(os.environ)
        env.pop("PPC_ACCEPT", None)
        if mode == "device":
            env["PPC_ACCEPT"] = "device"
        print("library pipeline, %s acceptance: c1 flasher" % mode, flush=True)
        r = subprocess.run([sys.executable, "-c", _LIB_RUN, ROOT, c1_dir, "7.4e8"], capture_output=True, text=True,
                           timeout=300, env=env)
        assert r.returncode == 0, r.stderr[-2000:]
        out[mode] = json.loads(r.stdout.strip().splitlines()[-1])
    ho, de = out["host"], out["device"]
    assert ho["raw"] == de["raw"] > 20000 and ho["photons"] == de["photons"]  # the same raw hits
    nh, nd = len(ho["dom"]), len(de["dom"])
    print("accepted: host %d, device %d of %d raw hits" % (nh, nd, ho["raw"]), flush=True)
    assert abs(nh - nd) < 5 * np.sqrt(nh + nd)
    key = lambda o: np.array(o["str"]) * 1000 + np.array(o["dom"])
    doms = np.union1d(key(ho), key(de))
    ch = np.array([np.sum(key(ho) == d) for d in doms], dtype=float)
    cd = np.array([np.sum(key(de) == d) for d in doms], dtype=float)
    m = (ch + cd) >= 10
    c2 = np.sum((ch[m] - cd[m]) ** 2 / (ch[m] + cd[m]))
    pchi = stats.chi2.sf(c2, m.sum() - 1)
    pks = stats.ks_2samp(ho["wv"], de["wv"]).pvalue
    print("per-DOM chi2 p %.3f over %d DOMs, wavelength KS p %.3f" % (pchi, m.sum(), pks), flush=True)
    assert pchi > 0.01 and pks > 0.01


def _tube_transit(end, zz, rng):
    """the transit time through a wavelength-shifting tube of the reference's model
    (gpu/ini.cxx:211-243) for hits at axial coordinate zz [m] read out at end `end`, sampled
    with numpy's generator: the law the device's draws must follow"""
    cq, att, zmin, tau, L = 0.299792458 / 1.46, 0.678, 0.01, 1.6, 0.76
    a = L / (1 - np.exp(-L / att))
    z = np.where(end == 1, L / 2 - zz, L / 2 + zz)
    d = (z / L) * (1 - z / (2 * a)) / (1 - L / (2 * a))
    z = np.maximum(-att * np.log(1 - d * (1 - np.exp(-L / att))), zmin)
    x = np.log(z)
    s = np.exp(((0.0034112 * x + 0.041727) * x + 0.45550) * x - 1.9188)
    g = np.exp(((0.0012658 * x + 0.017228) * x - 0.012750) * x + 0.20924)
    t = np.full(len(z), np.nan)
    todo = np.ones(len(z), bool)
    while todo.any():
        k = np.flatnonzero(todo)
        u1, u2 = rng.random(len(k)), rng.random(len(k))
        tk = s[k] * np.sqrt(2 * g[k] * (1 - u1) ** (1 / (1 - g[k])) - 1)
        acc = tk <= -(3.0 / cq) * np.log(u2)
        t[k[acc]] = tk[acc]
        todo[k[acc]] = False
    return t + z / cq - tau * np.log(rng.random(len(z)))


def test_device_tube_sensors_follow_the_tube_model():
    """Wavelength-shifting tube sensors (beta -2, cylinders) on strings 1 and 2 of the 7-string
    ice, C1 flasher, acceptance on the device: per accepted tube hit the end is 1 with
    probability 1/2 + zz/(2a) (zz the impact point's axial coordinate scaled to the 0.76 m
    tube, a = 0.76/(1 - exp(-0.76/0.678))); the sum of (end - p) zz over the hits, which a
    swapped or missing axial term moves by tens of sigma, is within 5 sigma of 0; the transit
    times follow the model's law given each hit's end and zz (two-sample KS against numpy
    draws, p > 0.01); the efficiency alone decides acceptance (binomial, 5 sigma); hits on
    the one-PMT DOMs get no added time; a tube of zero length is refused at upload"""
    from ppc_amd import api, configs

    det = api.Detector(configs.ensure("c1wom"), {"WFLA": "405", "FWID": "10", "DTMN": "1"})
    T = AcceptTables.from_address(det.lib.xppc_c_accept_tables())
    tube = [k for k in range(T.nsensor) if T.sensors[k].beta == -2]
    assert len(tube) == 1
    S = T.sensors[tube[0]]
    dev = api.Device(0, deterministic=True)
    dev.upload(det)
    dev.upload_accept(det)
    b = det.flasher_bunch(1, 30, num=200_000_000)
    print("device acceptance: c1 flasher on the tube ice, %d photons" % b.num, flush=True)
    dev.run([b])
    h = dev.hits()
    a = dev.accept(seed=777)
    # a tube without length is refused at upload (its transit times would be NaN)
    Rz = S.Rz
    S.Rz = 0.0
    assert dev.lib.ppc_hip_upload_accept(dev.ctx, ctypes.byref(T)) != 0
    S.Rz = Rz
    dev.close()
    dsens = np.array([T.doms[i].sensor for i in range(T.gsize)])
    dacc = np.array([T.doms[i].acc for i in range(T.gsize)])
    drde = np.array([T.doms[i].rde for i in range(T.gsize)], dtype=np.float64)
    dF = np.array([T.doms[i].F for i in range(T.gsize)], dtype=np.float64)
    ist = dsens[h["i"]] == tube[0]
    ok = a["pmt"] >= 0
    assert ist.sum() > 20000
    assert np.all(a["dt"][ok & ~ist] == 0) and np.all(a["pmt"][ok & ~ist] == 0)
    rat = np.ctypeslib.as_array(T.rat, shape=(T.nacc * WNUM,)).reshape(T.nacc, WNUM)
    pe = np.minimum(1.0, rat[dacc[h["i"][ist]], h["z"][ist]] * drde[h["i"][ist]])
    want, sd = pe.sum(), np.sqrt((pe * (1 - pe)).sum())
    print("tube hits accepted %d of %d, expected %.1f +- %.1f" % ((ok & ist).sum(), ist.sum(), want, sd), flush=True)
    assert abs((ok & ist).sum() - want) < 5 * sd + 1
    m = ok & ist
    rz = dF[h["i"][m]] * np.cos(h["dth"][m].astype(np.float64))
    L, att = 0.76, 0.678
    A = L / (1 - np.exp(-L / att))
    zz = S.Rr * rz * L / abs(2 * S.Rz)
    p1 = np.clip(0.5 + zz / (2 * A), 0, 1)
    end = a["pmt"][m]
    Z = np.sum((end - p1) * zz) / np.sqrt(np.sum(p1 * (1 - p1) * zz * zz))
    Zs = np.sum(((1 - end) - p1) * zz) / np.sqrt(np.sum(p1 * (1 - p1) * zz * zz))
    print("ends: %d at 0, %d at 1; axial statistic %.2f sigma (ends swapped: %.1f)" % ((end == 0).sum(),
                                                                                      (end == 1).sum(), Z, Zs), flush=True)
    assert abs(Z) < 5 and abs(Zs) > 20
    ref = _tube_transit(end, zz, np.random.default_rng(5))
    dt = a["dt"][m].astype(np.float64)
    pks = stats.ks_2samp(dt, ref).pvalue
    print("transit time: mean %.3f vs %.3f ns, median %.3f vs %.3f, KS p %.3f" % (dt.mean(), ref.mean(), np.median(dt),
                                                                                  np.median(ref), pks), flush=True)
    assert np.all(dt > 0) and pks > 0.01


def _curve(beta, x):
    """a multi-PMT sensor's PMT acceptance against the cosine x (gpu/ini.cxx:169-188)"""
    x = np.clip(np.asarray(x, np.float64), -1, 1)
    if beta < -1:
        return np.sqrt(1 - x * x)
    s = np.maximum(x, 0)
    if beta < 1:
        y = np.sqrt(1 - x * x) / beta
        c = 1 - beta * beta
        big = y > 1
        yi = np.where(big, 1 / np.where(big, y, 1), 0.5)
        aS = lambda v: np.arccos(v) - np.sin(2 * np.arccos(v)) / 2
        s = s + np.where(big, (aS(yi) / c - aS(yi * np.abs(x) / np.sqrt(c)) * np.abs(x)) / np.pi, 0)
    return s


def test_device_multipmt_choice_follows_the_pmt_layout():
    """Multi-PMT modules (a 12-PMT and a 2-PMT module on strings 1 and 2 of the 7-string ice),
    C1 flasher, acceptance on the device: per PMT, the number of accepted hits against the
    sum over the hits of efficiency x the PMT's share of the module's acceptance for the
    photon's direction (gpu/ini.cxx:169-192, 258-277: the first PMT whose cumulative share
    exceeds the draw), chi-square p > 0.01 over the PMTs"""
    from ppc_amd import api, configs

    det = api.Detector(configs.ensure("c1ng"), {"WFLA": "405", "FWID": "10", "DTMN": "1"})
    T = AcceptTables.from_address(det.lib.xppc_c_accept_tables())
    multi = [k for k in range(T.nsensor) if T.sensors[k].def_ == 0 and T.sensors[k].beta != -2]
    assert len(multi) == 2
    dev = api.Device(0, deterministic=True)
    dev.upload(det)
    dev.upload_accept(det)
    b = det.flasher_bunch(1, 30, num=200_000_000)
    print("device acceptance: c1 flasher on the multi-PMT ice, %d photons" % b.num, flush=True)
    dev.run([b])
    h = dev.hits()
    a = dev.accept(seed=4242)
    dev.close()
    dsens = np.array([T.doms[i].sensor for i in range(T.gsize)])
    dacc = np.array([T.doms[i].acc for i in range(T.gsize)])
    drde = np.array([T.doms[i].rde for i in range(T.gsize)], dtype=np.float64)
    assert all(T.doms[i].azi < 0 for i in range(T.gsize))  # no cable in this ice
    rat = np.ctypeslib.as_array(T.rat, shape=(T.nacc * WNUM,)).reshape(T.nacc, WNUM)
    dirs = np.ctypeslib.as_array(T.pmt_dirs, shape=(T.npmt * 3,)).reshape(T.npmt, 3).astype(np.float64)
    pth, pph = h["pth"].astype(np.float64), h["pph"].astype(np.float64)
    n = np.stack([np.sin(pth) * np.cos(pph), np.sin(pth) * np.sin(pph), np.cos(pth)], 1)
    obs, exp = [], []
    for k in multi:
        S = T.sensors[k]
        m = dsens[h["i"]] == k
        pe = np.minimum(1.0, rat[dacc[h["i"][m]], h["z"][m]] * drde[h["i"][m]])
        nz = n[m, 2]
        xa = np.pi * S.Rr * np.sqrt(S.Rz ** 2 - nz * nz * (S.Rz ** 2 - S.Rr ** 2)) if S.Rz > 0 else \
            -4 * S.Rr * S.Rz * np.sqrt(1 - nz * nz)
        crx = S.fx * xa
        share = np.stack([_curve(S.beta, -(n[m] @ dirs[S.pmt0 + j])) / crx for j in range(S.npmt)], 1)
        cum = np.minimum(np.cumsum(share, axis=1), 1.0)
        p = np.diff(np.concatenate([np.zeros((len(cum), 1)), cum], 1), axis=1)
        exp.append((pe[:, None] * p).sum(axis=0))
        obs.append(np.array([np.sum(a["pmt"][m] == j) for j in range(S.npmt)], dtype=np.float64))
        print("sensor %d (%d PMTs): %d hits, accepted %d, expected %.1f; per PMT %s vs %s" % (
            k, S.npmt, m.sum(), obs[-1].sum(), exp[-1].sum(), obs[-1].astype(int).tolist(),
            np.round(exp[-1], 1).tolist()), flush=True)
    o, e = np.concatenate(obs), np.concatenate(exp)
    assert e.min() > 20
    chi2 = float(np.sum((o - e) ** 2 / e))
    p = stats.chi2.sf(chi2, len(o))
    print("chi2 %.1f / %d, p = %.3f" % (chi2, len(o), p), flush=True)
    assert p > 0.01


def test_device_cable_shadow_on_full_feature_ice():
    """Cable positions (dx.dat of the 86-string ice with every optional feature), a flasher on
    string 36, acceptance on the device: a hit whose photon passes the cable before reaching
    the DOM (gpu/f2k.cxx:301-322: the line back from the impact point meets the cable's
    circle, radius 0.023 m beside the DOM at its azimuth) is never accepted, and the other
    hits are accepted at efficiency x angular sensitivity (binomial, 5 sigma)"""
    from ppc_amd import api, configs

    det = api.Detector(configs.ensure("c2x"), {"WFLA": "405", "FWID": "10", "DTMN": "1"})
    T = AcceptTables.from_address(det.lib.xppc_c_accept_tables())
    dev = api.Device(0, deterministic=True)
    dev.upload(det)
    dev.upload_accept(det)
    b = det.flasher_bunch(36, 30, num=60_000_000)
    print("device acceptance: flasher 36-30 on the full-feature ice, %d photons" % b.num, flush=True)
    dev.run([b])
    h = dev.hits()
    a = dev.accept(seed=99)
    dev.close()
    f64 = np.float64
    D = [T.doms[i] for i in range(T.gsize)]
    sens, acc = np.array([d.sensor for d in D]), np.array([d.acc for d in D])
    rde, azi = np.array([d.rde for d in D], f64), np.array([d.azi for d in D], f64)
    R, F = np.array([d.R for d in D], f64), np.array([d.F for d in D], f64)
    tilt = np.array([list(d.tilt) for d in D], f64)
    i = h["i"].astype(np.int64)
    pth, pph, dth, dph = (h[k].astype(f64) for k in ("pth", "pph", "dth", "dph"))
    nx, ny, nz = np.sin(pth) * np.cos(pph), np.sin(pth) * np.sin(pph), np.cos(pth)
    dts = np.where(F[i] < 0, 1.0, np.sin(dth))
    rx, ry = dts * np.cos(dph), dts * np.sin(dph)
    # the cable's shadow, where the DOM has a cable azimuth
    omr, rc = R[i], 0.023
    dc = omr + rc
    drx = -omr * rx - dc * np.cos(np.radians(azi[i]))
    dry = -omr * ry - dc * np.sin(np.radians(azi[i]))
    qa = nx * nx + ny * ny
    qb = -2 * (nx * drx + ny * dry)
    qc = drx * drx + dry * dry - rc * rc
    disc = qb * qb - 4 * qa * qc
    with np.errstate(invalid="ignore", divide="ignore"):
        far = (-qb + np.sqrt(np.maximum(disc, 0))) / (2 * qa)
    shadow = (azi[i] >= 0) & (qa > 0) & (disc >= 0) & (far > 0)
    # a margin for the device's float arithmetic at the circle's edge
    edge = (np.abs(disc) < 1e-6) | (np.abs(far) < 1e-6)
    ok = a["pmt"] >= 0
    assert (azi >= 0).sum() > 5000 and shadow.sum() > 200
    bad = ok & shadow & ~edge
    print("raw hits %d, in the cable's shadow %d (%d near its edge), accepted there %d" % (
        len(h), shadow.sum(), (shadow & edge).sum(), bad.sum()), flush=True)
    assert bad.sum() == 0
    # the others: efficiency x angular sensitivity of the one-PMT sensor
    rat = np.ctypeslib.as_array(T.rat, shape=(T.nacc * WNUM,)).reshape(T.nacc, WNUM)
    p = np.zeros(len(h))
    live = (sens[i] >= 0) & ~shadow
    for k in np.unique(sens[i][live]):
        S = T.sensors[int(k)]
        assert S.def_ == 1
        m = live & (sens[i] == k)
        pe = np.minimum(1.0, rat[acc[i[m]], h["z"][m]] * rde[i[m]])
        x = nx[m] * tilt[i[m], 0] + ny[m] * tilt[i[m], 1] + nz[m] * tilt[i[m], 2]
        poly = sum(S.s[j] * x ** j for j in range(S.ns))
        pa = np.clip(poly / S.mas, 0, 1) if S.mas > 0 else None
        assert pa is not None
        p[m] = pe * pa
    want, sd = p[live].sum(), np.sqrt((p * (1 - p))[live].sum())
    got = (ok & live).sum()
    print("accepted outside the shadow %d, expected %.1f +- %.1f" % (got, want, sd), flush=True)
    assert abs(got - want) < 5 * sd + 1
