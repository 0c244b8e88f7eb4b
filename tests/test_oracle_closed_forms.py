"""The oracle's building blocks against closed forms and brute force.

Each check is chosen so that a plausible slip in the restatement fails it:
a wrong carry or shift in the MWC step, a swapped Box-Muller draw, a wrong
rotation sign or axis swap, an HG/SAM formula with a dropped term.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy import stats

import oracle_util as U


def _rng(x=362436069, c=1234567, a=4294967118):
    return U.Rng(x, c, a)


def test_mwc_matches_big_integer_recurrence():
    """t = a*x + c; x = t mod 2^32; c = t >> 32; u = 2 - float(0x3f800000 | x>>9)"""
    lib = U.oracle()
    for x0, c0, a in [(362436069, 1234567, 4294967118), (1, 0, 4199122803), (0xFFFFFFFF, 0x7FFFFFFF, 4294966893)]:
        s = U.Rng(x0, c0, a)
        x, c = x0, c0
        for _ in range(2000):
            u = lib.orc_xrnd(ctypes.byref(s))
            t = a * x + c
            x, c = t & 0xFFFFFFFF, t >> 32
            exp = 2.0 - np.frombuffer(np.uint32(0x3F800000 | (x >> 9)).tobytes(), dtype=np.float32)[0]
            assert (s.x, s.c) == (x, c)
            assert u == exp
            assert 0.0 < u <= 1.0


def test_mwc_uniformity():
    lib = U.oracle()
    s = _rng()
    u = np.array([lib.orc_xrnd(ctypes.byref(s)) for _ in range(100000)])
    assert stats.kstest(u, "uniform").pvalue > 1e-3


@pytest.mark.parametrize("k", [0.5, 2.5, 9.29])
def test_gamma_sampler(k):
    """Weibull (k<1) and Cheng (k>=1) rejection samplers against Gamma(k, 1)."""
    lib = U.oracle()
    s = _rng(a=4294966830)
    x = np.array([lib.orc_mrnd(ctypes.c_float(k), ctypes.byref(s)) for _ in range(60000)])
    assert abs(x.mean() - k) < 5 * math.sqrt(k / len(x))
    assert abs(x.var() - k) < 0.05 * k + 5 * math.sqrt(2 * k * k / len(x))
    assert stats.kstest(x, stats.gamma(k).cdf).pvalue > 1e-3


def test_gauss_sampler():
    lib = U.oracle()
    s = _rng(a=4294966893)
    x = np.array([lib.orc_grnd(ctypes.byref(s)) for _ in range(60000)])
    assert stats.kstest(x, "norm").pvalue > 1e-3


def test_rotate_polar_angle_and_azimuth():
    """rotate(cs, si): new direction has n.n' = 1 - cs, stays unit, azimuth about n uniform."""
    lib = U.oracle()
    s = _rng(a=4294966830)
    rng = np.random.default_rng(1)
    phis = []
    for _ in range(20000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        cs = float(rng.uniform(0, 2))
        si = math.sqrt(cs * (2 - cs))
        v = (ctypes.c_float * 4)(*n.astype(np.float32), 0.0)
        lib.orc_rotate(ctypes.c_float(cs), ctypes.c_float(si), v, ctypes.byref(s))
        m = np.array(v[:3], dtype=np.float64)
        assert abs(np.linalg.norm(m) - 1) < 1e-5
        assert abs(float(n @ m) - (1 - cs)) < 2e-4
        # azimuth of m around n relative to a fixed perpendicular
        e1 = np.cross(n, [0.0, 0.0, 1.0] if abs(n[2]) < 0.9 else [1.0, 0.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(n, e1)
        p = m - (n @ m) * n
        if np.linalg.norm(p) > 1e-3:
            phis.append(math.atan2(p @ e2, p @ e1))
    assert stats.kstest(np.array(phis), stats.uniform(-math.pi, 2 * math.pi).cdf).pvalue > 1e-3


@pytest.mark.parametrize("sf", [0.0, 0.45, 1.0])
def test_scattering_mean_cosine(sf):
    """Henyey-Greenstein and simplified-Liu both have <cos theta> = g (pro.cu:848-862)."""
    lib = U.oracle()
    g = 0.9
    gr = (1 - g) / (1 + g)
    s = _rng(a=4294966893)
    xi = np.array([lib.orc_scatter(ctypes.c_float(sf), ctypes.c_float(g), ctypes.c_float(gr), ctypes.byref(s))
                   for _ in range(200000)])
    cos = 1 - xi
    assert abs(cos.mean() - g) < 5 * cos.std() / math.sqrt(len(cos))
    if sf == 0.0:  # HG: P(cos) = (1-g^2) / (2 (1+g^2-2 g cos)^{3/2})
        cdf = lambda c: (1 - g * g) / (2 * g) * ((1 + g * g - 2 * g * c) ** -0.5 - 1 / (1 + g))
        assert stats.kstest(cos, cdf).pvalue > 1e-3
    if sf == 1.0:  # SAM: 1 - cos = 2 (1 - u^gr) -> P(cos <= c) = 1 - ((1+c)/2)^(1/gr)
        cdf = lambda c: ((1 + c) / 2) ** (1 / gr)
        assert stats.kstest(cos, cdf).pvalue > 1e-3


def _homogeneous_c1(a0, s0):
    """the 7-string ice with every layer and wavelength set to absorption a0, scattering s0"""
    from ppc_amd import api, configs

    det = api.Detector(configs.ensure("c1"), {"WFLA": None, "FWID": None})
    T = U.product_tables(det)
    for name, v in (("sca", s0), ("abs_", a0)):
        arr = np.full(U.WNUM * U.MAXLYS, v, dtype=np.float32)
        ctypes.memmove(T.field(name), arr.tobytes(), arr.nbytes)
    wv = np.frombuffer(det.field("wv"), dtype=np.float32).reshape(U.WNUM, 8)
    return det, T, wv


def test_survival_in_homogeneous_non_scattering_ice():
    """survival exp(-a L): without scattering, the path lengths to absorption (from the
    absorption scores) follow Exp(a) (pro.cu:496, 546-572)"""
    det, T, wv = _homogeneous_c1(0.03, 1e-7)
    q = det.ice_parameters()
    rm = det.array("rm", np.uint32)
    rng = np.random.default_rng(8)
    states = [(int(rng.integers(1, 2**32)), int(rng.integers(0, rm[k % 64])), int(rm[k % 64])) for k in range(8000)]
    b = U.point_bunch(2.0, 1.0, 3.0, num=len(states))
    GA, GB, rows, nsc = U.oracle_path_scores(T.ptr, b, states, wv[:, 4], wv[:, 3], q["dpda"], q["ds1db"])
    L = (-GA / (wv[rows, 4][:, None].astype(np.float64) * q["dpda"][None, :].astype(np.float64))).sum(axis=1)
    assert (nsc == 0).mean() > 0.999
    assert stats.kstest(L, stats.expon(scale=1 / 0.03).cdf).pvalue > 1e-3
    # a dropped or doubled absorption term moves the mean by a factor, far beyond its error
    assert abs(L.mean() * 0.03 - 1) < 4 / math.sqrt(len(L))


def test_hit_time_is_path_length_times_group_index():
    """t = L n_g / c: in ice without scattering every hit lies on the straight line from the
    source at path length t / ocm (ini.cxx:1693-1694), on the surface the DOM search
    intersects: the DOM oversized by xR across the photon's direction only (pro.cu:634-651),
    so with par the offset from the DOM centre along the direction and perp across it,
    (par / R)^2 + (perp / (xR R))^2 = 1 and par <= 0 (the photon enters)"""
    det, T, wv = _homogeneous_c1(0.003, 1e-7)
    rm = det.array("rm", np.uint32)
    oms = np.frombuffer(bytes(T.field("oms")), np.float32).reshape(-1, 5)
    xR = float(np.frombuffer(bytes(T.field("xR")), np.float32)[0])
    assert xR > 1
    src = np.array([2.0, 1.0, 3.0])
    b = U.point_bunch(*src, num=1)
    lib = U.oracle()
    H = (U.Hit * 64)()
    rng = np.random.default_rng(9)
    hits = []
    for k in range(30000):
        s = U.Rng(int(rng.integers(1, 2**32)), int(rng.integers(0, rm[k % 64])), int(rm[k % 64]))
        m = lib.orc_photon(T.ptr, ctypes.byref(b), 1, 0, ctypes.byref(s), H, 64, None)
        hits.append(U.hits_array(H, min(m, 64)))
    h = np.concatenate(hits)
    assert len(h) > 50
    L = h["t"].astype(np.float64) / wv[h["z"], 0]
    n = np.column_stack([np.sin(h["pth"]) * np.cos(h["pph"]), np.sin(h["pth"]) * np.sin(h["pph"]), np.cos(h["pth"])])
    v = src[None, :] + L[:, None] * n - oms[h["i"]][:, 2:5].astype(np.float64)
    par = np.sum(v * n, axis=1)
    perp2 = np.sum(v * v, axis=1) - par ** 2
    R = oms[h["i"], 0].astype(np.float64)
    q = (par / R) ** 2 + perp2 / (xR * R) ** 2
    # the direction comes back as FP32 angles: acos loses precision near the poles, and an
    # angle error grows with the path into a position error; so the equation is tested where
    # that error stays below 1e-4 m (paths under 150 m, |cos| < 0.95), the sign everywhere
    sel = (L < 150) & (np.abs(np.cos(h["pth"])) < 0.95)
    assert sel.sum() > 30
    assert np.abs(q[sel] - 1).max() < 5e-3
    assert np.abs(q - 1).max() < 0.05
    assert (par <= 1e-3).all()
