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
