"""The HIP kernel against the reference's own outputs (tests/golden/).

Replay: every lane of a deterministic launch starts from the MWC state the
reference started the same photon from (photon g runs on lane g with
multiplier g), so each lane must end in the reference's final state and
produce the reference's hits.  The kernel uses hardware transcendentals and
FMA contraction: at least 99 % of the photons must end in exactly the
reference's MWC state, and hits of matching photons agree within 1e-4 in time.

Statistics: raw hits of many photons against the reference's raw hits on the
same ice and source: hit rate (Poisson), per-DOM hit distribution (two-sample
chi-square) and arrival times and directions (two-sample KS), each at p > 0.01.
Launches are deterministic (fixed seed), so a result is reproducible.
"""
import ctypes
import math
import os
import sys

import numpy as np
import pytest
from scipy import stats

import oracle_util as U

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as G  # noqa: E402

pytestmark = pytest.mark.gpu


def _setup(case, **kw):
    from ppc_amd import api, configs

    cfg, env, src = G.CASES[case]
    env = dict({k: None for k in G.ENV_KEYS}, **env)
    env["DTMN"] = "1"
    det = api.Detector(configs.ensure(cfg), env)
    dev = api.Device(0, deterministic=True, **kw)
    dev.upload(det)
    tab = U.product_tables(det)
    return det, dev, tab, src


@pytest.mark.parametrize("case", list(G.CASES))
def test_replay_reference_vectors(case):
    gold = np.load(os.path.join(HERE, "golden", "replay_%s.npz" % case))
    n, procs = int(gold["n"]), int(gold["procs"])
    det, dev, tab, src = _setup(case)
    total = n * procs
    assert dev.lanes >= total
    st = np.zeros((dev.lanes, 2), dtype=np.uint32)
    st[:, 0], st[:, 1] = 362436069, 1234567
    for p in range(procs):
        st[p * n:(p + 1) * n] = G.replay_states(p, n)
    dev.set_states(st[:, 0].astype(np.uint64) | (st[:, 1].astype(np.uint64) << np.uint64(32)))
    b = U.to_api_bunch(G._bunch(tab, src, total, det=det))
    s = dev.run([b])
    assert s.photons == total
    hits, hph = dev.hits(with_photon=True)
    fin = dev.states()[:total]
    fin2 = np.column_stack([(fin & np.uint64(0xFFFFFFFF)).astype(np.uint32), (fin >> np.uint64(32)).astype(np.uint32)])
    same = G.fold(fin2) == gold["digest"]
    gh, gp = gold["hits"], gold["hit_photon"]
    agree = 0
    phot = np.unique(np.concatenate([hph, gp]))
    for g in phot:
        a = np.sort(hits[hph == g], order="t")
        r = np.sort(gh[gp == g], order="t")
        if len(a) == len(r) and (a["i"] == r["i"]).all() and np.allclose(a["t"], r["t"], rtol=1e-4):
            agree += 1
    print("%s: %d photons, %d end in the reference state; hits %d vs %d; %d of %d photons with hits agree"
          % (case, total, same.sum(), len(hits), len(gh), agree, len(phot)))
    dev.close()
    assert same.sum() >= 0.99 * total
    assert agree >= 0.97 * len(phot) - 1


def _two_sample_chi2(a, b, min_exp=10):
    """chi-square of two histograms with different totals, bins merged until
    both expected counts reach min_exp; returns (chi2, dof, p)"""
    a, b = np.asarray(a, float), np.asarray(b, float)
    A, B = a.sum(), b.sum()
    order = np.argsort(-(a + b))
    xa, xb, ca, cb = [], [], 0.0, 0.0
    for i in order:
        ca += a[i]
        cb += b[i]
        tot = ca + cb
        if tot * A / (A + B) >= min_exp and tot * B / (A + B) >= min_exp:
            xa.append(ca), xb.append(cb)
            ca = cb = 0.0
    if ca + cb > 0 and xa:
        xa[-1] += ca
        xb[-1] += cb
    xa, xb = np.array(xa), np.array(xb)
    k1, k2 = math.sqrt(B / A), math.sqrt(A / B)
    chi2 = float((((k1 * xa - k2 * xb) ** 2) / (xa + xb)).sum())
    dof = len(xa) - 1
    return chi2, dof, float(stats.chi2.sf(chi2, dof))


SOURCES = {"c1": 200_000_000, "c2": 100_000_000, "c2x": 400_000_000, "c3": 10,  # photons; c3: event repeats
           "c2t9": 100_000_000, "c1iso": 200_000_000, "c1lab": 200_000_000, "c1cone": 200_000_000,
           "c2vthk": 100_000_000, "cbig": 100_000_000, "c2xho": 100_000_000, "c2xhe": 400_000_000, "c3km": 1_000_000_000,
           "c3clst": 200_000_000, "c3mopo": 200_000_000, "c3mopob": 200_000_000, "c1t2lab": 200_000_000,
           "c1sclab": 1_000_000_000}


@pytest.mark.parametrize("case", list(G.CASES))
def test_statistics_match_reference(case):
    path = os.path.join(HERE, "golden", "stats_%s.npz" % case)
    if not os.path.exists(path):
        pytest.skip("no statistics fixture")
    gold = np.load(path)
    det, dev, tab, src = _setup(case, max_hits=1 << 24)
    dev.seed(12345)
    if src[0] == "muon":  # the whole muon event: 1200 segments and its cascades, each bunch repeated
        bunches = G.muon_event_bunches(
            lambda x, y, z, t, E, dr, ty, n: det.particle_bunch(x, y, z, t, E, dr, ty, n), 1, 0)
        for b in bunches:
            b.num *= SOURCES[case]
        N = sum(b.num for b in bunches)
    else:
        N = SOURCES[case]
        bunches = [U.to_api_bunch(G._bunch(tab, src, N, det=det))]
    s = dev.run(bunches)
    assert s.photons == N
    hits = dev.hits()
    dev.close()
    assert s.stored == s.hits
    Nr, hr = int(gold["photons"]), int(gold["nhits"])
    rate_g, rate_r = len(hits) / N, hr / Nr
    z = (rate_g - rate_r) / math.sqrt(hr / Nr ** 2 + len(hits) / N ** 2)
    p_rate = 2 * stats.norm.sf(abs(z))
    counts = np.bincount(hits["i"], minlength=len(gold["counts"]))
    chi2, dof, p_dom = _two_sample_chi2(counts, gold["counts"])
    p_t = stats.ks_2samp(hits["t"], gold["t"]).pvalue
    p_c = stats.ks_2samp(np.cos(hits["pth"]), gold["cth"]).pvalue
    print("%s: rate %.4e vs %.4e (p=%.3f); DOM chi2 %.1f/%d (p=%.3f); KS time p=%.3f, direction p=%.3f"
          % (case, rate_g, rate_r, p_rate, chi2, dof, p_dom, p_t, p_c))
    assert p_rate > 0.01
    assert p_dom > 0.01
    assert p_t > 0.01
    assert p_c > 0.01
