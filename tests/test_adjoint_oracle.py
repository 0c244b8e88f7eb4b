"""The oracle's score-function gradient (orc_adjoint_photon) against what the
geometry of the paths fixes.

An isotropic point source emits every photon at one place and time, so a hit's
path length is (t_hit - t0) / ocm of its wavelength bin.  Hence, exactly and
photon by photon:
  * the absorption scores summed over layers are -sum_h w_h lk_h L_h (no
    absorption anisotropy: nr = 1);
  * in ice without scattering every path is the straight line from the source to
    the hit point, so the scattering scores sum to -sum_h w_h la_h L_h and the
    path-length-weighted mean layer of the absorption scores is the mean of the
    source's and the hit DOM's layer coordinates (to within the oversized DOM
    radius).
A dropped score, a wrong sign or a layer index off by one fails one of these.
"""
import ctypes

import numpy as np
import pytest

import oracle_util as U


@pytest.fixture(scope="module")
def c1(c1_dir):
    from ppc_amd import api

    det = api.Detector(c1_dir, {"WFLA": None, "FWID": None})
    T = U.product_tables(det)
    wv = np.frombuffer(det.field("wv"), dtype=np.float32).reshape(U.WNUM, 8)
    q = det.ice_parameters()
    return {"det": det, "T": T, "ocm": wv[:, 0].copy(), "la": wv[:, 3].copy(), "lk": wv[:, 4].copy(), **q,
            "rm": det.array("rm", np.uint32), "gsize": int(det.scalar("gsize", np.int32)),
            "size": int(det.scalar("size", np.int32)), "hmin": float(det.scalar("hmin")),
            "rdh": float(det.scalar("rdh")), "oms": det.array("oms", np.float32).reshape(-1, 5)}


SRC = (2.0, 1.0, 3.0)


def _run(c, tables, n, seed):
    rng = np.random.default_rng(seed)
    b = U.point_bunch(*SRC, num=n)
    weight = rng.uniform(0.5, 1.5, size=(1, c["gsize"])).astype(np.float32)
    states = [(int(rng.integers(1, 2**32)), int(rng.integers(0, c["rm"][k % 64])), int(c["rm"][k % 64]))
              for k in range(n)]
    ga, gb, nh = U.oracle_adjoint(tables, [b], range(n), states, weight, c["lk"], c["la"], c["dpda"], c["ds1db"])
    # the hits of the same photons, for the path lengths
    lib = U.oracle()
    H = (U.Hit * 64)()
    hits = []
    arr = (U.Bunch * 1)(b)
    for i, st in enumerate(states):
        s = U.Rng(*st)
        m = lib.orc_photon(tables, arr, 1, i, ctypes.byref(s), H, 64, None)
        hits.append(U.hits_array(H, min(m, 64)))
        assert m == nh[i]
    hits = np.concatenate(hits)
    return ga, gb, hits, weight[0]


def test_absorption_scores_sum_to_path_lengths(c1):
    assert c1["det"].scalar("k1") == 1 and c1["det"].scalar("k2") == 1 and c1["det"].scalar("kz") == 1
    ga, gb, hits, w = _run(c1, c1["T"].ptr, 20000, 1)
    assert len(hits) > 200
    L = hits["t"].astype(np.float64) / c1["ocm"][hits["z"]]
    want = -np.sum(w[hits["i"]] * c1["lk"][hits["z"]] * L)
    got = np.sum(ga / c1["dpda"])
    assert got == pytest.approx(want, rel=2e-4)
    # scattering scores: the path terms alone would be -sum w la L; the scatterings add a
    # positive 1/sca per event, so the sum lies above that
    assert np.sum(gb / c1["ds1db"]) > -np.sum(w[hits["i"]] * c1["la"][hits["z"]] * L)


def test_straight_paths_without_scattering(c1):
    T = U.product_tables(c1["det"])
    sca = np.frombuffer(bytes(T.field("sca")), dtype=np.float32).copy()
    sca[:] = 1e-7  # mean free path 1e7 m: photons fly straight
    ctypes.memmove(T.field("sca"), sca.tobytes(), sca.nbytes)
    ga, gb, hits, w = _run(c1, T.ptr, 20000, 2)
    assert len(hits) > 100
    wh = w[hits["i"]].astype(np.float64)
    L = hits["t"].astype(np.float64) / c1["ocm"][hits["z"]]
    assert np.sum(gb / c1["ds1db"]) == pytest.approx(-np.sum(wh * c1["la"][hits["z"]] * L), rel=2e-4)
    assert np.sum(ga / c1["dpda"]) == pytest.approx(-np.sum(wh * c1["lk"][hits["z"]] * L), rel=2e-4)
    # path-weighted mean layer coordinate of the scores vs the straight lines' midpoints
    size = c1["size"]
    zs = (SRC[2] - c1["hmin"]) * c1["rdh"]
    zd = (c1["oms"][hits["i"], 4] - c1["hmin"]) * c1["rdh"]
    wl = wh * c1["lk"][hits["z"]] * L
    want = np.sum(wl * (zs + zd) / 2) / np.sum(wl)
    got = np.sum(ga * np.arange(size)) / np.sum(ga)
    assert abs(got - want) < 0.05  # oversized DOM radius 0.83 m = 0.08 layers at the far end
    # and no score outside the layers the straight lines cross
    lo, hi = int(np.floor(min(zs, zd.min()) - 0.5)), int(np.ceil(max(zs, zd.max()) + 0.5))
    assert np.all(ga[:max(lo, 0)] == 0) and np.all(ga[hi + 1:] == 0)
