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


# ---- complete paths: the scattering-event term against the Poisson law of the scatterings ----
#
# Along a complete path (emission to absorption) the scattering score of layer k is
# (n_k / sca_k - l_k) * la * ds1db[k] and the absorption score -l_k * lk * dpda[k] (nr = 1),
# with l_k the path length and n_k the scatterings in layer k.  So, photon by photon, the two
# scores give l_k and n_k, and n_k must be a whole number that sums to the photon's counted
# scatterings; over many photons the scatterings in a layer are a Poisson process along the
# path, so E[n_k - sca_k l_k] = 0 in every layer.  An event term with a wrong factor breaks
# the whole numbers; one attributed to a neighbouring layer breaks the per-layer means.


def _path_scores(c, tables, n, seed):
    rng = np.random.default_rng(seed)
    b = U.point_bunch(*SRC, num=n)
    states = [(int(rng.integers(1, 2**32)), int(rng.integers(0, c["rm"][k % 64])), int(c["rm"][k % 64]))
              for k in range(n)]
    GA, GB, rows, nsc = U.oracle_path_scores(tables, b, states, c["lk"], c["la"], c["dpda"], c["ds1db"])
    lk, la = c["lk"][rows].astype(np.float64), c["la"][rows].astype(np.float64)
    L = -GA / (lk[:, None] * c["dpda"][None, :].astype(np.float64))       # path length per layer
    S = GB / (la[:, None] * c["ds1db"][None, :].astype(np.float64)) + L   # n_k / sca_k
    return L, S, rows, nsc


def _sca_table(tables):
    T = tables
    return np.frombuffer(bytes(T.field("sca")), dtype=np.float32).reshape(U.WNUM, U.MAXLYS).astype(np.float64)


def test_scattering_event_scores_count_whole_scatterings(c1):
    L, S, rows, nsc = _path_scores(c1, c1["T"].ptr, 4000, 3)
    sca = _sca_table(c1["T"])[rows, :c1["size"]]
    nk = S * sca
    assert np.all(L >= -1e-9)
    assert np.abs(nk - np.rint(nk)).max() < 1e-3
    assert np.all(np.rint(nk) >= 0)
    assert (np.rint(nk).sum(axis=1) == nsc).all()
    assert nsc.mean() > 5


def test_scatterings_per_layer_follow_path_length(c1):
    L, S, rows, nsc = _path_scores(c1, c1["T"].ptr, 20000, 4)
    sca = _sca_table(c1["T"])[rows, :c1["size"]]
    nk = np.rint(S * sca)
    d = nk - sca * L                      # per photon and layer: n_k - sca_k l_k, mean 0
    used = (L > 0).sum(axis=0) >= 200
    assert used.sum() >= 5
    z = d[:, used].sum(axis=0) / np.sqrt((d[:, used] ** 2).sum(axis=0))
    assert np.abs(z).max() < 4.5
    assert np.mean(z ** 2) < 2.0
    # the same statistic with the scatterings credited to the layer above fails: the test can
    # tell a neighbouring layer's events from a layer's own
    ks = np.flatnonzero(used)
    ks = ks[ks + 1 < c1["size"]]
    dz = nk[:, ks + 1] - sca[:, ks] * L[:, ks]
    zs = dz.sum(axis=0) / np.sqrt((dz ** 2).sum(axis=0))
    assert np.abs(zs).max() > 6


def test_one_layer_closed_form():
    """homogeneous ice (every layer and wavelength: absorption a, scattering s): a complete
    path has length L ~ Exp(a) and n ~ Poisson(s L) scatterings, so E[L] = 1/a, and
    n - s L has mean 0 and variance s/a"""
    from ppc_amd import api, configs

    det = api.Detector(configs.ensure("c1"), {"WFLA": None, "FWID": None})
    wv = np.frombuffer(det.field("wv"), dtype=np.float32).reshape(U.WNUM, 8)
    q = det.ice_parameters()
    c = {"lk": wv[:, 4].copy(), "la": wv[:, 3].copy(), "rm": det.array("rm", np.uint32), **q,
         "size": int(det.scalar("size", np.int32))}
    T = U.product_tables(det)
    a0, s0 = 0.02, 0.5
    for name, v in (("sca", s0), ("abs_", a0)):
        arr = np.full(U.WNUM * U.MAXLYS, v, dtype=np.float32)
        ctypes.memmove(T.field(name), arr.tobytes(), arr.nbytes)
    n = 20000
    L, S, rows, nsc = _path_scores(c, T.ptr, n, 5)
    Lp = L.sum(axis=1)
    nk = S * s0
    assert np.abs(nk - np.rint(nk)).max() < 1e-3
    Np = np.rint(nk).sum(axis=1)
    assert (Np == nsc).all()
    # E[L] = 1/a (sd 1/a per photon), E[n - s L] = 0 (variance s/a per photon)
    assert abs(Lp.mean() - 1 / a0) < 4 * (1 / a0) / np.sqrt(n)
    assert abs((Np - s0 * Lp).mean()) < 4 * np.sqrt(s0 / a0 / n)
    assert np.var(Np - s0 * Lp) == pytest.approx(s0 / a0, rel=0.1)
