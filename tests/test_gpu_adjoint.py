"""The GPU adjoint (ppc_hip_adjoint, ppc_amd.adjoint): gradient of the weighted
hit counts by per-layer absorption (adust400) and scattering (be400).

  * replayed photon by photon against the oracle's score-function gradient
    (same photons, same MWC states), to FP32 accumulation tolerance;
  * at 2e7 photons, the absorption scores against the hits' path lengths (an
    isotropic point source starts every photon at one place and time, so a
    hit's path length is t / ocm);
  * against common-random-number central differences of the simulated counts
    (deterministic launches from the same lane states), within the statistical
    error of both estimators over independent seeds;
  * through torch.autograd: the backward of a Poisson log-likelihood equals the
    adjoint with weight k/mu - 1.
"""
import os

import numpy as np
import pytest

import oracle_util as U

pytestmark = pytest.mark.gpu

SRC = (2.0, 1.0, 3.0)


@pytest.fixture(scope="module")
def env(c1_dir):
    from ppc_amd import api

    det = api.Detector(c1_dir, {"WFLA": None, "FWID": None, "DTMN": "1"})
    wv = np.frombuffer(det.field("wv"), dtype=np.float32).reshape(U.WNUM, 8)
    return {"det": det, "ocm": wv[:, 0].copy(), "la": wv[:, 3].copy(), "lk": wv[:, 4].copy(),
            "gsize": int(det.scalar("gsize", np.int32)), "size": int(det.scalar("size", np.int32)),
            **det.ice_parameters()}


def point_bunch(num):
    return U.to_api_bunch(U.point_bunch(*SRC, num=num))


def _states(lanes, seed):
    rng = np.random.default_rng(seed)
    s = rng.integers(1, 2 ** 32, size=lanes, dtype=np.uint64)
    return s | (rng.integers(0, 2 ** 30, size=lanes, dtype=np.uint64) << np.uint64(32))


def test_adjoint_matches_oracle_photon_by_photon(env):
    from ppc_amd import adjoint, api

    det = env["det"]
    print("adjoint vs oracle: c1 point source, 300000 photons", flush=True)
    sim = adjoint.HitSimulation(det, point_bunch(300000), deterministic=True)
    sim.dev.set_states(_states(sim.dev.lanes, 5))
    mu, launch = sim.forward(sim.par_a, sim.par_b)
    rng = np.random.default_rng(7)
    grad_mu = rng.uniform(0.5, 1.5, size=mu.shape)
    ga, gb = sim.backward(launch, grad_mu)
    rec = sim.dev.hit_photons()
    assert len(rec) == launch["stats"].hit_photons > 1000
    assert sim.last_adjoint["missed"] == 0  # both runs of every photon agree
    assert sim.last_adjoint["replays"] == 2 * len(rec)
    rm = det.array("rm", np.uint32)
    T = U.product_tables(det)
    states = [(int(r["x0"]), int(r["c0"]), int(rm[r["lane"]])) for r in rec]
    ga_o, gb_o, nh = U.oracle_adjoint(T.ptr, [U.point_bunch(*SRC, num=300000)], rec["photon"], states,
                                      grad_mu.astype(np.float32), env["lk"], env["la"], env["dpda"], env["ds1db"])
    sim.close()
    print("records %d  |ga| %.4g  |gb| %.4g" % (len(rec), np.abs(ga_o).sum(), np.abs(gb_o).sum()), flush=True)
    assert np.abs(ga - ga_o).sum() <= 3e-3 * np.abs(ga_o).sum()
    assert np.abs(gb - gb_o).sum() <= 3e-3 * np.abs(gb_o).sum()
    # the layers the photons cross carry the gradient, layer by layer
    big = np.abs(ga_o) > 0.05 * np.abs(ga_o).max()
    assert np.all(np.abs(ga - ga_o)[big] <= 0.02 * np.abs(ga_o)[big])


def test_absorption_scores_sum_to_path_lengths_at_scale(env):
    from ppc_amd import adjoint

    det = env["det"]
    n = 20_000_000
    print("adjoint path-length sum: c1 point source, %d photons" % n, flush=True)
    sim = adjoint.HitSimulation(det, point_bunch(n), max_hits=4 << 20)
    mu, launch = sim.forward(sim.par_a, sim.par_b)
    w = np.random.default_rng(3).uniform(0.5, 1.5, size=mu.shape)
    ga, gb = sim.backward(launch, w)
    sim.close()
    h = launch["hits"]
    assert len(h) > 100000
    L = h["t"].astype(np.float64) / env["ocm"][h["z"]]
    want = -np.sum(w[h["n"], h["i"]] * env["lk"][h["z"]] * L)
    assert np.sum(ga / env["dpda"]) == pytest.approx(want, rel=2e-3)
    assert sim.last_adjoint["missed"] == 0


# launches of one photon per lane per comparison; more (PPC_CRN_REPEATS=240) for a tighter
# figure of the relative difference, printed with -s
CRN_REPEATS = int(os.environ.get("PPC_CRN_REPEATS", "24"))


def _crn_check(env, which, direction, eps, repeats=CRN_REPEATS):
    """LR directional derivative vs CRN central differences of sum_d w_d N_d.

    Every launch runs one photon per lane from chosen lane states, so photon k of
    the plus, minus and nominal launches starts from the same MWC state (common
    random numbers photon by photon); the repeats give independent estimates."""
    from ppc_amd import adjoint

    det = env["det"]
    sim = adjoint.HitSimulation(det, point_bunch(1), deterministic=True)
    sim.bunches[0].num = sim.num = sim.dev.lanes
    a0, b0 = sim.par_a.astype(np.float64), sim.par_b.astype(np.float64)
    w = np.random.default_rng(11).uniform(0.5, 1.5, size=(1, env["gsize"]))
    lr, fd = [], []
    for k in range(repeats):
        S = _states(sim.dev.lanes, 100 + k)
        sim.dev.set_states(S)
        mu, launch = sim.forward(a0, b0)
        ga, gb = sim.backward(launch, w)
        g = ga if which == "a" else gb
        base = a0 if which == "a" else b0
        lr.append(np.sum(g * direction * base))
        val = []
        for sgn in (1, -1):
            p = base * (1 + sgn * eps * direction)
            sim.dev.set_states(S)
            m, _ = sim.forward(p, b0) if which == "a" else sim.forward(a0, p)
            val.append(np.sum(w * m))
        fd.append((val[0] - val[1]) / (2 * eps))
    lanes = sim.dev.lanes
    sim.close()
    lr, fd = np.array(lr), np.array(fd)
    err = np.sqrt((lr.var(ddof=1) + fd.var(ddof=1)) / repeats)
    d = lr - fd  # paired: both estimates from the same photons
    print("%s: %d x %d photons: LR %.5g +- %.3g  FD %.5g +- %.3g  LR-FD relative %.2f%% +- %.2f%%" % (
        which, repeats, lanes, lr.mean(), lr.std(ddof=1) / np.sqrt(repeats), fd.mean(),
        fd.std(ddof=1) / np.sqrt(repeats), 100 * d.mean() / abs(fd.mean()),
        100 * d.std(ddof=1) / np.sqrt(repeats) / abs(fd.mean())), flush=True)
    assert abs(lr.mean()) > 5 * err  # the comparison is informative
    # paired: the standard error of the photon-by-photon difference (the 5 % accuracy
    # itself is asserted on the C5 workload, tests/test_gpu_c5.py)
    assert abs(d.mean()) <= 4 * d.std(ddof=1) / np.sqrt(repeats) + 0.01 * abs(fd.mean())


def test_absorption_gradient_vs_crn_differences(env):
    print("adjoint vs CRN differences: absorption, all layers", flush=True)
    _crn_check(env, "a", np.ones(env["size"]), 0.05)


def test_absorption_gradient_vs_crn_differences_upper_layers(env):
    print("adjoint vs CRN differences: absorption, layers above the source", flush=True)
    det = env["det"]
    z = (SRC[2] - float(det.scalar("hmin"))) * float(det.scalar("rdh"))
    _crn_check(env, "a", (np.arange(env["size"]) > z + 0.5).astype(np.float64), 0.05)


def test_scattering_gradient_vs_crn_differences(env):
    print("adjoint vs CRN differences: scattering, all layers", flush=True)
    _crn_check(env, "b", np.ones(env["size"]), 0.05)


def test_autograd_poisson_llh(env):
    import torch

    from ppc_amd import adjoint

    det = env["det"]
    print("autograd: Poisson LLH of a c1 point source", flush=True)
    sim = adjoint.HitSimulation(det, point_bunch(2_000_000), deterministic=True, scale=0.5)
    sim.dev.set_states(_states(sim.dev.lanes, 21))
    data, _ = sim.forward(sim.par_a * 1.1, sim.par_b)
    k = np.round(data)
    S = _states(sim.dev.lanes, 22)
    sim.dev.set_states(S)
    a = torch.tensor(sim.par_a, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(sim.par_b, dtype=torch.float64, requires_grad=True)
    mu = adjoint.ExpectedHits.apply(a, b, sim)
    llh = adjoint.poisson_llh(mu, k)
    llh.backward()
    # the same launch by hand: d llh / d mu = k / mu - 1 where mu > 0 (-1 where the floor clamps)
    sim.dev.set_states(S)
    mu2, launch = sim.forward(sim.par_a, sim.par_b)
    assert np.array_equal(mu.detach().numpy(), mu2)
    wmu = np.where(mu2 > 0, k / np.maximum(mu2, 1e-9) - 1, -1.0)
    ga, gb = sim.backward(launch, wmu)
    sim.close()
    assert np.abs(ga).max() > 0 and np.abs(gb).max() > 0
    assert np.allclose(a.grad.numpy(), ga, rtol=1e-4, atol=1e-6 * np.abs(ga).max())
    assert np.allclose(b.grad.numpy(), gb, rtol=1e-4, atol=1e-6 * np.abs(gb).max())
