"""The differentiation workload (12 flasher events on the 7-string ice, data at an ice
perturbed by up to 5 % per layer, llh's likelihood llf) on the GPU: the adjoint gradient
against central differences at 1 % steps with common random numbers.

Every launch seeds each photon from (seed, photon index) (ppc_hip_seed_photons), so the
nominal and the +-1 % launches of one repeat follow photon i with the same random numbers
whatever the other photons do; the repeats use independent seeds.  Per repeat: the adjoint's
directional derivative along "all layers" (absorption a, scattering b) and the central
difference of the same weighted hit sum, paired photon for photon.  The mean paired
difference and its standard error over the repeats give the relative error of the gradient:

  * weights 1 per DOM (the expected total of hits, the positive-weight building block of any
    likelihood's gradient, which is linear in the weights): |relative error| + 2 sigma <= 5 %
    for absorption and for scattering;
  * weights 1 on the DOMs within 60 m of the event's flasher, and on the others: the paired
    difference within 4 sigma of 0 for each (no bias in the near or the far light);
  * weights d llf / d mu at the nominal ice (llh's likelihood of the data): the paired
    difference within 4 sigma of 0 (its terms cancel between DOMs, so differences at 1 %
    resolve this gradient only to some tens of percent at this number of photons).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EVENTS = 12
PER_FLASHER = 10_000_000                  # photons per flasher per launch of a repeat
REPEATS = int(os.environ.get("PPC_C5_REPEATS", "100"))
EPS = 0.01


def _c5(c1_dir):
    import torch

    from ppc_amd import adjoint, api

    det = api.Detector(c1_dir, {"WFLA": "405", "FWID": "10", "DTMN": "1"})
    rng = np.random.default_rng(6)
    strs, doms = det.array("str", np.int32), det.array("dom", np.int32)
    pick = rng.choice(len(strs), EVENTS, replace=False)
    bunches = [det.flasher_bunch(int(strs[d]), int(doms[d]), num=PER_FLASHER * 10, seg=k) for k, d in enumerate(pick)]
    sim = adjoint.HitSimulation(det, bunches, deterministic=True, max_hits=1 << 23)
    a0, b0 = sim.par_a.astype(np.float64), sim.par_b.astype(np.float64)
    size = len(a0)
    # data: the C5 events (1e8 photons per flasher) at an ice perturbed by up to 5 % per layer
    sim.dev.seed_photons(11)
    mu_data, _ = sim.forward(a0 * rng.uniform(0.95, 1.05, size), b0 * rng.uniform(0.95, 1.05, size))
    k = rng.poisson(mu_data)
    sim.dev.seed_photons(12)
    mu0, _ = sim.forward(a0, b0)
    mu = torch.tensor(mu0, requires_grad=True)
    llh = adjoint.llf(mu, torch.tensor(k, dtype=torch.float64)).sum()
    llh.backward()
    scale = 10.0  # the repeats' launches have a tenth of the photons
    # DOMs within 60 m of an event's flasher and the others, each a functional with positive
    # weights of its own: a bias in the far light would show in "far"
    oms = det.array("oms", np.float32).reshape(-1, 5)[:, 2:5].astype(np.float64)
    dist = np.linalg.norm(oms[None, :, :] - oms[pick][:, None, :], axis=2)
    near = (dist < 60.0).astype(np.float64)
    weights = {"hits": np.ones_like(mu0), "near": near, "far": 1.0 - near, "llh": mu.grad.numpy() * scale}
    for b in sim.bunches:
        b.num = PER_FLASHER
    sim.num = PER_FLASHER * EVENTS
    print("C5: %d flashers, %d layers; llf %.1f over %d DOM bins with data; %d repeats of %d photons"
          % (EVENTS, size, float(llh.detach()), int((k > 0).sum()), REPEATS, sim.num), flush=True)
    res = {(wn, x): ([], []) for wn in weights for x in "ab"}
    for r in range(REPEATS):
        sim.dev.seed_photons(1000 + r)
        _, launch = sim.forward(a0, b0)
        grads = {wn: sim.backward(launch, w) for wn, w in weights.items()}
        assert sim.last_adjoint["missed"] == 0
        for x, base in (("a", a0), ("b", b0)):
            m = []
            for sgn in (1, -1):
                p = base * (1 + sgn * EPS)
                m.append(sim.forward(p, b0)[0] if x == "a" else sim.forward(a0, p)[0])
            for wn, w in weights.items():
                g = grads[wn][0] if x == "a" else grads[wn][1]
                res[(wn, x)][0].append(np.sum(g * base))
                res[(wn, x)][1].append((np.sum(w * m[0]) - np.sum(w * m[1])) / (2 * EPS))
    sim.close()
    out = {}
    for key, (lr, fd) in res.items():
        lr, fd = np.array(lr), np.array(fd)
        d = lr - fd
        rel, sig = d.mean() / abs(fd.mean()), d.std(ddof=1) / np.sqrt(len(d)) / abs(fd.mean())
        out[key] = (rel, sig, d.mean(), d.std(ddof=1) / np.sqrt(len(d)))
        print("%s, %s: adjoint %.5g  differences %.5g  relative error %.2f%% +- %.2f%%"
              % (key[0], "absorption" if key[1] == "a" else "scattering", lr.mean(), fd.mean(), 100 * rel, 100 * sig),
              flush=True)
    return out


def test_c5_gradient_vs_crn_differences(c1_dir):
    out = _c5(c1_dir)
    for x in "ab":
        rel, sig, _, _ = out[("hits", x)]
        assert abs(rel) + 2 * sig <= 0.05
        for wn in ("near", "far", "llh"):
            _, _, dm, ds = out[(wn, x)]
            assert abs(dm) <= 4 * ds
