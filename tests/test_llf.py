"""llh's likelihood llf (llh/llh.cxx:637-662) restated in ppc_amd.adjoint.llf, against
what its definition fixes:

  * it is the minimum over the two rates (ls, ld) of the negative log-likelihood of
    Poisson counts s ~ nsim*ls, d ~ ndat*ld with a Gaussian prior of width sig on
    log(ld/ls), less the same at the saturated rates (s/nsim, d/ndat) -- checked by a
    general-purpose minimiser on the objective itself;
  * sig -> 0 gives the two-sample saturated Poisson likelihood in closed form
    (ls = ld = (s+d)/(nsim+ndat)); a very wide sig gives 0 where both samples have counts
    (each at its own rate);
  * its float64 autograd gradient equals central finite differences.
"""
import math

import numpy as np
import pytest
import torch
from scipy import optimize

from ppc_amd.adjoint import LSIG, llf


def _objective_min(s, d, nsim, ndat, sig):
    def F(x):
        ls, ld = math.exp(x[0]), math.exp(x[1])
        v = nsim * ls - s * math.log(ls) + ndat * ld - d * math.log(ld)
        return v + (x[1] - x[0]) ** 2 / (2 * sig * sig)

    x0 = [math.log(max(s, 0.5) / nsim), math.log(max(d, 0.5) / ndat)]
    r = optimize.minimize(F, x0, method="Nelder-Mead", options={"xatol": 1e-12, "fatol": 1e-13, "maxiter": 20000})
    sat = (s - s * math.log(s / nsim) if s > 0 else 0) + (d - d * math.log(d / ndat) if d > 0 else 0)
    return r.fun - sat


CASES = [(5.0, 3.0, 1.0, 1.0), (0.0, 4.0, 1.0, 1.0), (3.0, 0.0, 2.0, 1.0), (120.0, 100.0, 10.0, 1.0),
         (7.3, 2.0, 1.0, 3.0), (40.0, 80.0, 1.0, 1.0)]


@pytest.mark.parametrize("s,d,nsim,ndat", CASES)
def test_llf_is_the_minimum_of_its_objective(s, d, nsim, ndat):
    got = float(llf(torch.tensor([s]), torch.tensor([d]), nsim, ndat)[0])
    want = _objective_min(s, d, nsim, ndat, LSIG)
    assert got == pytest.approx(want, rel=1e-6, abs=1e-8)


def test_llf_zero_counts_and_limits():
    s = torch.tensor([0.0, 5.0, 0.0, 12.0, 3.0], dtype=torch.float64)
    d = torch.tensor([0.0, 0.0, 6.0, 9.0, 11.0], dtype=torch.float64)
    assert float(llf(s, d)[0]) == 0.0
    for nsim, ndat in ((1.0, 1.0), (4.0, 1.0)):
        r = (s + d) / (nsim + ndat)
        sat = torch.where(s > 0, s * torch.log(s / nsim / r), torch.zeros_like(s)) + \
            torch.where(d > 0, d * torch.log(d / ndat / r), torch.zeros_like(d))
        sat = torch.where((s + d) > 0, sat, torch.zeros_like(s))
        assert torch.allclose(llf(s, d, nsim, ndat, sig=0), sat, rtol=1e-12, atol=1e-12)
        assert torch.allclose(llf(s, d, nsim, ndat, sig=1e-4), sat, rtol=1e-6, atol=1e-9)
        both = (s > 0) & (d > 0)
        assert float(llf(s, d, nsim, ndat, sig=1e4)[both].abs().max()) < 1e-6
    # the model error only lowers the mismatch cost
    assert bool((llf(s, d) <= llf(s, d, sig=0) + 1e-12).all())


def test_llf_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    s = torch.tensor(rng.uniform(0.5, 30, 12), dtype=torch.float64, requires_grad=True)
    d = torch.tensor(rng.poisson(10, 12).astype(float) + 0.5, dtype=torch.float64, requires_grad=True)
    for sig in (LSIG, 0.0, 0.5):
        assert torch.autograd.gradcheck(lambda a, b: llf(a, b, 2.0, 1.0, sig=sig), (s, d), eps=1e-6, atol=1e-6,
                                        rtol=1e-5)
    # the gradient by s is the log ratio of the simulated rate to the fitted one
    s0, d0 = torch.tensor([8.0]), torch.tensor([20.0])
    x = s0.clone().double().requires_grad_(True)
    llf(x, d0).sum().backward()
    assert float(x.grad[0]) < 0  # more simulated light where the data have more: lower llf
