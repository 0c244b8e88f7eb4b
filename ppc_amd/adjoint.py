"""Gradient of a likelihood of DOM hit counts by the per-layer ice parameters.

The forward simulation propagates photons in ice whose layer i has absorption
parameter a_i (icemodel.dat adust(400)) and scattering parameter b_i
(be(400)); the layer factors of the kernel depend on them affinely
(p_i = Dc*a_i + ..., s1_i = b_i/(1-g) + ..., gpu/ini.cxx:1630-1640).  The
expected hit count of DOM d in track segment s is mu[s, d] = scale * n[s, d],
n the simulated count.

The gradient is the score-function (likelihood-ratio) estimator of the path
density: a hit's weight times d log p(path up to the hit)/d theta, where along
every straight segment of length l_k in layer k the absorption score is
-nr*lk*dp/da*l_k and the scattering score -la*ds1/db*l_k, plus
la*ds1/db / sca_J at every scattering in layer J.  The backward pass replays
only the photons that hit, each from the MWC state it started from (kept by the
forward launch on the device): once to sum the weights of its hits, once to
score its path with the weight of the hits it still has to make.  It runs on
the GPU (ppc_hip_adjoint) and accumulates in FP64.

    sim = HitSimulation(det, bunches, scale=1.0)
    a = torch.tensor(par_a, requires_grad=True); b = ...
    mu = ExpectedHits.apply(a, b, sim)          # [nseg, gsize]
    llh = poisson_llh(mu, data)                 # or llf(mu, data).sum(), llh's own likelihood
    llh.backward()                              # a.grad, b.grad
"""
import numpy as np
import torch

from . import api


class HitSimulation:
    """A device context, the bunches of one event and the bookkeeping between a
    forward launch and its adjoint.  The ice parameters enter through
    Detector.layers_for(a, b); everything else of the ice stays as loaded."""

    def __init__(self, det, bunches, scale=1.0, nseg=None, device=0, **device_kw):
        self.det = det
        self.bunches = [bunches] if isinstance(bunches, api.Bunch) else list(bunches)
        self.num = sum(b.num for b in self.bunches)
        self.scale = float(scale)
        self.gsize = int(det.scalar("gsize", np.int32))
        self.nseg = nseg if nseg is not None else 1 + max(b.q for b in self.bunches)
        q = det.ice_parameters()
        self.dpda, self.ds1db = q["dpda"], q["ds1db"]
        self.par_a, self.par_b = q["par_a"], q["par_b"]
        self.dev = api.Device(device, record_starts=True, **device_kw)
        self.dev.upload(det)

    def close(self):
        self.dev.close()

    def forward(self, a, b):
        """simulate with per-layer a, b; returns (mu [nseg, gsize], record of the launch)"""
        self.dev.set_layers(self.det.layers_for(a, b))
        self.dev.launch(self.bunches, self.num)
        st = self.dev.wait()
        if st.hits > st.stored or st.hit_photons > self.dev.max_hits:
            raise api.PPCError("%d hits of %d photons exceed the buffers (%d); open the device with a larger "
                               "max_hits" % (st.hits, st.hit_photons, self.dev.max_hits))
        hits = self.dev.hits()
        seg = hits["n"].astype(np.int64)
        ok = seg < self.nseg
        n = np.zeros((self.nseg, self.gsize), dtype=np.float64)
        np.add.at(n, (seg[ok], hits["i"][ok].astype(np.int64)), 1.0)
        self._launches = getattr(self, "_launches", 0) + 1
        return n * self.scale, {"id": self._launches, "hits": hits, "stats": st}

    def backward(self, launch, grad_mu):
        """d(L)/d(a), d(L)/d(b) given dL/dmu [nseg, gsize] of the launch's mu.  The
        adjoint replays the hitting photons the launch recorded on the device, so
        it must follow its own forward launch."""
        if launch["id"] != self._launches:
            raise api.PPCError("the adjoint replays the most recent forward launch only")
        weight = np.asarray(grad_mu, dtype=np.float64) * self.scale
        ga, gb, st = self.dev.adjoint(self.bunches, self.num, weight.astype(np.float32), self.dpda, self.ds1db)
        self.last_adjoint = {"photons": launch["stats"].hit_photons, "replays": st.photons,
                             "missed": st.nonfinite, "ms": st.device_ms}
        return ga, gb


class ExpectedHits(torch.autograd.Function):
    """mu[s, d] = scale * (simulated hits of DOM d in segment s) as a function of the
    per-layer ice parameters a, b (float tensors [size]); the backward pass is the
    GPU adjoint of that very launch (same photons, same random numbers)."""

    @staticmethod
    def forward(ctx, a, b, sim):
        mu, launch = sim.forward(a.detach().cpu().double().numpy(), b.detach().cpu().double().numpy())
        ctx.sim, ctx.launch = sim, launch
        ctx.dtypes = (a.dtype, b.dtype)
        return torch.from_numpy(mu)

    @staticmethod
    def backward(ctx, grad_mu):
        ga, gb = ctx.sim.backward(ctx.launch, grad_mu.detach().cpu().double().numpy())
        return torch.from_numpy(ga).to(ctx.dtypes[0]), torch.from_numpy(gb).to(ctx.dtypes[1]), None


def poisson_llh(mu, k, floor=1e-9):
    """sum over bins of k log mu - mu (the hit-count log-likelihood up to a constant)"""
    k = torch.as_tensor(k, dtype=mu.dtype)
    return (k * torch.log(mu.clamp_min(floor)) - mu).sum()


# model error of llh's likelihood: LSSL::SIG (llh/lssl.cxx:4), set from LSIG (llh/llh.cxx:2446-2448)
LSIG = 0.10


def _llf_solve(s, d, nsim, ndat, sig, tol=1e-8, max_iter=200):
    """rates (ls, ld) per simulated and per data repetition that minimise
    nsim ls - s log ls + ndat ld - d log ld + log(ld/ls)^2 / (2 sig^2): Newton's method on
    ls with nsim ls + ndat ld = s + d (llh/llh.cxx:643-656), all bins at once"""
    sd = s + d
    ntot = nsim + ndat
    ls = sd / ntot
    ld = ls.clone()
    if sig <= 0:
        return ls, ld
    s2 = sig * sig
    act = sd > 0
    for _ in range(max_iter):
        if not bool(act.any()):
            break
        f = nsim * ls - s - torch.log(ld / ls) / s2
        df = 1 + (1 / (ndat * ld) + 1 / (nsim * ls)) / s2
        dls = torch.where(act, -f / (nsim * df), torch.zeros_like(ls))
        ls = ls + dls
        ld = torch.where(act, (sd - nsim * ls) / ndat, ld)
        if bool(((ls < 0) | (ld < 0)).any()):
            raise ValueError("llf: no solution for the rates (negative rate in Newton's method)")
        act = act & (dls.abs() > tol * ls)
    else:
        if bool(act.any()):
            raise ValueError("llf: Newton's method did not converge")
    return ls, ld


class _Llf(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s, d, nsim, ndat, sig):
        if bool((s < 0).any()) or bool((d < 0).any()):
            raise ValueError("llf: negative counts")
        ls, ld = _llf_solve(s, d, nsim, ndat, sig)
        zero = torch.zeros_like(s)
        llx = torch.log(ld / ls) ** 2 / (2 * sig * sig) if sig > 0 else zero
        lls = torch.where(s > 0, s * torch.log((s / nsim) / ls), zero)
        lld = torch.where(d > 0, d * torch.log((d / ndat) / ld), zero)
        both0 = (s == 0) & (d == 0)
        ctx.save_for_backward(s, d, ls, ld)
        ctx.nsim, ctx.ndat = nsim, ndat
        return torch.where(both0, zero, llx + lls + lld)

    @staticmethod
    def backward(ctx, g):
        # at the minimising rates the value's derivative is the explicit one (envelope):
        # d/ds = log((s/nsim)/ls), d/dd = log((d/ndat)/ld); a bin with no counts on a side
        # has no finite derivative there and gets 0 (a bin without simulated hits has no
        # photons for the adjoint to weight)
        s, d, ls, ld = ctx.saved_tensors
        zero = torch.zeros_like(s)
        gs = torch.where(s > 0, torch.log((s / ctx.nsim) / ls.clamp_min(1e-300)), zero)
        gd = torch.where(d > 0, torch.log((d / ctx.ndat) / ld.clamp_min(1e-300)), zero)
        return g * gs, g * gd, None, None, None


def llf(s, d, nsim=1.0, ndat=1.0, sig=None):
    """llh's likelihood of simulated counts s (summed over nsim repetitions) against data
    counts d (over ndat repetitions), bin by bin (llh/llh.cxx:637-662): the saturated
    two-sample Poisson likelihood with a Gaussian model error sig on log(ld/ls) between
    the rates of the two samples (sig = LSIG, 10 %; sig = 0: no model error).  Smaller is
    better; 0 where both are 0.  Differentiable in s and d (float64)."""
    sig = LSIG if sig is None else float(sig)
    s = torch.as_tensor(s, dtype=torch.float64)
    d = torch.as_tensor(d, dtype=torch.float64)
    s, d = torch.broadcast_tensors(s, d)
    return _Llf.apply(s, d, float(nsim), float(ndat), sig)
