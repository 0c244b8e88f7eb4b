"""Photon replay: the HIP kernel and the oracle propagate the same photon from
the same MWC state and must produce the same hits.

Each lane of a deterministic launch gets a chosen state and propagates exactly
one photon; the oracle (CPU restatement of gpu/pro.cu) replays each photon from
the state the kernel recorded at its emission.  Tolerance: the kernel uses the
hardware transcendental instructions and FMA contraction, so a photon may take
a different path after a near-tie (FP32 reordering); at least 99 % of photons
must give the same hit DOM sequence, and matching hits must agree in time to
1e-4 relative and in direction to 2e-3 rad.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_util as U

pytestmark = pytest.mark.gpu

CASES = {
    "c1-flasher": ("c1", {"WFLA": "405", "FWID": "10"}),
    "c2-cascade": ("c2", {"WFLA": None, "FWID": None}),
}


def run_case(name, nphot, seed):
    """Run in this process: returns (n, n_same, n_hits, details)."""
    from ppc_amd import api, configs

    cfgname, env = CASES[name]
    env = dict(env, DTMN="1")
    det = api.Detector(configs.ensure(cfgname), env)
    dev = api.Device(0, deterministic=True, rec_stride=1, rec_max=nphot)
    dev.upload(det)
    rng = np.random.default_rng(seed)
    states = rng.integers(0, 2 ** 32, size=dev.lanes, dtype=np.uint64)
    states |= rng.integers(0, 2 ** 31, size=dev.lanes, dtype=np.uint64) << np.uint64(32)
    dev.set_states(states)
    if cfgname == "c1":
        b = det.flasher_bunch(1, 30, num=nphot)
    else:
        b, _ = det.particle_bunch(10.0, 20.0, -51.93, 0.0, 1e5, 0.0, 11, (0.3, 0.4, -np.sqrt(0.75)), num=nphot)
    st = dev.run([b])
    assert st.photons == nphot
    hits, hph = dev.hits(with_photon=True)
    rec = dev.records()
    assert len(rec) == nphot
    dev.close()
    tab = U.product_tables(det)
    orc = U.oracle()
    ob = U.to_orc_bunch(b)
    cap = 256
    H = (U.Hit * cap)()
    same = 0
    nh_orc = 0
    by_photon = {}
    for k, p in zip(hits, hph):
        by_photon.setdefault(int(p), []).append(k)
    tmax = 0.0
    for r in rec:
        s = U.Rng(int(r["x"]), int(r["c"]), int(r["a"]))
        n = orc.orc_photon(tab.ptr, ctypes.byref(ob), 1, int(r["photon"]), ctypes.byref(s), H, cap, None)
        ho = U.hits_array(H, min(n, cap))
        hg = np.array(by_photon.get(int(r["photon"]), []), dtype=U.HIT_DTYPE)
        hg = np.sort(hg, order="t") if len(hg) else hg
        ho = np.sort(ho, order="t") if len(ho) else ho
        nh_orc += n
        if len(hg) == len(ho) and (hg["i"] == ho["i"]).all():
            ok = True
            if len(hg):
                dt = np.abs(hg["t"] - ho["t"]) / np.maximum(np.abs(ho["t"]), 1.0)
                da = np.abs(np.cos(hg["pth"]) - np.cos(ho["pth"]))
                tmax = max(tmax, float(dt.max()))
                ok = bool((dt < 1e-4).all() and (da < 2e-3).all())
            same += ok
    return len(rec), same, len(hits), nh_orc, tmax


@pytest.mark.parametrize("name", list(CASES))
def test_replay(name):
    n, same, nh_gpu, nh_orc, tmax = run_case(name, 65536, 11)
    print("%s: %d photons, %d identical, hits gpu %d oracle %d, max rel dt %.2e" % (name, n, same, nh_gpu, nh_orc, tmax))
    assert nh_gpu > 10
    assert same >= 0.99 * n
    assert abs(nh_gpu - nh_orc) <= 0.02 * nh_orc + 3
