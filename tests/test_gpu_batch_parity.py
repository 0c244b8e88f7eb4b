"""Sources of a batch over ice variants against the oracle, source by source.

Three sources of the batched workload (a flasher, an EM cascade and a 10 m muon
segment on the 7-string ice) run in one launch, each in its own ice variant
(per-layer adust(400) and be(400) scaled by factors in [0.9, 1.1]).  Each source
is also propagated by the oracle (CPU restatement of gpu/pro.cu) in the dense
tables of its variant; per source: hit rate (Poisson), per-DOM hits (two-sample
chi-square) and arrival times (two-sample KS), each at p > 0.01.
"""
import ctypes
import math
import multiprocessing
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import pytest
from scipy import stats

import oracle_util as U
from test_gpu_golden import _two_sample_chi2

pytestmark = pytest.mark.gpu

ENV = {"WFLA": "405", "FWID": "10", "DTMN": "1"}
N_GPU = 10_000_000
N_ORC = 4_000_000
WORKERS = 8


def _sources(det):
    from ppc_amd import configs

    bunches, variant, layers = configs.c4_batch(det, seed=5)
    pick = [next(k for k in range(len(bunches)) if k % 3 == kind) for kind in range(3)]
    # three different variants for the three sources
    var = [0, 3, 6]
    return [bunches[k] for k in pick], var, layers


def _variant_tables(det, lay):
    """oracle tables of the loaded ice with its layer factors replaced by lay"""
    T = U.product_tables(det)
    wv = np.frombuffer(det.field("wv"), dtype=np.float32).reshape(U.WNUM, 8)
    la, lk, abl = wv[:, 3:4], wv[:, 4:5], wv[:, 5:6]
    size = len(lay)
    sca = np.zeros((U.WNUM, U.MAXLYS), dtype=np.float32)
    ab = np.zeros((U.WNUM, U.MAXLYS), dtype=np.float32)
    sca[:, :size] = lay["s0"][None] + lay["s1"][None] * la
    ab[:, :size] = lay["p"][None] * lk + lay["t"][None] * abl
    ctypes.memmove(T.field("sca"), sca.tobytes(), sca.nbytes)
    ctypes.memmove(T.field("abs_"), ab.tobytes(), ab.nbytes)
    return T


def _oracle_part(args):
    """one worker: num photons of source k in variant v from MWC stream `seed`"""
    icedir, k, v, num, seed = args
    from ppc_amd import api

    det = api.Detector(icedir, ENV)
    bunches, var, layers = _sources(det)
    T = _variant_tables(det, layers[var[k]])
    b = U.to_orc_bunch(bunches[k])
    b.num = num
    rm = det.array("rm", np.uint32)
    s = U.Rng(362436069 + seed, 1234567, int(rm[seed]))
    cap = 200000
    H = (U.Hit * cap)()
    n = U.oracle().orc_run_xcpu(T.ptr, ctypes.byref(b), 1, num, 512, ctypes.byref(s), H, cap, None)
    return U.hits_array(H, min(n, cap))


def test_batch_sources_match_oracle(c1_dir):
    from ppc_amd import api

    det = api.Detector(c1_dir, ENV)
    srcs, var, layers = _sources(det)
    for b in srcs:
        b.num = N_GPU
    dev = api.Device(0, deterministic=True, max_bunches=8)
    dev.upload(det)
    dev.set_layers(layers)
    dev.seed(777)
    print("batch parity: 3 sources x %d photons in variants %s" % (N_GPU, var), flush=True)
    dev.launch_batch(srcs, var)
    st = dev.wait()
    assert st.photons == 3 * N_GPU and st.stored == st.hits
    hits, photon = dev.hits(with_photon=True)
    dev.close()
    jobs = [(c1_dir, k, var[k], N_ORC // WORKERS, 100 * k + w) for k in range(3) for w in range(WORKERS)]
    # fresh interpreters for the CPU workers (this process has initialised the GPU)
    with ProcessPoolExecutor(WORKERS, mp_context=multiprocessing.get_context("spawn")) as ex:
        parts = list(ex.map(_oracle_part, jobs))
    gsize = int(det.scalar("gsize", np.int32))
    for k in range(3):
        hg = hits[photon // N_GPU == k]
        ho = np.concatenate(parts[k * WORKERS:(k + 1) * WORKERS])
        rate_g, rate_o = len(hg) / N_GPU, len(ho) / N_ORC
        z = (rate_g - rate_o) / math.sqrt(len(ho) / N_ORC ** 2 + len(hg) / N_GPU ** 2)
        p_rate = 2 * stats.norm.sf(abs(z))
        chi2, dof, p_dom = _two_sample_chi2(np.bincount(hg["i"], minlength=gsize), np.bincount(ho["i"], minlength=gsize))
        p_t = stats.ks_2samp(hg["t"], ho["t"]).pvalue
        print("source %d (type %d, variant %d): rate %.4e vs %.4e (p=%.3f); DOM chi2 %.1f/%d (p=%.3f); KS time "
              "p=%.3f" % (k, srcs[k].type, var[k], rate_g, rate_o, p_rate, chi2, dof, p_dom, p_t), flush=True)
        assert len(ho) > 150
        assert p_rate > 0.01 and p_dom > 0.01 and p_t > 0.01
