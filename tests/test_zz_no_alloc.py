"""ppc_hip_adjoint_noalloc (the adjoint that allocates no device memory: its staging,
gradient accumulators and counters sit in a fixed scratch area of the GPU) against
ppc_hip_adjoint on the same forward launch, through the C-ABI: two calls in a row, the
second with fewer replay records and fewer weights, each equal to the allocating entry
point's gradient to FP32 accumulation-order tolerance and with the same replay counts;
calls beyond the scratch area's limits are refused before any device work."""
import numpy as np
import pytest

import oracle_util as U

pytestmark = pytest.mark.gpu

SRC = (2.0, 1.0, 3.0)
SCRATCH_RECORDS, SCRATCH_WEIGHTS = 1 << 16, 1 << 18


def _same(g_alloc, g_noalloc, st_alloc, st_noalloc):
    scale = np.abs(g_alloc).max()
    assert scale > 0
    assert np.allclose(g_noalloc, g_alloc, rtol=1e-4, atol=1e-6 * scale)
    assert st_noalloc.photons == st_alloc.photons > 0
    assert st_noalloc.nonfinite == st_alloc.nonfinite == 0


def test_adjoint_noalloc_matches_adjoint(c1_dir):
    from ppc_amd import api

    det = api.Detector(c1_dir, {"WFLA": None, "FWID": None, "DTMN": "1"})
    q = det.ice_parameters()
    gsize = int(det.scalar("gsize", np.int32))
    dev = api.Device(0, record_starts=True, deterministic=True)
    dev.upload(det)
    print("no-alloc adjoint: c1 point source, 300000 photons", flush=True)
    b = U.to_api_bunch(U.point_bunch(*SRC, num=300000))
    dev.launch([b], 300000)
    st = dev.wait()
    rec = dev.hit_photons()
    assert st.hit_photons == len(rec) > 500
    rng = np.random.default_rng(9)
    w = rng.uniform(0.5, 1.5, size=(1, gsize)).astype(np.float32)
    # call 1: every photon that hit, recorded on the device
    ga, gb, sa = dev.adjoint([b], 300000, w, q["dpda"], q["ds1db"])
    na, nb_, sn = dev.adjoint([b], 300000, w, q["dpda"], q["ds1db"], noalloc=True)
    _same(np.concatenate([ga, gb]), np.concatenate([na, nb_]), sa, sn)
    print("call 1: %d records, replays %d" % (len(rec), sn.photons), flush=True)
    # call 2, smaller: a third of the records passed by the caller, half the weights zero
    sub = rec[: len(rec) // 3]
    w2 = w.copy()
    w2[:, ::2] = 0
    ga, gb, sa = dev.adjoint([b], 300000, w2, q["dpda"], q["ds1db"], records=sub)
    na, nb_, sn = dev.adjoint([b], 300000, w2, q["dpda"], q["ds1db"], records=sub, noalloc=True)
    _same(np.concatenate([ga, gb]), np.concatenate([na, nb_]), sa, sn)
    print("call 2: %d records, replays %d" % (len(sub), sn.photons), flush=True)
    # beyond the scratch area: refused (host-side checks, no device work)
    many = np.repeat(rec[:1], SCRATCH_RECORDS + 1)
    with pytest.raises(api.PPCError, match="scratch"):
        dev.adjoint([b], 300000, w, q["dpda"], q["ds1db"], records=many, noalloc=True)
    wide = np.zeros((SCRATCH_WEIGHTS // gsize + 1, gsize), dtype=np.float32)
    with pytest.raises(api.PPCError, match="scratch"):
        dev.adjoint([b], 300000, wide, q["dpda"], q["ds1db"], noalloc=True)
    # the allocating entry point still takes the large record list
    ga, gb, sa = dev.adjoint([b], 300000, w, q["dpda"], q["ds1db"], records=many)
    assert sa.photons == 2 * len(many)
    dev.close()
