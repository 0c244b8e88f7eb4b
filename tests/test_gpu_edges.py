"""Edge cases of a launch on the GPU: an empty launch, bunches of zero photons
between others, ragged bunch sizes under the dynamic work distribution (every
photon started exactly once), and a hit buffer too small for the launch (the
reference counts the overflow and keeps what fits, gpu/ppc.cu:262-270)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ENV = {"WFLA": "405", "FWID": "10", "DTMN": "1"}


@pytest.fixture(scope="module")
def det(c1_dir):
    from ppc_amd import api

    return api.Detector(c1_dir, ENV)


def test_empty_launch_then_normal_launch(det):
    from ppc_amd import api

    dev = api.Device(0)
    dev.upload(det)
    print("empty launch", flush=True)
    st = dev.run([])
    assert st.photons == 0 and st.hits == 0 and st.stored == 0 and st.steps == 0
    assert st.span_min_ms == st.span_max_ms == 0  # no kernel ran
    assert len(dev.hits()) == 0
    st = dev.run([det.flasher_bunch(1, 30, num=1_000_000)])
    assert st.photons == 1_000_000 and st.hits == st.stored > 0
    # the in-kernel span: first wave end <= last wave end, within the launch's event time
    print("device %.3f ms, in-kernel %.3f...%.3f ms" % (st.device_ms, st.span_min_ms, st.span_max_ms), flush=True)
    assert 0 < st.span_min_ms <= st.span_max_ms <= st.device_ms * 1.05 + 0.05
    dev.close()


def test_zero_photon_bunches_between_others(det):
    from ppc_amd import api

    sizes = [0, 3_000_000, 0, 0, 1_000_000, 0]
    bunches = [det.flasher_bunch(1, 30, num=n, seg=k) for k, n in enumerate(sizes)]
    dev = api.Device(0)
    dev.upload(det)
    print("bunch sizes %s" % sizes, flush=True)
    st = dev.run(bunches)
    h = dev.hits()
    dev.close()
    assert st.photons == sum(sizes) and st.stored == st.hits == len(h)
    segs = np.bincount(h["n"], minlength=len(sizes))
    print("hits per bunch %s" % segs.tolist(), flush=True)
    assert all(segs[k] == 0 for k, n in enumerate(sizes) if n == 0)
    # the same source in both bunches: hits in proportion to the photons (binomial, 5 sigma)
    n1, n4 = segs[1], segs[4]
    f = 3 / 4
    assert abs(n1 - f * (n1 + n4)) < 5 * np.sqrt((n1 + n4) * f * (1 - f)) + 1


@pytest.mark.parametrize("chunk", [0, 64, 1024])
def test_ragged_bunches_every_photon_started_once(det, chunk):
    from ppc_amd import api

    rng = np.random.default_rng(7 + chunk)
    sizes = [1, 0, 7, 1000, 0, 123457, 5, 1_999_999] + rng.integers(0, 50_000, 40).tolist()
    total = int(sum(sizes))
    bunches = [det.flasher_bunch(1 + k % 7, 1 + (7 * k) % 60, num=int(n), seg=k) for k, n in enumerate(sizes)]
    dev = api.Device(0, chunk=chunk, rec_stride=1, rec_max=total, max_bunches=len(sizes))
    dev.upload(det)
    print("ragged bunches: %d bunches, %d photons, chunk %d" % (len(sizes), total, chunk), flush=True)
    st = dev.run(bunches)
    rec = dev.records()
    h = dev.hits()
    dev.close()
    assert st.photons == total
    # every photon index wrote its own start record: all started, none twice (total starts)
    assert len(rec) == total and np.array_equal(rec["photon"], np.arange(total, dtype=np.uint32))
    # hits carry the segment of their own bunch, none of an empty bunch
    assert np.all(np.isin(h["n"], [k for k, n in enumerate(sizes) if n > 0]))


def test_hit_buffer_overflow_keeps_what_fits(det):
    from ppc_amd import api

    b = det.flasher_bunch(1, 30, num=4_000_000)
    full = api.Device(0, deterministic=True)
    full.upload(det)
    full.seed(11)
    st_full = full.run([b])
    h_full = full.hits()
    full.close()
    small = api.Device(0, deterministic=True, max_hits=1000)
    small.upload(det)
    small.seed(11)
    print("overflow: %d hits into a buffer of 1000" % st_full.hits, flush=True)
    st = small.run([b])
    h = small.hits()
    small.close()
    assert st_full.hits > 2000 and st_full.stored == st_full.hits
    assert st.hits == st_full.hits and st.stored == 1000 and len(h) == 1000
    # the kept hits are hits of the same (deterministic) launch
    key = lambda a: set(map(bytes, a.view(np.dtype((np.void, a.dtype.itemsize)))))
    assert key(h) <= key(h_full)


def test_bad_launch_shapes_are_rejected():
    from ppc_amd import api

    for kw in ({"threads": 100}, {"threads": 2048}, {"chunk": 2048}):
        print("launch shape %s" % kw, flush=True)
        with pytest.raises(api.PPCError):
            api.Device(0, **kw)
    dev = api.Device(0)  # and a good one still opens afterwards
    assert dev.lanes > 0
    dev.close()


def test_nonfinite_layer_guard_like_the_oracle(det):
    """A NaN forced from the host into one layer's absorption factor: a photon whose optical
    depth turns non-finite ends in that step and is counted (the guard of the reference's GPU
    build, gpu/pro.cu:727-729), every launch completes, and photon by photon the lanes end in
    the oracle's MWC states (same draws, the guard at the same step)."""
    import ctypes

    import oracle_util as U
    from ppc_amd import api

    src = (2.0, 1.0, 3.0)
    hmin, rdh = float(det.scalar("hmin")), float(det.scalar("rdh"))
    J = int(round((src[2] - hmin) * rdh)) + 3  # 30 m above the source
    lay = det.layers()
    lay["p"][J] = np.nan
    dev = api.Device(0, deterministic=True)
    dev.upload(det)
    dev.set_layers(lay)
    rng = np.random.default_rng(31)
    x = rng.integers(1, 2 ** 32, size=dev.lanes, dtype=np.uint64)
    c = rng.integers(0, 2 ** 30, size=dev.lanes, dtype=np.uint64)
    dev.set_states(x | (c << np.uint64(32)))
    N = dev.lanes  # one photon per lane
    print("NaN absorption in layer %d, %d photons" % (J, N), flush=True)
    st = dev.run([U.to_api_bunch(U.point_bunch(*src, num=N))])
    fin = dev.states()
    dev.close()
    assert st.photons == N
    # the oracle on the first photons, with the same NaN in its tables
    T = U.product_tables(det)
    ab = np.frombuffer(bytes(T.field("abs_")), np.float32).reshape(U.WNUM, U.MAXLYS).copy()
    ab[:, J] = np.nan
    ctypes.memmove(T.field("abs_"), ab.tobytes(), ab.nbytes)
    rm = det.array("rm", np.uint32)
    lib = U.oracle(nan=True)
    b = U.point_bunch(*src, num=N)
    H = (U.Hit * 64)()
    n = 20000
    cnt = U.Counters()
    same = 0
    for k in range(n):
        s = U.Rng(int(x[k]), int(c[k]), int(rm[k]))
        lib.orc_photon(T.ptr, ctypes.byref(b), 1, k, ctypes.byref(s), H, 64, ctypes.byref(cnt))
        same += (s.x, s.c) == (int(fin[k]) & 0xFFFFFFFF, int(fin[k]) >> 32)
    f_gpu, f_orc = st.nonfinite / N, cnt.nonfinite / n
    print("non-finite: GPU %d of %d (%.4f), oracle %d of %d (%.4f); %d of %d lanes in the oracle's state"
          % (st.nonfinite, N, f_gpu, cnt.nonfinite, n, f_orc, same, n), flush=True)
    assert cnt.nonfinite > 100 and st.nonfinite > 0
    assert same >= 0.99 * n
    assert abs(f_gpu - f_orc) < 5 * np.sqrt(f_orc * (1 - f_orc) / n)
