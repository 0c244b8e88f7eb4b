"""Batches over ice variants and captured launches.

Deterministic launches (photon i on lane i mod lanes) from the same lane states
must give bit-identical hits whichever way the work is submitted:
  * a batch whose bunches use two ice variants equals, bunch by bunch, plain
    launches in each variant's ice;
  * every replay of a captured graph equals an eager launch from the lane
    states the replay started from, with eager launches of other bunches and
    synchronisations between the replays.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def det(c2_dir):
    from ppc_amd import api

    return api.Detector(c2_dir, {"WFLA": None, "FWID": None, "DTMN": "1"})


def _states(lanes, seed):
    rng = np.random.default_rng(seed)
    s = rng.integers(1, 2 ** 32, size=lanes, dtype=np.uint64)
    return s | (rng.integers(0, 2 ** 30, size=lanes, dtype=np.uint64) << np.uint64(32))


def _sorted_hits(dev):
    h, p = dev.hits(with_photon=True)
    order = np.lexsort((h["t"], h["i"], p))
    return h[order], p[order]


def _cascade(det, num, seg=0, x=10.0):
    b, _ = det.particle_bunch(x, 20.0, -51.93, 0.0, 1e5, 0.0, 11, (0.3, 0.4, -np.sqrt(0.75)), num=num, seg=seg)
    return b


def test_variants_equal_separate_launches(det):
    from ppc_amd import api

    dev = api.Device(0, deterministic=True)
    dev.upload(det)
    n = dev.lanes // 2
    q = det.ice_parameters()
    lay0 = det.layers()
    lay1 = det.layers_for(q["par_a"] * 1.3, q["par_b"] * 0.8)
    b0, b1 = _cascade(det, n, seg=0), _cascade(det, n, seg=1, x=-30.0)
    S = _states(dev.lanes, 1)
    print("variants: batch of 2 bunches x %d photons in 2 ice variants" % n, flush=True)
    dev.set_layers(np.stack([lay0, lay1]))
    dev.set_states(S)
    dev.launch_batch([b0, b1], [0, 1])
    st = dev.wait()
    assert st.photons == 2 * n
    hb, pb = _sorted_hits(dev)
    ref = []
    for lay in (lay0, lay1):
        dev.set_layers(lay)
        dev.set_states(S)
        dev.launch([b0, b1])
        ref.append(_sorted_hits(dev))
    dev.close()
    for v, (h, p) in enumerate(ref):
        sel = (p >= v * n) & (p < (v + 1) * n)
        selb = (pb >= v * n) & (pb < (v + 1) * n)
        assert selb.sum() > 50
        assert np.array_equal(pb[selb], p[sel])
        assert np.array_equal(hb[selb], h[sel])
    # the two variants differ
    assert not np.array_equal(ref[0][0]["i"], ref[1][0]["i"])


def test_variant_out_of_range_is_rejected(det):
    from ppc_amd import api

    dev = api.Device(0)
    dev.upload(det)
    with pytest.raises(api.PPCError):
        dev.launch_batch([_cascade(det, 1000)], [1])
    dev.close()


def test_graph_replays_equal_eager_launches(det):
    from ppc_amd import api

    dev = api.Device(0, deterministic=True)
    dev.upload(det)
    b = _cascade(det, 3 * dev.lanes // 2)
    other = _cascade(det, 5000, x=-50.0)
    dev.set_states(_states(dev.lanes, 2))
    print("graph: capture one cascade bunch of %d photons, 3 replays" % b.num, flush=True)
    dev.graph_capture([b])
    for k in range(3):
        S = dev.states()
        dev.graph_launch()
        st = dev.wait()
        assert st.photons == b.num
        hg, pg = _sorted_hits(dev)
        after = dev.states()
        dev.set_states(S)
        dev.launch([b])
        dev.wait()
        he, pe = _sorted_hits(dev)
        assert len(hg) > 100
        assert np.array_equal(pg, pe) and np.array_equal(hg, he)
        assert np.array_equal(dev.states(), after)
        dev.set_states(after)
        dev.run([other])  # eager work between replays
        print("replay %d: %d hits equal" % (k, len(hg)), flush=True)
    dev.close()


def test_clearance_skip_changes_no_hit(det):
    """Steps that cannot reach any string skip the cell-grid search; deterministic
    launches with and without that skip give identical hits."""
    from ppc_amd import api

    out = []
    for off in (False, True):
        dev = api.Device(0, deterministic=True, no_clearance=off)
        dev.upload(det)
        dev.set_states(_states(dev.lanes, 3))
        print("clearance skip %s: cascade of %d photons" % ("off" if off else "on", 2 * dev.lanes), flush=True)
        dev.run([_cascade(det, 2 * dev.lanes)])
        out.append(_sorted_hits(dev))
        dev.close()
    assert len(out[0][0]) > 100
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][0], out[1][0])


def test_full_c4_batch_graph_equals_eager(c1_dir):
    """The whole batched workload on the 7-string ice -- 512 sources (flashers, cascades, 10 m
    muon segments) with photon counts log-uniform in [1e5, 1e8], 8 ice variants, 8.1e9 photons
    in groups of at most 3e9 (32-bit photon indices), one context per group -- launched once
    eagerly and once as a replay of its captured graph from the same lane states: the hits
    are identical, hit for hit, every photon of every source runs, and each source's hits
    carry its own segment only."""
    from ppc_amd import api, configs

    det = api.Detector(c1_dir, {"WFLA": "405", "FWID": "10", "DTMN": "1"})
    bunches, variant, layers = configs.c4_batch(det, seed=5)
    groups, cur, tot = [], [], 0
    for k, b in enumerate(bunches):
        if cur and tot + b.num > 3e9:
            groups.append(cur)
            cur, tot = [], 0
        cur.append(k)
        tot += b.num
    groups.append(cur)
    total_hits = 0
    for gi, g in enumerate(groups):
        bs = [bunches[k] for k in g]
        num = sum(b.num for b in bs)
        print("C4 group %d: %d sources, %.3g photons" % (gi, len(g), num), flush=True)
        dev = api.Device(0, deterministic=True, max_bunches=len(g), max_hits=1 << 25)
        dev.upload(det)
        dev.set_layers(layers)
        S = _states(dev.lanes, 40 + gi)
        dev.set_states(S)
        dev.launch_batch(bs, variant[g])
        st = dev.wait()
        assert st.photons == num and st.stored == st.hits
        he, pe = _sorted_hits(dev)
        dev.graph_capture(bs, variant[g])
        dev.set_states(S)
        dev.graph_launch()
        st2 = dev.wait()
        hg, pg = _sorted_hits(dev)
        dev.close()
        assert st2.photons == num and st2.hits == st.hits
        assert np.array_equal(pe, pg) and np.array_equal(he, hg)
        assert set(np.unique(he["n"]).tolist()) <= set(g)
        total_hits += st.hits
    assert total_hits > 100000
    print("C4: %d groups, %d hits, graph replays equal to eager launches" % (len(groups), total_hits), flush=True)
