"""The oracle against the reference's own outputs, photon by photon.

* golden replay vectors (tests/golden/replay_*.npz, written by the reference's
  CPU build from random MWC states): the oracle, fed with the product's tables,
  must end every photon in the same MWC state (so it consumed the same draws in
  the same order) and produce the same raw hits;
* where oracle/_ref is built: fresh random states on the real ice models of the
  reference (birefringence, 1-d and 2-d tilt, anisotropy), oracle vs reference.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_util as U

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as G  # noqa: E402

REFICE = "/root/reference/ice"

_WORK = r'''
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[2]); sys.path.insert(0, sys.argv[2] + "/golden")
import oracle_util as U
import make_golden as G
from ppc_amd import api, configs
case = sys.argv[1]
cfg, env, src = G.CASES[case]
D = api.Detector(configs.ensure(cfg), env)
T = U.product_tables(D)
gold = np.load(sys.argv[2] + "/golden/replay_%s.npz" % case)
n, procs = int(gold["n"]), int(gold["procs"])
rm = D.array("rm", np.uint32)
b = G._bunch(T, src, 1, det=D)
orc = U.oracle()
cap = 1000
H = (U.Hit * cap)()
fin = np.zeros((n * procs, 2), dtype=np.uint32)
hp, hh = [], []
for p in range(procs):
    st = G.replay_states(p, n)
    for k in range(n):
        g = p * n + k
        s = U.Rng(int(st[k, 0]), int(st[k, 1]), int(rm[g]))
        m = orc.orc_photon(T.ptr, ctypes.byref(b), 1, 0, ctypes.byref(s), H, cap, None)
        fin[g] = s.x, s.c
        h = U.hits_array(H, min(m, cap))
        hh.append(h); hp.append(np.full(len(h), g, dtype=np.uint32))
hp, hh = np.concatenate(hp), np.concatenate(hh)
same_state = (G.fold(fin) == gold["digest"])
print("photons", n * procs)
print("same_state", int(same_state.sum()))
gh, gp = gold["hits"], gold["hit_photon"]
print("hits", len(hh), len(gh))
ok = 0
for g in np.unique(np.concatenate([hp, gp])):
    a, r = hh[hp == g], gh[gp == g]
    if len(a) == len(r) and (a["i"] == r["i"]).all() and np.allclose(a["t"], r["t"], rtol=1e-5) \
       and np.allclose(a["pth"], r["pth"], atol=1e-3):
        ok += 1
print("photons_with_hits", len(np.unique(np.concatenate([hp, gp]))), ok)
'''


def _run(code, *args, env=None):
    e = dict(os.environ)
    for k in G.ENV_KEYS:
        e.pop(k, None)
    e.update(env or {})
    e["DTMN"] = "1"
    e["PYTHONPATH"] = ROOT + os.pathsep + HERE
    r = subprocess.run([sys.executable, "-c", code] + list(args), env=e, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return dict((l.split()[0], [int(v) for v in l.split()[1:]]) for l in r.stdout.splitlines() if l.strip())


@pytest.mark.parametrize("case", list(G.CASES))
def test_oracle_replays_reference_vectors(case):
    out = _run(_WORK, case, HERE, env=G.CASES[case][1])
    n = out["photons"][0]
    assert out["same_state"][0] >= 0.999 * n
    nh_o, nh_g = out["hits"]
    assert abs(nh_o - nh_g) <= 2
    pwh, agree = out["photons_with_hits"]
    print(case, out)
    assert agree >= 0.98 * pwh - 1


_LIVE = r'''
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[2])
import oracle_util as U
T = U.ref_tables(sys.argv[1])
ref, orc = U.reference(), U.oracle()
# cascades next to a string of the real geometry, aimed in random directions
om = np.frombuffer(bytes(T.field("oms")), np.float32).reshape(-1, 5)
rng = np.random.default_rng(int(sys.argv[3]))
cap = 1000
HB, HO = (U.Hit * cap)(), (U.Hit * cap)()
n = same = domseq = hits = 0
for trial in range(8):
    d = om[rng.integers(0, int(T.get("gsize", np.int32)[0]))]
    v = rng.normal(size=3); v /= np.linalg.norm(v)
    b = U.cascade_bunch(float(d[2]) + 6.0, float(d[3]) - 4.0, float(d[4]) + 3.0, *v, 1, E=1e4)
    for k in range(600):
        x, c, a = int(rng.integers(0, 2**32)), int(rng.integers(0, 2**31)), int(ref.ref_multiplier(k))
        ox, oc = ctypes.c_uint32(), ctypes.c_uint32()
        n1 = ref.ref_run(ctypes.byref(b), 1, 1, 1, 0, a, x, c, HB, cap, ctypes.byref(ox), ctypes.byref(oc))
        s = U.Rng(x, c, a)
        n2 = orc.orc_photon(T.ptr, ctypes.byref(b), 1, 0, ctypes.byref(s), HO, cap, None)
        h1, h2 = U.hits_array(HB, min(n1, cap)), U.hits_array(HO, min(n2, cap))
        n += 1; hits += n1
        same += (s.x, s.c) == (ox.value, oc.value)
        domseq += n1 == n2 and (h1["i"] == h2["i"]).all()
print("n", n); print("same", same); print("domseq", domseq); print("hits", hits)
'''


@pytest.mark.ref
@pytest.mark.parametrize("ice", ["spice_bfr-v2", "spice_ftp-v3m"])
def test_oracle_vs_reference_real_ice(ice):
    if not U.ref_available() or not os.path.isdir(os.path.join(REFICE, ice)):
        pytest.skip("reference build or ice directories not present")
    out = _run(_LIVE, os.path.join(REFICE, ice), HERE, "7")
    n = out["n"][0]
    assert out["hits"][0] >= 5
    assert out["same"][0] >= 0.995 * n
    assert out["domseq"][0] >= 0.995 * n
