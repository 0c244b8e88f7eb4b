"""The drop-in paths on the GPU against the reference's own command line.

The reference's CPU program ran the C1 flasher (`ppc 1 30 N dev`, WFLA=405,
FWID=10) and its HIT lines are in tests/golden/cli_c1.npz.  Here the same
command of this package (bin/ppc, the HIP kernel behind the xppc front-end
and print()'s acceptance) and the library pipeline (initialize, choose, ini,
flset, sett, flone, eout -> hitz, the calls llh makes) must reproduce the
HIT rate per propagated photon (Poisson, p > 0.01), the distribution over
DOMs (two-sample chi-square, p > 0.01) and the arrival-time and wavelength
distributions (two-sample KS, p > 0.01).
"""
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest
from scipy import stats

from test_gpu_golden import G, _two_sample_chi2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _env():
    from ppc_amd import configs

    e = dict(os.environ)
    e.update({"WFLA": "405", "FWID": "10", "PPCTABLESDIR": configs.ensure("c1"), "DTMN": "1"})
    return e


def _compare(photons, strs, doms, t, wv, fixture="cli_c1.npz"):
    gold = np.load(os.path.join(HERE, "golden", fixture))
    Nr, nr = int(gold["photons"]), len(gold["t"])
    z = (len(t) / photons - nr / Nr) / math.sqrt(nr / Nr ** 2 + len(t) / photons ** 2)
    p_rate = 2 * stats.norm.sf(abs(z))
    key = lambda s, d: np.asarray(s, np.int64) * 100 + np.asarray(d, np.int64)
    kg, kp = key(gold["str"], gold["dom"]), key(strs, doms)
    allk = np.union1d(kg, kp)
    cg = np.array([np.sum(kg == k) for k in allk])
    cp = np.array([np.sum(kp == k) for k in allk])
    chi2, dof, p_dom = _two_sample_chi2(cp, cg)
    p_t = stats.ks_2samp(t, gold["t"]).pvalue
    p_w = stats.ks_2samp(wv, gold["wv"]).pvalue
    print("HIT rate %.4e vs %.4e (p=%.3f), DOM chi2 %.1f/%d (p=%.3f), KS time p=%.3f, wavelength p=%.3f"
          % (len(t) / photons, nr / Nr, p_rate, chi2, dof, p_dom, p_t, p_w))
    assert p_rate > 0.01 and p_dom > 0.01 and p_t > 0.01 and p_w > 0.01


def test_cli_flasher_iterations():
    """`ppc 1 30 N*4 0`: four flashes of N photons (gpu/f2k.cxx:768-783, ppc.cu:584-588),
    each closed by an empty segment line after its HIT lines; the four flashes have
    Poisson-compatible hit counts and the same HIT rate as the reference CLI's single flash"""
    from ppc_amd import _native

    r = subprocess.run([_native.ppc_exe(), "1", "30", "4e9*4", "0"], env=_env(), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    groups, cur = [], 0
    for l in r.stdout.splitlines():
        if l.startswith("HIT "):
            cur += 1
        else:
            assert l == "", l
            groups.append(cur)
            cur = 0
    assert cur == 0 and len(groups) == 4, (groups, cur)
    photons = sum(int(m) for m in re.findall(r"photons: (\d+)", r.stderr))
    g = np.array(groups, dtype=float)
    chi2 = np.sum((g - g.mean()) ** 2 / g.mean())
    gold = np.load(os.path.join(HERE, "golden", "cli_c1.npz"))
    Nr, nr = int(gold["photons"]), len(gold["t"])
    z = (g.sum() / photons - nr / Nr) / math.sqrt(nr / Nr ** 2 + g.sum() / photons ** 2)
    print("hits per flash %s, chi2 %.1f/3, rate z %.2f" % (groups, chi2, z), flush=True)
    assert g.min() > 1000 and stats.chi2.sf(chi2, 3) > 0.01 and 2 * stats.norm.sf(abs(z)) > 0.01


def test_cli_small_launches_and_hit_quota():
    """NPHO=8 (photons per lane per launch, gpu/ppc.cu:117-131): the flashes of `ppc 1 30 N*3 0`
    go out in tens of launches, each printed while the next one runs, and keep the reference's
    HIT rate and the grouping by flash; with HQUO (photons per hit slot, gpu/ppc.cu:41-43) set so
    that a launch's hits overflow the buffer, every launch reports the overflow and prints the
    hits that fit"""
    from ppc_amd import _native

    env = _env()
    env["NPHO"] = "8"
    r = subprocess.run([_native.ppc_exe(), "1", "30", "4e9*3", "0"], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    launches = re.findall(r"photons: (\d+)  hits: (\d+)", r.stderr)
    photons = sum(int(p) for p, _ in launches)
    lines = r.stdout.splitlines()
    nh = sum(l.startswith("HIT ") for l in lines)
    assert len(launches) > 30 and lines.count("") == 3 and nh + 3 == len(lines)
    gold = np.load(os.path.join(HERE, "golden", "cli_c1.npz"))
    Nr, nr = int(gold["photons"]), len(gold["t"])
    z = (nh / photons - nr / Nr) / math.sqrt(nr / Nr ** 2 + nh / photons ** 2)
    print("NPHO=8: %d launches, %d photons, %d HIT lines, rate z %.2f" % (len(launches), photons, nh, z), flush=True)
    assert 2 * stats.norm.sf(abs(z)) > 0.01

    env["HQUO"] = "20000"
    r = subprocess.run([_native.ppc_exe(), "1", "30", "4e9", "0"], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    over = re.findall(r"Error: data buffer overflow occurred: (\d+)>(\d+)!", r.stderr)
    launches = re.findall(r"photons: (\d+)  hits: (\d+)", r.stderr)
    nh = sum(l.startswith("HIT ") for l in r.stdout.splitlines())
    cap = int(over[0][1]) if over else 0
    print("HQUO=20000: %d launches, %d overflowing (buffer %d hits), %d HIT lines" % (len(launches), len(over), cap,
                                                                                 nh), flush=True)
    assert len(over) >= len(launches) - 1 > 10
    assert all(int(h) <= cap for _, h in launches) and 0 < nh <= cap * len(launches)


def test_cli_without_arguments_lists_devices():
    """`ppc` alone lists the devices and prints the usage (gpu/ppc.cu:404-425, 525-530)"""
    from ppc_amd import _native

    r = subprocess.run([_native.ppc_exe()], env=_env(), capture_output=True, text=True, timeout=120)
    print(r.stderr.strip(), flush=True)
    assert r.returncode == 0 and r.stdout == ""
    assert re.search(r"Found 1 devices", r.stderr) and re.search(r"^0: .* CUs\(256\) G\(\d+\)$", r.stderr, re.M)
    assert "[device] (f2k muons)" in r.stderr and "[str] [om] [num] [device] (flasher)" in r.stderr


@pytest.mark.parametrize("var", ["NPHO", "NPHO_0"])
def test_cli_npho_zero_drops_the_device(var):
    """NPHO=0 (or NPHO_<device>=0) takes the device out of use, and with no device left the
    command line stops with the reference's message and exit code 5 (gpu/ppc.cu:120-131)"""
    from ppc_amd import _native

    env = _env()
    env[var] = "0"
    r = subprocess.run([_native.ppc_exe(), "1", "30", "1e6", "0"], env=env, capture_output=True, text=True,
                       timeout=120)
    print("%s=0: exit %d" % (var, r.returncode), flush=True)
    assert r.returncode == 5 and "Not using device # 0!" in r.stderr and "No active GPU(s) selected!" in r.stderr
    assert "HIT" not in r.stdout


def test_cli_flasher_matches_reference_cli():
    from ppc_amd import _native

    r = subprocess.run([_native.ppc_exe(), "1", "30", "1.6e10", "0"], env=_env(), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    photons = sum(int(m) for m in re.findall(r"photons: (\d+)", r.stderr))
    rows = [l.split() for l in r.stdout.splitlines() if l.startswith("HIT ")]
    arr = np.array([[float(v) for v in row[1:5]] for row in rows])
    assert photons > 1e8
    _compare(photons, arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3])


@pytest.mark.parametrize("cfg", ["c2x", "c2xb"])
@pytest.mark.parametrize("accept", ["host", "device"])
def test_cli_flasher_full_feature_ice_matches_reference_cli(accept, cfg):
    """`ppc 36 30 N 0` on the 86-string ice with every optional feature (2-d tilt,
    birefringence, hole ice, both anisotropies) and cable positions (dx.dat, whose shadow
    print() applies, gpu/f2k.cxx:301-322) against the reference command line's HIT lines
    (tests/golden/cli_<cfg>.npz, 1.1e8 photons); c2xb adds tilted DOM axes (cx.dat: the
    flasher's beam, gpu/f2k.cxx:730-740, and the angular acceptance of the hits) and fixed
    hole-ice columns (str-f2k); with the acceptance on the host and on the device: HIT rate,
    per-DOM counts, arrival times and wavelengths"""
    from ppc_amd import _native, configs

    env = _env()
    env["PPCTABLESDIR"] = configs.ensure(cfg)
    if accept == "device":
        env["PPC_ACCEPT"] = "device"
    print("%s flasher 36-30, acceptance on the %s" % (cfg, accept), flush=True)
    r = subprocess.run([_native.ppc_exe(), "36", "30", "3.2e8", "0"], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Loaded 5160 cable positions" in r.stderr
    photons = sum(int(m) for m in re.findall(r"photons: (\d+)", r.stderr))
    arr = np.array([[float(v) for v in l.split()[1:5]] for l in r.stdout.splitlines() if l.startswith("HIT ")])
    assert photons > 2e8
    _compare(photons, arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3], fixture="cli_%s.npz" % cfg)


@pytest.mark.parametrize("cfg", ["c1ng", "c1wom"])
@pytest.mark.parametrize("accept", ["host", "device"])
def test_cli_nextgen_flasher_matches_reference_cli(accept, cfg):
    """`ppc 1 30 N 0` in the 7-string ice with nextgen sensors against the reference command
    line's "HIT str dom_pmt" lines (tests/golden/cli_<cfg>.npz): HIT rate, counts per (string,
    DOM, PMT), arrival times and wavelengths; with print()'s acceptance on the host (the
    reference's) and on the device (PPC_ACCEPT=device).  c1ng: multi-PMT modules (a 12-PMT
    module on string 1, a 2-PMT module on string 2, ellipsoidal; ini.cxx:731-761); c1wom:
    wavelength-shifting tubes on strings 1 and 2 (cylinders; their two ends and the transit
    time through the tube, ini.cxx:211-243), whose arrival times on the two DOMs next to the
    flasher are compared on their own as well (the transit adds a few ns to times of ~100 ns)"""
    from ppc_amd import _native, configs

    env = _env()
    env["PPCTABLESDIR"] = configs.ensure(cfg)
    if accept == "device":
        env["PPC_ACCEPT"] = "device"
    print("%s flasher, acceptance on the %s" % (cfg, accept), flush=True)
    r = subprocess.run([_native.ppc_exe(), "1", "30", "1.6e10", "0"], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    photons = sum(int(m) for m in re.findall(r"photons: (\d+)", r.stderr))
    rows = []
    for l in r.stdout.splitlines():
        if l.startswith("HIT "):
            f = l.split()
            dom, pmt = f[2].split("_")
            rows.append((int(f[1]), int(dom), int(pmt), float(f[3]), float(f[4])))
    arr = np.array(rows)
    gold = np.load(os.path.join(HERE, "golden", "cli_%s.npz" % cfg))
    Nr, nr = int(gold["photons"]), len(gold["t"])
    z = (len(arr) / photons - nr / Nr) / math.sqrt(nr / Nr ** 2 + len(arr) / photons ** 2)
    p_rate = 2 * stats.norm.sf(abs(z))
    key = lambda s, d, p: (np.asarray(s, np.int64) * 100 + np.asarray(d, np.int64)) * 100 + np.asarray(p, np.int64)
    kg, kp = key(gold["str"], gold["dom"], gold["pmt"]), key(arr[:, 0], arr[:, 1], arr[:, 2])
    allk = np.union1d(kg, kp)
    chi2, dof, p_bin = _two_sample_chi2(np.array([np.sum(kp == k) for k in allk]),
                                        np.array([np.sum(kg == k) for k in allk]))
    p_t = stats.ks_2samp(arr[:, 3], gold["t"]).pvalue
    p_w = stats.ks_2samp(arr[:, 4], gold["wv"]).pvalue
    multi = arr[:, 0] <= 2
    print("HIT rate %.4e vs %.4e (p=%.3f), (DOM, PMT) chi2 %.1f/%d (p=%.3f), KS time p=%.3f, wavelength p=%.3f; "
          "%d hits on multi-PMT modules, PMTs %s" % (len(arr) / photons, nr / Nr, p_rate, chi2, dof, p_bin, p_t, p_w,
                                                    multi.sum(), np.unique(arr[multi, 2]).tolist()))
    assert multi.sum() > 1000
    if cfg == "c1ng":
        assert len(np.unique(arr[arr[:, 0] == 1, 2])) > 6
    else:
        assert np.unique(arr[multi, 2]).tolist() == [0, 1]
        near = lambda s, d: (np.asarray(s) == 1) & np.isin(np.asarray(d), (29, 31))
        tn, gn = arr[near(arr[:, 0], arr[:, 1]), 3], gold["t"][near(gold["str"], gold["dom"])]
        p_near = stats.ks_2samp(tn, gn).pvalue
        print("tubes next to the flasher: %d vs %d hits, median time %.2f vs %.2f ns, KS p=%.3f"
              % (len(tn), len(gn), np.median(tn), np.median(gn), p_near))
        assert p_near > 0.01
    assert p_rate > 0.01 and p_bin > 0.01 and p_t > 0.01 and p_w > 0.01


def _event_counts(ev, strs, doms, nev, gsize, index):
    """hits per event and DOM (events x DOMs) from the HIT lines' events and (string, DOM)"""
    C = np.zeros((nev, gsize), dtype=np.int64)
    np.add.at(C, (np.asarray(ev, np.int64), index(strs, doms)), 1)
    return C


def _permutation_chi2(E1, E2, nperm=2000, seed=1, min_exp=10):
    """Two-sample chi-square of the per-DOM counts of two sets of events (bins merged as in
    _two_sample_chi2), and its p-value under the permutations of whole events between the
    two sets: a photon can hit several DOMs, so DOM counts are not independent Poisson
    variables and the chi-square distribution understates the statistic's spread; events are
    independent and exchangeable under the null, so their permutations give its law"""
    a, b = E1.sum(axis=0).astype(float), E2.sum(axis=0).astype(float)
    A, B = a.sum(), b.sum()
    groups, cur, ca, cb = [], [], 0.0, 0.0
    for i in np.argsort(-(a + b)):
        cur.append(i)
        ca += a[i]
        cb += b[i]
        tot = ca + cb
        if tot * A / (A + B) >= min_exp and tot * B / (A + B) >= min_exp:
            groups.append(cur)
            cur, ca, cb = [], 0.0, 0.0
    if cur and groups:
        groups[-1] += cur
    E = np.concatenate([E1, E2])
    M = np.column_stack([E[:, q].sum(axis=1) for q in groups]).astype(float)
    total = M.sum(axis=0)

    def stat(sel):
        xa = M[sel].sum(axis=0)
        xb = total - xa
        k1, k2 = math.sqrt(xb.sum() / xa.sum()), math.sqrt(xa.sum() / xb.sum())
        return float((((k1 * xa - k2 * xb) ** 2) / (xa + xb)).sum())

    sel0 = np.zeros(len(E), dtype=bool)
    sel0[:len(E1)] = True
    obs = stat(sel0)
    rng = np.random.default_rng(seed)
    null = np.array([stat(rng.permutation(sel0)) for _ in range(nperm)])
    return obs, len(groups) - 1, float((1 + (null >= obs).sum()) / (nperm + 1))


def _run_f2k(lines, extra=None):
    from ppc_amd import _native, configs

    env = {k: v for k, v in os.environ.items() if k not in G.ENV_KEYS and k != "PPC_ACCEPT"}
    env.update({"PPCTABLESDIR": configs.ensure("c2"), "DTMN": "1"})
    env.update(extra or {})
    r = subprocess.run([_native.ppc_exe(), "0"], env=env, input="".join(lines), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    photons = sum(int(m) for m in re.findall(r"photons: (\d+)", r.stderr))
    out = r.stdout.splitlines()
    # every input line is echoed, and the HIT lines follow the events
    assert sum(l.startswith("TR ") for l in out) == len(lines)
    rows, ev, e = [], [], -1
    for l in out:
        if l.startswith("TR "):
            e += 1
        elif l.startswith("HIT "):
            rows.append([float(v) for v in l.split()[1:5]])
            ev.append(e)
    return photons, np.array(rows), np.array(ev)


def test_cli_f2k_echoes_every_line():
    """`ppc 0 < events.f2k`: every input line, TR or not (headers, event brackets, comments),
    is echoed in input order, with the HIT lines of a line's photons after it (gpu/f2k.cxx:
    555-583, 260-267: one segment per line)"""
    from ppc_amd import _native, configs

    tr = G.cli_f2k_lines(events=1, E=100.0)[0]
    inp = ["V 2000.1", "HI some header", "EM 1 1 1970 0 0 0", tr.strip(), "EE", "EM 2 1 1970 0 0 0", tr.strip(),
           "# a comment", "EE", "END"]
    env = {k: v for k, v in os.environ.items() if k not in G.ENV_KEYS}
    env.update({"PPCTABLESDIR": configs.ensure("c2"), "DTMN": "1"})
    r = subprocess.run([_native.ppc_exe(), "0"], env=env, input="\n".join(inp) + "\n", capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = r.stdout.splitlines()
    echo = [l for l in out if not l.startswith("HIT ")]
    assert echo == inp, echo
    # the hits follow the TR lines that made them: between a TR line and its event's EE
    pos = [k for k, l in enumerate(out) if not l.startswith("HIT ")]
    hit_after = [sum(1 for l in out[a + 1:b] if l.startswith("HIT ")) for a, b in zip(pos, pos[1:] + [len(out)])]
    print("hits after each line: %s" % hit_after, flush=True)
    assert hit_after[3] > 0 and hit_after[6] > 0
    assert sum(hit_after) == hit_after[3] + hit_after[6]


@pytest.mark.parametrize("accept", ["host", "device"])
def test_cli_f2k_cascade_stream_matches_reference_cli(accept):
    """`ppc 0 < events.f2k` on the 86-string ice against the reference command line's HIT lines
    on the same TR lines: 1 TeV EM cascades (tests/golden/cli_f2k_c2.npz, 8 events, 8e7
    photons); 8e8 photons here, in one process; with the acceptance on the host and on the
    device (the full emission spectrum: hits in every wavelength bin, each with its own
    acceptance and wavelength range)"""
    lines = G.cli_f2k_lines(events=80)
    print("f2k stream: %d cascade TR lines on c2, acceptance on the %s" % (len(lines), accept), flush=True)
    photons, arr, _ = _run_f2k(lines, {"PPC_ACCEPT": "device"} if accept == "device" else None)
    assert photons > 5e8
    _compare(photons, arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3], fixture="cli_f2k_c2.npz")


def test_cli_f2k_muon_stream_matches_reference_cli():
    """`ppc 0 < events.f2k` with 1 km down-going 1 TeV muons (amu TR lines) on the 86-string ice
    against the reference command line on the same TR lines (tests/golden/cli_f2k_muon_c2.npz:
    16 processes reading 16 TR lines each, 6.5e8 photons); 320 events, 8e8 photons here, in
    one process.  HIT rate (Poisson), per-DOM counts (two-sample chi-square, p-value from
    permutations of whole events: DOM counts of a muon are correlated, since 6 % of its hits
    come from photons that hit more than once), arrival-time and wavelength distributions (KS)"""
    from ppc_amd import api, configs

    lines = G.cli_f2k_muon_lines(events=320)
    print("f2k stream: %d muon TR lines on c2" % len(lines), flush=True)
    photons, arr, ev = _run_f2k(lines)
    assert photons > 5e8
    gold = np.load(os.path.join(HERE, "golden", "cli_f2k_muon_c2.npz"))
    Nr, nr = int(gold["photons"]), len(gold["ev"])
    z = (len(arr) / photons - nr / Nr) / math.sqrt(nr / Nr ** 2 + len(arr) / photons ** 2)
    p_rate = 2 * stats.norm.sf(abs(z))
    det = api.Detector(configs.ensure("c2"), {"DTMN": "1"})
    key = np.asarray(det.array("str", np.int32), np.int64) * 100 + det.array("dom", np.int32)
    pos = {int(k): i for i, k in enumerate(key)}
    index = lambda s, d: np.array([pos[int(a) * 100 + int(b)] for a, b in zip(s, d)], dtype=np.int64)
    E1 = _event_counts(ev, arr[:, 0], arr[:, 1], len(lines), len(key), index)
    E2 = _event_counts(gold["ev"], gold["str"], gold["dom"], int(gold["events"]), len(key), index)
    chi2, dof, p_dom = _permutation_chi2(E1, E2)
    p_t = stats.ks_2samp(arr[:, 2], gold["t"]).pvalue
    p_w = stats.ks_2samp(arr[:, 3], gold["wv"]).pvalue
    print("HIT rate %.4e vs %.4e (p=%.3f), DOM chi2 %.1f/%d (permutation p=%.3f), KS time p=%.3f, wavelength p=%.3f"
          % (len(arr) / photons, nr / Nr, p_rate, chi2, dof, p_dom, p_t, p_w))
    assert p_rate > 0.01 and p_dom > 0.01 and p_t > 0.01 and p_w > 0.01


def test_cli_f2k_mixed_particles_match_reference_cli():
    """`ppc 0 < events.f2k` with the other particle names of the f2k reader (gpu/f2k.cxx:555-583)
    on the 86-string ice: a hadronic cascade (hadr, yield with its gaussian draw), an electron,
    a positron, a munu hadronic cascade, a bremsstrahlung cascade and a point flash of 4e6
    photons (ph, addf: isotropic), against the reference command line on the same six TR lines
    (tests/golden/cli_f2k_mixed_c2.npz, 8 processes); 16 repeats here, in one process.  The
    hadronic yields vary from event to event (12 % rms for hadr), so the tests treat events as
    the units: per particle, hits per event (Welch t-test), arrival times and wavelengths (KS:
    within a particle the shapes do not depend on its yield, while the particles' shares of the
    pooled hits do, and the wavelength of a hit depends on its distance to the source); per-DOM
    counts (chi-square, p-value from permutations of whole events)"""
    from ppc_amd import api, configs

    six = G.cli_f2k_mixed_lines()
    lines = six * 16
    print("f2k stream: 16 x %d mixed TR lines on c2" % len(six), flush=True)
    photons, arr, ev = _run_f2k(lines)
    gold = np.load(os.path.join(HERE, "golden", "cli_f2k_mixed_c2.npz"))
    gev = gold["ev"].astype(np.int64)
    names = [l.split()[3] for l in six]
    cp = np.bincount(ev, minlength=len(lines)).reshape(16, 6)
    cg = np.bincount(gev, minlength=int(gold["events"])).reshape(-1, 6)
    ps = []
    for k, name in enumerate(names):
        p = stats.ttest_ind(cp[:, k], cg[:, k], equal_var=False).pvalue
        tk = stats.ks_2samp(arr[ev % 6 == k, 2], gold["t"][gev % 6 == k]).pvalue
        wk = stats.ks_2samp(arr[ev % 6 == k, 3], gold["wv"][gev % 6 == k]).pvalue
        ps.append((p, tk, wk))
        print("%-5s hits per event %.1f +- %.1f vs %.1f +- %.1f (p=%.3f), KS time p=%.3f, wavelength p=%.3f"
              % (name, cp[:, k].mean(), cp[:, k].std(ddof=1), cg[:, k].mean(), cg[:, k].std(ddof=1), p, tk, wk),
              flush=True)
    det = api.Detector(configs.ensure("c2"), {"DTMN": "1"})
    key = np.asarray(det.array("str", np.int32), np.int64) * 100 + det.array("dom", np.int32)
    pos = {int(k): i for i, k in enumerate(key)}
    index = lambda s, d: np.array([pos[int(a) * 100 + int(b)] for a, b in zip(s, d)], dtype=np.int64)
    E1 = _event_counts(ev // 6, arr[:, 0], arr[:, 1], 16, len(key), index)
    E2 = _event_counts(gev // 6, gold["str"], gold["dom"], len(cg), len(key), index)
    chi2, dof, p_dom = _permutation_chi2(E1, E2)
    print("photons %d vs %d; DOM chi2 %.1f/%d (permutation p=%.3f)" % (photons, int(gold["photons"]), chi2, dof, p_dom),
          flush=True)
    # six particles, three tests each: each at p > 0.01/18
    assert min(min(p) for p in ps) > 0.01 / 18 and p_dom > 0.01


_LIB = r'''
import sys, json
import numpy as np
from ppc_amd import api, configs
D = api.Detector(sys.argv[1])
D.start(0)
D.lib.xppc_c_flset(1, 30)
D.lib.xppc_c_sett(0.0, 0.0, 1.0, 1, 0, 0, 0.0)
N = 16000000000
D.lib.xppc_c_flone(N)
D.lib.xppc_c_eout()
h = D.hitz()
photons = D.flasher_photons(N)
np.savez(sys.argv[2], photons=photons, **h)
D.stop()
'''


_LIB_RES = r'''
import sys
import numpy as np
from ppc_amd import api
D = api.Detector(sys.argv[1])
D.start(0)
D.lib.xppc_c_set_res(float(sys.argv[3]), float(sys.argv[4]))
D.lib.xppc_c_flset(1, 30)
D.lib.xppc_c_sett(0.0, 0.0, 1.0, 1, 0, 0, 0.0)
D.lib.xppc_c_flone(4000000000)
D.lib.xppc_c_eout()
h, k = D.hitz(), D.hitz_keys()
np.savez(sys.argv[2], **h, **{"key_" + n: v for n, v in k.items()})
D.stop()
'''


def test_library_hitz_time_and_direction_bins(tmp_path):
    """set_res(t, c) (gpu/f2k.cxx:200-214, 326-331): the hitz keys carry the hit time in bins of
    t ns and 1 + cos(photon zenith) in bins of c, each at its bin's centre, and photons that
    share a key share its entry; the photon records themselves (unbinned time, wavelength) are
    those of the same deterministic run without bins, photon for photon"""
    from ppc_amd import configs

    runs = {}
    for name, trz, crz in (("plain", 0.0, 0.0), ("binned", 5.0, 0.1)):
        out = str(tmp_path / (name + ".npz"))
        e = _env()
        e["PYTHONPATH"] = ROOT
        print("library flasher, set_res(%g, %g)" % (trz, crz), flush=True)
        r = subprocess.run([sys.executable, "-c", _LIB_RES, configs.ensure("c1"), out, str(trz), str(crz)], env=e,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        runs[name] = dict(np.load(out))
    a, b = runs["plain"], runs["binned"]
    rec = lambda x: np.sort(np.rec.fromarrays([x["str"], x["dom"], x["t"], x["wv"]]))
    assert len(a["t"]) > 50000 and np.array_equal(rec(a), rec(b))
    for x in (a, b):
        assert int(x["key_count"].sum()) == len(x["t"])
    kt, kd = b["key_time"].astype(np.float64), b["key_dir"].astype(np.float64)
    assert np.all(np.abs(kt / 5.0 - 0.5 - np.round(kt / 5.0 - 0.5)) < 1e-6)
    assert np.all(np.abs((kd + 1) / 0.1 - 0.5 - np.round((kd + 1) / 0.1 - 0.5)) < 1e-3)
    # each photon record lies in the time bin of its key
    t_of_key = np.repeat(kt, b["key_count"])
    q = b["t"].astype(np.float64) / 5.0
    edge = np.abs(q - np.round(q)) < 1e-5  # float rounding may put a time on a bin edge either side
    assert np.all((np.floor(q) == np.floor(t_of_key / 5.0)) | edge)
    print("photon records %d; keys %d without bins, %d with 5 ns x 0.1 bins" % (len(a["t"]), len(a["key_time"]),
                                                                            len(kt)), flush=True)
    assert len(kt) < 0.9 * len(a["key_time"])


_LIB_EVENTS = r'''
import sys
import numpy as np
from ppc_amd import api
D = api.Detector(sys.argv[1])
oms = D.array("oms", np.float32).reshape(-1, 5)[:, 2:5]
key = D.array("str", np.int32) * 100 + D.array("dom", np.int32)
pos = {s: oms[int(np.flatnonzero(key == s * 100 + 30)[0])] + np.float32(5.0) for s in (36, 80)}
D.start(0)
# three segments, as llh's sett() calls give them: two 10 TeV cascades next to strings 36 and 80,
# and a segment without light
D.lib.xppc_c_sett(0.0, 0.0, -1.0, 1, 10, 0, 0.0)
D.lib.xppc_c_addp(*[float(v) for v in pos[36]], 0.0, 1e4, 0.0, 11, 1.0)
D.lib.xppc_c_sett(0.0, 0.0, -1.0, 2, 20, 1, 5.0)
D.lib.xppc_c_addp(*[float(v) for v in pos[80]], 0.0, 1e4, 0.0, 11, 1.0)
D.lib.xppc_c_sett(0.0, 0.0, -1.0, 3, 30, 2, 0.0)
D.lib.xppc_c_eout()
h, k, tr = D.hitz(), D.hitz_keys(), D.hitz_tracks()
np.savez(sys.argv[2], **h, **{"key_" + n: v for n, v in k.items()}, **{"tr_" + n: v for n, v in tr.items()},
         p36=pos[36], p80=pos[80], oms=oms, omkey=key)
D.stop()
'''


def test_library_segments_keep_their_hits(tmp_path):
    """Library mode with three segments (sett with distinct ids, frames and t0, gpu/f2k.cxx:548-552):
    a cascade next to string 36, one next to string 80, and a segment without light.  Each hitz
    key carries the segment its photons came from, keys are ordered by segment (frame first,
    mcid's order), no key belongs to the empty segment, and each segment's hits lie around
    its own cascade"""
    from ppc_amd import configs

    out = str(tmp_path / "events.npz")
    e = _env()
    e.pop("WFLA", None), e.pop("FWID", None)
    e["PYTHONPATH"] = ROOT
    r = subprocess.run([sys.executable, "-c", _LIB_EVENTS, configs.ensure("c2"), out], env=e, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    x = np.load(out)
    segs = list(zip(x["tr_frame"].tolist(), x["tr_id"].tolist(), x["tr_sub"].tolist(), x["tr_t0"].tolist()))
    assert set(segs) == {(0, 1, 10, 0.0), (1, 2, 20, 5.0)}
    order = [0 if s[0] == 0 else 1 for s in segs]
    assert order == sorted(order)  # frame 0's keys first
    seg = np.repeat(np.array(order), x["key_count"])
    idx = {int(k): i for i, k in enumerate(x["omkey"])}
    where = x["oms"][[idx[int(s) * 100 + int(d)] for s, d in zip(x["str"], x["dom"])]]
    d36 = np.linalg.norm(where - x["p36"], axis=1)
    d80 = np.linalg.norm(where - x["p80"], axis=1)
    n0, n1 = int((seg == 0).sum()), int((seg == 1).sum())
    print("hits: segment 1 %d, segment 2 %d; median distance to own / other cascade %.0f / %.0f m, %.0f / %.0f m"
          % (n0, n1, np.median(d36[seg == 0]), np.median(d80[seg == 0]), np.median(d80[seg == 1]),
             np.median(d36[seg == 1])), flush=True)
    assert n0 > 1000 and n1 > 1000
    assert np.mean(d36[seg == 0] < d80[seg == 0]) > 0.99 and np.mean(d80[seg == 1] < d36[seg == 1]) > 0.99


def test_library_flasher_hitz(tmp_path):
    from ppc_amd import configs

    out = str(tmp_path / "hitz.npz")
    e = _env()
    e["PYTHONPATH"] = ROOT
    r = subprocess.run([sys.executable, "-c", _LIB, configs.ensure("c1"), out], env=e, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    h = np.load(out)
    assert (h["pmt"] == 0).all()
    _compare(int(h["photons"]), h["str"], h["dom"], h["t"], h["wv"])
