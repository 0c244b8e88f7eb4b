"""Writes the golden fixtures of tests/golden/ from the reference's own CPU build.

Runs only where the reference tree and oracle/_ref are present (this container):
    python tests/golden/make_golden.py

Fixtures (data only: inputs and the reference's outputs on them):
  mwc_rnd.json            first/last entries, count and SHA-256 of the reference's rnd.txt
  table_<cfg>.txt         output of `ppc -` (ice table at the middle wavelength bin)
  zshift_<cfg>.txt        output of `ppc =` on a grid of points (tilt)
  stats_<case>.npz        raw hits of the reference kernel (gpu/pro.cu in its CPU form)
                          on many photons: photons, per-DOM hit counts, hit times and DOMs
  replay_<case>.npz       single photons from random MWC states: state before emission,
                          state after transport, and every raw hit of each photon
  cli_c1.npz              HIT lines of the reference command line `ppc 1 30 N 0` (C1 flasher)
  cli_c1ng.npz            the same in the 7-string ice with nextgen multi-PMT modules (HIT str dom_pmt)
  cli_f2k_c2.npz          HIT lines of `ppc k < TR line`, one 1 TeV EM cascade per process (86 strings)
  cli_f2k_muon_c2.npz     HIT lines of `ppc k < f2k` for 1 km down-going 1 TeV muons (amu TR lines):
                          16 processes reading 16 TR lines each, with each HIT line's event
The synthetic ice directories are regenerated from their seeds by ppc_amd.configs.
"""
import ctypes
import hashlib
import json
import pickle
import tempfile
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_util as U  # noqa: E402
from ppc_amd import configs  # noqa: E402

REF_RND = "/root/reference/ice/spice_bfr-v2/rnd.txt"

# name -> (ice dir, environment, source)
CASES = {
    "c1": ("c1", {"WFLA": "405", "FWID": "10"}, ("flasher", 1, 30)),
    "c2": ("c2", {}, ("cascade", 8.0, 5.0, -51.93, -0.6, -0.3)),
    "c2x": ("c2x", {}, ("cascade", 5.0, -6.0, 120.0, -0.5, 0.6)),
    "c3": ("c2", {}, ("muon",)),
    # the cascade of c2 on the 86-string ice with nine 1-d tilt reference strings
    "c2t9": ("c2t9", {}, ("cascade", 8.0, 5.0, -51.93, -0.6, -0.3)),
    # flasher beams set by the environment (the bunch flset() builds for each)
    "c1iso": ("c1", {"WFLA": "405", "FWID": "-1", "FLDR": "60"}, ("flasher_env", 1, 30)),
    "c1lab": ("c1", {"WFLA": "405", "FWID": "1090", "FLDR": "30", "FTLT": "0.1 0.2 0.97"}, ("flasher_env", 1, 30)),
    "c1cone": ("c1", {"WFLA": "405", "FWID": "-2", "ELEV": "20", "FZCR": "3", "OFLA": "0"}, ("flasher_env", 1, 30)),
    # the c2 cascade with the tilt's layer-thickness correction
    "c2vthk": ("c2", {"VTHK": "1"}, ("cascade", 8.0, 5.0, -51.93, -0.6, -0.3)),
    # a cascade in a detector near the reference's capacity (200 strings, 14280 DOMs)
    "cbig": ("cbig", {}, ("cascade", 8.0, 5.0, -51.93, -0.6, -0.3)),
    # the full-feature ice with hole ice only for photons later than 300 ns and DOMs 3x oversized,
    # and with hole ice only for photons earlier than 300 ns (a negative HIFL, gpu/pro.cu:655)
    "c2xho": ("c2x", {"HIFL": "300", "OVSZ": "3"}, ("cascade", 5.0, -6.0, 120.0, -0.5, 0.6)),
    "c2xhe": ("c2x", {"HIFL": "-300"}, ("cascade", 5.0, -6.0, 120.0, -0.5, 0.6)),
    # one whole 1 km down-going 1 TeV muon (the amu TR line of the command line's f2k stream test)
    "c3km": ("c2", {}, ("muon_km",)),
    # the IceTray entry points (gpu/f2k.cxx:507-528): a 5 m segment of a particle at beta = 0.77,
    # near the Cherenkov threshold (addp_clst); a 5 m segment with delayed isotropic
    # luminescence, tau = 40 ns (addp_mopo); a point with beta = 0.85 and 30 % of its light in
    # cascade directions (addp_mopo)
    "c3clst": ("c2", {}, ("clst", dict(dr=5.0, beta=0.77))),
    "c3mopo": ("c2", {}, ("mopo", dict(length=5.0, beta=1.0, tau=40.0, fr=0.5))),
    "c3mopob": ("c2", {}, ("mopo", dict(length=0.0, beta=0.85, tau=0.0, fr=0.7))),
    # the lab-measured beam of the tilted-LED flasher type (a negative string number: type 2)
    # and of the second standard candle (string 0, DOM 2: type 4), each with its own beam
    # constants (gpu/pro.cu:362-372) and LED elevation (gpu/f2k.cxx:640-670)
    "c1t2lab": ("c1", {"WFLA": "405", "FWID": "1090", "FLDR": "30"}, ("flasher_env", -1, 30)),
    "c1sclab": ("c1", {"WFLA": "405", "FWID": "1090"}, ("flasher_env", 0, 2)),
}
# single-bunch source for the replay of the muon case: a 1 m segment of the 1 TeV muon
C3_SEGMENT = dict(x=6.0, y=-4.0, z=-30.0, E=1000.0, dr=1.0, n=(0.6, -0.48, -0.64))


def c3km_track():
    """the 1 km amu track of cli_f2k_muon_lines(): start L/2 above the benchmark cascade's
    position, zenith 0 (down-going), 1 TeV"""
    cas = configs.c2_cascade()
    return dict(x=cas["x"], y=cas["y"], z=cas["z"] + 500.0, E=1000.0, dr=1000.0, n=configs.f2k_direction(0.0, 0.0))


ENV_KEYS = ["WFLA", "FWID", "DTMN", "FLDR", "FTLT", "ELEV", "FZCR", "OFLA", "VTHK", "HIFL", "OVSZ"]


def _env(env):
    e = dict(os.environ)
    for k in ENV_KEYS:
        e.pop(k, None)
    e.update(env)
    e["DTMN"] = "1"
    return e


def _bunch(T, src, num, det=None):
    """the single bunch of a case; for the muon case the segment bunch is made by
    the product's front-end (det) or, here, by the reference's own addp()"""
    if src[0] == "flasher":
        return U.flasher_bunch(T, src[1], src[2], num)
    if src[0] == "flasher_env":  # flset() of the product's front-end (det) or of the reference
        if det is not None:
            return U.to_orc_bunch(det.flasher_bunch(src[1], src[2], num=num))
        b = U.ref_flasher(src[1], src[2])
        b.num = num
        return b
    if src[0] in ("muon", "muon_km"):
        s = C3_SEGMENT if src[0] == "muon" else c3km_track()
        n = np.array(s["n"]) / np.linalg.norm(s["n"])
        if det is not None:
            b, _ = det.particle_bunch(s["x"], s["y"], s["z"], 0.0, s["E"], s["dr"], 0, n, num=num)
            return U.to_orc_bunch(b)
        b, _ = U.ref_particle(s["x"], s["y"], s["z"], 0.0, s["E"], s["dr"], 0, n)
        b.num = num
        return b
    if src[0] in ("clst", "mopo"):  # at the 1 m muon segment's position and direction
        s, a = C3_SEGMENT, src[1]
        n = np.array(s["n"]) / np.linalg.norm(s["n"])
        if src[0] == "clst":
            if det is not None:
                return U.to_orc_bunch(det.clst_bunch(s["x"], s["y"], s["z"], 0.0, n, a["dr"], a["beta"], num=num))
            b = U.ref_clst(s["x"], s["y"], s["z"], 0.0, n, a["dr"], a["beta"])
        else:
            args = (s["x"], s["y"], s["z"], 0.0, n, 1.0, a["length"], a["beta"], a["tau"], a["fr"])
            if det is not None:
                return U.to_orc_bunch(det.mopo_bunch(*args, num=num)[0])
            b = U.ref_mopo(*args)[0]
        b.num = num
        return b
    x, y, z, nx, ny = src[1:]
    nz = -np.sqrt(1 - nx * nx - ny * ny)
    return U.cascade_bunch(x, y, z, nx, ny, nz, num)


def _fanout(kind, args_list, timeout=1800):
    """run _<kind>_worker(args) in separate processes (the reference holds global state)"""
    tmp = tempfile.mkdtemp(prefix="golden_")
    procs = []
    for k, a in enumerate(args_list):
        inp, out = os.path.join(tmp, "in%d.pkl" % k), os.path.join(tmp, "out%d.pkl" % k)
        pickle.dump(a, open(inp, "wb"))
        procs.append((subprocess.Popen([sys.executable, __file__, "_worker", kind, inp, out],
                                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL), out))
    res = []
    for p, out in procs:
        if p.wait(timeout=timeout) != 0:
            raise RuntimeError("worker failed")
        res.append(pickle.load(open(out, "rb")))
    return res


def muon_event_bunches(make, procs=1, part=0):
    """bunches of the muon event (records part, part+procs, ...): make(x,y,z,t,E,dr,type,n) -> (bunch, mean);
    each bunch gets the rounded mean photon count and its record index as segment"""
    out = []
    for k, r in enumerate(configs.c3_muon_event()):
        if k % procs != part:
            continue
        n = configs.f2k_direction(r["th"], r["ph"])
        b, mean = make(r["x"], r["y"], r["z"], r["t"], r["E"], r["l"], configs.F2K_TYPE[r["name"]], n)
        b.num = int(round(mean))
        b.q = k
        out.append(b)
    return out


def _muon_stats(k, procs, T):
    ref = U.reference()
    pmax = ref.ref_pmax()
    bunches = muon_event_bunches(U.ref_particle, procs, k)
    cap = 400000
    H = (U.Hit * cap)()
    a = ref.ref_multiplier(k)
    x, c = 362436069 + 7919 * k, 1234567
    ox, oc = ctypes.c_uint32(), ctypes.c_uint32()
    hits, photons, i = [], 0, 0
    while i < len(bunches):
        grp, tot = [], 0
        while i < len(bunches) and tot + bunches[i].num <= pmax:
            grp.append(bunches[i])
            tot += bunches[i].num
            i += 1
        if not grp:  # a bunch larger than one launch: split it
            b = bunches[i]
            part = U.Bunch.from_buffer_copy(b)
            part.num = pmax
            b.num -= pmax
            grp, tot = [part], pmax
        arr = (U.Bunch * len(grp))(*grp)
        n = ref.ref_run(arr, len(grp), tot, 512, k, a, x, c, H, cap, ctypes.byref(ox), ctypes.byref(oc))
        hits.append(U.hits_array(H, min(n, cap)))
        x, c = ox.value, oc.value
        photons += tot
    return photons, np.concatenate(hits), int(T.get("gsize", np.int32)[0])


def _stats_worker(args):
    case, k, batches = args[:3]
    cfg, env, src = CASES[case]
    for kk in ENV_KEYS:
        os.environ.pop(kk, None)
    T = U.ref_tables(configs.ensure(cfg), dict(env, DTMN="1"))
    if src[0] == "muon":
        return _muon_stats(k, args[3], T)
    ref = U.reference()
    pmax = ref.ref_pmax()
    b = _bunch(T, src, pmax)
    cap = 200000
    H = (U.Hit * cap)()
    a = ref.ref_multiplier(k)
    x, c = 362436069 + 7919 * k, 1234567
    hits = []
    ox, oc = ctypes.c_uint32(), ctypes.c_uint32()
    for _ in range(batches):
        n = ref.ref_run(ctypes.byref(b), 1, pmax, 512, k, a, x, c, H, cap, ctypes.byref(ox), ctypes.byref(oc))
        hits.append(U.hits_array(H, min(n, cap)))
        x, c = ox.value, oc.value
    return batches * pmax, np.concatenate(hits), int(T.get("gsize", np.int32)[0])


# launches of 524288 photons per worker (8 workers) where the default of 8 is too few: the 1 km
# track spreads its light over the whole string, so its per-DOM comparison needs more photons;
# the standard candle sits 55 m outside the strings and few of its photons hit; the
# full-feature ice's rate is compared at 0.1 %
STATS_BATCHES = {"c3km": 32, "c1sclab": 64, "c2x": 32, "c2xhe": 32}
STATS_PROCS = {"c2x": 16, "c2xhe": 16}  # 16 workers: 2.7e8 reference photons
STATS_SAMPLE = 40000  # hits kept for the time and direction distributions


def make_stats(case, procs=8, batches=8):
    res = _fanout("stats", [(case, k, batches, procs) for k in range(procs)])
    photons = sum(r[0] for r in res)
    hits = np.concatenate([r[1] for r in res])
    gsize = res[0][2]
    counts = np.bincount(hits["i"], minlength=gsize).astype(np.int64)
    sub = np.sort(np.random.default_rng(7).permutation(len(hits))[:STATS_SAMPLE])
    np.savez_compressed(os.path.join(HERE, "stats_%s.npz" % case), photons=np.int64(photons), counts=counts,
                        nhits=np.int64(len(hits)), t=hits["t"][sub].astype(np.float32),
                        dom=hits["i"][sub].astype(np.uint16), cth=np.cos(hits["pth"][sub]).astype(np.float32))
    print("stats %s: %d photons, %d hits" % (case, photons, len(hits)))


def _replay_worker(args):
    case, p, n = args
    cfg, env, src = CASES[case]
    for kk in ENV_KEYS:
        os.environ.pop(kk, None)
    T = U.ref_tables(configs.ensure(cfg), dict(env, DTMN="1"))
    ref = U.reference()
    b = _bunch(T, src, 1)
    states = replay_states(p, n)
    cap = 1000
    H = (U.Hit * cap)()
    st0, st1, mult, hp, hh = [], [], [], [], []
    for k in range(n):
        x, c = int(states[k, 0]), int(states[k, 1])
        a = ref.ref_multiplier(p * n + k)  # photon p*n+k runs on lane p*n+k of a deterministic launch
        ox, oc = ctypes.c_uint32(), ctypes.c_uint32()
        nh = ref.ref_run(ctypes.byref(b), 1, 1, 1, 0, a, x, c, H, cap, ctypes.byref(ox), ctypes.byref(oc))
        st0.append((x, c)), st1.append((ox.value, oc.value)), mult.append(a)
        h = U.hits_array(H, min(nh, cap))
        hh.append(h)
        hp.append(np.full(len(h), k, dtype=np.uint32))
    return np.array(st0, dtype=np.uint32), np.array(st1, dtype=np.uint32), np.array(mult, dtype=np.uint32), \
        np.concatenate(hp), np.concatenate(hh)


def replay_states(p, n):
    """MWC states (x, c) of the replayed photons of worker p, regenerated from the seed"""
    rng = np.random.default_rng(1000 + p)
    out = np.zeros((n, 2), dtype=np.uint32)
    for k in range(n):
        out[k] = int(rng.integers(0, 2 ** 32)), int(rng.integers(0, 2 ** 31))
    return out


def fold(st):
    """32-bit digest of final MWC states (x, c)"""
    x, c = st[:, 0].astype(np.uint32), st[:, 1].astype(np.uint32)
    return x ^ ((c << np.uint32(16)) | (c >> np.uint32(16)))


REPLAY_N, REPLAY_PROCS = 4096, 8


def make_replay(case, n=REPLAY_N, procs=REPLAY_PROCS):
    """Single photons from random states; photon k uses the multiplier of lane k.
    Stored: the digest of each photon's final state and all its raw hits."""
    res = _fanout("replay", [(case, p, n) for p in range(procs)])
    st1 = np.concatenate([r[1] for r in res])
    hp = np.concatenate([r[3] + p * n for p, r in enumerate(res)])
    hh = np.concatenate([r[4] for r in res])
    np.savez_compressed(os.path.join(HERE, "replay_%s.npz" % case), n=np.int64(n), procs=np.int64(procs),
                        digest=fold(st1), hit_photon=hp.astype(np.uint32), hits=hh)
    print("replay %s: %d photons, %d with hits, %d hits" % (case, len(st1), len(np.unique(hp)), len(hh)))


GRID_LINES = 25100  # 100 columns of the `ppc _` dump


def make_tables():
    for cfg, env in [("c1", {"WFLA": "405"}), ("c2", {}), ("c2x", {})]:
        out = subprocess.run([U.REF_EXE, "-"], env=_env(dict(env, PPCTABLESDIR=configs.ensure(cfg))),
                             capture_output=True, text=True, check=True).stdout
        open(os.path.join(HERE, "table_%s.txt" % cfg), "w").write(out)
    rng = np.random.default_rng(5)
    pts = np.column_stack([rng.uniform(-700, 700, 400), rng.uniform(-700, 700, 400), rng.uniform(-600, 600, 400)])
    inp = "\n".join("%.3f %.3f %.3f" % tuple(p) for p in pts) + "\n"
    for cfg in ["c2", "c2x"]:
        out = subprocess.run([U.REF_EXE, "="], env=_env(dict(PPCTABLESDIR=configs.ensure(cfg), VTHK="1")), input=inp,
                             capture_output=True, text=True, check=True).stdout
        open(os.path.join(HERE, "zshift_%s.txt" % cfg), "w").write(out)
    # `ppc - x y`: the layer table at a horizontal position, and the first lines of the
    # `ppc _` grid dump (x = -750, y from -750 up, z innermost)
    for cfg in ["c2", "c2x"]:
        out = subprocess.run([U.REF_EXE, "-", "310.5", "-205.25"], env=_env(dict(PPCTABLESDIR=configs.ensure(cfg))),
                             capture_output=True, text=True, check=True).stdout
        open(os.path.join(HERE, "table_%s_xy.txt" % cfg), "w").write(out)
        p = subprocess.Popen([U.REF_EXE, "_"], env=_env(dict(PPCTABLESDIR=configs.ensure(cfg))), stdout=subprocess.PIPE,
                             stderr=subprocess.DEVNULL, text=True)
        lines = [p.stdout.readline() for _ in range(GRID_LINES)]
        p.kill()
        p.wait()
        open(os.path.join(HERE, "grid_%s.txt" % cfg), "w").write("".join(lines))


def make_mwc():
    vals = [int(v) for v in open(REF_RND).read().split()]
    text = "\n".join(str(v) for v in vals) + "\n"
    json.dump({"count": len(vals), "first": vals[:16], "last": vals[-4:],
               "sha256": hashlib.sha256(text.encode()).hexdigest()},
              open(os.path.join(HERE, "mwc_rnd.json"), "w"), indent=1)


def make_cli(n=int(1.6e9), cfg="c1", flasher=(1, 30), procs=8):
    """accepted HIT lines of the reference command line on a flasher: the C1 flasher, or a
    flasher of the 86-string ice with every optional feature and cable positions (c2x), also
    with tilted DOM axes and fixed hole-ice columns (c2xb)"""
    env = _env({"WFLA": "405", "FWID": "10", "PPCTABLESDIR": configs.ensure(cfg)})
    procs = [subprocess.Popen([U.REF_EXE, str(flasher[0]), str(flasher[1]), "%d" % (n // 8), str(k)], env=env,
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for k in range(procs)]
    # one thread per process, so that the processes run side by side and none blocks on a full pipe
    with ThreadPoolExecutor(len(procs)) as ex:
        res = list(ex.map(lambda p: p.communicate(), procs))
    photons = sum(int(l.split()[1]) for _, err in res for l in err.splitlines() if l.startswith("photons:"))
    rows = [l.split() for out, _ in res for l in out.splitlines() if l.startswith("HIT")]
    arr = np.array([[float(v) for v in r[1:]] for r in rows], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "cli_%s.npz" % cfg), photons=np.int64(photons), str=arr[:, 0].astype(np.int16),
                        dom=arr[:, 1].astype(np.int16), t=arr[:, 2].astype(np.float32),
                        wv=arr[:, 3].astype(np.float32))
    print("cli %s: %d photons, %d HIT lines" % (cfg, photons, len(rows)))


def make_cli_nextgen(n=int(8e9), cfg="c1ng"):
    """accepted HIT lines ("HIT str dom_pmt ...") of the reference command line on the C1 flasher
    in the 7-string ice with nextgen sensors on strings 1 and 2: multi-PMT modules
    (configs.write_c1ng) or wavelength-shifting tubes (configs.write_c1wom)"""
    env = _env({"WFLA": "405", "FWID": "10", "PPCTABLESDIR": configs.ensure(cfg)})
    procs = [subprocess.Popen([U.REF_EXE, "1", "30", "%d" % (n // 8), str(k)], env=env, stdout=subprocess.PIPE,
                              stderr=subprocess.PIPE, text=True) for k in range(8)]
    # one thread per process, so that the processes run side by side and none blocks on a full pipe
    with ThreadPoolExecutor(len(procs)) as ex:
        res = list(ex.map(lambda p: p.communicate(), procs))
    photons, rows = 0, []
    for out, err in res:
        photons += sum(int(l.split()[1]) for l in err.splitlines() if l.startswith("photons:"))
        for l in out.splitlines():
            if l.startswith("HIT"):
                f = l.split()
                dom, pmt = f[2].split("_")
                rows.append([float(f[1]), float(dom), float(pmt), float(f[3]), float(f[4])])
    arr = np.array(rows, dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "cli_%s.npz" % cfg), photons=np.int64(photons), str=arr[:, 0].astype(np.int16),
                        dom=arr[:, 1].astype(np.int16), pmt=arr[:, 2].astype(np.int16),
                        t=arr[:, 3].astype(np.float32), wv=arr[:, 4].astype(np.float32))
    print("cli %s: %d photons, %d HIT lines" % (cfg, photons, len(rows)))


def cli_f2k_lines(events=8, E=1000.0):
    """f2k TR lines of EM cascades on the 86-string ice (the benchmark cascade's position and
    direction, energy E GeV): the input of the f2k mode of the command line"""
    cas = configs.c2_cascade()
    return ["TR 1 0 e %.3f %.3f %.3f %.4f %.4f 0 %.6g 0\n" % (cas["x"], cas["y"], cas["z"], cas["th"], cas["ph"], E)
            for _ in range(events)]


def cli_f2k_muon_lines(events=128, E=1000.0, L=1000.0):
    """f2k TR lines of down-going muons of energy E GeV and track length L m, starting L/2
    above the benchmark cascade's position on the 86-string ice"""
    cas = configs.c2_cascade()
    return ["TR 1 0 amu %.3f %.3f %.3f 0 0 %.1f %.6g 0\n" % (cas["x"], cas["y"], cas["z"] + L / 2, L, E)
            for _ in range(events)]


def cli_f2k_mixed_lines():
    """f2k TR lines of the other particle names the command line reads (gpu/f2k.cxx:555-583):
    a hadronic cascade (hadr, munu: yield with its gaussian draw), an electron and a positron
    (e-, e+: their longitudinal profiles), a bremsstrahlung cascade and a point flash of
    photons (ph: addf, isotropic, E photons), around the benchmark cascade's position"""
    cas = configs.c2_cascade()
    x, y, z = cas["x"], cas["y"], cas["z"]
    return ["TR 1 0 hadr %.3f %.3f %.3f 40.0 30.0 0 800 0\n" % (x, y, z),
            "TR 1 0 e- %.3f %.3f %.3f 120.0 200.0 0 500 3\n" % (x + 20, y, z + 10),
            "TR 1 0 e+ %.3f %.3f %.3f 80.0 300.0 0 500 5\n" % (x - 15, y + 10, z - 20),
            "TR 1 0 munu %.3f %.3f %.3f 150.0 90.0 0 300 1\n" % (x, y - 25, z + 30),
            "TR 1 0 brems %.3f %.3f %.3f 60.0 10.0 0 400 2\n" % (x + 5, y + 5, z),
            "TR 1 0 ph %.3f %.3f %.3f 0 0 0 4000000 0\n" % (x + 30, y - 10, z - 40)]


def make_cli_f2k(name="cli_f2k_c2.npz", lines=None):
    """accepted HIT lines of the reference command line in its f2k mode (`ppc dev < f2k`),
    one process per event, on the 86-string ice"""
    env = _env({"PPCTABLESDIR": configs.ensure("c2")})
    lines = lines or cli_f2k_lines()
    procs = [subprocess.Popen([U.REF_EXE, str(k)], env=env, stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                              stderr=subprocess.PIPE, text=True) for k in range(len(lines))]
    # one thread per process, so that the processes run side by side and none blocks on a full pipe
    with ThreadPoolExecutor(len(procs)) as ex:
        res = list(ex.map(lambda pl: pl[0].communicate(pl[1]), zip(procs, lines)))
    outs = [out for out, _ in res]
    photons = sum(int(x.split()[1]) for _, err in res for x in err.splitlines() if x.startswith("photons:"))
    rows = [l.split() for o in outs for l in o.splitlines() if l.startswith("HIT")]
    arr = np.array([[float(v) for v in r[1:]] for r in rows], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, name), photons=np.int64(photons),
                        str=arr[:, 0].astype(np.int16), dom=arr[:, 1].astype(np.int16),
                        t=arr[:, 2].astype(np.float32), wv=arr[:, 3].astype(np.float32))
    print("%s: %d photons, %d HIT lines" % (name, photons, len(rows)))


def make_cli_f2k_events(name, lines, procs, sample=200000):
    """accepted HIT lines of the reference command line in its f2k mode, `procs` processes
    each reading the same stream of TR lines (`ppc k < f2k`, k = 0..procs-1), with the event
    (TR line) of every HIT line, so that tests can compare samples event by event"""
    env = _env({"PPCTABLESDIR": configs.ensure("c2")})
    ps = [subprocess.Popen([U.REF_EXE, str(k)], env=env, stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True) for k in range(procs)]
    with ThreadPoolExecutor(len(ps)) as ex:
        res = list(ex.map(lambda p: p.communicate("".join(lines)), ps))
    photons = sum(int(x.split()[1]) for _, err in res for x in err.splitlines() if x.startswith("photons:"))
    rows, ev, e = [], [], -1
    for out, _ in res:
        for l in out.splitlines():
            if l.startswith("TR"):
                e += 1
            elif l.startswith("HIT"):
                rows.append(l.split()[1:5])
                ev.append(e)
    arr = np.array(rows, dtype=np.float64)
    sub = np.sort(np.random.default_rng(3).permutation(len(arr))[:sample])
    np.savez_compressed(os.path.join(HERE, name), photons=np.int64(photons), events=np.int64(e + 1),
                        ev=np.array(ev, dtype=np.int16), str=arr[:, 0].astype(np.int16), dom=arr[:, 1].astype(np.int16),
                        t=arr[sub, 2].astype(np.float32), wv=arr[sub, 3].astype(np.float32))
    print("%s: %d photons, %d events, %d HIT lines" % (name, photons, e + 1, len(arr)))


if __name__ == "__main__":
    if len(sys.argv) == 5 and sys.argv[1] == "_worker":
        fn = {"stats": _stats_worker, "replay": _replay_worker}[sys.argv[2]]
        pickle.dump(fn(pickle.load(open(sys.argv[3], "rb"))), open(sys.argv[4], "wb"))
        sys.exit(0)
    what = sys.argv[1:] or ["mwc", "tables", "replay", "stats", "cli"]
    if "mwc" in what:
        make_mwc()
    if "tables" in what:
        make_tables()
    if "replay" in what:
        for c in CASES:
            make_replay(c)
    if "stats" in what:
        for c in CASES:
            make_stats(c, procs=STATS_PROCS.get(c, 8), batches=STATS_BATCHES.get(c, 8))
    for c in CASES:
        if "stats_" + c in what:
            make_stats(c, procs=STATS_PROCS.get(c, 8), batches=STATS_BATCHES.get(c, 8))
        if "replay_" + c in what:
            make_replay(c)
    if "cli" in what:
        make_cli()
    if "cli" in what or "cli_nextgen" in what:
        make_cli_nextgen()
    if "cli" in what or "cli_c2x" in what:
        make_cli(n=int(1.6e8), cfg="c2x", flasher=(36, 30))
    if "cli" in what or "cli_c2xb" in what:
        make_cli(n=int(1.6e8), cfg="c2xb", flasher=(36, 30), procs=16)
    if "cli" in what or "cli_wom" in what:
        make_cli_nextgen(cfg="c1wom")
    if "cli" in what or "cli_f2k" in what:
        make_cli_f2k()
    if "cli" in what or "cli_f2k_mixed" in what:
        # 8 processes, each reading the six TR lines (every HIT line kept, with its TR line)
        make_cli_f2k_events("cli_f2k_mixed_c2.npz", cli_f2k_mixed_lines(), 8, sample=10 ** 9)
    if "cli" in what or "cli_f2k_muon" in what:
        # 16 processes, each reading a stream of 16 muon TR lines
        make_cli_f2k_events("cli_f2k_muon_c2.npz", cli_f2k_muon_lines(events=16), 16)
