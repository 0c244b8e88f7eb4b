"""Synthetic ice/detector directories in ppc's own file formats.

Each generator writes a complete ice directory (the files ``ini.cxx:578-1700``
reads: ``cfg.txt``, ``icemodel.par``, ``icemodel.dat``, ``geo-f2k``, ``wv.dat``,
``as.dat``, ``rnd.txt`` and, for the tilted ice, ``tilt.par``/``tilt.dat``) from
a seed, so that the same workload can be given to the reference's CPU build and
to this package.  Formats follow ``doc/index.rst:516-697``.

Workloads
---------
``c1``  7 strings on a 125 m hexagon (centre string 1), 60 DOMs at 17 m spacing,
        100 layers of 10 m (1450-2450 m), per-layer b_e(400) log-uniform in
        [0.02, 0.1] 1/m and a(400) log-uniform in [0.005, 0.03] 1/m, HG g=0.9
        with SAM fraction 0.45, no tilt, no anisotropy.  The 405 nm flasher is
        selected with ``WFLA=405`` and ``FWID=10`` in the environment.
``c2``  86 strings of the radius-5 hex lattice x 60 DOMs, 170 layers of 10 m
        (1100-2800 m), 1-D tilt with 6 reference strings along 225 deg and
        offsets within +-20 m, 8 % absorption anisotropy at a seeded azimuth,
        wavelength spectrum ~ 1/lambda^2 x smooth acceptance over 300-600 nm.
        ``eff`` is scaled so that a 100 TeV EM cascade gives a Poisson mean of
        1e9 propagated photons at xR=5.
"""
import ctypes
import math
import os

import numpy as np

# constants of the reference front-end (ini.cxx:31-36, ini.cxx:553, f2k.cxx:12,68-69)
OMR = 0.16510
DPPM = 2450.08
OMAV = 0.335822
ZOFF = 1948.07
DIR1 = 9.3
SPACING = 125.0
DOM_DEPTH0 = 1448.5
DOM_STEP = 17.0
NDOM = 60
EM_YIELD = 5.321 * 0.910 / 0.9216  # photons per GeV of EM cascade before efficiency

_HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(_HERE, "ice", "_gen")


def mwc_multipliers(n, below=0xFFFFFFFF):
    """The first n safe-prime MWC multipliers below ``below`` (descending)."""
    from . import _native

    lib = _native.xppc_lib()
    out = (ctypes.c_uint32 * n)()
    k = lib.ppc_mwc_multipliers(ctypes.c_uint32(below), ctypes.c_uint32(n), out)
    if k != n:
        raise RuntimeError("multiplier search returned %d of %d" % (k, n))
    return np.frombuffer(out, dtype=np.uint32).copy()


def _hex_sites(radius):
    """Hex lattice sites (i, j) within hex distance ``radius``; lattice vectors at
    9.3 and 69.3 deg so that the rows align with the 9.3/129.3 deg cell axes
    (ini.cxx:6-7)."""
    a1 = np.array([math.cos(math.radians(DIR1)), math.sin(math.radians(DIR1))])
    a2 = np.array([math.cos(math.radians(DIR1 + 60)), math.sin(math.radians(DIR1 + 60))])
    sites = []
    for i in range(-radius, radius + 1):
        for j in range(-radius, radius + 1):
            k = -i - j
            if max(abs(i), abs(j), abs(k)) <= radius:
                p = SPACING * (i * a1 + j * a2)
                sites.append((float(p[0]), float(p[1])))
    return sites


def _string_positions(nstr):
    if nstr == 7:
        sites = _hex_sites(1)
    elif nstr <= 91:
        sites = _hex_sites(5)
    else:
        sites = _hex_sites(8)
    r = [math.hypot(x, y) for x, y in sites]
    ang = [math.degrees(math.atan2(y, x)) % 360.0 for x, y in sites]
    order = sorted(range(len(sites)), key=lambda k: (round(r[k], 3), round(ang[k], 3)))
    return [sites[k] for k in order[:nstr]]


def _write_geo(path, strings, ndom=NDOM, step=DOM_STEP):
    with open(path, "w") as f:
        for s, (x, y) in enumerate(strings, start=1):
            for d in range(1, ndom + 1):
                depth = DOM_DEPTH0 + step * (d - 1)
                omid = (s << 16) | d
                f.write("S%03dD%02d\t0x%012x\t%.2f\t%.2f\t%.2f\t%d\t%d\n" % (s, d, omid, x, y, -depth, s, d))


def _write_ice(path, depths, be, ba):
    with open(os.path.join(path, "icemodel.dat"), "w") as f:
        for dp, b, a in zip(depths, be, ba):
            f.write("%.2f %.8g %.8g 0\n" % (dp, b, a))
    # alpha, kappa as in SPICE; A = 0 so that abs(400) = a(400) exactly (ini.cxx:1688)
    with open(os.path.join(path, "icemodel.par"), "w") as f:
        f.write("0.898608505726 0.027638472617\n")
        f.write("1.084106802940 0.014470303431\n")
        f.write("0.0 0.0\n")
        f.write("6954.090332031250 973.426452636719\n")


def _write_cfg(path, eff, azimuth=0.0, lk=(0.0, 0.0, 0.0), xr=5, sf=0.45, g=0.9):
    lines = [
        "# ppc configuration file: follow strict order below",
        "%d     # over-R: DOM radius oversize scaling factor" % xr,
        "%.9g  # overall DOM efficiency correction" % eff,
        "%.6g  # 0=HG; 1=SAM" % sf,
        "%.6g  # g=<cos(theta)>" % g,
        "",
        "%.6f # azimuth of major anisotropy axis (deg)" % azimuth,
        "0.0   # magnitude of major anisotropy coefficient k1",
        "0.0   # magnitude of minor anisotropy coefficient k2",
        "",
        "0.0   # hole ice radius in units of [DOM radius]",
        "0.03  # hole ice effective scattering length [m]",
        "100   # hole ice absorption length [m]",
        "%.6g  # hole ice 0=HG; 1=SAM" % sf,
        "%.6g  # hole ice g=<cos(theta)>" % g,
        "",
        "%.9f # new absorption anisotropy log k1" % lk[0],
        "%.9f # new absorption anisotropy log k2" % lk[1],
        "%.9f # new absorption anisotropy log kz" % lk[2],
        "",
        "0.0   # scaling for old absorption anisotropy",
    ]
    with open(os.path.join(path, "cfg.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


def _write_wv(path):
    """CDF of 1/lambda^2 x a smooth acceptance bump over 300-600 nm."""
    wl = np.arange(300.0, 601.0, 10.0)
    fine = np.linspace(300.0, 600.0, 3001)
    acc = np.exp(-0.5 * ((fine - 410.0) / 70.0) ** 2) + 0.05
    dens = acc / fine ** 2
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(fine))])
    cdf /= cdf[-1]
    vals = np.interp(wl, fine, cdf)
    vals[0], vals[-1] = 0.0, 1.0
    with open(os.path.join(path, "wv.dat"), "w") as f:
        for c, w in zip(vals, wl):
            f.write("%.8f %.1f\n" % (c, w))


def _write_common(path, nrnd):
    with open(os.path.join(path, "as.dat"), "w") as f:
        f.write("1\n1\n")
    rnd = mwc_multipliers(nrnd)
    with open(os.path.join(path, "rnd.txt"), "w") as f:
        f.write("\n".join(str(int(a)) for a in rnd) + "\n")
    _write_wv(path)


def write_c1(path, seed=1, nrnd=131072):
    os.makedirs(path, exist_ok=True)
    rng = np.random.default_rng(seed)
    depths = 1455.0 + 10.0 * np.arange(100)
    be = np.exp(rng.uniform(math.log(0.02), math.log(0.1), 100))
    ba = np.exp(rng.uniform(math.log(0.005), math.log(0.03), 100))
    _write_geo(os.path.join(path, "geo-f2k"), _string_positions(7))
    _write_ice(path, depths, be, ba)
    _write_cfg(path, eff=1.0)
    _write_common(path, nrnd)
    return path


def c2_tilt_strings(strings, n=6):
    """n reference strings (six by default) along the 225 deg tilt gradient (tilt.par)."""
    lnx, lny = math.cos(math.radians(225.0)), math.sin(math.radians(225.0))
    proj = [lnx * x + lny * y for x, y in strings]
    perp = [abs(-lny * x + lnx * y) for x, y in strings]
    cand = sorted(range(len(strings)), key=lambda k: (perp[k] > 70.0, proj[k]))
    near = [k for k in cand if perp[k] <= 70.0]
    near.sort(key=lambda k: proj[k])
    pick_idx = np.linspace(0, len(near) - 1, n).round().astype(int)
    pick = [near[i] for i in pick_idx]
    zero = min(pick, key=lambda k: abs(proj[k]))
    return [(k + 1, proj[k] - proj[zero]) for k in pick], zero + 1


def write_c2(path, seed=2, nrnd=131072, ntilt=6):
    os.makedirs(path, exist_ok=True)
    rng = np.random.default_rng(seed)
    strings = _string_positions(86)
    depths = 1105.0 + 10.0 * np.arange(170)
    be = np.exp(rng.uniform(math.log(0.02), math.log(0.1), 170))
    ba = np.exp(rng.uniform(math.log(0.005), math.log(0.03), 170))
    _write_geo(os.path.join(path, "geo-f2k"), strings)
    _write_ice(path, depths, be, ba)
    azimuth = float(rng.uniform(0.0, 360.0))
    lk = math.log(1.08)
    # Mean photons of a 1e5 GeV EM cascade: eff * ppm(~DPPM) * EM_YIELD * E / xR^2 (f2k.cxx:68-71,
    # f2k.cxx:502, ini.cxx:1488); eff is set so that the mean is 1e9.
    eff = 1e9 * 25.0 / (EM_YIELD * 1e5 * DPPM)
    _write_cfg(path, eff=eff, azimuth=azimuth, lk=(2.0 * lk / 3.0, -lk / 3.0, -lk / 3.0))
    tilt, _ = c2_tilt_strings(strings, ntilt)
    with open(os.path.join(path, "tilt.par"), "w") as f:
        for s, d in tilt:
            f.write("%d %.3f\n" % (s, d))
    rows = 1100.0 + 10.0 * np.arange(171)
    phase = rng.uniform(0, 2 * math.pi, ntilt)
    amp = rng.uniform(8.0, 20.0, ntilt)
    with open(os.path.join(path, "tilt.dat"), "w") as f:
        for dp in rows:
            offs = [amp[k] * math.sin(2 * math.pi * (dp - 1100.0) / 850.0 + phase[k]) *
                    (0.6 + 0.4 * k / (ntilt - 1)) for k in range(ntilt)]
            f.write("%.2f " % dp + " ".join("%.4f" % o for o in offs) + "\n")
    _write_common(path, nrnd)
    return path


def f2k_direction(th, ph):
    """propagation direction of an f2k track given its zenith/azimuth in degrees
    (the source direction, gpu/f2k.cxx:567-569)"""
    fcv = math.pi / 180
    t = fcv * (180 - th)
    p = fcv * (ph + 180 if ph < 180 else ph - 180)
    return (math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t))


def c2_cascade(seed=2):
    """The 100 TeV EM cascade of the 86-string workload: f2k TR line fields
    (x, y, z in the detector frame, zenith/azimuth in degrees, E in GeV)."""
    rng = np.random.default_rng(seed + 1000)
    depth = 2000.0
    x, y = rng.uniform(-60.0, 60.0, 2)
    cz = rng.uniform(-1.0, 1.0)
    th = math.degrees(math.acos(cz))
    ph = rng.uniform(0.0, 360.0)
    return dict(x=float(x), y=float(y), z=float(ZOFF - depth), th=float(th), ph=float(ph), E=1e5, t=0.0)


# birefringence parameters of a published ice model (cfg.txt format, doc/index.rst:548-561)
BFR12 = [0.076795, 544284.545166, 2.229494, 0.002624, 0.077381, 1547618.883526, 2.449589, 0.002505,
         0.000995, 0.248264, 2.354436, 1.680717]


def write_c2x(path, seed=3, nrnd=131072):
    """The 86-string ice with every optional feature of the kernel switched on:
    2-D tilt map (tilt.set/tilt.map), birefringence, hole ice, per-layer
    scattering anisotropy and birefringence strength, old absorption
    anisotropy, air bubbles and the dust/ice absorption split (A, B of
    icemodel.par); plus RDE, high-voltage, cable-position and high-QE
    acceptance files for the hit post-processing."""
    os.makedirs(path, exist_ok=True)
    rng = np.random.default_rng(seed)
    strings = _string_positions(86)
    _write_geo(os.path.join(path, "geo-f2k"), strings)
    n = 170
    depths = 1105.0 + 10.0 * np.arange(n)
    be = np.exp(rng.uniform(math.log(0.02), math.log(0.1), n))
    ba = np.exp(rng.uniform(math.log(0.005), math.log(0.03), n))
    td = rng.uniform(-10.0, 10.0, n)
    k1 = rng.uniform(0.0, 0.08, n)
    k2 = -k1 / 2
    bfra = rng.uniform(0.8, 1.2, n)
    bfrb = rng.uniform(0.8, 1.2, n)
    with open(os.path.join(path, "icemodel.dat"), "w") as f:
        for row in zip(depths, be, ba, td, k1, k2, bfra, bfrb):
            f.write("%.2f %.8g %.8g %.5g %.5g %.5g %.5g %.5g\n" % row)
    with open(os.path.join(path, "icemodel.par"), "w") as f:
        f.write("0.898608505726 0.027638472617\n1.084106802940 0.014470303431\n")
        f.write("6954.090332031250 973.426452636719\n6617.754394531250 71.282264709473\n")
    with open(os.path.join(path, "icemodel.bbl"), "w") as f:
        f.write("11.7 1350 1400\n")
    lines = ["5", "%.9g" % (1e9 * 25.0 / (EM_YIELD * 1e5 * DPPM)), "0.35", "0.9", "", "130.0", "0.04", "-0.02", "",
             "0.5", "0.03", "100", "0.35", "0.9", "", "0.6", "-0.3", "-0.3", "", "1.0", ""]
    lines += ["%.9g" % v for v in BFR12]
    with open(os.path.join(path, "cfg.txt"), "w") as f:
        f.write("# ppc configuration file: follow strict order below\n" + "\n".join(lines) + "\n")
    # 2-d tilt: 11 x 10 hexagonal grid along 9.3 / 129.3 deg with the SE and NW corners cut
    mnum, mcut = (11, 10), (6, 5)
    pts = [(i, j) for j in range(mnum[1]) for i in range(mnum[0]) if i < j + mcut[0] and j < i + mcut[1]]
    with open(os.path.join(path, "tilt.set"), "w") as f:
        f.write("9.3 -609.975 125.154 %d %d\n129.3 -549.188 125.402 %d %d\n" % (mnum[0], mcut[0], mnum[1], mcut[1]))
    ph = rng.uniform(0, 2 * math.pi, len(pts))
    amp = rng.uniform(3.0, 15.0, len(pts))
    with open(os.path.join(path, "tilt.map"), "w") as f:
        for dp in 1100.0 + 10.0 * np.arange(171):
            vals = [amp[k] * math.sin(2 * math.pi * (dp - 1100.0) / 700.0 + ph[k]) for k in range(len(pts))]
            f.write("%.2f " % dp + " ".join("%.4f" % v for v in vals) + "\n")
    # post-processing inputs: RDE / high-QE, DOMs switched off, cable positions, high-QE acceptance
    with open(os.path.join(path, "eff-f2k"), "w") as f:
        for s in range(36, 45):
            for d in range(20, 60, 3):
                f.write("%d %d %.3f %d\n" % (s, d, rng.uniform(1.2, 1.4), int(s % 2)))
    with open(os.path.join(path, "hvs-f2k"), "w") as f:
        for s in range(1, 87):
            for d in range(1, NDOM + 1):
                f.write("%d %d %d\n" % (s, d, 0 if rng.uniform() < 0.02 else 1300))
    with open(os.path.join(path, "dx.dat"), "w") as f:
        for s in range(1, 87):
            for d in range(1, NDOM + 1):
                f.write("%d %d %.2f 10\n" % (s, d, rng.uniform(0.0, 360.0)))
    with open(os.path.join(path, "wv.rde"), "w") as f:
        for w in range(260, 700, 20):
            f.write("%d %.4f\n" % (w, 1.0 + 0.35 * math.exp(-0.5 * ((w - 400.0) / 60.0) ** 2)))
    with open(os.path.join(path, "as.dat"), "w") as f:
        f.write("0.68\n0.34\n0.21\n0.08\n-0.03\n-0.01\n")
    rnd = mwc_multipliers(nrnd)
    with open(os.path.join(path, "rnd.txt"), "w") as f:
        f.write("\n".join(str(int(a)) for a in rnd) + "\n")
    _write_wv(path)
    return path


def write_c2xb(path, nrnd=131072):
    """The full-feature 86-string ice (write_c2x) with the two remaining per-DOM and per-string
    inputs of ini.cxx: DOM orientations (cx.dat: the axis of every DOM on strings 36-44, tilted
    by up to 10 deg) and fixed positions of the hole-ice columns (str-f2k: strings 36-44 moved
    by up to 0.5 m from their DOMs' average)."""
    write_c2x(path, nrnd=nrnd)
    rng = np.random.default_rng(11)
    with open(os.path.join(path, "cx.dat"), "w") as f:
        for st in range(36, 45):
            for d in range(1, NDOM + 1):
                th, ph = math.radians(rng.uniform(0.0, 10.0)), rng.uniform(0.0, 2 * math.pi)
                f.write("%d %d %.6f %.6f %.6f 0.5\n" % (st, d, math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph),
                                                       -math.cos(th)))
    strings = _string_positions(86)
    with open(os.path.join(path, "str-f2k"), "w") as f:
        for st in range(36, 45):
            x, y = strings[st - 1]
            f.write("%d %.3f %.3f\n" % (st, x + rng.uniform(-0.5, 0.5), y + rng.uniform(-0.5, 0.5)))
    return path


def c3_muon_event(seed=4, length=1200.0, step=1.0, E=1000.0):
    """The muon workload on the 86-string ice: a 1 TeV muon on a straight 1.2 km
    track through the array centre at a zenith in [0, 80] deg, as f2k track
    records (gpu/f2k.cxx:555-583) -- 1 m "amu" segments with continuous light
    and "brems" cascades at exponentially spaced points (mean 1 per 10 m) with
    energies log-uniform in [1, 100] GeV.

    Returns a list of dicts (name, x, y, z, th, ph, l, E, t) in the detector
    frame; th/ph are the f2k zenith/azimuth of the track (source direction)."""
    rng = np.random.default_rng(seed)
    zen = math.radians(rng.uniform(0.0, 80.0))
    azi = rng.uniform(0.0, 2 * math.pi)
    # propagation direction: down-going from zenith (zen, azi)
    d = np.array([-math.sin(zen) * math.cos(azi), -math.sin(zen) * math.sin(azi), -math.cos(zen)])
    centre = np.array([rng.uniform(-50, 50), rng.uniform(-50, 50), ZOFF - 1950.0])
    start = centre - d * length / 2
    th, ph = math.degrees(zen), math.degrees(azi)
    c = 0.299792458
    recs = []
    nseg = int(round(length / step))
    for k in range(nseg):
        p = start + d * (k * step)
        recs.append(dict(name="amu", x=float(p[0]), y=float(p[1]), z=float(p[2]), th=th, ph=ph, l=step, E=E,
                         t=k * step / c))
    s = rng.exponential(10.0)
    while s < length:
        p = start + d * s
        e = float(math.exp(rng.uniform(math.log(1.0), math.log(100.0))))
        recs.append(dict(name="brems", x=float(p[0]), y=float(p[1]), z=float(p[2]), th=th, ph=ph, l=0.0, E=e, t=s / c))
        s += rng.exponential(10.0)
    return recs


def c4_batch(det, seed=5, nsrc=512, nvar=8, lo=1e5, hi=1e8, scale=1.0):
    """The batched workload on the 7-string ice (det must hold it): nsrc sources,
    a third each flashers at random DOMs, EM cascades and 10 m muon segments at
    random points of the array with random directions, photon counts log-uniform
    in [lo, hi] (times scale), in nvar ice variants whose per-layer adust(400)
    and be(400) are the loaded ice's scaled by factors drawn in [0.9, 1.1].

    Returns (bunches, variant per bunch, layer factors [nvar][size])."""
    rng = np.random.default_rng(seed)
    q = det.ice_parameters()
    size = len(q["par_a"])
    layers = np.stack([det.layers_for(q["par_a"] * rng.uniform(0.9, 1.1, size),
                                      q["par_b"] * rng.uniform(0.9, 1.1, size)) for _ in range(nvar)])
    strs = det.array("str", np.int32)
    doms = det.array("dom", np.int32)
    oms = det.array("oms", np.float32).reshape(-1, 5)
    bunches, variant = [], []
    for k in range(nsrc):
        num = max(1, int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))) * scale)))
        kind = k % 3
        if kind == 0:
            d = int(rng.integers(len(strs)))
            b = det.flasher_bunch(int(strs[d]), int(doms[d]), num=num, seg=k)
        else:
            d = oms[int(rng.integers(len(oms)))]
            x, y, z = float(d[2] + rng.uniform(-60, 60)), float(d[3] + rng.uniform(-60, 60)), float(d[4])
            ct = rng.uniform(-1, 1)
            st = math.sqrt(1 - ct * ct)
            phi = rng.uniform(0, 2 * math.pi)
            n = (st * math.cos(phi), st * math.sin(phi), ct)
            if kind == 1:
                b, _ = det.particle_bunch(x, y, z, 0.0, float(math.exp(rng.uniform(math.log(1e3), math.log(1e6)))),
                                          0.0, 11, n, num=num, seg=k)
            else:
                b, _ = det.particle_bunch(x, y, z, 0.0, 1e3, 10.0, 0, n, num=num, seg=k)
        bunches.append(b)
        variant.append(int(rng.integers(nvar)))
    return bunches, np.array(variant, dtype=np.uint16), layers


def f2k_lines(recs):
    """the records as f2k TR lines"""
    return ["TR 1 0 %s %.4f %.4f %.4f %.5f %.5f %.4f %.6g %.4f" % (r["name"], r["x"], r["y"], r["z"], r["th"], r["ph"],
                                                                  r["l"], r["E"], r["t"]) for r in recs]


F2K_TYPE = {"amu": 0, "brems": 11, "e": 11, "delta": 11, "epair": 11, "e-": 12, "e+": 13, "hadr": 101, "munu": 101}


def write_cbig(path, seed=4, nrnd=131072):
    """A detector near the reference's capacity: 200 strings x 80 DOM positions at 12 m
    spacing, 14280 of them in ice (on strings 1-86 the reference counts DOMs above 60 as
    surface positions), 208 lines of at most 240 (strings 79-86 split), a cell grid at its
    21 x 19 maximum, on the 170 layers of the 86-string ice without tilt or anisotropy."""
    os.makedirs(path, exist_ok=True)
    rng = np.random.default_rng(seed)
    strings = _string_positions(200)
    depths = 1105.0 + 10.0 * np.arange(170)
    be = np.exp(rng.uniform(math.log(0.02), math.log(0.1), 170))
    ba = np.exp(rng.uniform(math.log(0.005), math.log(0.03), 170))
    _write_geo(os.path.join(path, "geo-f2k"), strings, ndom=80, step=12.0)
    _write_ice(path, depths, be, ba)
    eff = 1e9 * 25.0 / (EM_YIELD * 1e5 * DPPM)
    _write_cfg(path, eff=eff)
    _write_common(path, nrnd)
    return path


def write_c2t9(path, nrnd=131072):
    """The 86-string ice with nine 1-d tilt reference strings instead of six (a tilt
    table of more points than IceCube's tilt.par)."""
    return write_c2(path, nrnd=nrnd, ntilt=9)


def _fibonacci_dirs(n):
    """n nearly uniform unit directions (Fibonacci lattice on the sphere)"""
    k = np.arange(n) + 0.5
    ct = 1 - 2 * k / n
    st = np.sqrt(1 - ct * ct)
    ph = math.pi * (1 + math.sqrt(5.0)) * k
    return np.column_stack([st * np.cos(ph), st * np.sin(ph), ct])


def write_c1ng(path, nrnd=131072):
    """The 7-string ice with nextgen sensors (multi-PMT modules, ini.cxx:731-761): om.conf
    with a 12-PMT module (rings of 4 PMTs at zenith 30, 90 and 150 deg; an ellipsoid of
    radii 0.16/0.205 m) and a 2-PMT module (up and down; 0.15/0.27 m), om.map placing them on
    strings 1 and 2 (the other strings keep the one-PMT DOM), their wavelength acceptances
    om.wv_<type>.0, and the direction mesh om.dirs of the PMT-to-module area ratio."""
    write_c1(path, nrnd=nrnd)
    with open(os.path.join(path, "om.conf"), "w") as f:
        pmts = [(30.0, 90.0 * k) for k in range(4)] + [(90.0, 45.0 + 90.0 * k) for k in range(4)] + \
               [(150.0, 90.0 * k) for k in range(4)]
        f.write("mdom 1 0.35 0.50 0.16 0.205 12 %.1f %.1f\n" % pmts[0])
        for th, ph in pmts[1:]:
            f.write("%.1f %.1f\n" % (th, ph))
        f.write("degg 2 0.9 0.25 0.15 0.27 2 0.0 0.0\n180.0 0.0\n")
    with open(os.path.join(path, "om.dirs"), "w") as f:
        for k, d in enumerate(_fibonacci_dirs(2000)):
            f.write("%d %.7f %.7f %.7f\n" % (k, d[0], d[1], d[2]))
    with open(os.path.join(path, "om.map"), "w") as f:
        for s, m in ((1, 1), (2, 2)):
            for d in range(1, 61):
                f.write("%d %d %d\n" % (s, d, m))
    # acceptance per nm on the scale of the one-PMT DOM's (dppm * doma * omav, gpu/ini.cxx:1425-1433)
    for m, (lo, hi, amp) in ((1, (290.0, 620.0, 60.0)), (2, (300.0, 600.0, 40.0))):
        with open(os.path.join(path, "om.wv_%d.0" % m), "w") as f:
            for w in np.arange(lo, hi + 1e-3, 10.0):
                x = (w - lo) / (hi - lo)
                f.write("%.1f %.6f\n" % (w, amp * math.sin(math.pi * x) ** 2))
    return path


def write_c1wom(path, nrnd=131072):
    """The 7-string ice with wavelength-shifting tube sensors (gpu/ini.cxx:211-243, 731-761):
    om.conf with a tube of radius 0.057 m and length 0.76 m (beta -2, Rz < 0: a cylinder, read
    out at its two ends, "PMTs" 0 and 1 along the axis), om.map placing tubes on string 2 and
    on string 1 except its DOM 30 (the flasher of the C1 case keeps the one-PMT DOM), the
    tube's wavelength acceptance om.wv_3.0, and the direction mesh om.dirs."""
    write_c1(path, nrnd=nrnd)
    with open(os.path.join(path, "om.conf"), "w") as f:
        f.write("wom 3 0.5 -2 0.057 -0.38 2 0.0 0.0\n180.0 0.0\n")
    with open(os.path.join(path, "om.dirs"), "w") as f:
        for k, d in enumerate(_fibonacci_dirs(2000)):
            f.write("%d %.7f %.7f %.7f\n" % (k, d[0], d[1], d[2]))
    with open(os.path.join(path, "om.map"), "w") as f:
        for s in (1, 2):
            for d in range(1, 61):
                if (s, d) != (1, 30):
                    f.write("%d %d 3\n" % (s, d))
    with open(os.path.join(path, "om.wv_3.0"), "w") as f:
        lo, hi = 280.0, 560.0
        for w in np.arange(lo, hi + 1e-3, 10.0):
            x = (w - lo) / (hi - lo)
            f.write("%.1f %.6f\n" % (w, 50.0 * math.sin(math.pi * x) ** 2))
    return path


_WRITERS = {"c1": write_c1, "c2": write_c2, "c2x": write_c2x, "c2t9": write_c2t9, "cbig": write_cbig,
            "c1ng": write_c1ng, "c1wom": write_c1wom, "c2xb": write_c2xb}


def ensure(name, nrnd=131072):
    """Path of the generated ice directory ``name`` (generated on first use)."""
    path = os.path.join(CACHE, name)
    stamp = os.path.join(path, ".complete")
    if not os.path.exists(stamp):
        _WRITERS[name](path, nrnd=nrnd)
        with open(stamp, "w") as f:
            f.write("ok\n")
    return path
