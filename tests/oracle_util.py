"""ctypes access to the oracle (oracle/liboracle.so) and to the reference's own
CPU build (oracle/_ref/libppcref.so), for the tests only."""
import ctypes
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
# the same restatement built keeping NaN semantics, for tests that force a NaN into a table
ORACLE_NAN_SO = os.path.join(ROOT, "oracle", "liboracle_nan.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libppcref.so")
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "ppc")

FPI = 3.141592653589793
FCV = FPI / 180.0
OMR = 0.16510

WNUM = 512
MAXLYS = 172


class Bunch(ctypes.Structure):
    _fields_ = [("r", ctypes.c_float * 4), ("n", ctypes.c_float * 4), ("q", ctypes.c_uint32),
                ("num", ctypes.c_uint32), ("type", ctypes.c_int32), ("f", ctypes.c_float),
                ("a", ctypes.c_float), ("b", ctypes.c_float), ("beta", ctypes.c_float),
                ("tau", ctypes.c_float), ("ka", ctypes.c_float), ("up", ctypes.c_float),
                ("fldr", ctypes.c_float), ("fla", ctypes.c_int32), ("ofla", ctypes.c_int32)]


class Hit(ctypes.Structure):
    _fields_ = [("i", ctypes.c_uint32), ("n", ctypes.c_uint32), ("z", ctypes.c_uint32),
                ("t", ctypes.c_float), ("pth", ctypes.c_float), ("pph", ctypes.c_float),
                ("dth", ctypes.c_float), ("dph", ctypes.c_float)]


class Rng(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("c", ctypes.c_uint32), ("a", ctypes.c_uint32)]


class Counters(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in
                ("photons", "steps", "crossings", "cells", "strings", "doms", "scatters", "hits", "draws",
                 "killed", "nonfinite")]


HIT_DTYPE = np.dtype([("i", "<u4"), ("n", "<u4"), ("z", "<u4"), ("t", "<f4"), ("pth", "<f4"),
                      ("pph", "<f4"), ("dth", "<f4"), ("dph", "<f4")])

TABLE_FIELDS = ["size", "dh", "rdh", "hmin", "ocv", "sf", "g", "gr", "xR", "SF", "G", "GR", "hr", "hr2",
                "hs", "ha", "hifl", "azx", "azy", "k1", "k2", "kz", "fr", "sum", "bfr", "az_k1", "az_k2",
                "az_ra", "az_rb", "rx", "cn", "cl", "crst", "cb", "ls", "is_", "nsc", "sc", "tmod", "vthk",
                "lnum", "lpts", "l0", "lmin", "lrdz", "r0", "lnx", "lny", "lr", "mmin", "mstp", "mnum",
                "mcut", "mcol", "lp_x", "lp_y", "wvl", "ocm", "coschr", "sinchr", "abs_", "sca", "wvs_w",
                "wvs_i", "wvs_f", "eff", "gsize", "oms", "str", "dom", "rsize", "l0"]


def _setup(lib):
    lib.orc_tables_new.restype = ctypes.c_void_p
    lib.orc_tables_free.argtypes = [ctypes.c_void_p]
    lib.orc_field.restype = ctypes.c_void_p
    lib.orc_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
    return lib


_oracle = {}
_ref = None


def oracle(nan=False):
    """the oracle library; nan=True: its build that keeps NaN semantics"""
    if nan not in _oracle:
        path = ORACLE_NAN_SO if nan else ORACLE_SO
        if not os.path.exists(path):
            raise RuntimeError("%s is not built" % path)
        lib = _setup(ctypes.CDLL(path))
        lib.orc_xrnd.restype = ctypes.c_float
        lib.orc_xrnd.argtypes = [ctypes.POINTER(Rng)]
        lib.orc_mrnd.restype = ctypes.c_float
        lib.orc_mrnd.argtypes = [ctypes.c_float, ctypes.POINTER(Rng)]
        lib.orc_grnd.restype = ctypes.c_float
        lib.orc_grnd.argtypes = [ctypes.POINTER(Rng)]
        lib.orc_rotate.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                                   ctypes.POINTER(Rng)]
        lib.orc_scatter.restype = ctypes.c_float
        lib.orc_scatter.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.POINTER(Rng)]
        lib.orc_zshift.restype = ctypes.c_float
        lib.orc_zshift.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        lib.orc_photon.restype = ctypes.c_int
        lib.orc_photon.argtypes = [ctypes.c_void_p, ctypes.POINTER(Bunch), ctypes.c_int, ctypes.c_uint32,
                                   ctypes.POINTER(Rng), ctypes.POINTER(Hit), ctypes.c_int,
                                   ctypes.POINTER(Counters)]
        lib.orc_run_xcpu.restype = ctypes.c_uint32
        lib.orc_run_xcpu.argtypes = [ctypes.c_void_p, ctypes.POINTER(Bunch), ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.POINTER(Rng), ctypes.POINTER(Hit), ctypes.c_uint32,
                                     ctypes.POINTER(Counters)]
        lib.orc_adjoint_photon.restype = ctypes.c_int
        lib.orc_adjoint_photon.argtypes = [ctypes.c_void_p, ctypes.POINTER(Bunch), ctypes.c_int, ctypes.c_uint32,
                                           ctypes.POINTER(Rng), ctypes.POINTER(Adjoint)]
        lib.orc_adjoint_path.restype = ctypes.c_int
        lib.orc_adjoint_path.argtypes = [ctypes.c_void_p, ctypes.POINTER(Bunch), ctypes.c_int, ctypes.c_uint32,
                                         ctypes.POINTER(Rng), ctypes.POINTER(Adjoint), ctypes.POINTER(Counters),
                                         ctypes.POINTER(ctypes.c_uint32)]
        _oracle[nan] = lib
    return _oracle[nan]


class Adjoint(ctypes.Structure):
    _fields_ = [("lk", ctypes.POINTER(ctypes.c_float)), ("la", ctypes.POINTER(ctypes.c_float)),
                ("dpda", ctypes.POINTER(ctypes.c_float)), ("ds1db", ctypes.POINTER(ctypes.c_float)),
                ("weight", ctypes.POINTER(ctypes.c_float)), ("nseg", ctypes.c_uint32), ("gsize", ctypes.c_uint32),
                ("ga", ctypes.POINTER(ctypes.c_double)), ("gb", ctypes.POINTER(ctypes.c_double)),
                ("G", ctypes.c_double)]


def oracle_adjoint(tables, bunches, photons, states, weight, lk, la, dpda, ds1db):
    """Score-function gradient of the oracle for the given photons (index, MWC
    state (x, c, a)): per-layer sums (ga, gb) in float64 and the hits per photon."""
    lib = oracle()
    size = len(dpda)
    f32 = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    lk, la, dpda, ds1db, weight = f32(lk), f32(la), f32(dpda), f32(ds1db), f32(weight)
    ga = np.zeros(size, dtype=np.float64)
    gb = np.zeros(size, dtype=np.float64)
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    A = Adjoint(P(lk, ctypes.c_float), P(la, ctypes.c_float), P(dpda, ctypes.c_float), P(ds1db, ctypes.c_float),
                P(weight, ctypes.c_float), weight.shape[0], weight.shape[1], P(ga, ctypes.c_double),
                P(gb, ctypes.c_double), 0.0)
    arr = (Bunch * len(bunches))(*bunches)
    nh = np.zeros(len(photons), dtype=np.int32)
    for k, (i, st) in enumerate(zip(photons, states)):
        s = Rng(int(st[0]), int(st[1]), int(st[2]))
        nh[k] = lib.orc_adjoint_photon(tables, arr, len(bunches), int(i), ctypes.byref(s), ctypes.byref(A))
    return ga, gb, nh


def oracle_path_scores(tables, bunch, states, lk, la, dpda, ds1db):
    """Scores of complete photon paths (orc_adjoint_path), photon by photon: per photon the
    per-layer absorption and scattering scores (float64 [n, size]), its wavelength row and
    its number of scatterings."""
    lib = oracle()
    size = len(dpda)
    f32 = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    lk, la, dpda, ds1db = f32(lk), f32(la), f32(dpda), f32(ds1db)
    weight = np.zeros((1, 1), dtype=np.float32)
    n = len(states)
    GA = np.zeros((n, size), dtype=np.float64)
    GB = np.zeros((n, size), dtype=np.float64)
    rows = np.zeros(n, dtype=np.uint32)
    nsc = np.zeros(n, dtype=np.int64)
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    arr = (Bunch * 1)(bunch)
    row = ctypes.c_uint32()
    for k, st in enumerate(states):
        ga, gb = GA[k], GB[k]
        A = Adjoint(P(lk, ctypes.c_float), P(la, ctypes.c_float), P(dpda, ctypes.c_float), P(ds1db, ctypes.c_float),
                    P(weight, ctypes.c_float), 0, 1, P(ga, ctypes.c_double), P(gb, ctypes.c_double), 0.0)
        cnt = Counters()
        s = Rng(int(st[0]), int(st[1]), int(st[2]))
        lib.orc_adjoint_path(tables, arr, 1, k, ctypes.byref(s), ctypes.byref(A), ctypes.byref(cnt), ctypes.byref(row))
        rows[k] = row.value
        nsc[k] = cnt.scatters
    return GA, GB, rows, nsc


def point_bunch(x, y, z, num, t=0.0, seg=0):
    """an isotropic point source (a flasher-type bunch with ka = -1 and no DOM):
    every photon starts at (x, y, z) at time t (gpu/pro.cu:355-378)"""
    b = Bunch()
    b.r[:] = [x, y, z, t]
    b.n[:] = [0.0, 0.0, 1.0, 1.0]
    b.q, b.num, b.type, b.f = seg, num, 1, 0.0
    b.ka, b.up, b.fldr = -1.0, 0.0, 0.0
    b.fla, b.ofla = -1, -1
    return b


def ref_available():
    return os.path.exists(REF_SO)


def reference():
    """The reference's CPU build (one per process: it holds global state)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            raise RuntimeError("oracle/_ref/libppcref.so is not built")
        lib = ctypes.CDLL(REF_SO)
        lib.ref_export.argtypes = [ctypes.c_void_p]
        lib.ref_zshift.restype = ctypes.c_float
        lib.ref_zshift.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
        lib.ref_run.restype = ctypes.c_uint32
        lib.ref_run.argtypes = [ctypes.POINTER(Bunch), ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.POINTER(Hit), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                ctypes.POINTER(ctypes.c_uint32)]
        lib.ref_multiplier.restype = ctypes.c_uint32
        lib.ref_pmax.restype = ctypes.c_uint32
        _ref = lib
    return _ref


class Tables:
    """An orc_tables instance owned by Python."""

    def __init__(self):
        self.lib = oracle()
        self.ptr = self.lib.orc_tables_new()

    def __del__(self):
        try:
            self.lib.orc_tables_free(self.ptr)
        except Exception:
            pass

    def field(self, name):
        nb = ctypes.c_uint64()
        p = self.lib.orc_field(self.ptr, name.encode(), ctypes.byref(nb))
        if not p:
            raise KeyError(name)
        return (ctypes.c_char * nb.value).from_address(p)

    def as_dict(self):
        return {k: bytes(self.field(k)) for k in TABLE_FIELDS}

    def load_dict(self, d):
        for k, v in d.items():
            buf = self.field(k)
            if len(v) != len(buf):
                raise ValueError("field %s: %d bytes vs %d" % (k, len(v), len(buf)))
            ctypes.memmove(buf, v, len(v))

    def get(self, name, dtype):
        return np.frombuffer(bytes(self.field(name)), dtype=dtype)


def ref_tables(icedir, env=None):
    """Initialise the reference from ``icedir`` and export its tables."""
    os.environ["PPCTABLESDIR"] = icedir
    for k, v in (env or {}).items():
        os.environ[k] = v
    lib = reference()
    lib.ref_initialize()
    lib.ref_start(0)
    t = Tables()
    lib.ref_export(t.ptr)
    return t


def flasher_bunch(tables, string, dom, num, fwid=10.0, seg=0):
    """The bunch flset()/flone() build for an in-ice flasher (f2k.cxx:630-759):
    azimuthally symmetric LED ring (FLDR unset), elevation -0.2 deg, 2-d gaussian
    beam of width fwid degrees, starting on the DOM surface."""
    strs = tables.get("str", np.int32)
    doms = tables.get("dom", np.int32)
    gsize = int(tables.get("gsize", np.int32)[0])
    idx = [k for k in range(gsize) if strs[k] == string and doms[k] == dom]
    fla = idx[0]
    om = np.frombuffer(bytes(tables.field("oms")), dtype=np.float32).reshape(-1, 5)[fla]
    f32 = np.float32
    fcv = f32(FPI) / f32(180.0)
    up = f32(-0.2) * fcv
    az = f32(-1.0) * fcv
    b = Bunch()
    b.r[:] = [om[2], om[3], om[4], 0.0]
    b.n[:] = [np.cos(up) * np.cos(az), np.cos(up) * np.sin(az), np.sin(up), 0.0]
    b.q, b.num, b.type, b.f = seg, num, 1, 0.0
    b.ka = float((fcv * f32(fwid)) * (fcv * f32(fwid)))
    b.up = float(up)
    b.fldr = -1.0
    b.fla = fla
    b.ofla = fla
    return b


def cascade_bunch(x, y, z, nx, ny, nz, num, E=1e5, seg=0, t=0.0):
    """An EM cascade bunch as addp(type 11) builds it (f2k.cxx:59-79,484-503)."""
    b = Bunch()
    b.r[:] = [x, y, z, t]
    b.n[:] = [nx, ny, nz, 0.0]
    logE = math.log(max(0.105658389, E))
    Lrad = 0.39652 * 0.910 / 0.9216
    b.q, b.num, b.type, b.f = seg, num, -11, 0.0
    b.a = float(np.float32(2.01849 + 0.63176 * logE))
    b.b = float(np.float32(Lrad / 0.63207))
    b.beta, b.tau = 1.0, 0.0
    return b


def hits_array(buf, n):
    return np.frombuffer(bytes(buf), dtype=HIT_DTYPE, count=n).copy()


# ---------------------------------------------------------------- product tables -> oracle tables

_SAME = ["size", "dh", "rdh", "hmin", "ocv", "sf", "g", "gr", "xR", "SF", "G", "GR", "hr", "hr2", "hs", "ha", "hifl",
         "azx", "azy", "k1", "k2", "kz", "fr", "sum", "bfr", "rx", "cn", "cl", "crst", "cb", "ls", "nsc", "tmod",
         "vthk", "lnum", "lpts", "lmin", "lrdz", "r0", "lnx", "lny", "lr", "mmin", "mstp", "mnum", "mcut", "mcol",
         "abs_", "sca", "eff", "gsize", "rsize", "l0"]


def product_tables(det):
    """orc_tables filled from the product's table builder (ppc_amd.api.Detector)."""
    t = Tables()

    def put(name, data):
        buf = t.field(name)
        data = bytes(data)
        if len(data) > len(buf):
            raise ValueError("%s: %d > %d bytes" % (name, len(data), len(buf)))
        ctypes.memmove(buf, data + b"\0" * (len(buf) - len(data)), len(buf))

    for k in _SAME:
        put(k, det.field(k))
    put("is_", det.field("is"))
    put("sc", det.field("sc"))
    size = int(det.scalar("size", np.int32))
    az = np.frombuffer(det.field("az"), dtype=np.float32).reshape(size, 4)
    for c, k in enumerate(["az_k1", "az_k2", "az_ra", "az_rb"]):
        put(k, np.ascontiguousarray(az[:, c]).tobytes())
    lp = np.frombuffer(det.field("lp"), dtype=np.float32).reshape(-1, 2)
    put("lp_x", np.ascontiguousarray(lp[:, 0]).tobytes())
    put("lp_y", np.ascontiguousarray(lp[:, 1]).tobytes())
    wv = np.frombuffer(det.field("wv"), dtype=np.float32).reshape(WNUM, 8)
    put("wvl", np.frombuffer(det.field("wv"), dtype=np.int32).reshape(WNUM, 8)[:, 6].copy().tobytes())
    put("ocm", np.ascontiguousarray(wv[:, 0]).tobytes())
    put("coschr", np.ascontiguousarray(wv[:, 1]).tobytes())
    put("sinchr", np.ascontiguousarray(wv[:, 2]).tobytes())
    wvs = np.frombuffer(det.field("wvs"), dtype=np.float32).reshape(WNUM, 3)
    for c, k in enumerate(["wvs_w", "wvs_i", "wvs_f"]):
        put(k, np.ascontiguousarray(wvs[:, c]).tobytes())
    put("oms", det.field("oms"))
    put("str", det.field("str"))
    put("dom", det.field("dom"))
    return t


def to_orc_bunch(b):
    """ppc_amd.api.Bunch -> oracle Bunch"""
    o = Bunch()
    o.r[:] = list(b.r)
    o.n[:] = list(b.n)
    o.q, o.num, o.type, o.f = b.q, b.num, b.type, b.f
    if b.type > 0:
        o.ka, o.up, o.fldr, o.fla, o.ofla = b.u.l.ka, b.u.l.up, b.u.l.fldr, b.u.l.fla, b.u.l.ofla
    else:
        o.a, o.b, o.beta, o.tau = b.u.p.a, b.u.p.b, b.u.p.beta, b.u.p.tau
    return o


def to_api_bunch(o):
    """oracle Bunch -> ppc_amd.api.Bunch (the 64-byte C-ABI bunch)"""
    from ppc_amd import api

    b = api.Bunch()
    b.r[:] = list(o.r)
    b.n[:] = list(o.n)
    b.q, b.num, b.type, b.f = o.q, o.num, o.type, o.f
    if o.type > 0:
        b.u.l.ka, b.u.l.up, b.u.l.fldr, b.u.l.fla, b.u.l.ofla = o.ka, o.up, o.fldr, o.fla, o.ofla
    else:
        b.u.p.a, b.u.p.b, b.u.p.beta, b.u.p.tau = o.a, o.b, o.beta, o.tau
    return b


def ref_particle(x, y, z, t, E, dr, ptype, n):
    """the reference's addp() bunch (without its Poisson draw) and its mean photon count"""
    lib = reference()
    lib.ref_particle.restype = ctypes.c_double
    lib.ref_particle.argtypes = [ctypes.c_float] * 6 + [ctypes.c_int] + [ctypes.c_float] * 3 + [ctypes.POINTER(Bunch)]
    b = Bunch()
    mean = lib.ref_particle(x, y, z, t, E, dr, ptype, n[0], n[1], n[2], ctypes.byref(b))
    return b, mean


def ref_clst(x, y, z, t, n, dr, beta):
    """the reference's addp_clst() bunch for direction n"""
    lib = reference()
    lib.ref_clst.argtypes = [ctypes.c_float] * 9 + [ctypes.POINTER(Bunch)]
    b = Bunch()
    lib.ref_clst(x, y, z, t, n[0], n[1], n[2], dr, beta, ctypes.byref(b))
    return b


def ref_mopo(x, y, z, t, n, light, length, beta, tau, fr):
    """the reference's addp_mopo() bunch for direction n and its mean photon count"""
    lib = reference()
    lib.ref_mopo.restype = ctypes.c_double
    lib.ref_mopo.argtypes = [ctypes.c_float] * 12 + [ctypes.POINTER(Bunch)]
    b = Bunch()
    mean = lib.ref_mopo(x, y, z, t, n[0], n[1], n[2], light, length, beta, tau, fr, ctypes.byref(b))
    return b, mean


def ref_led(x, y, z, t, E, ptype, n):
    """the reference's addp() bunch for light of the configured flasher (ptype < 0) with
    direction n, and its mean photon count"""
    lib = reference()
    lib.ref_led.restype = ctypes.c_double
    lib.ref_led.argtypes = [ctypes.c_float] * 5 + [ctypes.c_int] + [ctypes.c_float] * 3 + [ctypes.POINTER(Bunch)]
    b = Bunch()
    mean = lib.ref_led(x, y, z, t, E, ptype, n[0], n[1], n[2], ctypes.byref(b))
    return b, mean


def ref_flasher(string, dom):
    lib = reference()
    lib.ref_flasher.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(Bunch)]
    b = Bunch()
    lib.ref_flasher(string, dom, ctypes.byref(b))
    return b


def bunch_fields(b):
    """comparable tuple of an oracle Bunch (the fields the kernel reads for its type)"""
    base = tuple(np.float32(v) for v in list(b.r) + list(b.n)) + (b.q, b.type, np.float32(b.f))
    if b.type > 0:
        return base + (np.float32(b.ka), np.float32(b.up), np.float32(b.fldr), b.fla, b.ofla)
    return base + (np.float32(b.a), np.float32(b.b), np.float32(b.beta), np.float32(b.tau))
