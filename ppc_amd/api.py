"""Python binding of the C-ABI (include/ppc_hip.h) and of the xppc C facade
(include/xppc_c.h).

``Detector`` builds the tables from an ice directory (what ``xppc::initialize``
does, gpu/ppc.cu:36) and makes photon bunches the way the reference front-ends
do.  ``Device`` owns one GPU context of libppc_hip.so: upload, launch, wait,
hits, replay records.  There is no fallback: without the HIP library or a GPU
the calls raise.
"""
import ctypes
import os

import numpy as np

from . import _native

PPC_HIP_WNUM = 512


class _LP(ctypes.Structure):
    _fields_ = [("a", ctypes.c_float), ("b", ctypes.c_float), ("beta", ctypes.c_float), ("tau", ctypes.c_float)]


class _LL(ctypes.Structure):
    _fields_ = [("ka", ctypes.c_float), ("up", ctypes.c_float), ("fldr", ctypes.c_float),
                ("fla", ctypes.c_int16), ("ofla", ctypes.c_int16)]


class _LU(ctypes.Union):
    _fields_ = [("p", _LP), ("l", _LL)]


class Bunch(ctypes.Structure):
    """ppc_hip_bunch: the reference's photon bunch (gpu/ini.cxx:399-419)."""
    _fields_ = [("r", ctypes.c_float * 4), ("n", ctypes.c_float * 4), ("q", ctypes.c_uint32),
                ("num", ctypes.c_uint32), ("type", ctypes.c_int32), ("f", ctypes.c_float), ("u", _LU)]


class Hit(ctypes.Structure):
    _fields_ = [("i", ctypes.c_uint32), ("n", ctypes.c_uint32), ("z", ctypes.c_uint32), ("t", ctypes.c_float),
                ("pth", ctypes.c_float), ("pph", ctypes.c_float), ("dth", ctypes.c_float), ("dph", ctypes.c_float)]


HIT_DTYPE = np.dtype([("i", "<u4"), ("n", "<u4"), ("z", "<u4"), ("t", "<f4"), ("pth", "<f4"), ("pph", "<f4"),
                      ("dth", "<f4"), ("dph", "<f4")])
RECORD_DTYPE = np.dtype([("photon", "<u4"), ("x", "<u4"), ("c", "<u4"), ("a", "<u4")])
ACCEPTED_DTYPE = np.dtype([("pmt", "<i4"), ("wv", "<f4"), ("r", "<f4", (3,)), ("n", "<f4", (3,)), ("dt", "<f4")])
ADJ_RECORD_DTYPE = np.dtype([("photon", "<u4"), ("x0", "<u4"), ("c0", "<u4"), ("lane", "<u4")])
LAYER_DTYPE = np.dtype([("s0", "<f4"), ("s1", "<f4"), ("p", "<f4"), ("t", "<f4")])


class Config(ctypes.Structure):
    _fields_ = [("blocks", ctypes.c_int32), ("threads", ctypes.c_int32), ("refill", ctypes.c_int32),
                ("deterministic", ctypes.c_int32), ("max_bunches", ctypes.c_uint32), ("max_hits", ctypes.c_uint32),
                ("rec_stride", ctypes.c_uint32), ("rec_max", ctypes.c_uint32), ("count_events", ctypes.c_int32),
                ("record_starts", ctypes.c_int32), ("chunk", ctypes.c_uint32), ("tilt_lds", ctypes.c_int32),
                ("no_clearance", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("photons", ctypes.c_uint64), ("steps", ctypes.c_uint64), ("hits", ctypes.c_uint64),
                ("stored", ctypes.c_uint64), ("killed", ctypes.c_uint64), ("nonfinite", ctypes.c_uint64),
                ("device_ms", ctypes.c_float), ("lanes", ctypes.c_int32), ("crossings", ctypes.c_uint64),
                ("strings", ctypes.c_uint64), ("doms", ctypes.c_uint64), ("scatters", ctypes.c_uint64),
                ("hit_photons", ctypes.c_uint64), ("span_min_ms", ctypes.c_float), ("span_max_ms", ctypes.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PPCError(RuntimeError):
    pass


_hip = None
_xppc = None


def hip():
    global _hip
    if _hip is None:
        lib = _native.hip_lib()
        c_p = ctypes.c_void_p
        lib.ppc_hip_last_error.restype = ctypes.c_char_p
        lib.ppc_hip_device_count.argtypes = [ctypes.POINTER(ctypes.c_int32)]
        lib.ppc_hip_device_info.argtypes = [ctypes.c_int32, ctypes.c_char_p, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint64)]
        lib.ppc_hip_open.argtypes = [ctypes.c_int32, ctypes.POINTER(Config), ctypes.POINTER(c_p)]
        lib.ppc_hip_close.argtypes = [c_p]
        lib.ppc_hip_lanes.argtypes = [c_p, ctypes.POINTER(ctypes.c_int32)]
        lib.ppc_hip_capacity.argtypes = [c_p, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
        lib.ppc_hip_upload.argtypes = [c_p, c_p]
        lib.ppc_hip_seed.argtypes = [c_p, ctypes.c_uint32]
        lib.ppc_hip_seed_photons.argtypes = [c_p, ctypes.c_uint64]
        lib.ppc_hip_get_states.argtypes = [c_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int32]
        lib.ppc_hip_set_states.argtypes = [c_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int32]
        lib.ppc_hip_launch.argtypes = [c_p, ctypes.POINTER(Bunch), ctypes.c_uint32, ctypes.c_uint32]
        lib.ppc_hip_wait.argtypes = [c_p, ctypes.POINTER(Stats)]
        lib.ppc_hip_hits.argtypes = [c_p, c_p, c_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        lib.ppc_hip_records.argtypes = [c_p, c_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        lib.ppc_hip_set_layers.argtypes = [c_p, c_p, ctypes.c_int32]
        lib.ppc_hip_launch_batch.argtypes = [c_p, ctypes.POINTER(Bunch), c_p, ctypes.c_uint32, ctypes.c_uint32]
        lib.ppc_hip_graph_capture.argtypes = [c_p, ctypes.POINTER(Bunch), c_p, ctypes.c_uint32, ctypes.c_uint32]
        lib.ppc_hip_graph_launch.argtypes = [c_p]
        lib.ppc_hip_hit_photons.argtypes = [c_p, c_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        lib.ppc_hip_upload_accept.argtypes = [c_p, c_p]
        lib.ppc_hip_accept.argtypes = [c_p, ctypes.c_uint64, c_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        lib.ppc_hip_adjoint.argtypes = [c_p, ctypes.POINTER(Bunch), ctypes.c_uint32, ctypes.c_uint32, c_p,
                                        ctypes.c_uint32, c_p, ctypes.c_uint32, c_p, c_p, c_p, ctypes.POINTER(Stats)]
        lib.ppc_hip_adjoint_noalloc.argtypes = lib.ppc_hip_adjoint.argtypes
        _hip = lib
    return _hip


def xppc():
    global _xppc
    if _xppc is None:
        hip()  # libxppc links libppc_hip
        lib = _native.xppc_lib()
        lib.xppc_c_initialize.argtypes = [ctypes.c_float]
        lib.xppc_c_tables.restype = ctypes.c_void_p
        lib.xppc_c_field.restype = ctypes.c_void_p
        lib.xppc_c_field.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
        lib.xppc_c_extend_multipliers.argtypes = [ctypes.c_int32]
        lib.xppc_c_flasher_bunch.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Bunch)]
        lib.xppc_c_particle_bunch.restype = ctypes.c_double
        lib.xppc_c_particle_bunch.argtypes = [ctypes.c_float] * 6 + [ctypes.c_int32] + [ctypes.c_float] * 3 + [
            ctypes.POINTER(Bunch)]
        lib.xppc_c_clst_bunch.argtypes = [ctypes.c_float] * 9 + [ctypes.POINTER(Bunch)]
        lib.xppc_c_mopo_bunch.restype = ctypes.c_double
        lib.xppc_c_mopo_bunch.argtypes = [ctypes.c_float] * 12 + [ctypes.POINTER(Bunch)]
        lib.xppc_c_flasher_photons.restype = ctypes.c_uint64
        lib.xppc_c_flasher_photons.argtypes = [ctypes.c_uint64]
        lib.xppc_c_zshift.restype = ctypes.c_float
        lib.xppc_c_zshift.argtypes = [ctypes.c_float] * 3 + [ctypes.POINTER(ctypes.c_float)]
        lib.xppc_c_seed.restype = ctypes.c_uint32
        lib.ppc_mwc_multipliers.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        lib.xppc_c_flone.argtypes = [ctypes.c_uint64]
        lib.xppc_c_sett.argtypes = [ctypes.c_float] * 3 + [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32,
                                                            ctypes.c_double]
        lib.xppc_c_addp.argtypes = [ctypes.c_float] * 6 + [ctypes.c_int32, ctypes.c_float]
        lib.xppc_c_hitz.restype = ctypes.c_uint64
        lib.xppc_c_hitz.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_uint64]
        lib.xppc_c_hitz_keys.restype = ctypes.c_uint64
        lib.xppc_c_hitz_keys.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_uint64]
        lib.xppc_c_set_res.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.xppc_c_hitz_tracks.restype = ctypes.c_uint64
        lib.xppc_c_hitz_tracks.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint64]
        lib.xppc_c_accept_tables.restype = ctypes.c_void_p
        lib.xppc_c_totals.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_double)]
        _xppc = lib
    return _xppc


MAX_PHOTONS = 2 ** 32 - 2 ** 24  # per launch: photon indices are 32-bit, the work counter overshoots by a chunk per wave


def _num(num):
    num = int(num)
    if num < 0 or num > MAX_PHOTONS:
        raise PPCError("%d photons in one launch: split the bunches over launches of at most %d" % (num, MAX_PHOTONS))
    return num


def _check(rc):
    if rc != 0:
        raise PPCError("ppc_hip error %d: %s" % (rc, hip().ppc_hip_last_error().decode()))


class Detector:
    """The process-wide detector of libxppc (one per process, as in the reference)."""

    def __init__(self, icedir, env=None, enh=1.0):
        os.environ["PPCTABLESDIR"] = icedir
        for k, v in (env or {}).items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = str(v)
        self.lib = xppc()
        self.lib.xppc_c_initialize(ctypes.c_float(enh))

    def field(self, name):
        nb = ctypes.c_uint64()
        p = self.lib.xppc_c_field(name.encode(), ctypes.byref(nb))
        if not p:
            raise KeyError(name)
        return ctypes.string_at(p, nb.value)

    def array(self, name, dtype):
        return np.frombuffer(self.field(name), dtype=dtype).copy()

    def scalar(self, name, dtype=np.float32):
        return self.array(name, dtype)[0]

    @property
    def tables(self):
        return self.lib.xppc_c_tables()

    def extend_multipliers(self, lanes):
        n = self.lib.xppc_c_extend_multipliers(lanes)
        if n < lanes:
            raise PPCError("could not extend the multiplier table to %d" % lanes)
        return n

    def flasher_bunch(self, string, dom, num=0, seg=0):
        b = Bunch()
        if self.lib.xppc_c_flasher_bunch(string, dom, ctypes.byref(b)) != 0:
            raise PPCError("flasher configuration failed")
        b.num, b.q = num, seg
        return b

    def clst_bunch(self, x, y, z, t, n, dr, beta, num=0, seg=0):
        """addp_clst's bunch: a segment of length dr of a particle of velocity beta"""
        b = Bunch()
        if self.lib.xppc_c_clst_bunch(x, y, z, t, n[0], n[1], n[2], dr, beta, ctypes.byref(b)) != 0:
            raise PPCError("clst bunch failed")
        b.num, b.q = num, seg
        return b

    def mopo_bunch(self, x, y, z, t, n, light, length, beta, tau, fr, num=None, seg=0):
        """addp_mopo's bunch and the Poisson mean of its photon count"""
        b = Bunch()
        mean = self.lib.xppc_c_mopo_bunch(x, y, z, t, n[0], n[1], n[2], light, length, beta, tau, fr, ctypes.byref(b))
        b.num = int(round(mean)) if num is None else num
        b.q = seg
        return b, mean

    def particle_bunch(self, x, y, z, t, E, dr, ptype, n, num=None, seg=0):
        """bunch of an f2k particle (type 0 muon, 11-14 EM, 101-107 hadron); num
        defaults to the mean photon count rounded."""
        b = Bunch()
        mean = self.lib.xppc_c_particle_bunch(x, y, z, t, E, dr, ptype, n[0], n[1], n[2], ctypes.byref(b))
        b.num = int(round(mean)) if num is None else num
        b.q = seg
        return b, mean

    def flasher_photons(self, num):
        return int(self.lib.xppc_c_flasher_photons(num))

    # ---- library-mode pipeline (xppc interface) ----
    def start(self, device=0):
        self.lib.xppc_c_start(device)

    def stop(self):
        self.lib.xppc_c_stop()

    def hitz(self):
        """the hitz map flattened: structured array (str, dom, pmt, t, wv)"""
        n = int(self.lib.xppc_c_hitz(None, None, None, None, None, 0))
        out = {k: np.zeros(n, dtype=d) for k, d in
               [("str", np.int32), ("dom", np.int32), ("pmt", np.int32), ("t", np.float32), ("wv", np.float32)]}
        self.lib.xppc_c_hitz(*[out[k].ctypes.data_as(ctypes.c_void_p) for k in ("str", "dom", "pmt", "t", "wv")], n)
        return out

    def hitz_keys(self):
        """the keys of the hitz map in order: binned time and direction, photon records under each"""
        n = int(self.lib.xppc_c_hitz_keys(None, None, None, 0))
        out = {"time": np.zeros(n, np.float32), "dir": np.zeros(n, np.float32), "count": np.zeros(n, np.uint32)}
        self.lib.xppc_c_hitz_keys(*[out[k].ctypes.data_as(ctypes.c_void_p) for k in ("time", "dir", "count")], n)
        return out

    def hitz_tracks(self):
        """the segment (mcid: frame, id, sub-id, t0) of each hitz key, in key order"""
        n = int(self.lib.xppc_c_hitz_tracks(None, None, None, None, 0))
        out = {"frame": np.zeros(n, np.int32), "id": np.zeros(n, np.int32), "sub": np.zeros(n, np.uint64),
               "t0": np.zeros(n, np.float64)}
        self.lib.xppc_c_hitz_tracks(*[out[k].ctypes.data_as(ctypes.c_void_p) for k in ("frame", "id", "sub", "t0")], n)
        return out

    def layers(self):
        """per-layer factors (s0, s1, p, t) of the ice tables, layer 0 deepest"""
        return self.array("lay", LAYER_DTYPE)

    def ice_parameters(self):
        """per layer: icemodel.dat adust(400) and be(400), and d(p)/d(adust), d(s1)/d(be)"""
        return {k: self.array(k, np.float32) for k in ("par_a", "par_b", "dpda", "ds1db")}

    def layers_for(self, a, b):
        """the layer factors of the same ice with per-layer adust(400) = a and be(400) = b
        (p and s1 are affine in them: gpu/ini.cxx:1630-1640)"""
        lay = self.layers()
        q = self.ice_parameters()
        a = np.asarray(a, dtype=np.float64)
        b = np.asarray(b, dtype=np.float64)
        lay["p"] = lay["p"] + q["dpda"] * (a - q["par_a"])
        lay["s1"] = lay["s1"] + q["ds1db"] * (b - q["par_b"])
        return lay

    def totals(self):
        """photons, raw hits and device time [ms] of the library-mode launches since start"""
        p, h, ms = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        self.lib.xppc_c_totals(ctypes.byref(p), ctypes.byref(h), ctypes.byref(ms))
        return p.value, h.value, ms.value

    def zshift(self, x, y, z):
        dz = ctypes.c_float()
        s = self.lib.xppc_c_zshift(x, y, z, ctypes.byref(dz))
        return s, dz.value


def device_count():
    n = ctypes.c_int32()
    _check(hip().ppc_hip_device_count(ctypes.byref(n)))
    return n.value


class Device:
    """One GPU context of libppc_hip.so."""

    def __init__(self, device=0, blocks=0, threads=0, refill=0, deterministic=False, max_bunches=0, max_hits=0,
                 rec_stride=0, rec_max=0, count_events=False, record_starts=False, chunk=0, tilt_lds=0,
                 no_clearance=False):
        self.lib = hip()
        self.cfg = Config(blocks, threads, refill, 1 if deterministic else 0, max_bunches, max_hits, rec_stride,
                          rec_max, 1 if count_events else 0, 1 if record_starts else 0, chunk, tilt_lds,
                          1 if no_clearance else 0)
        self.ctx = ctypes.c_void_p()
        _check(self.lib.ppc_hip_open(device, ctypes.byref(self.cfg), ctypes.byref(self.ctx)))
        n = ctypes.c_int32()
        _check(self.lib.ppc_hip_lanes(self.ctx, ctypes.byref(n)))
        self.lanes = n.value
        b, h = ctypes.c_uint32(), ctypes.c_uint32()
        _check(self.lib.ppc_hip_capacity(self.ctx, ctypes.byref(b), ctypes.byref(h)))
        self.max_bunches, self.max_hits = b.value, h.value

    def close(self):
        if self.ctx:
            self.lib.ppc_hip_close(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, det):
        det.extend_multipliers(self.lanes)
        _check(self.lib.ppc_hip_upload(self.ctx, det.tables))

    def seed(self, sv):
        _check(self.lib.ppc_hip_seed(self.ctx, sv))

    def seed_photons(self, seed):
        """seed != 0: every photon of later launches starts from a state of its own hashed
        from (seed, photon index) (common random numbers across launches); 0: off"""
        _check(self.lib.ppc_hip_seed_photons(self.ctx, seed))

    def states(self):
        a = np.zeros(self.lanes, dtype=np.uint64)
        _check(self.lib.ppc_hip_get_states(self.ctx, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), self.lanes))
        return a

    def set_states(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        _check(self.lib.ppc_hip_set_states(self.ctx, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), self.lanes))

    def launch(self, bunches, num=None):
        if isinstance(bunches, Bunch):
            bunches = [bunches]
        arr = (Bunch * len(bunches))(*bunches)
        if num is None:
            num = sum(b.num for b in bunches)
        _check(self.lib.ppc_hip_launch(self.ctx, arr, len(bunches), _num(num)))

    def wait(self):
        st = Stats()
        _check(self.lib.ppc_hip_wait(self.ctx, ctypes.byref(st)))
        return st

    def hits(self, with_photon=False, cap=None):
        cap = self.max_hits if cap is None else cap
        st = Stats()
        _check(self.lib.ppc_hip_wait(self.ctx, ctypes.byref(st)))
        m = int(min(st.stored, cap))
        h = np.zeros(m, dtype=HIT_DTYPE)
        ph = np.zeros(m, dtype=np.uint32)
        n = ctypes.c_uint32()
        _check(self.lib.ppc_hip_hits(self.ctx, h.ctypes.data_as(ctypes.c_void_p),
                                     ph.ctypes.data_as(ctypes.c_void_p) if with_photon else None, m,
                                     ctypes.byref(n)))
        h = h[:n.value]
        return (h, ph[:n.value]) if with_photon else h

    def records(self):
        cap = max(self.cfg.rec_max, 1)
        r = np.zeros(cap, dtype=RECORD_DTYPE)
        n = ctypes.c_uint32()
        _check(self.lib.ppc_hip_records(self.ctx, r.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(n)))
        return r[:n.value]

    def run(self, bunches, num=None):
        self.launch(bunches, num)
        return self.wait()

    # ---- ice variants, graphs ----
    def set_layers(self, lay):
        """lay: LAYER_DTYPE array [nvar][size] (or [size]) of per-layer factors"""
        lay = np.ascontiguousarray(lay, dtype=LAYER_DTYPE)
        nvar = 1 if lay.ndim == 1 else lay.shape[0]
        _check(self.lib.ppc_hip_set_layers(self.ctx, lay.ctypes.data_as(ctypes.c_void_p), nvar))

    @staticmethod
    def _bunches(bunches, variant):
        if isinstance(bunches, Bunch):
            bunches = [bunches]
        arr = (Bunch * len(bunches))(*bunches)
        v = None
        if variant is not None:
            v = np.ascontiguousarray(variant, dtype=np.uint16)
            if len(v) != len(bunches):
                raise ValueError("one ice variant per bunch")
        return arr, v

    def launch_batch(self, bunches, variant, num=None):
        arr, v = self._bunches(bunches, variant)
        if num is None:
            num = sum(b.num for b in arr)
        _check(self.lib.ppc_hip_launch_batch(self.ctx, arr, v.ctypes.data_as(ctypes.c_void_p) if v is not None
                                             else None, len(arr), _num(num)))

    def graph_capture(self, bunches, variant=None, num=None):
        arr, v = self._bunches(bunches, variant)
        if num is None:
            num = sum(b.num for b in arr)
        self._graph_keep = (arr, v)
        _check(self.lib.ppc_hip_graph_capture(self.ctx, arr, v.ctypes.data_as(ctypes.c_void_p) if v is not None
                                              else None, len(arr), _num(num)))

    def graph_launch(self):
        _check(self.lib.ppc_hip_graph_launch(self.ctx))

    # ---- hit post-processing ----
    def upload_accept(self, det):
        """the acceptance tables of the detector's print() (xppc_c_accept_tables)"""
        _check(self.lib.ppc_hip_upload_accept(self.ctx, det.lib.xppc_c_accept_tables()))

    def accept(self, seed):
        """device acceptance of the stored hits of the last launch: ACCEPTED_DTYPE per hit"""
        st = Stats()
        _check(self.lib.ppc_hip_wait(self.ctx, ctypes.byref(st)))
        m = int(st.stored)
        a = np.zeros(m, dtype=ACCEPTED_DTYPE)
        n = ctypes.c_uint32()
        _check(self.lib.ppc_hip_accept(self.ctx, seed, a.ctypes.data_as(ctypes.c_void_p), m, ctypes.byref(n)))
        return a[:n.value]

    # ---- adjoint ----
    def hit_photons(self):
        """the photons of the last launch that hit (record_starts): ADJ_RECORD_DTYPE"""
        st = Stats()
        _check(self.lib.ppc_hip_wait(self.ctx, ctypes.byref(st)))
        m = int(min(st.hit_photons, self.max_hits))
        a = np.zeros(m, dtype=ADJ_RECORD_DTYPE)
        n = ctypes.c_uint32()
        _check(self.lib.ppc_hip_hit_photons(self.ctx, a.ctypes.data_as(ctypes.c_void_p), m, ctypes.byref(n)))
        return a[:n.value]

    def adjoint(self, bunches, num, weight, dpda, ds1db, records=None, noalloc=False):
        """replay photons of a forward launch of these bunches -- ``records``
        (ADJ_RECORD_DTYPE), or by default the hitting photons the last launch
        recorded -- and return (d/d adust[size], d/d be[size], stats);
        weight [nseg][gsize] is d(LLH)/d(mean hits).  noalloc: the entry point that
        allocates no device memory (ppc_hip_adjoint_noalloc)"""
        arr, _ = self._bunches(bunches, None)
        weight = np.ascontiguousarray(weight, dtype=np.float32)
        nseg = weight.shape[0] if weight.ndim == 2 else 1
        dpda = np.ascontiguousarray(dpda, dtype=np.float32)
        ds1db = np.ascontiguousarray(ds1db, dtype=np.float32)
        size = len(dpda)
        grad = np.zeros(2 * size, dtype=np.float64)
        st = Stats()
        if records is not None:
            records = np.ascontiguousarray(records, dtype=ADJ_RECORD_DTYPE)
            rp, nrec = records.ctypes.data_as(ctypes.c_void_p), len(records)
        else:
            rp, nrec = None, 0
        fn = self.lib.ppc_hip_adjoint_noalloc if noalloc else self.lib.ppc_hip_adjoint
        _check(fn(self.ctx, arr, len(arr), num, rp, nrec,
                                        weight.ctypes.data_as(ctypes.c_void_p), nseg,
                                        dpda.ctypes.data_as(ctypes.c_void_p), ds1db.ctypes.data_as(ctypes.c_void_p),
                                        grad.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)))
        return grad[:size], grad[size:], st
