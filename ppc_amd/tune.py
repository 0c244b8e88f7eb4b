"""Launch-shape autotuner for a batch of sources.

The propagation kernel's speed depends on the launch shape in ways that differ
between workloads: lanes per workgroup (LDS image copies per CU), waves per CU
(latency hiding against VGPR pressure), the refill threshold (how many of a
wave's 64 lanes may idle before it takes new photons), the photons a wave
claims per atomic on the work counter, and whether the tilt table sits in LDS
(occupancy) or HBM (L2 reads).  ``tune`` times each candidate on the batch it
is given — the same bunches, ice variants and photon count — and returns the
fastest, so a production loop (many graph replays of one batch) runs at the
best shape for its own mix of sources.
"""
import copy
import itertools

import numpy as np

from . import api

DEFAULT_SPACE = {
    "threads": (256, 512),
    "waves_per_cu": (16, 28),
    "refill": (8, 10, 16),
    "chunk": (0, 256),
    "tilt_lds": (0, -1),
}


def candidates(space=None):
    space = dict(DEFAULT_SPACE, **(space or {}))
    keys = list(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, vals))


def sample(bunches, photons):
    """the bunches with their photon counts scaled so that they sum to about
    ``photons`` (each keeps at least one): the same mix of sources, smaller"""
    total = sum(b.num for b in bunches)
    f = min(1.0, photons / max(total, 1))
    out = []
    for b in bunches:
        c = copy.copy(b)
        c.num = max(1, int(b.num * f))
        out.append(c)
    return out


def tune(det, bunches, variant=None, layers=None, space=None, repeats=2, device=0, max_hits=0, photons=2e8,
         log=None):
    """Time every candidate shape on (bunches, variant) in the ice of det (with
    the ice variants ``layers`` [nvar][size] when given), the bunches scaled to
    about ``photons`` in total.  Returns (best shape, list of (shape, kernel ms))."""
    cus = device_cus(device)
    bunches = sample(bunches, photons)
    num = sum(b.num for b in bunches)
    results = []
    for shape in candidates(space):
        blocks = cus * shape["waves_per_cu"] * 64 // shape["threads"]
        if blocks < 1:
            continue
        dev = api.Device(device, blocks=blocks, threads=shape["threads"], refill=shape["refill"],
                         chunk=shape["chunk"], tilt_lds=shape["tilt_lds"], max_bunches=max(len(bunches), 1),
                         max_hits=max_hits)
        try:
            dev.upload(det)
            if layers is not None:
                dev.set_layers(layers)
            dev.launch_batch(bunches, variant if variant is not None else np.zeros(len(bunches)), num)
            dev.wait()
            ms = []
            for _ in range(repeats):
                dev.launch_batch(bunches, variant if variant is not None else np.zeros(len(bunches)), num)
                ms.append(dev.wait().device_ms)
            t = min(ms)
        finally:
            dev.close()
        results.append((shape, t))
        if log:
            log("%s: %.3f ms (%.3e photons/s)" % (shape, t, num / (t * 1e-3)))
    best = min(results, key=lambda r: r[1])[0]
    return best, results


def device_cus(device=0):
    import ctypes

    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int32()
    mem = ctypes.c_uint64()
    api._check(api.hip().ppc_hip_device_info(device, name, 256, ctypes.byref(cus), ctypes.byref(mem)))
    return cus.value


def device_kwargs(shape, cus):
    """ppc_amd.api.Device keyword arguments of a tuned shape"""
    return {"blocks": cus * shape["waves_per_cu"] * 64 // shape["threads"], "threads": shape["threads"],
            "refill": shape["refill"], "chunk": shape["chunk"], "tilt_lds": shape["tilt_lds"]}
