"""Benchmark: photons/s of the HIP propagation kernel on the 86-string cascade.

Workload: the 86-string synthetic array (86 strings x 60 DOMs on a 125 m hex
grid), 170 ice layers with 1-d tilt and 8 % absorption anisotropy, and a
100 TeV EM cascade at 2000 m depth whose Poisson mean is 1e9 photons
(ppc_amd.configs: c2).  One step = one launch that propagates all 1e9 photons
of the cascade; inputs (tables, bunch) are resident on the device, and the
hits stay in HBM (their count is read back).

The JSON line also carries
  roofline      algorithmic FP32 work of the propagation kernel per launch
                (event counts from a counting launch x the per-event operation
                counts of FLOPS below) over its measured duration, against the
                MI355X FP32 vector peak;
  cpu_baseline  the reference's own CPU program (oracle/_ref/ppc, single
                thread per process) on the same ice, one process per usable
                host core (CPU affinity, bounded by the cgroup's CPU quota),
                each propagating a cascade of at least 1e7 photons; the oracle
                restatement ("port") where the reference program is not built;
  adjoint       the gradient of llh's likelihood (llf) of 12 flasher events (1e8
                photons each, 7-string ice, 100 layers: 200 parameters): forward
                and adjoint time, their ratio, and the cost of the same gradient
                by central differences (2 x 200 forward passes);
  end_to_end    the same cascade through the library-mode pipeline llh drives
                (xppc::sett/addp/eout into hitz): bunch upload, launches, hit
                download and the host's hit post-processing, by wall clock;
  batch         512 sources of mixed types in 8 ice variants on the 7-string ice,
                one captured-graph replay per step (at most 2 steps), at the
                launch shape the autotuner picks (--no-batch skips it);
  full_feature_ice  (--ext) the same cascade (at most 2e8 photons per step) on
                the 86-string ice with 2-d tilt, birefringence, hole ice and
                both absorption anisotropies: the kernels with every option;
  adjoint_86_string  (--adjoint-86) the gradient leg on the 86-string ice:
                170 layers, 340 parameters.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--photons 1e9] [--no-cpu] [--cpu-photons 1e7]
                    [--no-adjoint] [--no-batch] [--ext] [--adjoint-86]
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32 vector (AMD published)

# HBM bytes per photon of the propagation kernel on this workload, from rocprofv3 PMC runs
# (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors a plain copy of known size gives):
# profiles/propagate_c2_cascade_pmc.md
HBM_BYTES_PER_PHOTON = 0.84

# Algorithmic FP32 operations of the reference algorithm per event, counted from
# gpu/pro.cu (add/mul = 1, FMA = 2; divide, sqrt, rsqrt, log, exp, sin, cos = 1).
FLOPS = {
    "step": 69,       # tilt-free step: anisotropy, final layer-walk iteration, cell box, advance
    "tilt1": 20,      # 1-d tilt interpolation, per step (pro.cu:169-191)
    "tilt2": 50,      # 2-d tilt interpolation, per step (pro.cu:192-262)
    "crossing": 10,   # one layer boundary in the layer walk (pro.cu:554-560)
    "string": 17,     # one string cylinder test (pro.cu:601-608)
    "dom": 30,        # one DOM ellipsoid test (pro.cu:631-651)
    "scatter": 136,   # HG/SAM angle, anisotropic frame in/out, rotate, next SCA draw (pro.cu:838-886, 501)
    "photon": 120,    # emission (source sampling, one rotate) and the TOT draw (pro.cu:334-477, 496)
    "hit": 25,        # hit record: angles and impact point (pro.cu:731-750)
}


def gpu_run(args):
    from ppc_amd import api, configs

    cas = configs.c2_cascade()
    det = api.Detector(configs.ensure("c2"), {"WFLA": None, "FWID": None, "DTMN": None})
    n = configs.f2k_direction(cas["th"], cas["ph"])
    photons = int(args.photons)

    # event counts per photon from one counting launch (separate kernel build)
    cal = api.Device(0, count_events=True)
    cal.upload(det)
    b, mean = det.particle_bunch(cas["x"], cas["y"], cas["z"], cas["t"], cas["E"], 0.0, 11, n, num=photons // 10)
    c = cal.run([b])
    cal.close()
    per = {k: getattr(c, k) / c.photons for k in ("steps", "crossings", "strings", "doms", "scatters", "hits")}
    fl_photon = (per["steps"] * (FLOPS["step"] + FLOPS["tilt1"]) + per["crossings"] * FLOPS["crossing"]
                 + per["strings"] * FLOPS["string"] + per["doms"] * FLOPS["dom"]
                 + per["scatters"] * FLOPS["scatter"] + FLOPS["photon"] + per["hits"] * FLOPS["hit"])

    dev = api.Device(0)
    dev.upload(det)
    b.num = photons
    for _ in range(args.warmup):
        dev.run([b])
    ms, spans = [], []
    steps = hits = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.launch([b])
        s = dev.wait()
        ms.append(s.device_ms)
        spans.append((s.span_min_ms, s.span_max_ms))
        steps += s.steps
        hits += s.hits
        assert s.photons == photons
    t1 = time.perf_counter()
    dev.close()
    wall = t1 - t0
    kern_ms = sum(ms) / len(ms)
    achieved = photons * fl_photon / (kern_ms * 1e-3) / 1e12
    return {
        "value": photons * args.steps / wall,
        "ms_per_step": wall / args.steps * 1e3,
        "kernel_ms": kern_ms,
        # in-kernel wall-clock span from the first wave start to the first and the last wave end
        "in_kernel_ms": [round(sum(a for a, _ in spans) / len(spans), 3), round(sum(b for _, b in spans) / len(spans), 3)],
        "mean_photons": mean,
        "steps_per_photon": steps / (photons * args.steps),
        "hits_per_photon": hits / (photons * args.steps),
        "flop_per_photon": fl_photon,
        "events_per_photon": per,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32, 4), "traffic": round(HBM_BYTES_PER_PHOTON * photons)},
        "cascade": cas,
    }


def end_to_end_run():
    """the library-mode pipeline llh drives (xppc: sett + addp of the cascade, eout, hitz):
    bunch upload, launches, hit download and the hit post-processing into hitz, by wall
    clock -- with the reference's host post-processing and with it on the device"""
    from ppc_amd import api, configs

    cas = configs.c2_cascade()
    det = api.Detector(configs.ensure("c2"), {"WFLA": None, "FWID": None, "DTMN": None})
    n = configs.f2k_direction(cas["th"], cas["ph"])
    lib = det.lib

    def event():
        lib.xppc_c_sett(n[0], n[1], n[2], 1, 0, 0, 0.0)
        lib.xppc_c_addp(cas["x"], cas["y"], cas["z"], cas["t"], cas["E"], 0.0, 11, 1.0)
        lib.xppc_c_eout()
        k = int(lib.xppc_c_hitz(None, None, None, None, None, 0))
        lib.xppc_c_efin()
        return k

    out = {"workload": "one 100 TeV cascade through xppc::addp/eout into hitz (86-string ice)"}
    for mode in ("host", "device"):
        if mode == "device":
            os.environ["PPC_ACCEPT"] = "device"
        else:
            os.environ.pop("PPC_ACCEPT", None)
        det.start(0)
        event()  # warm-up
        p0, h0, ms0 = det.totals()
        t0 = time.perf_counter()
        acc = event()
        wall = time.perf_counter() - t0
        p1, h1, ms1 = det.totals()
        det.stop()
        out[mode + "_acceptance"] = {"photons": p1 - p0, "raw_hits": h1 - h0, "accepted_hits": acc,
                                     "wall_s": round(wall, 4), "kernel_ms": round(ms1 - ms0, 2),
                                     "photons_per_s": round((p1 - p0) / wall, 1)}
    os.environ.pop("PPC_ACCEPT", None)
    return out


ICE_NAMES = {"c1": "7-string ice", "c2": "86-string ice (1-d tilt, absorption anisotropy)"}


def adjoint_run(events=12, photons=1e8, seed=6, ice="c1"):
    """C5: forward + adjoint of a Poisson LLH over 12 flasher events on the 7-string ice
    (ice="c2": on the 86-string ice, 170 layers)"""
    import numpy as np
    import torch

    from ppc_amd import adjoint, api, configs

    det = api.Detector(configs.ensure(ice), {"WFLA": "405", "FWID": "10", "DTMN": None})
    rng = np.random.default_rng(seed)
    strs, doms = det.array("str", np.int32), det.array("dom", np.int32)
    pick = rng.choice(len(strs), events, replace=False)
    bunches = [det.flasher_bunch(int(strs[d]), int(doms[d]), num=int(photons), seg=k) for k, d in enumerate(pick)]
    sim = adjoint.HitSimulation(det, bunches, max_hits=1 << 25)
    size = len(sim.par_a)
    # synthetic data: Poisson counts at an ice with every layer perturbed by up to 5 %
    mu_d, _ = sim.forward(sim.par_a * rng.uniform(0.95, 1.05, size), sim.par_b * rng.uniform(0.95, 1.05, size))
    k = torch.tensor(rng.poisson(mu_d), dtype=torch.float64)

    def weights(mu):  # d llf / d mu: llh's likelihood of the data (llh/llh.cxx:637-662)
        m = torch.tensor(mu, requires_grad=True)
        adjoint.llf(m, k).sum().backward()
        return m.grad.numpy()

    # warm-up of both passes (first calls allocate the adjoint's buffers); an optimiser
    # evaluates the gradient many times
    mu0, launch0 = sim.forward(sim.par_a, sim.par_b)
    sim.backward(launch0, weights(mu0))
    t0 = time.perf_counter()
    mu, launch = sim.forward(sim.par_a, sim.par_b)
    t1 = time.perf_counter()
    w = weights(mu)
    ga, gb = sim.backward(launch, w)
    t2 = time.perf_counter()
    sim.close()
    fwd, adj = t1 - t0, t2 - t1
    P = 2 * size
    return {"workload": "%d flasher events x %.0e photons, %s, %d layers, %d parameters; llh's likelihood llf"
                        % (events, photons, ICE_NAMES[ice], size, P),
            "forward_s": round(fwd, 4), "forward_kernel_ms": round(launch["stats"].device_ms, 2),
            "adjoint_s": round(adj, 4), "adjoint_kernel_ms": round(sim.last_adjoint["ms"], 2),
            "replayed_photons": int(sim.last_adjoint["photons"]),
            "gradient_cost_vs_forward": round((fwd + adj) / fwd, 3),
            "fd_cost_vs_forward": 2 * P,
            "gradient_evals_per_s": round(1.0 / (fwd + adj), 3),
            "fd_gradient_evals_per_s": round(1.0 / (2 * P * fwd), 5),
            "grad_norm": [float(np.abs(ga).sum()), float(np.abs(gb).sum())]}


def batch_run(steps, seed=5):
    """C4: 512 mixed sources in 8 ice variants, captured graphs, autotuned launch shape.
    Photon indices are 32-bit, so the 8e9 photons of the batch are split into consecutive
    groups of at most 3e9, each captured by its own context; a step replays all groups."""
    from ppc_amd import api, configs, tune

    det = api.Detector(configs.ensure("c1"), {"WFLA": "405", "FWID": "10", "DTMN": None})
    bunches, variant, layers = configs.c4_batch(det, seed=seed)
    num = sum(b.num for b in bunches)
    best, table = tune.tune(det, bunches, variant, layers, photons=2e8, repeats=1)
    cus = tune.device_cus()
    groups, cur, tot = [], [], 0
    for k, b in enumerate(bunches):
        if cur and tot + b.num > 3e9:
            groups.append(cur)
            cur, tot = [], 0
        cur.append(k)
        tot += b.num
    groups.append(cur)
    devs = []
    for g in groups:
        dev = api.Device(0, max_bunches=len(g), max_hits=1 << 25, **tune.device_kwargs(best, cus))
        dev.upload(det)
        dev.set_layers(layers)
        dev.graph_capture([bunches[k] for k in g], variant[g], sum(bunches[k].num for k in g))
        devs.append(dev)
    for dev in devs:
        dev.graph_launch()
    for dev in devs:
        dev.wait()
    t0 = time.perf_counter()
    photons = 0
    for _ in range(steps):
        for dev in devs:
            dev.graph_launch()
        for dev in devs:
            photons += dev.wait().photons
    wall = time.perf_counter() - t0
    for dev in devs:
        dev.close()
    assert photons == num * steps
    return {"workload": "512 sources (flashers, cascades, 10 m muon segments), photons log-uniform in [1e5, 1e8], "
                        "8 ice variants of the 7-string ice", "photons_per_step": num, "graphs": len(devs),
            "photons_per_s": round(num * steps / wall, 1), "ms_per_step": round(wall / steps * 1e3, 2),
            "tuned_shape": best, "candidates": len(table),
            "untuned_default_ms_at_2e8": next((round(t, 2) for sh, t in table
                                               if sh == {"threads": 256, "waves_per_cu": 28, "refill": 10, "chunk": 0,
                                                         "tilt_lds": 0}), None),
            "tuned_ms_at_2e8": round(min(t for _, t in table), 2)}


def usable_cores():
    """host cores this process may run on: its CPU affinity, bounded by the CPU quota of its
    cgroup where one is set (cgroup v2 cpu.max, v1 cfs quota/period)"""
    n = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        try:
            q = float(open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us").read())
            per = float(open("/sys/fs/cgroup/cpu/cpu.cfs_period_us").read())
            if q > 0:
                quota = q / per
        except (OSError, ValueError):
            pass
    if quota is not None:
        n = min(n, max(1, int(quota)))
    return n


def cpu_reference(cores, photons_per_core):
    """the reference's CPU program, one single-threaded process per core, each propagating one
    cascade of about photons_per_core photons read from stdin"""
    from ppc_amd import configs

    exe = os.path.join(ROOT, "oracle", "_ref", "ppc")
    ice = configs.ensure("c2")
    cas = configs.c2_cascade()
    env = dict(os.environ, PPCTABLESDIR=ice)
    for k in ("WFLA", "FWID"):
        env.pop(k, None)

    def run(E, k):
        line = "TR 1 0 e %.3f %.3f %.3f %.4f %.4f 0 %.6g 0\n" % (cas["x"], cas["y"], cas["z"], cas["th"], cas["ph"], E)
        return subprocess.Popen([exe, str(k)], env=env, stdin=subprocess.PIPE, stdout=subprocess.DEVNULL,
                                stderr=subprocess.PIPE, text=True), line

    # set-up time of one process (tables, 131072 multipliers) from an empty event
    t0 = time.perf_counter()
    p, line = run(1e-6, 0)
    p.communicate(line, timeout=120)
    init = time.perf_counter() - t0
    # an EM cascade yields about 1.0e4 photons per GeV on this ice (after the DOM-efficiency
    # scaling of the front-end): 5 % more energy keeps each process above photons_per_core
    E = 1.05 * photons_per_core / 1.0e4
    t0 = time.perf_counter()
    procs = [run(E, k) for k in range(cores)]
    for p, l in procs:
        p.stdin.write(l)
        p.stdin.close()
    per_proc = []
    for p, _ in procs:
        err = p.stderr.read()
        p.wait(timeout=1800)
        per_proc.append(sum(int(m) for m in re.findall(r"photons: (\d+)", err)))
    wall = time.perf_counter() - t0 - init
    photons = sum(per_proc)
    return {"value": photons / wall, "unit": "photons/s", "cores": cores, "kind": "reference",
            "per_core": photons / wall / cores, "photons_per_core_min": min(per_proc),
            "sample": "%d processes (one per usable host core) x one %.0f GeV EM cascade (%.3g photons in total, "
                      "at least %.3g per process) on the 86-string ice, reference CPU program gpu/ppc.cu built "
                      "with -O2 --fast-math -DDTMN; %.1f s after %.2f s set-up"
                      % (cores, E, photons, min(per_proc), wall, init)}


def cpu_port(cores, photons_per_core):
    """the oracle restatement (C), one process per core, when the reference program is absent"""
    code = r'''
import ctypes, sys, time
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/tests")
import numpy as np
import oracle_util as U
from ppc_amd import api, configs
cas = configs.c2_cascade()
D = api.Detector(configs.ensure("c2"), {"WFLA": None, "FWID": None})
T = U.product_tables(D)
n = configs.f2k_direction(cas["th"], cas["ph"])
b = U.cascade_bunch(cas["x"], cas["y"], cas["z"], n[0], n[1], n[2], int(sys.argv[2]))
k = int(sys.argv[3])
s = U.Rng(362436069 + k, 1234567, int(D.array("rm", np.uint32)[k]))
H = (U.Hit * 100000)()
t0 = time.perf_counter()
U.oracle().orc_run_xcpu(T.ptr, ctypes.byref(b), 1, b.num, 512, ctypes.byref(s), H, 100000, None)
print(b.num, time.perf_counter() - t0)
'''
    num = int(photons_per_core)
    procs = [subprocess.Popen([sys.executable, "-c", code, ROOT, str(num), str(k)], stdout=subprocess.PIPE,
                              stderr=subprocess.DEVNULL, text=True) for k in range(cores)]
    photons, wall = 0, 0.0
    for p in procs:
        out = p.communicate(timeout=900)[0].split()
        photons += int(out[0])
        wall = max(wall, float(out[1]))
    return {"value": photons / wall, "unit": "photons/s", "cores": cores, "kind": "port",
            "per_core": photons / wall / cores,
            "sample": "%d processes x %d photons of the 100 TeV cascade on the 86-string ice, oracle restatement "
                      "(oracle/ppc_oracle.c)" % (cores, num)}


def ext_run(photons, steps):
    """the same cascade on the 86-string ice with every optional feature of the kernel on (2-d tilt,
    birefringence, hole ice, per-layer anisotropy, old absorption anisotropy): the EXT kernels"""
    from ppc_amd import api, configs

    cas = configs.c2_cascade()
    det = api.Detector(configs.ensure("c2x"), {"WFLA": None, "FWID": None, "DTMN": None})
    n = configs.f2k_direction(cas["th"], cas["ph"])
    b, mean = det.particle_bunch(cas["x"], cas["y"], cas["z"], cas["t"], cas["E"], 0.0, 11, n, num=int(photons))
    dev = api.Device(0)
    dev.upload(det)
    dev.run([b])
    ms, steps_total = [], 0
    t0 = time.perf_counter()
    for _ in range(steps):
        s = dev.run([b])
        ms.append(s.device_ms)
        steps_total += s.steps
    wall = time.perf_counter() - t0
    dev.close()
    return {"workload": "100 TeV EM cascade, %.3g photons per step, on the 86-string ice with 2-d tilt, "
                        "birefringence, hole ice and both absorption anisotropies" % photons,
            "photons_per_s": round(photons * steps / wall, 1), "kernel_ms": round(sum(ms) / len(ms), 3),
            "steps_per_photon": round(steps_total / (photons * steps), 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e9)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-photons", type=float, default=1e7, help="photons per host core of the CPU baseline")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-workload leg")
    ap.add_argument("--ext", action="store_true", help="also time the full-feature ice (EXT kernels)")
    ap.add_argument("--adjoint-86", action="store_true", help="also time the gradient on the 86-string ice")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the propagator runs on a single MI355X (one photon stream per lane, "
              "no cross-GPU exchange); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return
    g = gpu_run(args)
    e2e = None if args.no_adjoint else end_to_end_run()
    adj = None if args.no_adjoint else adjoint_run()
    bat = None if args.no_batch else batch_run(min(args.steps, 2))
    ext = ext_run(min(args.photons, 2e8), args.steps) if args.ext else None
    adj86 = adjoint_run(ice="c2") if args.adjoint_86 else None
    cpu = None
    if not args.no_cpu:
        cores = usable_cores()
        if os.path.exists(os.path.join(ROOT, "oracle", "_ref", "ppc")):
            cpu = cpu_reference(cores, args.cpu_photons)
        else:
            cpu = cpu_port(cores, args.cpu_photons)
    out = {
        "metric": "photons/s on 1 GPU (86-string cascade) vs host CPU; LLH-gradient evals/s vs FD",
        "value": g["value"],
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "86-string cascade: 86x60 DOMs, 170 layers, 1-d tilt, 8%% absorption anisotropy, "
                               "100 TeV EM cascade, %.3g photons per step" % args.photons,
                   "photons_per_step": int(args.photons), "kernel_ms": round(g["kernel_ms"], 3),
                   "in_kernel_ms": g["in_kernel_ms"],
                   "steps_per_photon": round(g["steps_per_photon"], 2),
                   "hits_per_photon": g["hits_per_photon"], "flop_per_photon": round(g["flop_per_photon"], 1),
                   "events_per_photon": {k: round(v, 4) for k, v in g["events_per_photon"].items()}},
        "roofline": g["roofline"],
        "cpu_baseline": cpu,
    }
    if e2e:
        out["end_to_end"] = e2e
    if adj:
        out["adjoint"] = adj
    if bat:
        out["batch"] = bat
    if ext:
        out["full_feature_ice"] = ext
    if adj86:
        out["adjoint_86_string"] = adj86
    if cpu:
        out["config"]["gpu_over_cpu"] = round(g["value"] / cpu["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
