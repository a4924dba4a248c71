#!/usr/bin/env python3
"""Benchmark of the MI355X pseudo-spectral step (FP64) on one GPU.

Default workload: ps3d's plain FP64 time step on a 256^3 grid (2 pi x 2 pi x pi,
periodic x, y, free-slip lids), 4-mode Beltrami initial vorticity with seeded
phases plus 1e-3 seeded noise, impl-diff-rk4, 2/3-rule, nnu = 3, prediss = 30,
pretype vorch, adaptive CFL/strain dt.  A "step" is one ps3d advance(): vor2vel,
adapt, source and the four RK stages.  The grid is resident in HBM before the
timed region starts.

Prints one JSON line: steps/s (the metric), grid-point updates/s, the roofline
of the dominant kernel (HIP events on the handle's stream), the whole-step
algorithmic bandwidth, and the CPU baseline (the oracle, a C restatement of
ps3d's unfused algorithm, timed on the host cores on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM = 8.0e12          # MI355X HBM3E, B/s
B_STEP_VOLUMES = 218       # algorithmic full-volume FP64 touches per RK4 step (fused floor)
KIND_NAMES = ["zinv", "yinv", "xmain", "yfwd", "ztend", "yinv_adapt", "xadapt", "reduce", "mean", "ztend_zinv"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="grid size n^3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: the process's CPU share (OMP_NUM_THREADS if set, else all usable CPUs)")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--ic", choices=["beltrami", "random"], default="beltrami",
                    help="beltrami: 4 seeded Beltrami modes + noise; random: random-phase isotropic "
                         "E(k) ~ k^4 exp(-2 (k/4)^2) made on 128^3 and zero-padded to n^3 on the GPU")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 96-instance ensemble measurement")
    ap.add_argument("--no-closure", action="store_true", help="skip the closure-64^3 vs plain-256^3 measurement")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other grid sizes (32^3, 128^3, 512^3, 1024^3)")
    ap.add_argument("--closure-seed", type=int, default=9001, help="held-out seed of the closure measurement")
    ap.add_argument("--closure-steps", type=int, default=1000, help="fine 256^3 steps of the closure measurement")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--separate-steps", type=int, default=10)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy (y = x + 0) of MIB mebibytes of FP64, 5 times, to "
                         "calibrate the FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def copy_calibration(mib):
    import torch
    n = mib * 1024 * 1024 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    for _ in range(5):
        y = x + 0.0
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                      "launches": 5, "check": float(y[0].item())}))
    return 0


def workload_params(mod, n):
    import math
    return mod.make_params(n, n, n, lower=(0.0, 0.0, 0.0), extent=(2 * math.pi, 2 * math.pi, math.pi),
                           filtering="2/3-rule", stepper="impl-diff-rk4", nnu=3, prediss=30.0,
                           pretype="vorch", length_scale="Kolmogorov", alpha=0.1, dt_fixed=0.0,
                           limit=1.0e30)


def read_traffic(n):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    counter summary (profiles/traffic_<kernel>_<n>.txt: a line 'bytes_per_launch <B>'),
    or None."""
    import glob
    out = {}
    for path in glob.glob(os.path.join(HERE, "profiles", "traffic_*_%d.txt" % n)):
        kind = os.path.basename(path)[len("traffic_"):-len("_%d.txt" % n)]
        with open(path) as f:
            for line in f:
                if line.startswith("bytes_per_launch"):
                    out[kind] = float(line.split()[1])
    return out


ENSEMBLE_SIZES = [32, 48, 64, 96]
ENSEMBLE_COUNTS = [38, 29, 19, 10]     # 40/30/20/10 % of 96 by largest remainder


def ensemble_bench(G, steps, sep_steps):
    """96 instances of four grid sizes, random-phase isotropic ICs with k0 ~ U[2, 6]
    (seeded), the plain-workload physics.  One batched handle per size, groups on
    forked streams, one captured graph per step, z launch shapes tuned per size;
    against the same 96 instances as separate handles stepped one after another."""
    import numpy as np
    from ps3d_amd import ics
    rng = np.random.default_rng(2024)
    ntot = sum(ENSEMBLE_COUNTS)
    k0s = rng.uniform(2.0, 6.0, size=ntot)
    params = [workload_params(G, n) for n in ENSEMBLE_SIZES]
    ens = G.Ps3dEnsemble(params, ENSEMBLE_COUNTS)
    vors = []
    i = 0
    for gi, n in enumerate(ENSEMBLE_SIZES):
        for b in range(ENSEMBLE_COUNTS[gi]):
            v = ics.random_phase_isotropic(n, n, n, k0=float(k0s[i]), seed=1000 + i)
            ens.groups[gi].set_vorticity(b, v)
            vors.append((gi, v))
            i += 1
    shapes = ens.tune(steps=2)
    line_threads = ens.line_choices()
    fft_sched = ens.fft_schedules()
    ens.advance(2)                      # builds the graph
    t0 = time.perf_counter()
    ev_ms = ens.advance(steps, graph=True)
    el = time.perf_counter() - t0
    npts = sum(c * n * n * (n + 1) for n, c in zip(ENSEMBLE_SIZES, ENSEMBLE_COUNTS))
    ens.close()
    sep = []
    for gi, v in vors:
        h = G.Ps3dGpu(params[gi])
        h.set_vorticity(v)
        sep.append(h)
    for h in sep:
        h.advance_n(1)
    t0 = time.perf_counter()
    for h in sep:
        h.advance_n(sep_steps)
    el_sep = time.perf_counter() - t0
    for h in sep:
        h.close()
    b_step = B_STEP_VOLUMES * 8.0 * npts
    return {
        "instances": ntot, "sizes": ENSEMBLE_SIZES, "counts": ENSEMBLE_COUNTS, "npts_total": npts,
        "ic": "random-phase isotropic, k0 ~ U[2,6] per instance (numpy PCG64 seed 2024), IC seeds 1000+i",
        "steps": steps, "ms_per_step": 1e3 * el / steps, "event_ms_per_step": ev_ms / steps,
        "grid_point_updates_per_s": npts * steps / el,
        "z_shapes": {str(n): c for n, c in zip(ENSEMBLE_SIZES, shapes)},
        "line_threads": {str(n): list(c) for n, c in zip(ENSEMBLE_SIZES, line_threads)},
        "fft_schedule_xy": {str(n): list(c) for n, c in zip(ENSEMBLE_SIZES, fft_sched)},
        "step_roofline": {"bound": "hbm", "algo_bytes_per_step": b_step, "achieved": b_step * steps / el / 1e9,
                          "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": b_step * steps / el / PEAK_HBM},
        "separate": {"steps": sep_steps, "ms_per_step": 1e3 * el_sep / sep_steps,
                     "grid_point_updates_per_s": npts * sep_steps / el_sep},
        "speedup_vs_separate": (npts * steps / el) / (npts * sep_steps / el_sep),
    }


CLOSURE_WEIGHTS = os.path.join(HERE, "ps3d_amd", "data", "closure64.npz")


def closure_bench(G, seed, fine_steps, timing_steps):
    """The closure half of the metric on one held-out seed: the trained 64^3 net
    (2000 snapshots of twenty 256^3 runs) against the plain 256^3 run from the same
    random-phase initial state (k0 = 5) over `fine_steps` fine steps, and the plain
    64^3 run at the prediss tuned without the net.  Errors at the end against the
    fine run truncated to the modes a 64^3 step passes: the genspec spectrum below the 2/3 cutoff
    (relative L2) and the enstrophy (relative)."""
    from ps3d_amd import closure_run as CR
    net, gain, pd_c, meta = CR.load_net(CLOSURE_WEIGHTS, 64)
    pd_p = meta.get("coarse_prediss_plain_tuned", pd_c)
    cfg = [("no_closure_tuned", pd_p, None), ("closure", pd_c, gain)]
    ev = CR.evaluate(256, 64, net, [seed], fine_steps, cfg)[0]
    st = CR.step_timing(256, 64, net, steps=timing_steps)
    net.close()
    c = ev["closure"]
    return {
        "weights": os.path.relpath(CLOSURE_WEIGHTS, HERE), "training": meta, "seed": seed,
        "fine_steps": fine_steps, "t_end": ev["t_end"], "gain": gain, "coarse_prediss": pd_c,
        "fine_s": ev["fine_s"], "closure_s": c["s"], "closure_steps": c["steps"],
        "speedup_per_time": ev["fine_s"] / c["s"],
        "ms_per_step": st, "speedup_per_step": st["fine"] / st["coarse_closure"],
        "spectrum_error": {k: ev[k]["spectrum_error"] for k, _, _ in cfg},
        "enstrophy_error": {k: ev[k]["enstrophy_error"] for k, _, _ in cfg},
        "target": "closure 64^3 >= 20x faster than plain 256^3 at <= 5% spectrum error",
    }


# The other configurations of the workload, measured after the headline one:
# (label, n, initial condition, fixed dt (0: adaptive), timed steps)
CONFIG_RUNS = [("config 1: 32^3 Beltrami, fixed dt 0.01", 32, "beltrami", 0.01, 100),
               ("config 2: 128^3 Beltrami, adaptive dt", 128, "beltrami", 0.0, 100),
               ("config 3: 512^3 random-phase isotropic, adaptive dt", 512, "random", 0.0, 10),
               ("config 3: 1024^3 random-phase isotropic, adaptive dt", 1024, "random", 0.0, 3)]


def configs_bench(G, ics):
    """Steps/s, grid-point updates/s and the step roofline of every other grid size
    of the workload (one grid at a time, resident in HBM, graph-replayed steps)."""
    out = []
    for label, n, ic, dt, steps in CONFIG_RUNS:
        p = workload_params(G, n)
        p.dt_fixed = dt
        vor, coarse = initial_vorticity(G, ics, n, ic)
        d = G.Ps3dGpu(p)
        if coarse is None:
            d.set_vorticity(vor)
        else:
            G.resample(d, coarse)
            coarse.close()
        d.advance_n(2)
        d.sync()
        t0 = time.perf_counter()
        d.advance_n(steps)
        el = time.perf_counter() - t0
        d.close()
        npts = n * n * (n + 1)
        sps = steps / el
        b_step = B_STEP_VOLUMES * 8.0 * npts
        out.append({"config": label, "grid": [n, n, n], "steps": steps, "steps_per_s": sps,
                    "ms_per_step": 1e3 * el / steps, "grid_point_updates_per_s": npts * sps,
                    "step_roofline_frac": b_step * sps / PEAK_HBM})
    return out


def initial_vorticity(G, ics, n, kind):
    if kind == "beltrami":
        return ics.beltrami_multimode(n, n, n, seed=1), None
    # random-phase isotropic: band-limited (k0 = 4), exact on 128^3, zero-padded to n^3
    nc = min(n, 128)
    coarse = G.Ps3dGpu(workload_params(G, nc))
    coarse.set_vorticity(ics.random_phase_isotropic(nc, nc, nc, k0=4.0, seed=1))
    return None, coarse


def host_cpu():
    """Model name, logical CPUs usable by this process and physical cores of the host."""
    model, cores = None, set()
    phys = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    cores.add((phys, v))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": usable, "logical_cpus": os.cpu_count(),
            "physical_cores": len(cores) or None}


def cpu_baseline(vor, n, steps, threads):
    """The oracle (C FP64 restatement of ps3d's unfused step) on the host cores: the
    same workload on all usable cores, plus one core on the same workload at a
    quarter of the linear size (a bounded sample), scaled to the full grid by the
    operation count ratio N log N."""
    import math
    from oracle import oracle as O
    host = host_cpu()
    # the process's CPU share: OMP_NUM_THREADS where the host sets one, else every usable CPU
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or host["nproc"]
    ncores = threads if threads > 0 else min(share, host["nproc"])
    O.set_threads(ncores)
    o = O.Oracle(workload_params(O, n))
    o.set_vorticity(vor)
    t0 = time.perf_counter()
    o.advance_n(steps)
    el = time.perf_counter() - t0
    o.close()
    out = {"value": steps / el, "unit": "steps/s", "cores": ncores, "kind": "port",
           "sample": "%d step(s) of the same %d^3 workload through the C FP64 restatement of ps3d's "
                     "unfused step (oracle/ps3d_oracle.c), OpenMP over %d host threads" % (steps, n, ncores),
           "host": host}
    # one core: n/2 grid (one eighth of the points), 1 step
    from ps3d_amd import ics
    ns = max(32, n // 2)
    O.set_threads(1)
    o = O.Oracle(workload_params(O, ns))
    o.set_vorticity(ics.beltrami_multimode(ns, ns, ns, seed=1))
    t0 = time.perf_counter()
    o.advance_n(1)
    el1 = time.perf_counter() - t0
    o.close()
    O.set_threads(ncores)
    npts, nps = n * n * (n + 1), ns * ns * (ns + 1)
    ratio = (npts * math.log(npts)) / (nps * math.log(nps))
    out["one_core"] = {"value": 1.0 / (el1 * ratio), "unit": "steps/s", "cores": 1, "kind": "port",
                       "measured": {"grid": ns, "steps": 1, "seconds": el1},
                       "sample": "1 step of the same workload at %d^3 on one host thread, scaled to %d^3 by "
                                 "the operation-count ratio N log N = %.3f" % (ns, n, ratio)}
    return out


def main():
    a = parse()
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (the whole grid resident in one MI355X's HBM); "
              "--gpus %d is not supported" % a.gpus, file=sys.stderr)
        return 2
    from ps3d_amd import gpu as G
    from ps3d_amd import ics

    n = a.n
    vor, coarse = initial_vorticity(G, ics, n, a.ic)
    d = G.Ps3dGpu(workload_params(G, n))
    if coarse is None:
        d.set_vorticity(vor)
    else:
        G.resample(d, coarse)
        coarse.close()
        if n > 256:
            a.no_cpu_baseline = True     # the host oracle does not fit this grid in a bounded sample
        elif not a.no_cpu_baseline:
            vor = d.get(G.VOR)           # the oracle starts from the same physical vorticity
    if a.warmup:
        d.advance_n(a.warmup)
    d.sync()
    t0 = time.perf_counter()
    diag = d.advance_n(a.steps)       # synchronises at the end
    el = time.perf_counter() - t0
    ev_ms, _ = d.last_timing()
    steps_per_s = a.steps / el
    npts = n * n * (n + 1)

    # per-kernel durations (HIP events on the handle's stream), outside the timed region
    import ctypes as C
    import numpy as np
    ms = np.zeros(len(KIND_NAMES))
    launches = np.zeros(len(KIND_NAMES), dtype=np.int32)
    nbytes = np.zeros(len(KIND_NAMES))
    st = G.lib().ps3d_gpu_profile_kernels(d.h, a.profile_steps, ms.ctypes.data_as(C.c_void_p),
                                          launches.ctypes.data_as(C.c_void_p), nbytes.ctypes.data_as(C.c_void_p))
    kernels = {}
    if st == 0:
        for k, name in enumerate(KIND_NAMES):
            if launches[k]:
                avg = ms[k] / launches[k] * 1e-3
                kernels[name] = {"ms_per_step": ms[k] / a.profile_steps, "avg_launch_us": avg * 1e6,
                                 "algo_bytes": nbytes[k],
                                 "gbs": (nbytes[k] / avg / 1e9) if nbytes[k] > 0 else None}
    dom = max((k for k in kernels if kernels[k]["algo_bytes"] > 0), key=lambda k: kernels[k]["ms_per_step"],
              default=None)
    traffic = read_traffic(n)
    roofline = None
    if dom:
        achieved = kernels[dom]["gbs"] * 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                    "unit": "GB/s", "frac": achieved / PEAK_HBM,
                    "traffic": traffic.get(dom)}
    b_step = B_STEP_VOLUMES * 8.0 * npts
    out = {
        "metric": "RK steps/s at %d^3 FP64" % n,
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain FP64 %s %d^3, impl-diff-rk4, 2/3-rule, adaptive CFL dt"
                               % ("Beltrami" if a.ic == "beltrami" else "random-phase isotropic", n),
                   "grid": [n, n, n], "npts": npts,
                   "pressure_divergence": "output steps only (ps3d_gpu_derived); ps3d computes them in every "
                                          "adapt (src/stepper/advance.f90:278-282) although they never feed "
                                          "the step, and the algorithmic byte count leaves them out",
                   "ic": ("4 Beltrami modes, seeded phases, 1e-3 noise (seed 1)" if a.ic == "beltrami" else
                          "random-phase isotropic E(k)~k^4 exp(-2(k/4)^2), seed 1, made on 128^3, zero-padded")},
        "grid_point_updates_per_s": npts * steps_per_s,
        "event_ms_per_step": ev_ms / a.steps,
        "step_roofline": {"bound": "hbm", "algo_bytes_per_step": b_step,
                          "achieved": b_step * steps_per_s / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                          "frac": b_step * steps_per_s / PEAK_HBM},
        "roofline": roofline,
        "kernels": kernels,
        "t_end": float(diag[-1][0] + diag[-1][1]) if len(diag) else None,
    }
    d.close()
    if not a.no_ensemble:
        try:
            out["ensemble"] = ensemble_bench(G, a.ensemble_steps, a.separate_steps)
        except Exception as exc:       # reported, never silently dropped
            out["ensemble"] = {"error": repr(exc)}
    if not a.no_configs:
        try:
            out["configs"] = configs_bench(G, ics)
        except Exception as exc:       # reported, never silently dropped
            out["configs"] = {"error": repr(exc)}
    if not a.no_closure:
        try:
            out["closure"] = closure_bench(G, a.closure_seed, a.closure_steps, 20)
        except Exception as exc:       # reported, never silently dropped
            out["closure"] = {"error": repr(exc)}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(vor, n, a.cpu_steps, a.cpu_threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
