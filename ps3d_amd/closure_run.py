"""Closure experiment: train the 64^3 subgrid net on filtered 256^3 runs and
compare the coarse run with and without it against the filtered fine run.

  python -m ps3d_amd.closure_run [--fine 256] [--coarse 64] [--runs 4] [--snaps 25]

Fine runs: random-phase isotropic initial vorticity, E(k) ~ k^4 exp(-2 (k/k0)^2)
with k0 ~ U[3, 8] per run (seeded), made on 128^3 and zero-padded to the fine
grid; the plain FP64 step (impl-diff-rk4, 2/3-rule, nnu 3, prediss 30, vorch).
Snapshots every `every` steps after `spinup` steps; the training target keeps the
modes the coarse step's 2/3-rule mask passes.  On validation seeds the closure's
output gain and the coarse grid's prediss are chosen (and, for the comparison, the
best prediss without closure); held-out seeds are then run to the same physical
time on the fine grid (truncated to the coarse one for its spectrum) and on the
coarse grid from the truncated initial state: without closure at prediss 30,
without closure at its best prediss, and with the closure.  The genspec energy
spectra are compared below the coarse 2/3 cutoff.  Progress goes to stderr; the
result is one JSON object on stdout.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import closure as CL
from . import gpu as G
from . import ics

PI = math.pi


def progress(msg):
    """One progress line on stderr (the result is the JSON line on stdout)."""
    sys.stderr.write("[closure_run] %s\n" % msg)
    sys.stderr.flush()


def params(n, limit=1.0e30, prediss=30.0):
    return G.make_params(n, n, n, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI), filtering="2/3-rule",
                         stepper="impl-diff-rk4", nnu=3, prediss=prediss, pretype="vorch", length_scale="Kolmogorov",
                         alpha=0.1, dt_fixed=0.0, limit=limit)


def fine_run(nf, k0, seed, limit=1.0e30):
    h = G.Ps3dGpu(params(nf, limit))
    src = G.Ps3dGpu(params(128))
    src.set_vorticity(ics.random_phase_isotropic(128, 128, 128, k0=k0, seed=seed))
    G.resample(h, src)
    src.close()
    return h


def dataset(nf, nc, runs, snaps, every, spinup, seed0, diffusion=False, dealias=False):
    rng = np.random.default_rng(seed0)
    coarse, tco = G.Ps3dGpu(params(nc)), G.Ps3dGpu(params(nc))
    X, T = [], []
    for r in range(runs):
        f = fine_run(nf, float(rng.uniform(3.0, 8.0)), seed0 + r)
        f.advance_n(spinup)
        for s in range(snaps):
            x, t = CL.sample_from_fine(f, coarse, tco, diffusion, dealias)
            X.append(x)
            T.append(t)
            f.advance_n(every)
        f.close()
        progress("fine run %d/%d: %d snapshots" % (r + 1, runs, len(X)))
    coarse.close()
    tco.close()
    return X, T


def scales(X, T):
    v = math.sqrt(np.mean([np.mean(x[:, :3] ** 2) for x in X]))
    w = math.sqrt(np.mean([np.mean(x[:, 3:] ** 2) for x in X]))
    t = math.sqrt(np.mean([np.mean(tt ** 2) for tt in T]))
    return np.array([1 / v] * 3 + [1 / w] * 3, dtype=np.float32), np.array([t] * 3, dtype=np.float32)


def train(nc, X, T, epochs, lr, seed):
    net = CL.ClosureNet(nc, nc, nc)
    net.in_scale, net.out_scale = scales(X, T)
    net.set_weights(CL.init_weights(seed))
    net.reserve(len(X))
    for i, (x, t) in enumerate(zip(X, T)):
        net.set_sample(i, x * net.in_scale, t / net.out_scale)
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for e in range(epochs):
        lr_e = lr * 0.5 * (1.0 + math.cos(math.pi * e / max(1, epochs)))
        hist.append(net.train_epoch(rng.permutation(len(X)), lr_e))
        progress("epoch %d/%d: loss %.4f" % (e + 1, epochs, hist[-1]))
    return net, hist, time.perf_counter() - t0


def spectrum_of(h, t):
    h.vor2vel()
    _, row = h.adapt(t)
    return CL.energy_spectrum(h, row[G.DIAG_NAMES.index("ke")])


def evaluate(nf, nc, net, seeds, t_end, configs):
    """Coarse runs, one per config (tag, coarse prediss, closure output gain or None
    for no closure), against the truncated fine run at t_end."""
    out = []
    kcut = int(nc / 3)   # shells below the coarse 2/3 cutoff
    base_scale = np.array(net.out_scale, dtype=np.float32)
    for s in seeds:
        k0 = 5.0
        f = fine_run(nf, k0, s, limit=t_end)
        runs = {}
        for tag, pd, _ in configs:
            c = G.Ps3dGpu(params(nc, limit=t_end, prediss=pd))
            G.resample(c, f)                    # truncated initial state
            runs[tag] = c
        nfs = 0
        t0 = time.perf_counter()
        while f.t < t_end - 1e-12:
            f.advance_n(10)
            nfs += 10
        fine_s = time.perf_counter() - t0
        cref = G.Ps3dGpu(params(nc, limit=t_end))
        G.resample(cref, f)
        _, ef = spectrum_of(cref, t_end)
        res = {"seed": s, "fine_steps": nfs, "fine_s": fine_s}
        for (tag, c), (_, _, gain) in zip(runs.items(), configs):
            if gain is not None:
                net.out_scale = base_scale * np.float32(gain)
                net.attach(c)
            n = 0
            t0 = time.perf_counter()
            # a run whose step collapses (the closure destabilising it) is stopped at
            # 20 times the fine run's step count and reported as failed
            ok = True
            while c.t < t_end - 1e-12:
                diag = c.advance_n(10)
                n += 10
                if not np.all(np.isfinite(diag)) or n > 20 * nfs:
                    ok = False
                    break
            el = time.perf_counter() - t0
            if ok:
                _, ec = spectrum_of(c, t_end)
                err = CL.spectrum_error(ec, ef, kcut)
            else:
                err = float("inf")
            res[tag] = {"steps": n, "s": el, "spectrum_error": err, "completed": ok}
            progress("seed %d %s: %d steps, error %.4f" % (s, tag, n, err))
            if gain is not None:
                CL.ClosureNet.detach(c)
                net.out_scale = base_scale
            c.close()
        out.append(res)
        progress("seed %d: %s" % (s, {k: round(v["spectrum_error"], 4) for k, v in res.items() if isinstance(v, dict)}))
        f.close()
        cref.close()
    return out


def step_timing(nf, nc, net, steps=20):
    """ms per step: plain fine grid, plain coarse grid, coarse grid with the closure."""
    res = {}
    for tag, n, use in (("fine", nf, False), ("coarse", nc, False), ("coarse_closure", nc, True)):
        h = fine_run(nf, 5.0, 77) if n == nf else None
        if h is None:
            h = G.Ps3dGpu(params(n))
            src = G.Ps3dGpu(params(128 if n >= 128 else n))
            src.set_vorticity(ics.random_phase_isotropic(src.nx, src.ny, src.nz, k0=5.0, seed=77))
            G.resample(h, src)
            src.close()
        if use:
            net.attach(h)
        h.advance_n(3)
        t0 = time.perf_counter()
        h.advance_n(steps)
        res[tag] = 1e3 * (time.perf_counter() - t0) / steps
        if use:
            CL.ClosureNet.detach(h)
        h.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--fine", type=int, default=256)
    ap.add_argument("--coarse", type=int, default=64)
    ap.add_argument("--runs", type=int, default=4)
    ap.add_argument("--snaps", type=int, default=25)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--spinup", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-seeds", type=int, default=2)
    ap.add_argument("--t-end", type=float, default=0.5)
    ap.add_argument("--gains", default="0.25,0.5,1.0",
                    help="output gains tried on the validation seeds (comma separated); the best is evaluated")
    ap.add_argument("--val-seeds", type=int, default=2)
    ap.add_argument("--coarse-prediss", default="30,7.5,3.75",
                    help="coarse-grid hyperdiffusion prediss values tried on the validation seeds "
                         "(the closure may take over part of the coarse grid's dissipation)")
    ap.add_argument("--no-dealias-target", dest="dealias_target", action="store_false",
                    help="train on the whole truncated target, including the modes the coarse step's "
                         "2/3-rule mask removes (by default those are dropped)")
    ap.add_argument("--target", choices=["transfer", "transfer+diffusion"], default="transfer",
                    help="closure target: the subgrid nonlinear transfer, or also the difference of the "
                         "two grids' hyperdiffusion")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    X, T = dataset(a.fine, a.coarse, a.runs, a.snaps, a.every, a.spinup, 500, a.target == "transfer+diffusion",
                   a.dealias_target)
    t_data = time.perf_counter() - t0
    net, hist, t_train = train(a.coarse, X, T, a.epochs, a.lr, 0)
    fwd_ms, train_ms = net.timing(20)
    npts = a.coarse * a.coarse * (a.coarse + 1)
    gflop = CL.GFLOP_FWD_PER_POINT * npts
    gains = [float(g) for g in a.gains.split(",")]
    pds = [float(v) for v in a.coarse_prediss.split(",")]
    val = None
    gain, pd_c, pd_p = gains[0], pds[0], 30.0
    if len(gains) > 1 or len(pds) > 1:
        # closure gain and coarse prediss (and the best prediss without closure) chosen
        # on validation seeds, never on the evaluation seeds
        cfg = [("plain_p%g" % p, p, None) for p in pds] + \
              [("closure_p%g_g%g" % (p, g), p, g) for p in pds for g in gains]
        val = evaluate(a.fine, a.coarse, net, [8001 + i for i in range(a.val_seeds)], a.t_end, cfg)
        err = {tag: np.mean([v[tag]["spectrum_error"] for v in val]) for tag, _, _ in cfg}
        _, pd_c, gain = min((c for c in cfg if c[2] is not None), key=lambda c: err[c[0]])
        _, pd_p, _ = min((c for c in cfg if c[2] is None), key=lambda c: err[c[0]])
    cfg = [("no_closure", 30.0, None), ("no_closure_tuned", pd_p, None), ("closure", pd_c, gain)]
    ev = evaluate(a.fine, a.coarse, net, [9001 + i for i in range(a.eval_seeds)], a.t_end, cfg)
    st = step_timing(a.fine, a.coarse, net)
    out = {
        "fine": a.fine, "coarse": a.coarse, "target": a.target, "dealias_target": a.dealias_target,
        "snapshots": len(X), "data_s": t_data,
        "epochs": a.epochs, "train_s": t_train, "loss_first": hist[0], "loss_last": hist[-1],
        "forward_ms": fwd_ms, "forward_gflop": gflop, "forward_tflops": gflop / fwd_ms,
        "forward_mfma_frac": gflop / fwd_ms * 1e12 / 2.5e15,
        "train_step_ms": train_ms,
        "step_ms": st, "speedup_per_step": st["fine"] / st["coarse_closure"],
        "gain": gain, "coarse_prediss_closure": pd_c, "coarse_prediss_plain_tuned": pd_p, "validation": val,
        "eval": ev,
    }
    tf = [e["fine_s"] for e in ev]
    tc = [e["closure"]["s"] for e in ev]
    out["speedup_per_time"] = float(np.mean(tf) / np.mean(tc))
    out["spectrum_error_closure"] = float(np.mean([e["closure"]["spectrum_error"] for e in ev]))
    out["spectrum_error_no_closure"] = float(np.mean([e["no_closure"]["spectrum_error"] for e in ev]))
    out["spectrum_error_no_closure_tuned"] = float(np.mean([e["no_closure_tuned"]["spectrum_error"] for e in ev]))
    net.close()
    print(json.dumps(out))
    return out


if __name__ == "__main__":
    main()
