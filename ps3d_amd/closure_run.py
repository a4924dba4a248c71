"""Closure experiment: train the 64^3 subgrid net on filtered 256^3 runs and
compare the coarse run with and without it against the filtered fine run.

  python -m ps3d_amd.closure_run [--fine 256] [--coarse 64] [--runs 20] [--snaps 100]
                                 [--eval-steps 1000] [--eval-seeds 20] [--save FILE]
  python -m ps3d_amd.closure_run --load ps3d_amd/data/closure64.npz [--eval-seeds 20]

Fine runs: random-phase isotropic initial vorticity, E(k) ~ k^4 exp(-2 (k/k0)^2)
with k0 ~ U[3, 8] per run (seeded), made on 128^3 and zero-padded to the fine
grid; the plain FP64 step (impl-diff-rk4, 2/3-rule, nnu 3, prediss 30, vorch).
Snapshots every `every` steps after `spinup` steps, so that the training states
span the evaluation horizon; the training target keeps the modes the coarse
step's 2/3-rule mask passes.  On validation seeds the closure's output gain and the
coarse grid's prediss are chosen (and, for the comparison, the best prediss
without closure); held-out seeds are then run for `eval-steps` fine steps and the
coarse grids to the same physical time from the truncated initial state: without
closure at prediss 30, without closure at its best prediss, and with the closure.
Errors at the end against the fine run truncated to the modes the coarse step
passes (its 2/3-rule mask): the genspec energy spectrum below the coarse 2/3 cutoff
(relative L2) and the enstrophy (relative).  Progress goes to stderr; the result is
one JSON object on stdout; --save writes the trained net (weights, scales, gain,
coarse prediss) for later runs.
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
EN = G.DIAG_NAMES.index("en")
KE = G.DIAG_NAMES.index("ke")


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
    src = G.Ps3dGpu(params(min(nf, 128)))
    src.set_vorticity(ics.random_phase_isotropic(src.nx, src.ny, src.nz, k0=k0, seed=seed))
    G.resample(h, src)
    src.close()
    return h


def dataset(net, nf, nc, runs, snaps, every, spinup, seed0, diffusion=False, dealias=False):
    """Snapshots go straight to the net's device dataset; the input / output scales
    come from the first run's snapshots."""
    rng = np.random.default_rng(seed0)
    coarse, tco = G.Ps3dGpu(params(nc)), G.Ps3dGpu(params(nc))
    net.reserve(runs * snaps)
    first, i = [], 0
    for r in range(runs):
        f = fine_run(nf, float(rng.uniform(3.0, 8.0)), seed0 + r)
        f.advance_n(spinup)
        for s in range(snaps):
            x, t = CL.sample_from_fine(f, coarse, tco, diffusion, dealias)
            if r == 0:
                first.append((x, t))
            else:
                net.set_sample(i, x * net.in_scale, t / net.out_scale)
            i += 1
            f.advance_n(every)
        if r == 0:
            net.in_scale, net.out_scale = scales([x for x, _ in first], [t for _, t in first])
            for k, (x, t) in enumerate(first):
                net.set_sample(k, x * net.in_scale, t / net.out_scale)
            first = []
        f.close()
        progress("fine run %d/%d: %d snapshots, t = %.3f" % (r + 1, runs, i, f.t))
    coarse.close()
    tco.close()
    return i


def scales(X, T):
    v = math.sqrt(np.mean([np.mean(x[:, :3] ** 2) for x in X]))
    w = math.sqrt(np.mean([np.mean(x[:, 3:] ** 2) for x in X]))
    t = math.sqrt(np.mean([np.mean(tt ** 2) for tt in T]))
    return np.array([1 / v] * 3 + [1 / w] * 3, dtype=np.float32), np.array([t] * 3, dtype=np.float32)


def train(net, nsamples, epochs, lr, seed):
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for e in range(epochs):
        lr_e = lr * 0.5 * (1.0 + math.cos(math.pi * e / max(1, epochs)))
        hist.append(net.train_epoch(rng.permutation(nsamples), lr_e))
        progress("epoch %d/%d: loss %.4f" % (e + 1, epochs, hist[-1]))
    return hist, time.perf_counter() - t0


def spectrum_and_enstrophy(h, t):
    h.vor2vel()
    _, row = h.adapt(t)
    return CL.energy_spectrum(h)[1], row[EN]


def evaluate(nf, nc, net, seeds, nsteps, configs):
    """Per seed: the fine run for nsteps steps (to t_end), and the coarse runs (tag,
    coarse prediss, closure output gain or None for no closure) from the truncated
    initial state to t_end (their last step clipped to it by time%limit); spectrum
    and enstrophy errors against the fine state truncated to the modes the coarse
    step passes."""
    out = []
    kcut = int(nc / 3)   # shells below the coarse 2/3 cutoff
    base_scale = np.array(net.out_scale, dtype=np.float32)
    for s in seeds:
        f = fine_run(nf, 5.0, s)
        c0 = G.Ps3dGpu(params(nc))
        G.resample(c0, f)                       # truncated initial state
        t0 = time.perf_counter()
        f.advance_n(nsteps)
        fine_s = time.perf_counter() - t0
        t_end = f.t
        cref = G.Ps3dGpu(params(nc))
        G.resample(cref, f)
        # the reference keeps only the modes a coarse step can hold (its 2/3-rule
        # mask): the enstrophy of the modes between that cutoff and the coarse grid's
        # Nyquist limit is out of every coarse run's reach
        cref.set(G.SVOR, cref.get(G.SVOR) * CL.dealias_mask(cref)[None])
        spec_f, en_f = spectrum_and_enstrophy(cref, t_end)
        f.close()
        cref.close()
        res = {"seed": s, "fine_steps": nsteps, "fine_s": fine_s, "t_end": t_end}
        for tag, pd, gain in configs:
            c = G.Ps3dGpu(params(nc, limit=t_end, prediss=pd))
            G.resample(c, c0)
            if gain is not None:
                net.out_scale = base_scale * np.float32(gain)
                net.attach(c)
            n, ok = 0, True
            t0 = time.perf_counter()
            # a run whose step collapses (the closure destabilising it) is stopped at
            # 20 times the fine run's step count and reported as failed
            while c.t < t_end - 1e-12:
                diag = c.advance_n(10)
                n += 10
                if not np.all(np.isfinite(diag)) or n > 20 * nsteps:
                    ok = False
                    break
            el = time.perf_counter() - t0
            if ok:
                spec_c, en_c = spectrum_and_enstrophy(c, t_end)
                se, ee = CL.spectrum_error(spec_c, spec_f, kcut), abs(en_c - en_f) / en_f
            else:
                se = ee = float("inf")
            res[tag] = {"steps": n, "s": el, "completed": ok, "spectrum_error": se, "enstrophy_error": ee}
            progress("seed %d %s: %d steps, spectrum %.4f, enstrophy %.4f" % (s, tag, n, se, ee))
            if gain is not None:
                CL.ClosureNet.detach(c)
                net.out_scale = base_scale
            c.close()
        c0.close()
        out.append(res)
    return out


def step_timing(nf, nc, net, steps=20):
    """ms per step: plain fine grid, plain coarse grid, coarse grid with the closure."""
    res = {}
    for tag, n, use in (("fine", nf, False), ("coarse", nc, False), ("coarse_closure", nc, True)):
        h = fine_run(n, 5.0, 77)
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


def save_net(path, net, gain, pd_c, meta):
    np.savez(path, weights=net.get_weights(), in_scale=np.asarray(net.in_scale, dtype=np.float32),
             out_scale=np.asarray(net.out_scale, dtype=np.float32), gain=np.float64(gain),
             coarse_prediss=np.float64(pd_c), meta=np.array(json.dumps(meta)))


def load_net(path, n):
    """(net on an n^3 grid with the saved weights and scales, gain, coarse prediss,
    meta dict)."""
    z = np.load(path, allow_pickle=False)
    net = CL.ClosureNet(n, n, n)
    net.set_weights(z["weights"])
    net.in_scale = z["in_scale"].astype(np.float32)
    net.out_scale = z["out_scale"].astype(np.float32)
    return net, float(z["gain"]), float(z["coarse_prediss"]), json.loads(str(z["meta"]))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--fine", type=int, default=256)
    ap.add_argument("--coarse", type=int, default=64)
    ap.add_argument("--runs", type=int, default=20)
    ap.add_argument("--snaps", type=int, default=100)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--spinup", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-seeds", type=int, default=20)
    ap.add_argument("--eval-steps", type=int, default=1000, help="fine-grid steps of each held-out run")
    ap.add_argument("--gains", default="0.25,0.5,1.0",
                    help="output gains tried on the validation seeds (comma separated); the best is evaluated")
    ap.add_argument("--val-seeds", type=int, default=2)
    ap.add_argument("--val-steps", type=int, default=300, help="fine-grid steps of each validation run")
    ap.add_argument("--coarse-prediss", default="30,7.5,3.75",
                    help="coarse-grid hyperdiffusion prediss values tried on the validation seeds "
                         "(the closure may take over part of the coarse grid's dissipation)")
    ap.add_argument("--no-dealias-target", dest="dealias_target", action="store_false",
                    help="train on the whole truncated target, including the modes the coarse step's "
                         "2/3-rule mask removes (by default those are dropped)")
    ap.add_argument("--target", choices=["transfer", "transfer+diffusion"], default="transfer",
                    help="closure target: the subgrid nonlinear transfer, or also the difference of the "
                         "two grids' hyperdiffusion")
    ap.add_argument("--save", default=None, help="write the trained net here (.npz)")
    ap.add_argument("--load", default=None,
                    help="evaluate a saved net (.npz, with its gain and prediss choices) instead of training one")
    a = ap.parse_args(argv)
    if a.load:
        return evaluate_saved(a)
    net = CL.ClosureNet(a.coarse, a.coarse, a.coarse)
    net.set_weights(CL.init_weights(0))
    t0 = time.perf_counter()
    ns = dataset(net, a.fine, a.coarse, a.runs, a.snaps, a.every, a.spinup, 500, a.target == "transfer+diffusion",
                 a.dealias_target)
    t_data = time.perf_counter() - t0
    hist, t_train = train(net, ns, a.epochs, a.lr, 0)
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
        val = evaluate(a.fine, a.coarse, net, [8001 + i for i in range(a.val_seeds)], a.val_steps, cfg)
        err = {tag: np.mean([v[tag]["spectrum_error"] for v in val]) for tag, _, _ in cfg}
        _, pd_c, gain = min((c for c in cfg if c[2] is not None), key=lambda c: err[c[0]])
        _, pd_p, _ = min((c for c in cfg if c[2] is None), key=lambda c: err[c[0]])
    meta = {"fine": a.fine, "coarse": a.coarse, "snapshots": ns, "runs": a.runs, "every": a.every,
            "epochs": a.epochs, "target": a.target, "dealias_target": a.dealias_target,
            "coarse_prediss_plain_tuned": pd_p,
            "training_seeds": "fine runs 500..%d, k0 ~ U[3, 8] (numpy PCG64 seed 500)" % (499 + a.runs),
            "validation_seeds": [8001 + i for i in range(a.val_seeds)]}
    if a.save:
        save_net(a.save, net, gain, pd_c, meta)
        progress("saved %s" % a.save)
    cfg = [("no_closure", 30.0, None), ("no_closure_tuned", pd_p, None), ("closure", pd_c, gain)]
    ev = evaluate(a.fine, a.coarse, net, [9001 + i for i in range(a.eval_seeds)], a.eval_steps, cfg)
    st = step_timing(a.fine, a.coarse, net)
    out = dict(meta)
    out.update({
        "data_s": t_data, "train_s": t_train, "loss_first": hist[0], "loss_last": hist[-1],
        "forward_ms": fwd_ms, "forward_gflop": gflop, "forward_tflops": gflop / fwd_ms,
        "forward_mfma_frac": gflop / fwd_ms * 1e12 / 2.5e15,
        "train_step_ms": train_ms,
        "step_ms": st, "speedup_per_step": st["fine"] / st["coarse_closure"],
        "gain": gain, "coarse_prediss_closure": pd_c, "coarse_prediss_plain_tuned": pd_p, "validation": val,
        "eval_steps": a.eval_steps, "eval": ev,
    })
    summarise(out, ev)
    net.close()
    print(json.dumps(out))
    return out


def summarise(out, ev):
    tf = [e["fine_s"] for e in ev]
    tc = [e["closure"]["s"] for e in ev]
    out["speedup_per_time"] = float(np.mean(tf) / np.mean(tc))
    for tag in ("closure", "no_closure", "no_closure_tuned"):
        for q in ("spectrum_error", "enstrophy_error"):
            out["%s_%s" % (q, tag)] = float(np.mean([e[tag][q] for e in ev]))
    out["closure_better_spectrum"] = int(sum(e["closure"]["spectrum_error"] < e["no_closure_tuned"]["spectrum_error"]
                                             for e in ev))
    out["closure_better_enstrophy"] = int(sum(e["closure"]["enstrophy_error"] < e["no_closure_tuned"]["enstrophy_error"]
                                              for e in ev))


def evaluate_saved(a):
    """Held-out evaluation and step timing of a saved net, with the gain and coarse
    prediss stored beside its weights (chosen on validation seeds when it was trained)."""
    net, gain, pd_c, meta = load_net(a.load, a.coarse)
    pd_p = meta.get("coarse_prediss_plain_tuned", pd_c)
    fwd_ms, train_ms = net.timing(20)
    npts = a.coarse * a.coarse * (a.coarse + 1)
    gflop = CL.GFLOP_FWD_PER_POINT * npts
    cfg = [("no_closure", 30.0, None), ("no_closure_tuned", pd_p, None), ("closure", pd_c, gain)]
    ev = evaluate(a.fine, a.coarse, net, [9001 + i for i in range(a.eval_seeds)], a.eval_steps, cfg)
    st = step_timing(a.fine, a.coarse, net)
    out = dict(meta)
    out.update({
        "loaded": a.load, "forward_ms": fwd_ms, "forward_gflop": gflop, "forward_tflops": gflop / fwd_ms,
        "forward_mfma_frac": gflop / fwd_ms * 1e12 / 2.5e15, "train_step_ms": train_ms,
        "step_ms": st, "speedup_per_step": st["fine"] / st["coarse_closure"],
        "gain": gain, "coarse_prediss_closure": pd_c, "coarse_prediss_plain_tuned": pd_p,
        "eval_steps": a.eval_steps, "eval": ev,
    })
    summarise(out, ev)
    net.close()
    print(json.dumps(out))
    return out


if __name__ == "__main__":
    main()
