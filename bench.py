#!/usr/bin/env python3
"""Throughput of the JustPIC particle hot path on one MI355X.

Workload (the 2D temperature-advection case): 512^2 cells on [0,1]^2, README
stream function at amplitude 250, nxcell 24, max_xcell 48, min_xcell 12, one
particle field T = exp(-((x-0.5)^2 + (y-0.75)^2)/0.01) set by grid2particle!,
and per step: advection! (RK2), move_particles!, inject_particles!,
particle2grid! to the 513^2 nodal T.  A step is one pass of that sequence over
all particles; inputs are resident in HBM before timing starts.

The reported value is the FP32-split form (``--form split``); the FP64
drop-in runs the same workload beside it (``compare``), and a separate
advection + move run from one initial state measures the split form's
position error against FP64 and its cell-assignment disagreements.

Prints one JSON line (metric particle-steps/s).  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "justpic.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s (advect+move+p2g) FP32-split vs FP64 GPU, % HBM peak, 1 GPU"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes(S, P, D, M, F, Nn, Vb, injected=0, w=8):
    """Byte model per step: each kernel that touches an array counts it once.
    m = 1 B mask/slot, w = coordinate bytes (8 FP64, 4 FP32-split offset),
    wf = 8 B field; Vb = FP64 velocity bytes (the input in both forms)."""
    m, wf = 1, 8
    adv = S * m + 2 * P * D * w + Vb
    move = S * m + P * D * w + M * (2 * D * w + 3 * F * wf + 2 * m)
    inj = S * m + P * D * w + injected * (D * w + F * wf + m)
    p2g = S * m + P * (D * w + F * wf) + 8 * F * Nn
    return {"advection": adv, "move": move, "inject": inj, "p2g": p2g,
            "step": adv + move + inj + p2g}


# The two single-grid workloads: 2D temperature advection (RK2, with
# injection, Gaussian T from grid2particle!) and 3D RK4 advection of one
# uniform-random particle field (no injection).
WORKLOADS = {
    2: {"nxcell": 24, "max_xcell": 48, "min_xcell": 12, "method": "rk2", "inject": True},
    3: {"nxcell": 27, "max_xcell": 54, "min_xcell": 14, "method": "rk4", "inject": False},
}


def integrator(jp, spec):
    return jp.RungeKutta2() if spec["method"] == "rk2" else jp.RungeKutta4()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, spec, n_steps, seed):
    """The oracle on all host threads (its OpenMP build: the cells of one
    colour phase of move and inject, and the independent advection and
    particle2grid loops, in parallel, as the reference's threaded CPU backend
    runs them) on a bounded sample of the same workload: n_steps full steps at
    the same size.  Returns (particle-steps/s, seconds, threads)."""
    from oracle import oracle as O
    from justpic_hip.configs import gaussian_T
    O.set_threaded(True)
    xv, xi, V, dt = cfg["xv"], cfg["xi_vel"], cfg["V"], cfg["dt"]
    D = len(xv)
    nq = 2 ** D
    hp = O.HostParticles(xv, spec["max_xcell"], xi)
    O.init_particles(hp, spec["nxcell"] // nq, seed)
    if D == 2:
        pT = O.grid2particle(hp, gaussian_T(xv), np.full(hp.S, np.nan))
    else:
        pT = np.random.default_rng(seed).random(hp.S)
    method = O.JP_RK2 if spec["method"] == "rk2" else O.JP_RK4
    psteps = 0
    t0 = time.perf_counter()
    for it in range(n_steps):
        psteps += int(hp.index.sum())
        O.advection(hp, V, method, 0.5, dt)
        O.move_particles(hp, [pT])
        if spec["inject"]:
            O.inject_particles(hp, [pT], spec["min_xcell"], seed, it)
        O.particle2grid(hp, pT)
    el = time.perf_counter() - t0
    threads = O.threads()
    O.set_threaded(False)
    return psteps / el, el, threads


CALLS = ("advection", "move", "inject", "p2g")
# the kernels each call launches
KERNELS = {"advection": "one advection kernel (3D FP64: plus the flagged slot-path pass)",
           "move": "classify, colour phases (2D: one tiled launch; 3D: 3^D launches), "
                   "far-jump check, scatter",
           "inject": "scan, list join, one ordered launch of the four colours (2D; 3D: 2^D "
                     "colour launches)",
           "p2g": "per-cell partial sums and per-node gather (split) or occupancy words "
                  "and per-node gather (FP64)"}


def read_traffic(path):
    """Measured HBM bytes per call (all of the call's launches) from a
    committed counter summary (a line 'traffic_bytes_per_launch: N'), or None."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def run_form(jp, ctx, cfg, form, args, spec):
    """Warm up, then time args.steps full steps of one position form between
    two hipEvents on the context stream; then up to 50 more steps with every
    call bracketed by its own events, for the per-call times."""
    from justpic_hip.configs import gaussian_T
    xv, xi, V, dt = cfg["xv"], cfg["xi_vel"], cfg["V"], cfg["dt"]
    D = len(xv)
    p = jp.init_particles(ctx, spec["nxcell"], spec["max_xcell"], spec["min_xcell"], *xi,
                          seed=args.seed, form="split" if form == "split" else "f64")
    Vd = tuple(jp.DeviceArray.from_host(ctx, v) for v in V)
    pT = jp.CellArray(ctx, p.ni, p.max_xcell, fill=float("nan"))
    if D == 2:
        T = jp.DeviceArray.from_host(ctx, gaussian_T(xv))
        jp.grid2particle_(pT, T, p)
    else:
        T = jp.nodal_field(p)
        pT.copy_from(np.random.default_rng(args.seed).random(p.C * p.max_xcell))
    rk = integrator(jp, spec)
    ctx.sync()
    calls = (lambda: jp.advection_(p, rk, Vd, dt),
             lambda: jp.move_particles_(p, (pT,)),
             (lambda: jp.inject_particles_(p, (pT,))) if spec["inject"] else None,
             lambda: jp.particle2grid_(T, pT, p))

    def step(evs=None):
        for i, f in enumerate(calls):
            if evs is not None:
                ctx.record(evs[2 * i])
            if f is not None:
                f()
            if evs is not None:
                ctx.record(evs[2 * i + 1])

    log(form, "warmup", args.warmup)
    for _ in range(args.warmup):
        step()
    ctx.sync()
    P0 = p.active()
    e0, e1 = ctx.event(), ctx.event()
    log(form, "timing", args.steps, "steps")
    ctx.sync()
    t0 = time.perf_counter()
    ctx.record(e0)
    for s in range(args.steps):
        step()
    ctx.record(e1)
    ctx.sync()
    wall = time.perf_counter() - t0
    gpu_ms = ctx.elapsed_ms(e0, e1)
    P1 = p.active()
    # per-call breakdown: the same number of further steps, each call bracketed
    # by its own events (kept out of the timed loop above, whose only markers
    # are its two ends)
    nb = min(args.steps, 50)
    evs = [[ctx.event() for _ in range(2 * len(calls))] for _ in range(nb)]
    for s in range(nb):
        step(evs[s])
    ctx.sync()
    call_ms = {name: float(np.mean([ctx.elapsed_ms(ev[2 * i], ev[2 * i + 1]) for ev in evs]))
               for i, name in enumerate(CALLS)}
    ms_per_step = gpu_ms / args.steps
    P = 0.5 * (P0 + P1)
    value = P * args.steps / (gpu_ms * 1e-3)
    log(f"{form}: gpu {gpu_ms:.2f} ms (host wall {wall * 1e3:.2f} ms), {ms_per_step:.4f} ms/step,"
        f" P {P0}->{P1}, " + ", ".join(f"{k} {v * 1e3:.1f} us" for k, v in call_ms.items()))

    # one calibration step with outcome counters (not timed) for the byte model
    jp.advection_(p, rk, Vd, dt)
    st_m = jp.move_particles_(p, (pT,), stats=True)
    st_i = jp.inject_particles_(p, (pT,), stats=True) if spec["inject"] else {"injected": 0}
    ctx.sync()
    S = p.C * p.max_xcell
    Nn = int(np.prod([n + 1 for n in p.ni]))
    Vb = sum(v.nbytes for v in Vd)
    B = algorithmic_bytes(S, P, D, st_m["moved"] + st_m["deleted"], 1, Nn, Vb,
                          st_i["injected"], w=4 if form == "split" else 8)
    if not spec["inject"]:
        B["step"] -= B["inject"]
        B["inject"] = 0
    calls_out = {k: {"us": call_ms[k] * 1e3, "bytes": B[k],
                     "GBs": B[k] / (call_ms[k] * 1e-3) / 1e9} for k in CALLS
                 if B[k] > 0}
    return {"value": value, "ms_per_step": ms_per_step, "P": P, "bytes": B,
            "calls": calls_out, "moved": st_m["moved"], "injected": st_i["injected"],
            "hbm_frac_step": B["step"] / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS}


def split_error(jp, ctx, cfg, args, spec):
    """Advection + move for args.check_steps steps from one initial state in
    both forms, particles tracked by an id field: max position error of the
    FP32-split form against the FP64 form (bit-identical to the FP64 oracle in
    the parity tests), and the number of cell-assignment disagreements."""
    xi, V, dt = cfg["xi_vel"], cfg["V"], cfg["dt"]
    p = jp.init_particles(ctx, spec["nxcell"], spec["max_xcell"], spec["min_xcell"], *xi,
                          seed=args.seed)
    s = jp.to_split(p)
    Vd = tuple(jp.DeviceArray.from_host(ctx, v) for v in V)
    idx0 = p.index.data.flat()
    ids0 = np.where(idx0 == 1, np.arange(idx0.size, dtype=np.float64), np.nan)
    id_p = jp.CellArray(ctx, p.ni, p.max_xcell)
    id_p.copy_from(ids0)
    id_s = jp.CellArray(ctx, p.ni, p.max_xcell)
    id_s.copy_from(ids0)
    rk = integrator(jp, spec)
    for _ in range(args.check_steps):
        jp.advection_(p, rk, Vd, dt)
        jp.move_particles_(p, (id_p,))
        jp.advection_(s, rk, Vd, dt)
        jp.move_particles_(s, (id_s,))
    ctx.sync()
    C = p.C

    def table(coords, index, ids):
        e = np.nonzero(index == 1)[0]
        iv = ids[e].astype(np.int64)
        o = np.argsort(iv)
        return iv[o], (e % C)[o], np.stack([c[e] for c in coords], 1)[o]

    ca, ia = p.host_state()
    cb, ib = jp.to_f64(s).host_state()
    ka, cella, pa = table(ca, ia, id_p.data.flat())
    kb, cellb, pb = table(cb, ib, id_s.data.flat())
    common, ja, jb = np.intersect1d(ka, kb, assume_unique=True, return_indices=True)
    err = float(np.max(np.abs(pa[ja] - pb[jb])))
    dis = int(len(ka) + len(kb) - 2 * common.size + np.count_nonzero(cella[ja] != cellb[jb]))
    return {"steps": args.check_steps, "particles": int(common.size), "max_abs_err": err,
            "domain": 1.0, "cell_disagreements": dis}


def target_rows(jp, ctx, args):
    """The stated target: the FP32-split RK2 step (advect + move + p2g) at 2D
    256^2 (nxcell 24, max 48) and 3D 128^3 (nxcell 27, max 54), against the
    FP64 drop-in on the same workload: throughput of both forms, their ratio,
    the split step's algorithmic bytes against the HBM peak, and the split
    position error after a few steps."""
    from justpic_hip.configs import cellular3d
    rows = {}
    for D, n in ((2, 256), (3, 128)):
        spec = dict(WORKLOADS[D])
        spec["method"], spec["inject"] = "rk2", False
        cfg = stream2d_cfg(n) if D == 2 else cellular3d(n, amp=100.0)
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup, sub.check_steps = min(args.steps, 20), 3, 10
        log(f"target row {n}^{D} RK2")
        r_s = run_form(jp, ctx, cfg, "split", sub, spec)
        r_f = run_form(jp, ctx, cfg, "fp64", sub, spec)
        err = split_error(jp, ctx, cfg, sub, spec)
        rows[f"{D}d_{n}_rk2"] = {
            "workload": f"{n}^{D} cells, nxcell {spec['nxcell']}, max_xcell {spec['max_xcell']}, "
                        f"RK2 advection + move + p2g",
            "particles": int(r_s["P"]),
            "split": {"value": r_s["value"], "ms_per_step": r_s["ms_per_step"],
                      "hbm_frac_step": r_s["hbm_frac_step"], "calls": r_s["calls"]},
            "fp64": {"value": r_f["value"], "ms_per_step": r_f["ms_per_step"],
                     "hbm_frac_step": r_f["hbm_frac_step"], "calls": r_f["calls"]},
            "ratio_split_over_fp64": r_s["value"] / r_f["value"],
            "split_error": err,
            "target": {"hbm_frac": 0.6, "ratio": 1.7, "max_abs_err": 1e-6},
        }
        log("target", f"{n}^{D}", "ratio", rows[f"{D}d_{n}_rk2"]["ratio_split_over_fp64"],
            "hbm", r_s["hbm_frac_step"], "err", err["max_abs_err"])
    return rows


def stream2d_cfg(n):
    from justpic_hip.configs import stream2d
    return stream2d(n, amp=250.0, readme=False)


def run_ensemble(jp, args):
    """Batched ensemble: members with cell counts drawn from {64..512}^2, README
    stream function at a random amplitude and their own CFL dt, FP32-split
    positions; a step is the persistent advection launch plus per-member
    move_particles! and particle2grid!, captured once in a hipGraph."""
    ctx = jp.Context(0, synchronous=False)
    log("building ensemble of", args.members)
    ens, fields, nodal = jp.random_ensemble(ctx, args.members, seed=args.seed,
                                            streams=args.streams)
    ctx.sync()
    best, ms = ens.tune(reps=3)
    log("tuned variant", best, "ms per variant", [round(t, 3) for t in ms])
    ens.step(fields, nodal)  # eager step sizes the context scratch
    g = ens.capture_step(fields, nodal)
    for _ in range(args.warmup):
        ctx.graph_launch(g)
    ctx.sync()
    P0 = ens.particles()
    e0, e1 = ctx.event(), ctx.event()
    ctx.record(e0)
    for _ in range(args.steps):
        ctx.graph_launch(g)
    ctx.record(e1)
    ctx.sync()
    gpu_ms = ctx.elapsed_ms(e0, e1)
    P1 = ens.particles()
    P = 0.5 * (P0 + P1)
    cells = sum(p.C for p, _, _ in ens.members)
    # advection alone (tuned variant) inside eager steps, for the roofline line
    adv = []
    for _ in range(3):
        a0, a1 = ctx.event(), ctx.event()
        ctx.record(a0)
        ens.advect()
        ctx.record(a1)
        for (p, _, _), f, F in zip(ens.members, fields, nodal):
            jp.move_particles_(p, (f,))
            jp.particle2grid_(F, f, p)
        ctx.sync()
        adv.append(ctx.elapsed_ms(a0, a1))
    adv_ms = float(np.mean(adv))
    S = sum(p.C * p.max_xcell for p, _, _ in ens.members)
    Vb = sum(v.nbytes for _, V, _ in ens.members for v in V)
    adv_bytes = S * 1 + 2 * P * 2 * 4 + Vb
    adv_gbs = adv_bytes / (adv_ms * 1e-3) / 1e9
    out = {
        "metric": "particle-steps/s (batched ensemble, FP32-split, RK2 + move + p2g, "
                  "hipGraph step)",
        "value": P * args.steps / (gpu_ms * 1e-3),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gpu_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.members} independent 2D grids, cells drawn from "
                               f"{{64,128,192,256,384,512}}^2, README stream function at "
                               f"amplitude U[50,400], own CFL dt, nxcell 24, max_xcell 48",
                   "members": args.members, "cells": cells, "particles": int(P),
                   "streams": args.streams,
                   "tuned_variant": best, "variant_ms": ms},
        "roofline": {"bound": "hbm", "kernel": "ensemble advection (persistent, tuned)",
                     "achieved": adv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": adv_gbs / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": adv_bytes, "avg_launch_us": adv_ms * 1e3},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    ens.close()


def run_adjoint(jp, args):
    """Adjoint inverse problem: 256^2 cells, nxcell 16 (max 32, min 8), true
    velocity = rotating cell (amplitude 250) + 10% smooth divergence-free
    perturbation, guess = the rotating cell, Gaussian tracer (sigma 0.1 at
    (0.5, 0.7)), 100 RK2 steps, particle2grid! to the 257^2 nodes, L2 loss
    against the true field's result.  One step is one gradient: the forward
    pass (checkpoints every 10 steps), the loss cotangent, the pullback."""
    from justpic_hip.configs import inverse2d
    n, N, every = args.n or 256, 100, 10
    ctx = jp.Context(0, synchronous=False)
    cfg = inverse2d(n)
    p = jp.init_particles(ctx, 16, 32, 8, *cfg["xi_vel"], seed=args.seed)
    x0d, _ = jp.flatten(p)
    x0 = [x.to_host() for x in x0d]
    Tp = jp.DeviceArray.from_host(ctx, cfg["T"](x0[0], x0[1]))
    P = Tp.size
    Vg = tuple(jp.DeviceArray.from_host(ctx, v) for v in cfg["V"])
    Vt = tuple(jp.DeviceArray.from_host(ctx, v) for v in cfg["V_true"])
    rk2 = jp.RungeKutta2()
    Ft = jp.AdvectionP2G(p, rk2, cfg["dt"], N, every).forward(Vt, x0d, Tp).flat()
    op = jp.AdvectionP2G(p, rk2, cfg["dt"], N, every)
    Fb = jp.DeviceArray(ctx, tuple(k + 1 for k in p.ni))
    out = None
    fw_ms, bw_ms, losses = [], [], []
    e = [ctx.event() for _ in range(4)]
    log("adjoint: particles", P, "steps", N)
    for it in range(args.warmup + args.steps):
        ctx.record(e[0])
        F = op.forward(Vg, x0d, Tp)
        ctx.record(e[1])
        L, fb = jp.l2_loss(F, Ft)  # synchronises
        Fb.copy_from(fb)
        ctx.record(e[2])
        out = op.pullback(Fb, out)
        ctx.record(e[3])
        ctx.sync()
        if it >= args.warmup:
            fw_ms.append(ctx.elapsed_ms(e[0], e[1]))
            bw_ms.append(ctx.elapsed_ms(e[2], e[3]))
            losses.append(L)
    fw, bw = float(np.mean(fw_ms)), float(np.mean(bw_ms))
    # the forward pass in the FP32-split form beside it: time, and its F,
    # final positions and gradient against the FP64 forward's
    F64 = op.forward(Vg, x0d, Tp).flat()
    xN64 = [x.flat()[:P] for x in op.xN]
    al64 = op.alive.flat()[:P].astype(bool)
    L64, fb64 = jp.l2_loss(op._fw["F"], Ft)
    Fb.copy_from(fb64)
    g64 = [v.to_host().ravel(order="F") for v in op.pullback(Fb)[0]]
    ops = jp.AdvectionP2G(p, rk2, cfg["dt"], N, every, form="split")
    sw_ms = []
    for it in range(args.warmup + args.steps):
        ctx.record(e[0])
        Fs = ops.forward(Vg, x0d, Tp)
        ctx.record(e[1])
        ctx.sync()
        if it >= args.warmup:
            sw_ms.append(ctx.elapsed_ms(e[0], e[1]))
    Fs = Fs.flat()
    als = ops.alive.flat()[:P].astype(bool)
    both = als & al64
    xerr = max(float(np.abs(x.flat()[:P][both] - y[both]).max()) for x, y in zip(ops.xN, xN64))
    okF = np.isfinite(Fs) & np.isfinite(F64)
    Ls, fbs = jp.l2_loss(ops._fw["F"], Ft)
    Fb.copy_from(fbs)
    gs = [v.to_host().ravel(order="F") for v in ops.pullback(Fb)[0]]
    grel = np.concatenate([np.abs(a[np.abs(b) >= 1e-3 * np.abs(b).max()] -
                                  b[np.abs(b) >= 1e-3 * np.abs(b).max()]) /
                           np.abs(b[np.abs(b) >= 1e-3 * np.abs(b).max()])
                           for a, b in zip(gs, g64)])
    split = {"forward_ms": float(np.mean(sw_ms)), "loss": Ls, "loss_fp64": L64,
             "xN_max_abs_err_vs_fp64": xerr,
             "survivor_flips": int(np.count_nonzero(als != al64)),
             "F_max_abs_diff_vs_fp64": float(np.abs(Fs[okF] - F64[okF]).max()),
             "grad_rel_diff_vs_fp64": {"median": float(np.median(grel)),
                                       "p99": float(np.quantile(grel, 0.99)),
                                       "max": float(grel.max()), "entries": int(grel.size)}}
    log("split forward", split)
    # inverse problem: a few steepest-descent steps with backtracking
    V = [v.to_host() for v in Vg]
    L0 = losses[0]
    L, hist = L0, [L0]
    for _ in range(args.descent):
        g = [v.to_host() for v in out[0]]
        eta = 0.05 * max(np.abs(v).max() for v in V) / max(np.abs(x).max() for x in g)
        for _ in range(8):
            Vn = [v - eta * x for v, x in zip(V, g)]
            for d, v in zip(Vg, Vn):
                d.copy_from(v)
            Ln, fb = jp.l2_loss(op.forward(Vg, x0d, Tp), Ft)
            if Ln < L:
                break
            eta *= 0.5
        V, L = Vn, Ln
        hist.append(L)
        Fb.copy_from(fb)
        out = op.pullback(Fb, out)
    ctx.sync()
    err_true = max(float(np.abs(v - t).max()) for v, t in zip(V, cfg["V_true"]))
    log("loss history", hist)

    cpu = None
    if args.cpu_steps > 0:
        from oracle import oracle as O
        hp = O.HostParticles(cfg["xv"], 32, cfg["xi_vel"])
        stride = 64
        xs = [x[::stride].copy() for x in x0]
        ts = Tp.to_host()[::stride].copy()
        t0 = time.perf_counter()
        r = O.adjoint_forward(hp, cfg["V"], O.JP_RK2, 0.5, cfg["dt"], N, xs, ts)
        O.adjoint_backward(hp, cfg["V"], O.JP_RK2, 0.5, cfg["dt"], N, xs, ts, r["F"], r["W"],
                           np.nan_to_num(r["F"]))
        el = time.perf_counter() - t0
        cpu = {"value": xs[0].size * N / el, "unit": "gradient particle-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"every {stride}th particle ({xs[0].size}), {N} RK2 steps, forward + "
                         f"backward by the sequential C oracle, {el:.1f} s"}
    # forward-pass algorithmic bytes: x0, checkpoints, xN, alive, cell ids, bucket
    # lists, V once, nodal F and W
    nblk = -(-N // every)
    Vb = sum(v.nbytes for v in Vg)
    nodes = Fb.size
    fw_bytes = P * (2 * 8 + nblk * 2 * 8 + 2 * 8 + 1 + 4 + 4) + Vb + 2 * 8 * nodes
    res = {
        "metric": "gradient particle-steps/s (adjoint of RK2 advection + particle2grid, FP64)",
        "value": P * N / ((fw + bw) * 1e-3),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fw + bw,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"adjoint inverse problem {n}^2 cells, nxcell 16, max_xcell 32, "
                               f"{N} RK2 steps, checkpoints every {every}, L2 loss on the "
                               f"{n + 1}^2 nodes",
                   "particles": P, "forward_ms": fw, "backward_ms": bw,
                   "loss_history": hist, "max_abs_V_error_vs_true": err_true,
                   "initial_max_abs_V_error": max(float(np.abs(a - b).max()) for a, b in
                                                  zip(cfg["V"], cfg["V_true"])),
                   "split_forward": split},
        "roofline": {"bound": "hbm", "kernel": "adjoint forward pass (all launches)",
                     "achieved": fw_bytes / (fw * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": fw_bytes / (fw * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_launch": fw_bytes, "avg_launch_us": fw * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def time_op(ctx, fn, reps, prep=None):
    """GPU time of one call in microseconds: after a warm-up call, `reps`
    calls each bracketed by hipEvents on the context stream (`prep`, untimed,
    before each: restores a state the call changes)."""
    if prep:
        prep()
    fn()
    ctx.sync()
    tot = 0.0
    for _ in range(reps):
        if prep:
            prep()
        e0, e1 = ctx.event(), ctx.event()
        ctx.record(e0)
        fn()
        ctx.record(e1)
        ctx.sync()
        tot += ctx.elapsed_ms(e0, e1)
    return 1000.0 * tot / reps


def run_rows(jp, args):
    """The rows either side of the hot path (inject_particles_phase!,
    update_phase_ratios!, grid2particle_flip!, the centroid transfers, the
    CPU CellArray layout, advection_MQS!/advection_LinP!, the three
    semilagrangian_advection! forms, cell_index), FP64, at the benchmark's
    512^2 size: the state of the default workload after three RK2 advection +
    move steps (holes in the slot planes), built by the oracle and copied to
    the device.  Each row: GPU time per call (events around the call, state
    restored between calls where the call changes it), its algorithmic bytes
    and the fraction of the 8 TB/s HBM peak, and the oracle (C, OpenMP where
    it threads the loop) timed once on the same state."""
    from oracle import oracle as O
    from justpic_hip.configs import gaussian_T
    from justpic_hip.grids import compute_dx
    n = args.n or 512
    reps = max(3, min(args.steps, 20))
    cfg = stream2d_cfg(n)
    xv, xi, V, dt = cfg["xv"], cfg["xi_vel"], cfg["V"], cfg["dt"]
    spec = WORKLOADS[2]
    O.set_threaded(True)
    log("rows: building the state,", n, "^2")
    hp = O.HostParticles(xv, spec["max_xcell"], xi)
    O.init_particles(hp, spec["nxcell"] // 4, args.seed)
    for _ in range(3):
        O.advection(hp, V, O.JP_RK2, 0.5, dt)
        O.move_particles(hp, [])
    rng = np.random.default_rng(args.seed)
    act = hp.index == 1
    P, S, Cc, K = int(act.sum()), hp.S, hp.C, hp.K
    Nn = (n + 1) ** 2
    phases = np.where(act, rng.integers(1, 4, S), np.nan).astype(float)
    Fp = np.where(act, rng.random(S), np.nan)
    T = gaussian_T(xv)
    T0 = np.asfortranarray(T + 0.25 * np.sin(7.0 * T))
    xc = O.centre_axes(hp)
    mc = np.meshgrid(*[a.x for a in xc], indexing="ij")
    Fc_host = np.asfortranarray(np.sin(5 * mc[0]) * np.cos(4 * mc[1]))

    ctx = jp.Context(0, synchronous=False)
    p = jp.particles_from_host(ctx, xi, hp.coords, hp.index, spec["nxcell"], K,
                               spec["min_xcell"], seed=args.seed)

    def cells(a):
        ca = jp.CellArray(ctx, p.ni, K)
        ca.copy_from(a)
        return ca

    saved = [jp.CellArray(ctx, p.ni, K) for _ in range(2)] + \
        [jp.CellArray(ctx, p.ni, K, dtype=np.uint8)]
    for dst, src in zip(saved, list(p.coords) + [p.index]):
        dst.data.copy_from_device(src.data)
    ph_d, Fp_d, Fp2_d = cells(phases), cells(Fp), cells(Fp)
    ph_s, Fp_s = cells(phases), cells(Fp)
    Vd = tuple(jp.DeviceArray.from_host(ctx, v) for v in V)
    Td, T0d = jp.DeviceArray.from_host(ctx, T), jp.DeviceArray.from_host(ctx, T0)
    Fsl = jp.DeviceArray.from_host(ctx, T)
    Fcd = jp.DeviceArray.from_host(ctx, Fc_host)
    Fc_out = jp.centroid_field(p)
    pr = jp.PhaseRatios(ctx, 3, p.ni)
    bg = jp.BacktrackGrid(ctx, xi, xv)
    buf = jp.DeviceArray(ctx, (S,))
    xs = tuple(jp.DeviceArray.from_host(ctx, np.nan_to_num(c, nan=0.5)) for c in hp.coords)
    ci_out = tuple(jp.DeviceArray(ctx, (S,), np.int64) for _ in range(2))
    dax = [jp.particles.DeviceAxis(ctx, a) for a in xv]
    ci_axes = (jp._abi.jp_axis * 2)(*[a.abi() for a in dax])
    ctx.sync()

    def restore():
        for dst, src in zip(list(p.coords) + [p.index], saved):
            dst.data.copy_from_device(src.data)
        ph_d.data.copy_from_device(ph_s.data)
        Fp_d.data.copy_from_device(Fp_s.data)

    def host_copy(h):
        c = O.HostParticles(h.xv, h.K, h.xi_vel)
        for d in range(h.D):
            c.coords[d][:] = h.coords[d]
        c.index[:] = h.index
        return c

    def cpu(fn):
        t0 = time.perf_counter()
        fn()
        return 1000.0 * (time.perf_counter() - t0)

    Vb = sum(v.size for v in V) * 8
    ratio_bytes = sum(f.size for f in pr.fields.values()) * 8
    rk = jp.RungeKutta2()
    rows = {}

    def row(name, us, nbytes, cpu_ms, note):
        rows[name] = {"us": us, "bytes": float(nbytes), "GBs": nbytes / us / 1e3,
                      "hbm_frac": nbytes / us / 1e3 / 8000.0, "cpu_ms": cpu_ms,
                      "gpu_over_cpu": cpu_ms * 1e3 / us, "bytes_are": note}
        log(f"rows: {name:28s} {us:9.1f} us  {nbytes / us / 1e3:7.0f} GB/s  cpu {cpu_ms:9.2f} ms")

    us = time_op(ctx, lambda: jp.update_phase_ratios_(pr, p, ph_d), reps)
    c_ms = cpu(lambda: O.update_phase_ratios(hp, phases, 3))
    row("update_phase_ratios", us, S + P * 24 + ratio_bytes, c_ms,
        "mask + coordinates and phase of each particle + the ratio arrays written")

    us = time_op(ctx, lambda: jp.inject_particles_phase_(p, ph_d, (Fp_d,), (Td,)), reps,
                 prep=restore)
    h2, ph2, f2 = host_copy(hp), phases.copy(), Fp.copy()
    c_ms = cpu(lambda: O.inject_particles_phase(h2, ph2, [f2], [T], spec["min_xcell"],
                                                args.seed, 0))
    row("inject_particles_phase", us, S + P * 16, c_ms,
        "the scan (mask + coordinates of each particle); the listed cells' work is extra")

    us = time_op(ctx, lambda: jp.grid2particle_flip_(Fp2_d, Td, T0d, p, 0.5), reps)
    f2 = Fp.copy()
    c_ms = cpu(lambda: O.grid2particle_flip(hp, [f2], [T], [T0], 0.5))
    row("grid2particle_flip", us, S + P * 32 + 16 * Nn, c_ms,
        "mask + coordinates, field read and written per particle + F, F0")

    us = time_op(ctx, lambda: jp.particle2centroid_(Fc_out, Fp_s, p), reps)
    c_ms = cpu(lambda: O.particle2centroid(hp, [Fp]))
    row("particle2centroid", us, S + P * 24 + 8 * Cc, c_ms,
        "mask + coordinates and field per particle + the centroid field")

    us = time_op(ctx, lambda: jp.centroid2particle_(Fp2_d, Fcd, p), reps)
    f2 = Fp.copy()
    c_ms = cpu(lambda: O.centroid2particle(hp, [f2], [Fc_host]))
    row("centroid2particle", us, S + P * 24 + 8 * Cc, c_ms,
        "mask + coordinates per particle, field written + the centroid field")

    src = saved[0]
    us = time_op(ctx, lambda: ctx.check(ctx.lib.jp_cellarray_to_cpu_layout(
        ctx.h, src.ptr, buf.ptr, Cc, K, 8), "layout"), reps)
    a_host = hp.coords[0]
    c_ms = cpu(lambda: np.ascontiguousarray(a_host.reshape(K, Cc).T))
    row("cellarray_to_cpu_layout", us, 16 * S, c_ms,
        "one FP64 CellArray read and written (numpy transpose on one thread for the CPU)")

    for name, fn, sch in (("advection_MQS", jp.advection_MQS_, "MQS"),
                          ("advection_LinP", jp.advection_LinP_, "LinP")):
        us = time_op(ctx, lambda: fn(p, rk, Vd, dt), reps, prep=restore)
        h2 = host_copy(hp)
        c_ms = cpu(lambda: O.advection_scheme(h2, V, O.JP_RK2, 0.5, dt, sch))
        row(name + " (RK2)", us, S + 2 * P * 16 + Vb, c_ms,
            "mask + coordinates read and written + the velocity")

    grid = tuple(compute_dx(a) for a in xv)
    gvi = tuple(tuple(compute_dx(a) for a in g) for g in xi)
    for name, fn, sch in (("semilagrangian", jp.semilagrangian_advection_, "linear"),
                          ("semilagrangian_MQS", jp.semilagrangian_advection_MQS_, "MQS"),
                          ("semilagrangian_LinP", jp.semilagrangian_advection_LinP_, "LinP")):
        us = time_op(ctx, lambda: fn(Fsl, Td, rk, Vd, xi, bg, 0.5 * dt), reps)
        F2 = [T.copy()]
        c_ms = cpu(lambda: O.semilagrangian(grid, gvi, V, F2, [T], O.JP_RK2, 0.5, 0.5 * dt,
                                            sch))
        row(name + " (RK2)", us, 16 * Nn + Vb, c_ms,
            "one field read (at the node) and written per node + the velocity")

    us = time_op(ctx, lambda: ctx.check(ctx.lib.jp_cell_index(
        ctx.h, 2, ci_axes, S, jp.api._ptrs(xs), jp.api._ptrs(ci_out)), "cell_index"), reps)
    pts = [np.nan_to_num(c, nan=0.5) for c in hp.coords]
    xa = [np.asarray(a.x) for a in xv]
    # (the oracle's cell_index is one point per call; the CPU figure is its
    # formula vectorised by numpy, checked against it on a sample first)
    ref = [O.cell_index(float(v), xa[0]) for v in pts[0][:1000]]
    vec = lambda x, a: (np.abs(np.trunc((x - a[0]) / abs(a[1] - a[0]))) + 1).astype(np.int64)
    assert np.array_equal(vec(pts[0][:1000], xa[0]), ref)
    c_ms = cpu(lambda: [vec(x, a) for x, a in zip(pts, xa)])
    row("cell_index (2 x S points)", us, 2 * S * 16, c_ms,
        "two coordinate arrays read, two int64 index arrays written (CPU: the formula "
        "vectorised by numpy, one thread)")
    nthr = O.threads()
    O.set_threaded(False)
    out = {"case": "rows", "workload": f"2D {n}^2 cells, nxcell 24, max_xcell 48, FP64, the "
                                      "state after three RK2 advection + move steps",
           "particles": P, "slots": S, "reps": reps, "cpu_threads": nthr,
           "cpu": cpu_model(), "rows": rows}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=None,
                    help="cells per dimension (default 512 step, 128 step3d, 256 adjoint)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--form", choices=("split", "fp64"), default="split",
                    help="position form whose throughput is the reported value")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the other form and the error check")
    ap.add_argument("--check-steps", type=int, default=50)
    ap.add_argument("--case", choices=("step", "step3d", "ensemble", "adjoint", "rows"),
                    default="step",
                    help="step: the 2D 512^2 workload (default); step3d: 3D RK4 128^3; "
                         "ensemble: batched 2D grids; adjoint: gradient of the 256^2 "
                         "inverse problem")
    ap.add_argument("--descent", type=int, default=3,
                    help="steepest-descent steps of the inverse problem (--case adjoint)")
    ap.add_argument("--members", type=int, default=64, help="ensemble size (--case ensemble)")
    ap.add_argument("--streams", type=int, default=8,
                    help="streams the ensemble's per-grid calls are spread over")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--method", choices=("rk2", "rk4"), default=None,
                    help="integrator (default: RK2 for step, RK4 for step3d)")
    ap.add_argument("--no-targets", action="store_true",
                    help="skip the target rows (split vs FP64 RK2 at 256^2 and 128^3)")
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="oracle steps timed for the CPU baseline, about 10 s on 16 host "
                         "threads at 512^2 (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the multi-GPU sharding is a design note, not a runnable path", file=sys.stderr)
        sys.exit(2)

    import justpic_hip as jp
    from justpic_hip.configs import stream2d

    if args.calibrate_copy > 0:
        ctx = jp.Context(0)
        n = args.calibrate_copy * (1 << 20) // 8
        a = jp.DeviceArray(ctx, (n,)).fill(1.0)
        b = jp.DeviceArray(ctx, (n,))
        for _ in range(5):
            b.copy_from_device(a)
        ctx.sync()
        log(f"copied {args.calibrate_copy} MiB five times")
        return

    if args.case == "ensemble":
        run_ensemble(jp, args)
        return
    if args.case == "adjoint":
        run_adjoint(jp, args)
        return
    if args.case == "rows":
        run_rows(jp, args)
        return
    D = 3 if args.case == "step3d" else 2
    spec = dict(WORKLOADS[D])
    default_method = spec["method"]
    if args.method:
        spec["method"] = args.method
    args.n = args.n or (128 if D == 3 else 512)
    log("building workload", D, "D", args.n)
    if D == 2:
        cfg = stream2d(args.n, amp=250.0, readme=False)
    else:
        from justpic_hip.configs import cellular3d
        cfg = cellular3d(args.n, amp=100.0)
    ctx = jp.Context(0, synchronous=False)
    main_r = run_form(jp, ctx, cfg, args.form, args, spec)
    other = None
    err = None
    if not args.no_compare:
        other_form = "fp64" if args.form == "split" else "split"
        other = run_form(jp, ctx, cfg, other_form, args, spec)
        log("split vs fp64 error check,", args.check_steps, "steps")
        err = split_error(jp, ctx, cfg, args, spec)
        log("error", err)

    # roofline: the dominant call of the step (the longest; its launches are
    # bracketed by events on the context stream, which runs every kernel)
    dom = max(main_r["calls"], key=lambda k: main_r["calls"][k]["us"])
    adv = main_r["calls"][dom]
    tag = f"{args.n}" if D == 2 else f"{args.n}_3d"
    if spec["method"] != default_method:
        tag += f"_{spec['method']}"
    traffic = read_traffic(os.path.join(ROOT, "profiles",
                                        f"traffic_{dom}_{args.form}_{tag}.txt"))
    meth = spec["method"].upper()
    seq = f"advect {meth} + move" + (" + inject" if spec["inject"] else "") + " + p2g"
    size = f"{args.n}^{D}"

    cpu = None
    # 3D: one step of the full-size workload
    cpu_steps = args.cpu_steps if D == 2 else min(args.cpu_steps, 1)
    if cpu_steps > 0:
        ccfg, csize = cfg, size
        log("cpu baseline: threaded oracle,", cpu_steps, "steps at", csize)
        v_cpu, el, nthr = cpu_baseline(ccfg, spec, cpu_steps, args.seed)
        cpu = {"value": v_cpu, "unit": "particle-steps/s", "cores": nthr, "kind": "port",
               "sample": f"{cpu_steps} full FP64 steps ({seq}) of the {csize} workload by "
                         f"the C oracle on {nthr} OpenMP threads (cells of one colour "
                         f"phase in parallel), {el:.1f} s, CPU {cpu_model()}; Julia is not "
                         f"installed, so the reference itself cannot run"}

    names = {"split": "FP32-split positions", "fp64": "FP64 drop-in"}
    if D == 2:
        workload = (f"2D temperature advection {size} cells, nxcell 24, max_xcell 48, "
                    f"min_xcell 12, {meth} + move + inject + p2g, {names[args.form]}")
    else:
        workload = (f"3D {meth} advection {size} cells, cellular flow x100, nxcell 27, "
                    f"max_xcell 54, min_xcell 14, {meth} + move + p2g of one field, "
                    f"{names[args.form]}")
    out = {
        "metric": METRIC,
        "value": main_r["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.form == "split" else "f64",
        "data": "synthetic",
        "config": {"workload": workload, "cells": args.n ** D,
                   "particles": int(main_r["P"]), "form": args.form},
        "hbm_frac_step": main_r["hbm_frac_step"],
        "step_bytes": main_r["bytes"]["step"],
        "calls": main_r["calls"],
        "roofline": {"bound": "hbm", "kernel": f"{dom} call ({args.form}, "
                                               f"{KERNELS[dom]})",
                     "achieved": adv["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": adv["GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": adv["bytes"], "avg_launch_us": adv["us"]},
        "cpu_baseline": cpu,
        "movers_per_step": main_r["moved"],
        "injected_per_step": main_r["injected"],
    }
    if other is not None:
        other_form = "fp64" if args.form == "split" else "split"
        out["compare"] = {"form": other_form, "value": other["value"],
                          "ms_per_step": other["ms_per_step"], "calls": other["calls"],
                          "hbm_frac_step": other["hbm_frac_step"],
                          "ratio_split_over_fp64": (main_r["value"] / other["value"]
                                                    if args.form == "split"
                                                    else other["value"] / main_r["value"])}
        out["split_error"] = err
    if args.case == "step" and not args.no_targets:
        out["targets"] = target_rows(jp, ctx, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
