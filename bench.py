#!/usr/bin/env python3
"""Throughput of the JustPIC particle hot path on one MI355X.

Workload (the 2D temperature-advection case): 512^2 cells on [0,1]^2, README
stream function at amplitude 250, nxcell 24, max_xcell 48, min_xcell 12, one
particle field T = exp(-((x-0.5)^2 + (y-0.75)^2)/0.01) set by grid2particle!,
and per step: advection! (RK2), move_particles!, inject_particles!,
particle2grid! to the 513^2 nodal T.  A step is one pass of that sequence over
all particles; inputs are resident in HBM before timing starts.

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


def algorithmic_bytes(S, P, D, M, F, Nn, Vb, injected=0):
    """Byte model per step: each kernel that touches an array counts it once.
    m = 1 B mask/slot, w = 8 B coordinate, wf = 8 B field."""
    m, w, wf = 1, 8, 8
    adv = S * m + 2 * P * D * w + Vb
    move = S * m + P * D * w + M * (2 * D * w + 3 * F * wf + 2 * m)
    inj = S * m + P * D * w + injected * (D * w + F * wf + m)
    p2g = S * m + P * (D * w + F * wf) + 8 * F * Nn
    return {"advection": adv, "move": move, "inject": inj, "p2g": p2g,
            "step": adv + move + inj + p2g}


def cpu_baseline(cfg, n_steps, seed):
    """The oracle (sequential C, 1 core) on a bounded sample of the same
    workload: n_steps full steps at the same size."""
    from oracle import oracle as O
    from justpic_hip.configs import gaussian_T
    xv, xi, V, dt = cfg["xv"], cfg["xi_vel"], cfg["V"], cfg["dt"]
    hp = O.HostParticles(xv, 48, xi)
    O.init_particles(hp, 6, seed)
    T = gaussian_T(xv)
    pT = O.grid2particle(hp, T, np.full(hp.S, np.nan))
    psteps = 0
    t0 = time.perf_counter()
    for it in range(n_steps):
        psteps += int(hp.index.sum())
        O.advection(hp, V, O.JP_RK2, 0.5, dt)
        O.move_particles(hp, [pT])
        O.inject_particles(hp, [pT], 12, seed, it)
        O.particle2grid(hp, pT)
    el = time.perf_counter() - t0
    return psteps / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512, help="cells per dimension")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=6,
                    help="oracle steps timed for the CPU baseline (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "see the sharding design notes", file=sys.stderr)
        sys.exit(2)

    import justpic_hip as jp
    from justpic_hip.configs import gaussian_T, stream2d

    log("building workload", args.n)
    cfg = stream2d(args.n, amp=250.0, readme=False)
    xv, xi, V, dt = cfg["xv"], cfg["xi_vel"], cfg["V"], cfg["dt"]
    ctx = jp.Context(0, synchronous=False)
    p = jp.init_particles(ctx, 24, 48, 12, *xi, seed=args.seed)
    Vd = tuple(jp.DeviceArray.from_host(ctx, v) for v in V)
    T = jp.DeviceArray.from_host(ctx, gaussian_T(xv))
    pT = jp.CellArray(ctx, p.ni, p.max_xcell, fill=float("nan"))
    jp.grid2particle_(pT, T, p)
    rk2 = jp.RungeKutta2()
    ctx.sync()

    ev = [ctx.event() for _ in range(2)]

    def step(timed_adv=None):
        if timed_adv is not None:
            ctx.record(timed_adv[0])
        jp.advection_(p, rk2, Vd, dt)
        if timed_adv is not None:
            ctx.record(timed_adv[1])
        jp.move_particles_(p, (pT,))
        jp.inject_particles_(p, (pT,))
        jp.particle2grid_(T, pT, p)

    log("warmup", args.warmup)
    for _ in range(args.warmup):
        step()
    ctx.sync()
    P0 = p.active()
    # timed region: events bracket every advection launch (same stream)
    adv_ev = [(ctx.event(), ctx.event()) for _ in range(args.steps)]
    log("timing", args.steps, "steps")
    ctx.sync()
    t0 = time.perf_counter()
    ctx.record(ev[0])
    for s in range(args.steps):
        step(adv_ev[s])
    ctx.record(ev[1])
    ctx.sync()
    wall = time.perf_counter() - t0
    gpu_ms = ctx.elapsed_ms(ev[0], ev[1])
    P1 = p.active()
    adv_ms = float(np.mean([ctx.elapsed_ms(a, b) for a, b in adv_ev]))
    ms_per_step = gpu_ms / args.steps
    P = 0.5 * (P0 + P1)
    value = P * args.steps / (gpu_ms * 1e-3)
    log(f"gpu {gpu_ms:.2f} ms (host wall {wall * 1e3:.2f} ms), {ms_per_step:.4f} ms/step, "
        f"P {P0}->{P1}, advection {adv_ms * 1e3:.1f} us")

    # one calibration step with outcome counters (not timed) for the byte model
    jp.advection_(p, rk2, Vd, dt)
    st_m = jp.move_particles_(p, (pT,), stats=True)
    st_i = jp.inject_particles_(p, (pT,), stats=True)
    ctx.sync()
    S = p.C * p.max_xcell
    Nn = int(np.prod([n + 1 for n in p.ni]))
    Vb = sum(v.nbytes for v in Vd)
    B = algorithmic_bytes(S, P, 2, st_m["moved"] + st_m["deleted"], 1, Nn, Vb,
                          st_i["injected"])
    step_gbs = B["step"] / (ms_per_step * 1e-3) / 1e9
    adv_gbs = B["advection"] / (adv_ms * 1e-3) / 1e9

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_advection_f64_{args.n}.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if args.cpu_steps > 0:
        log("cpu baseline: oracle,", args.cpu_steps, "steps at the same size")
        v_cpu, el = cpu_baseline(cfg, args.cpu_steps, args.seed)
        cpu = {"value": v_cpu, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} full steps (advect RK2 + move + inject + p2g) of the "
                         f"same {args.n}^2 workload by the sequential C oracle, "
                         f"{el:.1f} s"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D temperature advection {args.n}^2 cells, nxcell 24, "
                               f"max_xcell 48, min_xcell 12, RK2 + move + inject + p2g, "
                               f"FP64 drop-in",
                   "cells": args.n * args.n, "particles": int(P), "form": "fp64"},
        "hbm_frac_step": step_gbs / HBM_PEAK_GBS,
        "step_bytes": B["step"],
        "roofline": {"bound": "hbm", "kernel": "advection (RK2)", "achieved": adv_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": adv_gbs / HBM_PEAK_GBS,
                     "traffic": traffic, "bytes_per_launch": B["advection"],
                     "avg_launch_us": adv_ms * 1e3},
        "cpu_baseline": cpu,
        "movers_per_step": st_m["moved"],
        "injected_per_step": st_i["injected"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
