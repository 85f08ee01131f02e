#!/usr/bin/env python3
"""Benchmark of the explicit finite-volume step on one MI355X.

Workload: the 3D two-phase pressure-non-equilibrium case (VELOCITYEQ,
stiffened-gas water + ideal-gas air, 512^3 cells, 4096 air pockets with radii
log-uniform over 2-10 cells, high-pressure slab drive, MUSCL minmod with the
explicit midpoint predictor-corrector, infinite pressure relaxation), plain
FP64.  A step is one full reference time step of the whole mesh (predictor,
corrector, relaxations, CFL reduction) with the state resident in HBM.

Prints one JSON line: Mcell-updates/s over the timed steps, the HBM roofline of
the dominant kernel (algorithmic bytes per launch over its event-timed average
duration) and the oracle timed on the host on a bounded replica.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def hip():
    for p in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        try:
            L = C.CDLL(p)
            break
        except OSError:
            continue
    else:
        raise RuntimeError("libamdhip64.so not found")
    L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    L.hipEventSynchronize.argtypes = [C.c_void_p]
    L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    L.hipEventDestroy.argtypes = [C.c_void_p]
    return L


def traffic_from_profiles(kernel):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE), if committed."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


WORKLOADS = {
    "c5": "3D two-phase VELOCITYEQ water/air {n}^3, {extra} air pockets (r 2-10 cells), 5e8 Pa slab drive, "
          "MUSCL minmod + explicit midpoint, infinite P-relaxation",
    "c4": "3D three-phase VELOCITYEQ water/epoxy/air {n}^3, {extra} inclusions, 1e9 Pa water shock, first order, "
          "infinite P-relaxation",
    "c3": "2D Kapila (PRESSUREVELOCITYEQ) water/air {n}^2, {extra} bubbles, 1e9 Pa water shock, MUSCL minmod + "
          "explicit midpoint",
    "c2": "2D Kapila (PRESSUREVELOCITYEQ) water/air {n}^2, {extra} bubbles, 1e9 Pa water shock, first order",
}
FORMS = {"fp64": "plain FP64",
         "fp32": "FP32 state, reconstruction and fluxes (compensated sums); FP64 cell update, EOS inversions "
                 "and relaxation Newton"}
DEFAULT_N = {"c5": 512, "c4": 256, "c3": 2048, "c2": 2048}


def make_case(workload, n):
    from ecogen_amd.case import kapila_bubbles_2d, ueq2_pockets_3d, ueq3_spheres_3d
    if workload == "c5":
        extra = 4096 if n == 512 else max(1, int(round(4096 * (n / 512.0) ** 3)))
        return ueq2_pockets_3d(n=n, npockets=extra), extra
    if workload == "c4":
        extra = 64 if n == 256 else max(2, int(round(64 * (n / 256.0) ** 3)))
        return ueq3_spheres_3d(n=n, nsph=extra), extra
    extra = 64
    return kapila_bubbles_2d(n=n, nbub=extra, order=2 if workload == "c3" else 1), extra


def cpu_baseline(workload="c5", steps=12):
    """The oracle (plain C, one thread) on a bounded replica of the benchmark case."""
    from oracle.pyoracle import Oracle
    n = {"c5": 64, "c4": 64, "c3": 192, "c2": 384}[workload]
    case, npk = make_case(workload, n)
    o = Oracle(case)
    dt = 1.e-15
    dt = case["cfl"] * o.step(dt)   # first step (dt0 = 1e-15) untimed
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = case["cfl"] * o.step(dt)
    el = time.perf_counter() - t0
    D = 3 if workload in ("c5", "c4") else 2
    return dict(value=round(o.ncell * steps / el / 1e6, 4), unit="Mcell-updates/s", cores=1, kind="port",
                sample=f"{n}^{D} replica ({npk} inclusions, same physics), {steps} steps after dt0, "
                       f"{el:.1f} s on one host core")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c5")
    ap.add_argument("--n", type=int, default=0, help="cells per axis (default: the workload's size)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate", action="store_true", help="also launch the byte-counter calibration copy")
    ap.add_argument("--form", choices=("fp64", "fp32"), default="fp64",
                    help="fp64: plain FP64; fp32: FP32 storage/fluxes with FP64 update and relaxation")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU decomposition is "
              f"not part of this build", file=sys.stderr)
        sys.exit(2)

    from ecogen_amd import Solver
    from ecogen_amd.case import variable_names

    n = args.n or DEFAULT_N[args.workload]
    case, npk = make_case(args.workload, n)
    case["precision"] = args.form
    esize = 8 if args.form == "fp64" else 4
    print(f"[bench] building {case['name']} ({args.workload}, {npk} inclusions)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    s = Solver(case)
    s.synchronize()
    print(f"[bench] initial condition ready in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    nvar = len(variable_names(case))
    ncell = s.ncell

    s.reset_clock()
    s.enqueue(args.warmup)
    s.synchronize()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    H = hip()
    e0, e1 = C.c_void_p(), C.c_void_p()
    H.hipEventCreate(C.byref(e0))
    H.hipEventCreate(C.byref(e1))
    s.set_profiling(True)
    stream = C.c_void_p(s.stream)
    w0 = time.perf_counter()
    H.hipEventRecord(e0, stream)
    s.enqueue(args.steps)
    H.hipEventRecord(e1, stream)
    H.hipEventSynchronize(e1)
    wall = time.perf_counter() - w0
    ms = C.c_float()
    H.hipEventElapsedTime(C.byref(ms), e0, e1)
    st = s.synchronize()
    kt = s.kernel_times()
    if args.calibrate:
        nb = esize * nvar * ncell // 16 * 16
        s.calibration_copy(nb)
        print(f"[bench] calibration copy: {nb} bytes read and {nb} bytes written in one launch", file=sys.stderr,
              flush=True)
    elapsed = ms.value / 1e3
    ms_per_step = ms.value / args.steps
    value = ncell * args.steps / elapsed / 1e6
    print(f"[bench] {args.steps} steps: {ms.value:.1f} ms (wall {wall * 1e3:.1f} ms), iteration {st['iteration']}, "
          f"newton iters {st['newton_iters']}, fallbacks {st['newton_fail']}", file=sys.stderr, flush=True)

    # Dominant kernel and its algorithmic bytes per launch: each stage reads and writes the
    # persistent SoA set once (predictor: W^n -> W*; corrector: W*, W^n -> W^{n+1}).
    transfers = {"predict": 2, "correct": 3, "step": 2}
    kind = max(("predict", "correct", "step"), key=lambda k: kt[k][0])
    kms, kn = kt[kind]
    avg_s = kms / kn / 1e3
    bytes_launch = transfers[kind] * nvar * esize * ncell
    achieved = bytes_launch / avg_s / 1e9
    kernel_name = {"predict": "k_step (second order, predictor)", "correct": "k_step (second order, corrector)",
                   "step": "k_step (first order)"}[kind]
    traffic = traffic_from_profiles(f"{args.workload}:{args.form}:{kind}")
    out = {
        "metric": "Mcell-updates/s: plain FP64 vs FP32+learned-relax at stated L1 err; HBM GB/s %",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32+f64",
        "data": "synthetic (seeded random air pockets, numpy default_rng 20261202)",
        "config": {
            "workload": WORKLOADS[args.workload].format(n=n, extra=npk),
            "cells": ncell,
            "state_vars": nvar,
            "form": FORMS[args.form],
            "bytes_per_cell_update": (5 if case["order"] == 2 else 2) * nvar * esize,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_name,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "avg_launch_ms": round(avg_s * 1e3, 3),
            "algorithmic_bytes_per_launch": bytes_launch,
        },
        "kernel_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in kt.items() if v[1]},
        "whole_step_hbm_gbs": round((5 if case["order"] == 2 else 2) * nvar * esize * value * 1e6 / 1e9, 2),
    }
    if not args.no_cpu_baseline:
        print("[bench] timing the oracle on the host (bounded replica)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.workload)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
