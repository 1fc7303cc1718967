#!/usr/bin/env python3
"""Benchmark of the particle-advection hot path on one MI355X.

Workload (one "step" of this bench = one pass of the hot path over one batch):
steady endpoint advection on a 256^3 float32 Gaussian-vortex field (64
vortex tubes + uniform drift 0.1, spacing 1), 2^20 seeds uniform in
[0, 255)^3, dpa's runge_kutta_4 with dt = 0.25 for 1000 iterations, endpoints
only, particles leaving the grid retire.  Field and seeds are resident in HBM
before the timed region; each pass reads the seeds and writes final positions,
remaining iterations and status.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "RK4 particle-steps/s on 1 GPU + p50/p99 query latency vs per-request launch"
BYTES_PER_STEP = 96          # 8 corners x 12 B gathered per dpa-mode step
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "advect_persistent_vortex256_traffic.txt"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def read_traffic():
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(field, seeds, iters, dt, target_s: float):
    """Oracle (CPU port of dpa's loop) on a bounded sample of the same
    workload, on the host's cores with OpenMP schedule(dynamic,1)."""
    from oracle import py_oracle as po
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    fld = po.Field(field, (0, 0, 0), (1, 1, 1))
    probe = seeds[: 2048]
    t = time.perf_counter()
    r = po.advect_endpoints(fld, probe, iters, dt, threads=threads)
    el = time.perf_counter() - t
    rate = r["total_steps"] / max(el, 1e-9)
    count = int(min(len(seeds), max(2048, rate * target_s / (r["total_steps"] / len(probe)))))
    sample = seeds[: count]
    t = time.perf_counter()
    r = po.advect_endpoints(fld, sample, iters, dt, threads=threads)
    el = time.perf_counter() - t
    return {"value": r["total_steps"] / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"first {count} of the 2^20 seeds, full 1000 iterations, same field "
                      f"({r['total_steps']} particle-steps in {el:.2f} s; host {cpu_model()})"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seeds", type=int, default=1 << 20)
    ap.add_argument("--iterations", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(seeds shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        return 2

    import dpa_amd
    from dpa_amd import workloads as W

    t = time.perf_counter()
    field = W.vortex_field(256)
    seeds = W.cfg2_seeds(args.seeds)
    log(f"inputs ready in {time.perf_counter() - t:.1f} s")
    dt = np.float32(0.25)
    n = len(seeds)

    ctx = dpa_amd.Context(0, kernel_timing=True)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    d_seeds = ctx.to_device(seeds)
    d_pos = ctx.alloc(seeds.nbytes)
    d_rem = ctx.alloc(8 * n)
    d_st = ctx.alloc(n)

    def one_pass(wait: bool):
        return ctx.advect_endpoints_device(fld, d_seeds, n, d_pos, uniform_remaining=args.iterations,
                                           step_size=dt, d_rem_out=d_rem, d_status_out=d_st,
                                           wait=wait)

    steps_per_pass = None
    for i in range(max(args.warmup, 1)):
        steps_per_pass = one_pass(True)
        log(f"warmup {i}: {steps_per_pass} particle-steps")
    ctx.synchronize()
    ctx.kernel_timing_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_pass(False)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    last = one_pass(True)
    if last != steps_per_pass:
        raise RuntimeError(f"step count changed between passes: {last} != {steps_per_pass}")
    launches, total_ms = ctx.kernel_timing("advect_persistent")
    # the check pass above is the last timed launch: average over all of them
    avg_ms = total_ms / max(launches, 1)
    value = args.steps * steps_per_pass / elapsed
    log(f"timed {args.steps} passes in {elapsed:.3f} s: {value:.4e} particle-steps/s; "
        f"kernel avg {avg_ms:.3f} ms over {launches} launches")
    achieved = steps_per_pass * BYTES_PER_STEP / (avg_ms * 1e-3) / 1e9
    traffic = read_traffic()

    for p in (d_seeds, d_pos, d_rem, d_st):
        ctx.free(p)
    fld.release()
    ctx.close()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle on a bounded sample)")
        cpu = cpu_baseline(field, seeds, args.iterations, dt, args.cpu_seconds)
        log(f"cpu: {cpu['value']:.4e} particle-steps/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian-vortex field and uniform seeds; no dataset)",
        "config": {
            "workload": "large steady endpoint advection: 256^3 f32 Gaussian-vortex field "
                        "(64 tubes + drift 0.1), 2^20 uniform seeds, dpa runge_kutta_4 "
                        "dt 0.25 x 1000 iterations, endpoints only, exits retire",
            "seeds": n, "iterations": args.iterations, "particle_steps_per_pass": steps_per_pass,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "k_advect_persistent", "kernel_avg_ms": avg_ms,
            "bytes_per_step": BYTES_PER_STEP,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
