#!/usr/bin/env python3
"""Benchmark of the particle-advection hot path on one MI355X.

Default workload (one "step" of this bench = one pass of the hot path over one
batch): steady endpoint advection on a 256^3 float32 Gaussian-vortex field (64
vortex tubes + uniform drift 0.1, spacing 1), 2^20 seeds uniform in
[0, 255)^3, dpa's runge_kutta_4 with dt = 0.25 for 1000 iterations, endpoints
only, particles leaving the grid retire.  Field and seeds are resident in HBM
before the timed region; each pass reads the seeds and writes final positions,
remaining iterations and status.  A short served-query leg (continuous
batching vs one launch per request on the same Poisson trace) adds p50/p99
latency.

Other workloads: --workload abc64-record | ftle-aux | ftle-lattice | batch64 | serve.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "RK4 particle-steps/s on 1 GPU + p50/p99 query latency vs per-request launch"
HBM_PEAK_GBS = 8000.0
# Roofline bytes of an endpoint launch: the register cell cache serves every
# step that stays in its cell, so what the launch must bring past the L2 is one
# 128-byte cell record per crossing into a new cell (96 bytes, a cell's eight
# nodes, for launches that read the node layout), counted on the device, plus
# each particle's seed read and final position, remaining count and status.
RECORD_BYTES = 128
NODE_GROUP_BYTES = 96
PARTICLE_IO_BYTES = 12 + 12 + 8 + 1
CORNER_BYTES_PER_STEP = 96      # eight 12-byte corners per dpa-mode step, cache-served included
COPY_BYTES = 1 << 30            # device-to-device copy that calibrates the byte counters

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


def host_threads() -> int:
    """Threads of the CPU baseline: OMP_NUM_THREADS when it is set (the GPU
    boxes allot each GPU's process its share of the host's cores through it),
    otherwise every CPU this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, len(os.sched_getaffinity(0)))


def host_info() -> dict:
    return {"cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _timed_sample(fn, fld, seeds, iters, dt, threads, target_s, runs=3):
    """Runs fn on the first seeds, sized from a 2048-seed probe (which also
    warms the caches) so that each of `runs` timed runs takes about
    target_s / runs seconds; returns the median run's rate."""
    probe = seeds[: 2048]
    t = time.perf_counter()
    r = fn(fld, probe, iters, dt, threads=threads)
    el = time.perf_counter() - t
    per_seed = max(r["total_steps"] / len(probe), 1e-9)
    rate = r["total_steps"] / max(el, 1e-9)
    count = int(min(len(seeds), max(2048, rate * target_s / runs / per_seed)))
    timed = []
    for _ in range(runs):
        t = time.perf_counter()
        r = fn(fld, seeds[: count], iters, dt, threads=threads)
        el = time.perf_counter() - t
        timed.append((r["total_steps"] / el, r["total_steps"], el))
    timed.sort()
    v, steps, el = timed[len(timed) // 2]
    return v, count, steps, el


def cpu_baseline_endpoints(field, offset, spacing, seeds, iters, dt, target_s: float,
                           what: str):
    """CPU restatements of dpa's loop on bounded samples of the same workload,
    OpenMP schedule(dynamic,1) over the host's cores: the dpa-structured form
    (a heap-allocated corner list and a recursive corner walk per
    interpolation, as dpa's regular_grid::interpolate) is the baseline; the
    plain form and a one-thread run are reported beside it."""
    from oracle import py_oracle as po
    threads = host_threads()
    fld = po.Field(field, offset, spacing)
    v, count, steps, el = _timed_sample(po.advect_endpoints_dpa_structured, fld, seeds, iters, dt,
                                        threads, target_s * 0.5)
    vp, cp, _, _ = _timed_sample(po.advect_endpoints, fld, seeds, iters, dt, threads,
                                 target_s * 0.25)
    v1, c1, _, _ = _timed_sample(po.advect_endpoints_dpa_structured, fld, seeds, iters, dt, 1,
                                 target_s * 0.25)
    return {"value": v, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "host": host_info(),
            "sample": f"first {count} of the {what} seeds, all {iters} iterations, same field, "
                      f"median of 3 warm runs ({steps} particle-steps in {el:.2f} s; host "
                      f"{cpu_model()}); dpa-structured interpolation (heap corner list, "
                      "recursive corner walk)",
            "plain_port": {"value": vp, "cores": threads, "seeds": cp},
            "one_thread": {"value": v1, "cores": 1, "seeds": c1}}


def recorded_advect_timing(ctx):
    """(launches, ms) of the advect kernels since the last timing reset: a
    recorded launch runs the persistent kernel or one thread per row, whichever
    the library chose for it."""
    n1, t1 = ctx.kernel_timing("advect_persistent")
    n2, t2 = ctx.kernel_timing("advect_simple")
    return n1 + n2, t1 + t2


def roofline(avg_ms, algorithmic_bytes, steps, traffic=None):
    """HBM roofline of one launch: the bytes it must bring past the L2 (one
    cell record per cell crossing plus the particles' own I/O, counted on the
    device) over its average duration.  `traffic` is what the byte counters
    measured past the L2 for the same launch (HBM and the Infinity Cache
    behind it); `corner_gbs` is the per-step corner figure (96 B per step,
    corners the register cache serves included), which is not a roofline."""
    achieved = algorithmic_bytes / (avg_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel_avg_ms": avg_ms,
           "algorithmic_bytes": algorithmic_bytes, "bytes_per_step": algorithmic_bytes / steps,
           "corner_gbs": steps * CORNER_BYTES_PER_STEP / (avg_ms * 1e-3) / 1e9}
    if traffic:
        out["traffic_gbs"] = traffic / (avg_ms * 1e-3) / 1e9
        out["traffic_frac"] = out["traffic_gbs"] / HBM_PEAK_GBS
    return out


def counted_bytes(ctx, one_pass, n, steps):
    """One more pass with the cell-load counting instance: the launch's
    algorithmic bytes (its results must not change)."""
    ctx.set_count_cell_loads(True)
    try:
        s = one_pass(True)
        rec, node = ctx.cell_loads()
    finally:
        ctx.set_count_cell_loads(False)
    if s != steps:
        raise RuntimeError(f"the load-counting pass took {s} steps, not {steps}")
    return rec * RECORD_BYTES + node * NODE_GROUP_BYTES + n * PARTICLE_IO_BYTES, rec, node


def endpoint_loads(ctx, fld, sets, steps):
    """Cell loads of endpoint passes over (seeds, iterations, dt) sets, run by
    the counting instance; their particle-steps must add up to `steps`."""
    total = 0
    ctx.set_count_cell_loads(True)
    try:
        for seeds, iters, dt in sets:
            n = len(seeds)
            d_s = ctx.to_device(np.ascontiguousarray(seeds, np.float32))
            d_p = ctx.alloc(12 * n)
            try:
                total += ctx.advect_endpoints_device(fld, d_s, n, d_p, uniform_remaining=iters,
                                                     step_size=dt, wait=True)
            finally:
                ctx.free(d_s)
                ctx.free(d_p)
        rec, node = ctx.cell_loads()
    finally:
        ctx.set_count_cell_loads(False)
    if total != steps:
        raise RuntimeError(f"counted endpoint passes took {total} steps, not {steps}")
    return rec, node


# ------------------------------------------------ traffic (rocprofv3 --pmc)

def under_profiler() -> bool:
    """True inside a rocprofv3 run (its preloaded library is in the
    environment), where a nested rocprofv3 must not start."""
    return ("rocprof" in os.environ.get("LD_PRELOAD", "") or
            any(k.startswith("ROCPROF") for k in os.environ))


def _counter_rows(outdir: Path):
    import csv
    rows = []
    for f in sorted(outdir.rglob("*counter_collection.csv")):
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _per_dispatch(rows, counter):
    """[(dispatch id, kernel name, value)] of one counter, in dispatch order."""
    acc = {}
    for r in rows:
        if r.get("Counter_Name") != counter:
            continue
        d = int(r.get("Dispatch_Id", 0))
        name, v = r.get("Kernel_Name", ""), float(r.get("Counter_Value", 0.0))
        prev = acc.get(d, (name, 0.0))
        acc[d] = (name, prev[1] + v)
    return [(d, acc[d][0], acc[d][1]) for d in sorted(acc)]


def measure_traffic(args, kernel="k_advect_persistent"):
    """Bytes past the L2 per launch of the advect kernel, measured in this run:
    two rocprofv3 --pmc runs of this script in --traffic-child mode (the same
    workload: a warm pass, then two measured passes, then a 1 GiB
    device-to-device copy), one for FETCH_SIZE and one for WRITE_SIZE, each
    scaled by what the same counter reads for the copy.  None if rocprofv3 is
    missing or fails."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not Path(prof).exists():
        return None, "rocprofv3 not found"
    if under_profiler():
        return None, "not measured: this run is itself under a profiler"
    out = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = Path(tempfile.mkdtemp(prefix="dpah_pmc_", dir="/tmp"))
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", str(d), "-o", "run", "--",
               sys.executable, "-u", str(ROOT / "bench.py"), "--traffic-child",
               "--seeds", str(args.seeds), "--iterations", str(args.iterations)]
        log(f"traffic: {counter} (rocprofv3 --pmc run of the same workload)")
        try:
            p = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=240)
            rows = _counter_rows(d)
        except subprocess.TimeoutExpired:
            shutil.rmtree(d, ignore_errors=True)
            return None, f"rocprofv3 --pmc {counter} timed out"
        shutil.rmtree(d, ignore_errors=True)
        disp = _per_dispatch(rows, counter)
        ours = [v for _, name, v in disp if kernel in name]
        if len(ours) < 3:
            tail = " | ".join(p.stdout.strip().splitlines()[-3:])
            return None, f"rocprofv3 --pmc {counter}: {len(ours)} advect dispatches (rc {p.returncode}: {tail})"
        last_ours = max(i for i, (_, name, _) in enumerate(disp) if kernel in name)
        copies = [v for _, name, v in disp[last_ours + 1:] if "dpah" not in name and v > 0]
        if not copies:
            return None, f"rocprofv3 --pmc {counter}: no copy dispatch after the advect passes"
        scale = COPY_BYTES / (copies[-1] * 1024.0)
        kib = sum(ours[1:3]) / 2.0                       # the two measured passes
        out[counter] = {"kib_per_dispatch": kib, "scale": scale, "bytes": kib * 1024.0 * scale,
                        "copy_kib": copies[-1]}
    total = out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]
    return total, out


def traffic_child(args) -> int:
    """The workload of run_vortex256 without timing, for a rocprofv3 --pmc run:
    a warm pass, two measured passes, then a device-to-device copy of
    COPY_BYTES that calibrates the byte counters."""
    import ctypes as C
    import dpa_amd
    from dpa_amd import workloads as W
    field = W.vortex_field(256)
    seeds = W.vortex_seeds(args.seeds)
    n = len(seeds)
    ctx = dpa_amd.Context(0)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    d_seeds = ctx.to_device(seeds)
    d_pos, d_rem, d_st = ctx.alloc(12 * n), ctx.alloc(8 * n), ctx.alloc(n)
    for _ in range(3):
        ctx.advect_endpoints_device(fld, d_seeds, n, d_pos, uniform_remaining=args.iterations,
                                    step_size=np.float32(0.25), d_rem_out=d_rem,
                                    d_status_out=d_st, wait=True)
    src, dst = ctx.alloc(COPY_BYTES), ctx.alloc(COPY_BYTES)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    if hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), COPY_BYTES, 3) != 0:   # device to device
        return 1
    if hip.hipDeviceSynchronize() != 0:
        return 1
    for p in (src, dst, d_seeds, d_pos, d_rem, d_st):
        ctx.free(p)
    fld.release()
    ctx.close()
    return 0


# ----------------------------------------------------------------- serving

def serve_trace(ctx, fields, trace, mode, steps_per_tick=0, saturate=False):
    """Replays a trace against a server (results left in device memory, freed
    as they are collected); returns latency / throughput figures.  saturate:
    every request is submitted at once (a backlog the server drains at its
    capacity) instead of at its Poisson arrival time."""
    import dpa_amd
    lat, steps = [], 0
    with dpa_amd.Server(ctx, mode=mode, steps_per_tick=steps_per_tick) as server:
        tickets = []
        done = threading.Event()
        results = {}
        failure = []

        def collector():
            # frees every result as soon as it is ready, in completion order
            try:
                open_ = {}
                seen = 0
                while len(results) < len(trace):
                    while seen < len(tickets):
                        open_[seen] = tickets[seen]
                        seen += 1
                    freed = False
                    for i, t in list(open_.items()):
                        if server.poll(t) >= dpa_amd.TICKET_DONE:
                            server.wait(t, -1).free()
                            results[i] = server.info(t)
                            server.forget(t)
                            del open_[i]
                            freed = True
                    if not freed:
                        time.sleep(0.0002)
            except Exception as e:   # surfaces in the caller
                failure.append(e)
            finally:
                done.set()

        th = threading.Thread(target=collector)
        th.start()
        t0 = time.perf_counter()
        for r in trace:
            if not saturate:
                delay = r["at"] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
            f = fields[r["field"]]
            if r["kind"] == "ftle":
                tickets.append(server.submit_ftle(f, r["origin"], r["spacing"], r["shape"],
                                                  r["delta"], r["iterations"], r["step_size"],
                                                  start_time=r["start_time"]))
            else:
                tickets.append(server.submit_advect(f, r["seeds"],
                                                    uniform_remaining=r["uniform_remaining"],
                                                    step_size=r["step_size"],
                                                    start_time=r["start_time"],
                                                    record=r["record"]))
        done.wait()
        th.join()
        t_end = time.perf_counter()
        if failure:
            raise failure[0]
    for i in range(len(trace)):
        lat.append(results[i]["latency_seconds"])
        steps += results[i]["particle_steps"]
    lat = np.array(lat)
    span = t_end - t0

    def p50(key):
        return float(np.percentile([results[i][key] for i in range(len(trace))], 50) * 1e3)

    return {"requests": len(trace), "particle_steps": int(steps),
            "throughput": steps / span, "makespan_s": span,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "p50_admitted_ms": p50("queue_seconds"), "p50_launched_ms": p50("launch_seconds"),
            "p50_device_done_ms": p50("device_seconds")}


SERVE_MODES = {"batched": "batched", "serial": "serial_library_calls",
               "per-request": "per_request_launch"}


def serving_leg(ctx, fields, trace, saturate=False, modes=tuple(SERVE_MODES)):
    """The trace through the batched server and both comparators (those of
    `modes`), results left in device memory in all of them."""
    import dpa_amd
    code = {"batched": dpa_amd.SERVE_BATCHED, "serial": dpa_amd.SERVE_SERIAL,
            "per-request": dpa_amd.SERVE_PER_REQUEST}
    out = {}
    for m in modes:
        key = SERVE_MODES[m]
        r = serve_trace(ctx, fields, trace, code[m], saturate=saturate)
        out[key] = r
        log(f"{'saturated' if saturate else 'poisson'} {key:22s} {r['throughput']:.3e} steps/s "
            f"makespan {r['makespan_s']:.2f} s p50 {r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms")
    return add_ratios(out)


def add_ratios(out):
    if "batched" in out and "per_request_launch" in out:
        out["batched_vs_per_request"] = (out["batched"]["throughput"] /
                                         max(out["per_request_launch"]["throughput"], 1e-9))
    if "batched" in out and "serial_library_calls" in out:
        out["batched_vs_serial"] = (out["batched"]["throughput"] /
                                    max(out["serial_library_calls"]["throughput"], 1e-9))
    return out


def serving_benchmark(ctx, base, args):
    """Four resident 256^3 x 16-slice time-variant fields (12.9 GB); a Poisson
    trace at args.serve_rate and the same number of requests submitted at once
    (saturation), each through the batched server, the serial library calls
    and the per-request launch form."""
    from dpa_amd import workloads as W
    t = time.perf_counter()
    hosts = W.serving_fields(base)
    fields = [ctx.field_unsteady(h, (0, 0, 0), (1, 1, 1), 0.0, 1.0) for h in hosts]
    del hosts
    log(f"served fields resident in {time.perf_counter() - t:.1f} s")
    try:
        t = time.perf_counter()
        trace = W.serving_trace(args.serve_rate, args.serve_requests)
        log(f"trace of {len(trace)} requests ready in {time.perf_counter() - t:.1f} s")
        out = {"trace": (f"{len(trace)} requests on 4 resident 256^3 x 16-slice time-variant "
                         "vortex fields: 80% recorded pathlines (1k-262k seeds log-uniform, "
                         "128-2048 steps, dt 0.005), 20% FTLE slabs (256x256x16 lattice, "
                         "+-0.5 cell, 200 steps, dt 0.02); results left in device memory "
                         "in every mode"),
               "ftle_requests": sum(r["kind"] == "ftle" for r in trace)}
        # one mode at a time through both traces: the per-request form's
        # hipMalloc / hipFree per request (and the context cache it trims)
        # would otherwise sit between the batched server's two traces
        poisson, saturated = {}, {}
        for m in args.serve_modes:
            poisson.update(serving_leg(ctx, fields, trace, modes=(m,)))
            if args.serve_saturate:
                saturated.update(serving_leg(ctx, fields, trace, saturate=True, modes=(m,)))
        out["poisson"] = dict(rate=args.serve_rate, **add_ratios(poisson))
        if args.serve_saturate:
            out["saturated"] = add_ratios(saturated)
        return out
    finally:
        for f in fields:
            f.release()


# --------------------------------------------------------------- workloads

def run_vortex256(args):
    import dpa_amd
    from dpa_amd import workloads as W

    t = time.perf_counter()
    field = W.vortex_field(256)
    seeds = W.vortex_seeds(args.seeds)
    log(f"inputs ready in {time.perf_counter() - t:.1f} s")
    dt = np.float32(0.25)
    n = len(seeds)

    ctx = make_context(args)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    d_seeds = ctx.to_device(seeds)
    d_pos = ctx.alloc(seeds.nbytes)
    d_rem = ctx.alloc(8 * n)
    d_st = ctx.alloc(n)

    def one_pass(wait: bool):
        return ctx.advect_endpoints_device(fld, d_seeds, n, d_pos,
                                           uniform_remaining=args.iterations, step_size=dt,
                                           d_rem_out=d_rem, d_status_out=d_st, wait=wait)

    tuned = None
    if args.autotune:
        # the autotuner times its candidates on this workload itself (a quarter
        # of the seeds picked the wrong queue order: ordering pays off with the
        # L2 pressure of the full set); untimed, like the warmup
        t = time.perf_counter()
        tuned, tuned_ms = ctx.autotune_endpoints(fld, d_seeds, n, args.iterations, dt)
        log(f"autotuned in {time.perf_counter() - t:.2f} s: {tuned} ({tuned_ms:.3f} ms per pass)")

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
    launches, total_ms = ctx.kernel_timing("advect_persistent")
    last = one_pass(True)
    if last != steps_per_pass:
        raise RuntimeError(f"step count changed between passes: {last} != {steps_per_pass}")
    avg_ms = total_ms / max(launches, 1)
    alg_bytes, rec_loads, node_loads = counted_bytes(ctx, one_pass, n, steps_per_pass)
    log(f"cell loads per pass: {rec_loads} records, {node_loads} node groups "
        f"({rec_loads / steps_per_pass:.3f} per step): {alg_bytes / 1e9:.2f} GB "
        f"= {alg_bytes / (avg_ms * 1e-3) / 1e9:.0f} GB/s")
    value = args.steps * steps_per_pass / elapsed
    log(f"timed {args.steps} passes in {elapsed:.3f} s: {value:.4e} particle-steps/s; "
        f"kernel avg {avg_ms:.3f} ms over {launches} launches")
    ol, oms = ctx.kernel_timing("order_queue")
    if ol:
        log(f"queue ordering avg {oms / ol:.3f} ms over {ol} launches")
    # the served-query leg comes before the hipGraph leg: run after a graph had
    # been instantiated and replayed in the process, its 500 req/s trace
    # measured p99 342 ms instead of 212 ms and 1.25e11 steps/s instead of 1.32e11
    serving = None
    if not args.no_serve:
        serving = serving_benchmark(ctx, field, args)
    # the same passes replayed from a hipGraph captured once
    graph = None
    if args.graph:
        with ctx.pass_endpoints(fld, d_seeds, n, d_pos, uniform_remaining=args.iterations,
                                step_size=dt, d_rem_out=d_rem, d_status_out=d_st) as gp:
            gp.launch()
            if gp.total_steps() != steps_per_pass:
                raise RuntimeError("graph replay took a different number of steps")
            ctx.synchronize()
            tg = time.perf_counter()
            for _ in range(args.steps):
                gp.launch()
            ctx.synchronize()
            g_elapsed = time.perf_counter() - tg
            if gp.total_steps() != steps_per_pass:
                raise RuntimeError("graph replay took a different number of steps")
        graph = {"particle_steps_per_s": args.steps * steps_per_pass / g_elapsed,
                 "ms_per_pass": g_elapsed / args.steps * 1e3}
        log(f"graph replays: {graph['particle_steps_per_s']:.4e} particle-steps/s "
            f"({graph['ms_per_pass']:.3f} ms per pass)")
    for p in (d_seeds, d_pos, d_rem, d_st):
        ctx.free(p)
    ctx_tuning = ctx.tuning()

    fld.release()
    ctx.close()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (CPU restatements on bounded samples)")
        cpu = cpu_baseline_endpoints(field, (0, 0, 0), (1, 1, 1), seeds, args.iterations, dt,
                                     args.cpu_seconds, "2^20")
        log(f"cpu: dpa-structured {cpu['value']:.4e} particle-steps/s on {cpu['cores']} threads, "
            f"plain {cpu['plain_port']['value']:.4e}, one thread {cpu['one_thread']['value']:.4e}")

    traffic, tsrc = (None, "not measured (--no-traffic)")
    if args.traffic:
        traffic, tsrc = measure_traffic(args)
        if traffic is None:
            log(f"traffic not measured: {tsrc}")
        else:
            log(f"traffic past the L2: {traffic / 1e9:.2f} GB per launch "
                f"({traffic / (avg_ms * 1e-3) / 1e9:.0f} GB/s)")
    rf = roofline(avg_ms, alg_bytes, steps_per_pass, traffic=traffic)
    rf["kernel"] = "k_advect_persistent"
    rf["cell_record_loads"] = rec_loads
    rf["node_group_loads"] = node_loads
    rf["traffic_source"] = tsrc
    out = base_line(args, value, elapsed, "f32")
    out["config"] = {
        "workload": "large steady endpoint advection: 256^3 f32 Gaussian-vortex field "
                    "(64 tubes + drift 0.1), 2^20 uniform seeds, dpa runge_kutta_4 "
                    "dt 0.25 x 1000 iterations, endpoints only, exits retire",
        "seeds": n, "iterations": args.iterations, "particle_steps_per_pass": steps_per_pass,
        "tuning": tuned or ctx_tuning}
    out["roofline"] = rf
    out["cpu_baseline"] = cpu
    out["serving"] = serving
    out["graph_replay"] = graph
    return out


def make_context(args):
    """Context with kernel timing and any launch-shape overrides from the command line."""
    import dpa_amd
    ctx = dpa_amd.Context(0, kernel_timing=True)
    over = {k: v for k, v in (("order", args.order), ("blocks_per_cu", args.blocks_per_cu),
                              ("block", args.block), ("refill", args.refill),
                              ("cells", args.cells), ("fastdiv", args.fastdiv),
                              ("cell_brick", args.cell_brick), ("order_brick", args.order_brick),
                              ("lattice_brick", args.lattice_brick),
                              ("lattice_group", args.lattice_group),
                              ("node_brick", args.node_brick))
            if v is not None}
    if over:
        ctx.set_tuning(**over)
        args.autotune = False   # explicit launch shapes are measured as given
    log(f"tuning {ctx.tuning()}")
    return ctx


def base_line(args, value, elapsed, dtype):
    return {"metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (seeded fields and seeds; no dataset)"}


def run_abc64(args):
    """ABC 64^3, 4096 seeds x 256 steps, record + prune + indices + colours."""
    import dpa_amd
    from dpa_amd import workloads as W
    c = W.abc64_streamlines()
    ctx = make_context(args)
    fld = ctx.field(c["field"], c["offset"], c["spacing"])

    def one():
        r = ctx.advect(fld, c["seeds"], uniform_remaining=c["iterations"],
                       particles_per_round=c["particles_per_round"], step_size=c["step_size"],
                       record=True)
        s = r.total_steps
        r.free()
        return s

    steps = None
    for _ in range(max(args.warmup, 1)):
        steps = one()
    ctx.kernel_timing_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one()
    elapsed = time.perf_counter() - t0
    na, ta = recorded_advect_timing(ctx)
    nc, tc = ctx.kernel_timing("compact_write")
    n = len(c["seeds"])
    # cell loads: the recorded particles follow the same trajectories as an
    # endpoint pass over the same seeds, which the counting instance runs
    rec, node = endpoint_loads(ctx, fld, [(c["seeds"], c["iterations"], c["step_size"])], steps)
    fld.release()
    ctx.close()
    value = args.steps * steps / elapsed
    adv_ms = ta / max(na, 1)
    # + every recorded vertex (steps, seeds, terminals) written once, 12 bytes
    alg = rec * RECORD_BYTES + node * NODE_GROUP_BYTES + n * PARTICLE_IO_BYTES + 12 * (steps + 2 * n)
    out = base_line(args, value, elapsed, "f32")
    out["config"] = {"workload": "ABC 64^3, 4096 seeds x 256 steps dt 0.05 h, record, u32; "
                                 "value includes host upload / download and allocation",
                     "particle_steps_per_pass": steps}
    out["kernels"] = {"advect_ms": adv_ms, "compact_write_ms": tc / max(nc, 1),
                      "device_steps_per_s": steps / (adv_ms * 1e-3)}
    out["roofline"] = roofline(adv_ms, alg, steps)
    # a single request's recorded round runs one thread per row
    out["roofline"]["kernel"] = "k_advect_simple"
    if not args.no_cpu:
        from oracle import py_oracle as po
        th = host_threads()
        t = time.perf_counter()
        r = po.run_rounds(po.Field(c["field"], c["offset"], c["spacing"]), c["seeds"],
                          c["iterations"], 4096, c["step_size"], record=True, threads=th)
        el = time.perf_counter() - t
        out["cpu_baseline"] = {"value": r["total_steps"] / el, "unit": "particle-steps/s",
                               "cores": th, "kind": "port",
                               "sample": "whole workload incl. prune / indices / colours"}
    return out


def run_ftle(args, dpa_mode=False):
    """FTLE on ABC 512x512x128: auxiliary-neighbour form (default) or dpa's
    lattice estimator."""
    import dpa_amd
    from dpa_amd import workloads as W
    c = W.ftle_abc512()
    ctx = make_context(args)
    fld = ctx.field(c["field"], c["offset"], c["spacing"])
    shape = c["lattice_shape"]
    points = int(np.prod(shape))
    d_out = ctx.alloc(4 * points)
    tuned = None
    if args.autotune:
        t = time.perf_counter()
        if dpa_mode:
            tuned, tms = ctx.autotune_lattice(fld, c["offset"], c["spacing"],
                                              c["field"].shape[:3], c["iterations"],
                                              c["step_size"])
        else:
            tuned, tms = ctx.autotune_lattice(fld, c["lattice_origin"], c["lattice_spacing"],
                                              shape, c["iterations"], c["step_size"],
                                              delta=c["delta"])
        log(f"autotuned in {time.perf_counter() - t:.1f} s: {tuned} ({tms:.3f} ms)")

    # dpa's stage returns the volume to host memory: one array, reused by
    # every pass, as a caller writing a volume per round would keep it
    volume = np.empty(tuple(shape), np.float32) if dpa_mode else None

    def one():
        if dpa_mode:
            v, s = ctx.ftle_lattice(fld, (1.0, 1.0, 1.0), c["iterations"], c["step_size"],
                                    out=volume)
            return s
        return ctx.ftle_aux_device(fld, c["lattice_origin"], c["lattice_spacing"], shape,
                                   c["delta"], c["iterations"], c["step_size"], d_out)

    steps = None
    # the first two passes of a process take longer (dpa-lattice passes of
    # 24.0 and 19.6 ms before a steady 12.5 ms; the second one's download
    # waited 6.6 ms inside its first hipMemcpyAsync), so at least three warm up
    for _ in range(max(args.warmup, 3)):
        steps = one()
    ctx.kernel_timing_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    na, ta = ctx.kernel_timing("advect_persistent")
    ctx.set_count_cell_loads(True)
    try:
        if one() != steps:
            raise RuntimeError("the load-counting pass took a different number of steps")
        rec, node = ctx.cell_loads()
    finally:
        ctx.set_count_cell_loads(False)
    particles = points if dpa_mode else dpa_amd.aux_particles(c["lattice_spacing"], c["delta"],
                                                              shape)
    # loads + each particle's final position (lattice seeds are computed, not read)
    alg = rec * RECORD_BYTES + node * NODE_GROUP_BYTES + 12 * particles
    log(f"cell loads per pass: {rec} records, {node} node groups ({(rec + node) / steps:.3f} "
        "per step)")
    ctx.free(d_out)
    fld.release()
    ctx.close()
    adv_ms = ta / max(na, 1)
    out = base_line(args, args.steps * steps / elapsed, elapsed, "f32")
    out["config"] = {"workload": ("FTLE dpa lattice mode" if dpa_mode else
                                  "FTLE auxiliary +-0.5 cell neighbours") +
                                 ", ABC 512x512x128, T=4 at dt 0.02",
                     "lattice_points": points, "particles_per_pass": particles,
                     "particle_steps_per_pass": steps,
                     "points_per_s": args.steps * points / elapsed,
                     "tuning": tuned or "library defaults"}
    out["roofline"] = roofline(adv_ms, alg, steps)
    out["roofline"].update(kernel="k_advect_persistent", cell_record_loads=rec,
                           node_group_loads=node)
    if out["roofline"]["frac"] > 1:
        # neighbouring lattice particles reload the same node groups, so most
        # of these loads hit in the L2 (hit rate about 0.93 under rocprofv3):
        # the count is what the lanes ask for, not what crosses from HBM, and
        # the launch is bound by VALU issue instead
        out["roofline"]["note"] = ("node-group loads mostly hit in the L2; frac above 1 is "
                                   "L2-served bytes, the launch is VALU-issue bound")
    return out


def run_batch(args):
    """64 uneven recorded jobs on the vortex field: one persistent launch
    (dpah_advect_batch) against the per-request form on the same jobs."""
    import dpa_amd
    from dpa_amd import workloads as W
    field = W.vortex_field(256)
    jobs = W.batch_jobs()
    ctx = make_context(args)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    for j in jobs:
        j["field"] = fld
    # per-request form first, while the context's block cache is empty: each
    # job alone, hipMalloc / one thread per particle / hipFree (a cache
    # filled by the tuner's candidates would otherwise be trimmed inside
    # this leg's timed region)
    with dpa_amd.Server(ctx, mode=dpa_amd.SERVE_PER_REQUEST) as server:
        # untimed warmup on the smallest job (module load, first allocations)
        j = min(jobs, key=lambda j: len(j["seeds"]))
        server.wait(server.submit_advect(fld, j["seeds"], uniform_remaining=j["uniform_remaining"],
                                         step_size=j["step_size"], record=True), -1).free()
    t0 = time.perf_counter()
    pr_steps = 0
    with dpa_amd.Server(ctx, mode=dpa_amd.SERVE_PER_REQUEST) as server:
        ts = [server.submit_advect(fld, j["seeds"], uniform_remaining=j["uniform_remaining"],
                                   step_size=j["step_size"], record=True) for j in jobs]
        for t in ts:
            r = server.wait(t, -1)
            pr_steps += r.total_steps
            r.free()
    per_req = time.perf_counter() - t0
    log(f"per-request: {pr_steps / per_req:.3e} steps/s")
    tuned = None
    if args.autotune:
        t = time.perf_counter()
        tuned, tms = ctx.autotune_batch(jobs)
        log(f"autotuned in {time.perf_counter() - t:.1f} s: {tuned} ({tms:.3f} ms)")

    def batched():
        res = ctx.advect_batch(jobs)
        s = sum(r.total_steps for r in res)
        for r in res:
            r.free()
        return s

    steps = batched()                    # warmup
    ctx.kernel_timing_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batched()
    elapsed = time.perf_counter() - t0
    na, ta = recorded_advect_timing(ctx)
    nc, tc = ctx.kernel_timing("compact_write")
    log(f"batched: {args.steps * steps / elapsed:.3e} steps/s")
    graph = None
    if args.graph:
        # the same batch captured once (dpah_batch_pass_*): replay + results per pass
        with ctx.batch_pass(jobs) as bp:
            def replayed():
                bp.launch()
                res = bp.results()
                s = sum(r.total_steps for r in res)
                for r in res:
                    r.free()
                return s
            gsteps = replayed()
            assert gsteps == steps, (gsteps, steps)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                replayed()
            gel = time.perf_counter() - t0
        graph = {"value": args.steps * steps / gel, "ms_per_pass": 1e3 * gel / args.steps}
        log(f"batched, graph replays: {graph['value']:.3e} steps/s")
    if pr_steps != steps:
        raise RuntimeError(f"per-request form took {pr_steps} steps, the batch {steps}")
    rec, node = endpoint_loads(ctx, fld, [(j["seeds"], j["uniform_remaining"], j["step_size"])
                                          for j in jobs], steps)
    nseeds = int(sum(len(j["seeds"]) for j in jobs))
    alg = (rec * RECORD_BYTES + node * NODE_GROUP_BYTES + nseeds * PARTICLE_IO_BYTES
           + 12 * (steps + 2 * nseeds))
    fld.release()
    ctx.close()
    value = args.steps * steps / elapsed
    out = base_line(args, value, elapsed, "f32")
    out["config"] = {"workload": "64 recorded jobs (256..262144 seeds, 64..4096 steps) on "
                                 "the 256^3 vortex field, one persistent launch + compaction; "
                                 "host upload / download included",
                     "jobs": len(jobs), "seeds": int(sum(len(j['seeds']) for j in jobs)),
                     "particle_steps_per_pass": steps, "tuning": tuned or "library defaults"}
    out["kernels"] = {"advect_ms": ta / max(na, 1),
                      "compact_write_ms_total_per_pass": tc / max(args.steps, 1)}
    if graph:
        out["graph_replay"] = graph
    out["per_request_launch"] = {"value": steps / per_req, "seconds": per_req}
    out["vs_per_request"] = value / (steps / per_req)
    out["roofline"] = roofline(ta / max(na, 1), alg, steps)
    out["roofline"].update(kernel="k_advect_persistent", cell_record_loads=rec,
                           node_group_loads=node)
    return out


def run_serve(args):
    """4 resident time-variant fields (16 slices of 256^3), Poisson traces at
    several rates, continuous batching vs per-request launch."""
    import dpa_amd
    from dpa_amd import workloads as W
    base = W.vortex_field(256)
    ctx = make_context(args)
    fields = [ctx.field_unsteady(h, (0, 0, 0), (1, 1, 1), 0.0, 1.0)
              for h in W.serving_fields(base)]
    rates = [float(r) for r in args.rates.split(",")]
    sweep = []
    for rate in rates:
        trace = W.serving_trace(rate, args.serve_requests)
        leg = serving_leg(ctx, fields, trace, modes=args.serve_modes)
        leg["rate"] = rate
        sweep.append(leg)
    if args.serve_saturate:
        leg = serving_leg(ctx, fields, W.serving_trace(rates[-1], args.serve_requests),
                          saturate=True, modes=args.serve_modes)
        leg["rate"] = "saturated"
        sweep.append(leg)
    for f in fields:
        f.release()
    ctx.close()
    best = max(max(v["throughput"] for k, v in s.items() if isinstance(v, dict))
               for s in sweep)
    out = base_line(args, best, 0.0, "f32")
    out["steps"] = 0
    out["config"] = {"workload": "served pathlines + FTLE slabs on 4 resident 256^3 x 16 "
                                 "time-variant fields (12.9 GB)", "requests_per_rate": args.serve_requests}
    out["sweep"] = sweep
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="vortex256",
                    choices=["abc64-record", "vortex256", "ftle-aux", "ftle-lattice", "batch64",
                             "serve"])
    ap.add_argument("--seeds", type=int, default=1 << 20)
    ap.add_argument("--iterations", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", dest="traffic", action="store_false",
                    help="skip the two rocprofv3 --pmc runs that measure the headline "
                         "kernel's bytes past the L2")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=500.0)
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--no-serve-saturate", dest="serve_saturate", action="store_false",
                    help="skip the saturation trace (every request submitted at once)")
    ap.add_argument("--serve-modes", type=lambda s: tuple(s.split(",")),
                    default=tuple(SERVE_MODES),
                    help="serving modes to run: batched,serial,per-request (default all)")
    ap.add_argument("--rates", default="50,100,200,500")
    ap.add_argument("--order", type=int, choices=[0, 1], default=None,
                    help="spatially ordered, XCD-split work queue (default: library's)")
    ap.add_argument("--blocks-per-cu", type=int, default=None)
    ap.add_argument("--block", type=int, default=None)
    ap.add_argument("--refill", type=int, default=None)
    ap.add_argument("--cells", type=int, choices=[0, 1, 2], default=None)
    ap.add_argument("--fastdiv", type=int, choices=[0, 1], default=None)
    ap.add_argument("--order-brick", type=int, choices=[0, 1, 2, 3, 4, 5], default=None,
                    help="queue ordering bricks of at least 2^k cells per axis (0: 4 cells)")
    ap.add_argument("--cell-brick", type=int, choices=[0, 2, 3, 4, 5], default=None,
                    help="cell layout record order of the fields (0: C order; k: 2^k-cell bricks)")
    ap.add_argument("--lattice-brick", type=int, choices=[0, 2, 4, 8], default=None,
                    help="whole-lattice launches: points per brick edge")
    ap.add_argument("--lattice-group", type=int, choices=[0, 1, 2, 4, 8, 16], default=None,
                    help="whole-lattice launches: bricks per group edge")
    ap.add_argument("--node-brick", type=int, choices=[0, 1, 2], default=None,
                    help="1: steady fields also keep their nodes in 4^3-node bricks (read by "
                         "launches that sample nodes); 2: time-variant fields too; 0: C-order "
                         "nodes only")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="skip the hipGraph replay leg")
    ap.add_argument("--autotune", dest="autotune", action="store_true", default=False,
                    help="run the launch-shape autotuner on the workload before the timed "
                         "passes (off by default, so that every dispatch of the advect kernel "
                         "in a profile of the default command is a measured pass)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="the default: time the library's launch shapes as they are")
    args = ap.parse_args()
    if args.traffic_child:
        return traffic_child(args)
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path here (with the field "
              "replicated, seeds would shard across GPUs with no exchange)", file=sys.stderr)
        return 2
    run = {"abc64-record": run_abc64, "vortex256": run_vortex256, "ftle-aux": run_ftle,
           "ftle-lattice": lambda a: run_ftle(a, dpa_mode=True), "batch64": run_batch,
           "serve": run_serve}[args.workload]
    out = run(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
