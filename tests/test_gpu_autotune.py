"""The launch-shape autotuner.

It times candidate tunings on a sample and keeps the fastest for that launch
key; the choice must not change any result (the oracle still agrees), must be
written to the context's cache file and be read back by a new context.
"""
import sys

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_parity import bits

pytestmark = pytest.mark.gpu
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def endpoints(ctx, fld, seeds, iters, dt):
    n = len(seeds)
    d_s = ctx.to_device(seeds)
    d_p = ctx.alloc(12 * n)
    try:
        steps = ctx.advect_endpoints_device(fld, d_s, n, d_p, uniform_remaining=iters, step_size=dt)
        return ctx.to_host(d_p, np.empty((n, 3), f32)), steps
    finally:
        ctx.free(d_s)
        ctx.free(d_p)


def test_autotune_keeps_results_and_persists(tmp_path):
    say("case autotune")
    cache = tmp_path / "tune.txt"
    field = W.vortex_field(64, count=24)
    seeds = W.uniform_seeds(60000, (0, 0, 0), (63, 63, 63), seed_offset=55)
    dt, iters = f32(0.25), 150
    ora = po.advect_endpoints(po.Field(field, (0, 0, 0), (1, 1, 1)), seeds, iters, dt)
    ctx = dpa_amd.Context(0, autotune_cache=cache)
    try:
        fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
        d_sample = ctx.to_device(seeds[::3].copy())
        try:
            best, ms = ctx.autotune_endpoints(fld, d_sample, len(seeds[::3]), iters, dt)
        finally:
            ctx.free(d_sample)
        say(f"  best {best} {ms:.3f} ms")
        assert ms > 0 and 1 <= best["refill"] <= 64 and best["block"] in (128, 256)
        assert best["block"] * best["blocks_per_cu"] // 64 in (12, 16, 20, 24, 32)
        assert best["order_brick"] in (2, 3)
        pos, steps = endpoints(ctx, fld, seeds, iters, dt)
        assert np.array_equal(bits(pos), bits(ora["position"])) and steps == ora["total_steps"]
        fld.release()
    finally:
        ctx.close()
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and "64x64x64x1:i0:s0:r0" in lines[0]
    ctx2 = dpa_amd.Context(0, autotune_cache=cache)
    try:
        fld = ctx2.field(field, (0, 0, 0), (1, 1, 1))
        pos, steps = endpoints(ctx2, fld, seeds, iters, dt)
        assert np.array_equal(bits(pos), bits(ora["position"])) and steps == ora["total_steps"]
        fld.release()
    finally:
        ctx2.close()
