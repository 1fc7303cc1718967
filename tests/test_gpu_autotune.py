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


def batch_arrays(results):
    out = []
    for res in results:
        rounds = [res.round(r) for r in range(res.round_count)]
        out.append((rounds, res.final_state(), res.total_steps))
        res.free()
    return out


def test_autotune_batch_keeps_results(tmp_path):
    # 2^16 recorded rows in all: large enough that the default stages slot
    # blocks in LDS, so the candidates cover both slot layouts
    say("case autotune batch")
    cache = tmp_path / "tune.txt"
    field = W.vortex_field(64, count=24)
    rng = np.random.default_rng(56)
    ctx = dpa_amd.Context(0, autotune_cache=cache)
    try:
        fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
        jobs = []
        for k in range(9):
            n = [12000, 300, 9000, 40, 15000, 2500, 8000, 700, 20000][k]
            lo = rng.uniform(0, 40, 3)
            jobs.append(dict(field=fld, seeds=W.uniform_seeds(n, lo, lo + rng.uniform(4, 20, 3),
                                                              seed_offset=600 + k),
                             uniform_remaining=int(rng.integers(16, 300)), step_size=f32(0.25),
                             record=True, particles_per_round=max(1, n // 2)))
        before = batch_arrays(ctx.advect_batch(jobs))
        best, ms = ctx.autotune_batch(jobs)
        say(f"  best {best} {ms:.3f} ms")
        assert ms > 0 and best["slot_stage"] in (1, 2, 3) and best["refill"] in (0, 16, 64)
        after = batch_arrays(ctx.advect_batch(jobs))
        for (ra, sa, ta), (rb, sb, tb) in zip(before, after):
            assert ta == tb and len(ra) == len(rb)
            for (va, ca, ia), (vb, cb, ib) in zip(ra, rb):
                assert np.array_equal(bits(va), bits(vb)) and np.array_equal(ia, ib)
                assert np.array_equal(bits(ca), bits(cb))
            for x, y in zip(sa, sb):
                assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
        # the other recorded launch forms, set explicitly, give the same rows
        # as well (persistent with one-slot runs or staged blocks, one thread
        # per row)
        for stage in {1, 2, 3} - {best["slot_stage"]}:
            ctx.set_tuning(slot_stage=stage)
            other = batch_arrays(ctx.advect_batch(jobs))
            for (ra, _, ta), (rb, _, tb) in zip(before, other):
                assert ta == tb
                for (va, _, ia), (vb, _, ib) in zip(ra, rb):
                    assert np.array_equal(bits(va), bits(vb)) and np.array_equal(ia, ib)
        fld.release()
    finally:
        ctx.close()
    assert "64x64x64x1:i0:s0:r1" in cache.read_text()


@pytest.mark.parametrize("aux", [True, False])
def test_autotune_lattice_keeps_results(ctx, aux):
    say(f"case autotune lattice, aux={aux}")
    import math
    field, h = W.abc_field((40, 36, 28), spacing=(2 * math.pi / 39,) * 3)
    off = np.zeros(3, f32)
    iters, dt = 80, f32(0.02)
    shape = field.shape[:3]
    n = int(np.prod(shape))
    delta = (f32(0.5) * h).astype(f32) if aux else None
    fld = ctx.field(field, off, h)
    before_t = ctx.tuning()

    def run():
        if not aux:
            return ctx.ftle_lattice(fld, (1.0, 1.0, 1.0), iters, dt)
        d = ctx.alloc(4 * n)
        try:
            s = ctx.ftle_aux_device(fld, off, h, shape, delta, iters, dt, d)
            return ctx.to_host(d, np.empty(n, f32)), s
        finally:
            ctx.free(d)

    try:
        vol0, s0 = run()
        best, ms = ctx.autotune_lattice(fld, off, h, shape, iters, dt, delta=delta)
        say(f"  best {best} {ms:.3f} ms")
        assert ms > 0 and best["lattice_brick"] in (0, 2, 4, 8)
        assert best["order"] == (1 if best["lattice_brick"] else 0)
        vol1, s1 = run()
        assert s0 == s1 and np.array_equal(bits(vol0), bits(vol1))
    finally:
        ctx.set_tuning(**before_t)
        fld.release()
