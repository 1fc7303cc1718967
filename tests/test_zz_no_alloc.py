"""The allocation-free endpoint entry point (dpah_advect_endpoints_device_noalloc).

It must give the results of dpah_advect_endpoints_device, bit for bit: final
positions, remaining iterations, statuses and the particle-step total.  Its
scratch is one fixed per-GPU array that keeps the last call's contents, so
calls run back to back, the later ones smaller (stale queue heads, keys and
histogram counts of a larger call must not leak into a smaller one), on both
sides of the ordering threshold (2^15 particles) and with per-particle
remaining iterations.  The noalloc call runs first in each pair, so the test
also holds in a fresh process where nothing else has run yet.
"""
import sys

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W

pytestmark = pytest.mark.gpu
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(ctx, fld, seeds, rem, iters, dt, noalloc):
    n = len(seeds)
    d_s = ctx.to_device(seeds)
    d_rem = ctx.to_device(rem) if rem is not None else None
    d_p, d_r, d_t = ctx.alloc(max(12 * n, 1)), ctx.alloc(max(8 * n, 1)), ctx.alloc(max(n, 1))
    try:
        fn = ctx.advect_endpoints_device_noalloc if noalloc else ctx.advect_endpoints_device
        steps = fn(fld, d_s, n, d_p, d_remaining=d_rem, uniform_remaining=iters, step_size=dt,
                   d_rem_out=d_r, d_status_out=d_t)
        return (ctx.to_host(d_p, np.empty((n, 3), f32)), ctx.to_host(d_r, np.empty(n, np.uint64)),
                ctx.to_host(d_t, np.empty(n, np.uint8)), steps)
    finally:
        for p in (d_s, d_p, d_r, d_t) + ((d_rem,) if d_rem is not None else ()):
            ctx.free(p)


def test_noalloc_equals_endpoints(ctx):
    field = W.vortex_field(64, count=24)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    try:
        big = W.uniform_seeds(150_000, (0, 0, 0), (63, 63, 63), seed_offset=31)
        big[::997] = (70.0, 1.0, 1.0)                 # outside: the ordering's last key
        rng = np.random.default_rng(32)
        cases = [("150000 ordered, uniform", big, None, 300),
                 ("40000 ordered, per-particle remaining", big[:40_000],
                  rng.integers(0, 400, 40_000).astype(np.uint64), 0),
                 ("5000 unordered", big[:5000], None, 250),
                 ("0 particles", big[:0], None, 10)]
        for name, seeds, rem, iters in cases:
            say(f"noalloc case: {name}")
            a = run(ctx, fld, seeds, rem, iters, f32(0.25), noalloc=True)
            b = run(ctx, fld, seeds, rem, iters, f32(0.25), noalloc=False)
            assert a[3] == b[3], (name, a[3], b[3])
            assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)), name
            assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), name
            if len(seeds):
                assert a[3] > 0
    finally:
        fld.release()


def test_noalloc_rejects_beyond_scratch(ctx):
    field = np.ones((8, 8, 8, 3), f32)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    d = ctx.alloc(12 * 16)
    try:
        # rejected on the host before any launch: the buffers are never touched
        with pytest.raises(dpa_amd.DpahError, match="DPAH_NOALLOC_MAX_PARTICLES"):
            ctx.advect_endpoints_device_noalloc(fld, d, dpa_amd.MAX_NOALLOC_PARTICLES + 1, d,
                                                uniform_remaining=4, step_size=0.1)
    finally:
        ctx.free(d)
        fld.release()
