"""The spatially ordered, XCD-split work queue of the persistent kernel.

Ordering changes which lane runs which particle when, never a result: with the
ordering on, off, on one queue part or eight, and at different launch shapes,
every output must equal the oracle's bit for bit.  Sizes are above the
library's ordering threshold (2^15 particles) so the ordering kernels run.
"""
import sys

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_parity import bits, compare_advect

pytestmark = pytest.mark.gpu
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vortex(ctx):
    field = W.vortex_field(64, count=24)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    yield field, fld
    fld.release()


def seeds_with_strays(n, seed_offset):
    s = W.uniform_seeds(n, (0, 0, 0), (63, 63, 63), seed_offset=seed_offset)
    s[::997] = (70.0, 1.0, 1.0)             # outside the grid: last key
    s[5::1999] = (np.nan, 3.0, 3.0)         # NaN: last key, exits at once
    s[7::1503] = (-0.5, 10.0, 10.0)         # negative cell
    return s


def endpoints(ctx, fld, seeds, iters, dt):
    n = len(seeds)
    d_s = ctx.to_device(seeds)
    d_p = ctx.alloc(12 * n)
    d_r = ctx.alloc(8 * n)
    d_t = ctx.alloc(n)
    try:
        steps = ctx.advect_endpoints_device(fld, d_s, n, d_p, uniform_remaining=iters,
                                            step_size=dt, d_rem_out=d_r, d_status_out=d_t)
        return (ctx.to_host(d_p, np.empty((n, 3), f32)), ctx.to_host(d_r, np.empty(n, np.uint64)),
                ctx.to_host(d_t, np.empty(n, np.uint8)), steps)
    finally:
        for p in (d_s, d_p, d_r, d_t):
            ctx.free(p)


@pytest.mark.parametrize("tuning", [dict(order=1, parts=8), dict(order=0, parts=8),
                                    dict(order=1, parts=1), dict(order=1, parts=8, blocks_per_cu=2),
                                    dict(order=1, parts=3, block=128, blocks_per_cu=5),
                                    dict(order=1, parts=8, order_brick=3),
                                    dict(order=1, parts=8, order_brick=5, block=64)])
def test_ordered_endpoints_match_oracle(ctx, vortex, tuning):
    field, fld = vortex
    say(f"case ordered endpoints {tuning}")
    seeds = seeds_with_strays(50000, 41)
    dt, iters = f32(0.25), 200
    ora = po.advect_endpoints(po.Field(field, (0, 0, 0), (1, 1, 1)), seeds, iters, dt)
    before = ctx.tuning()
    ctx.set_tuning(**tuning)
    try:
        pos, rem, st, steps = endpoints(ctx, fld, seeds, iters, dt)
    finally:
        ctx.set_tuning(**before)
    assert np.array_equal(bits(pos), bits(ora["position"]))
    assert np.array_equal(rem, ora["remaining"])
    assert np.array_equal(st, ora["status"])
    assert steps == ora["total_steps"]


def test_ordered_recorded_rounds_match_oracle(ctx, vortex):
    # recorded launches keep seed order even with ordering on (their slot rows
    # are written in seed order); the rounds must still match the oracle
    field, fld = vortex
    say("case recorded rounds with ordering on")
    seeds = seeds_with_strays(40000, 43)
    dt, iters = f32(0.25), 40
    rem = np.full(len(seeds), iters, np.uint64)
    rem[::7] = 13
    ora = po.run_rounds(po.Field(field, (0, 0, 0), (1, 1, 1)), seeds, rem, 36000, dt)
    before = ctx.tuning()
    ctx.set_tuning(order=1)
    try:
        res = ctx.advect(fld, seeds, remaining=rem, particles_per_round=36000, step_size=dt)
        try:
            compare_advect(res, ora, len(seeds))
        finally:
            res.free()
    finally:
        ctx.set_tuning(**before)


def test_set_tuning_rejects_bad_shapes(ctx):
    for bad in (dict(block=96), dict(block=512), dict(blocks_per_cu=-1), dict(blocks_per_cu=33),
                dict(order=2),
                dict(parts=9), dict(parts=0), dict(cell_brick=1), dict(cell_brick=6),
                dict(order_brick=-1), dict(order_brick=6), dict(lattice_group=3),
                dict(lattice_group=32), dict(node_brick=3), dict(node_brick=-1)):
        with pytest.raises(dpa_amd.DpahError):
            ctx.set_tuning(**bad)


def test_ordered_lattice_ftle_same_as_unordered(ctx, vortex):
    # dpa-mode FTLE seeds a plain lattice (64^3 = 262144 points, ordered when
    # ordering is on); the volume and the step count must not depend on it
    field, fld = vortex
    say("case lattice ftle ordered vs unordered")
    out = []
    before = ctx.tuning()
    try:
        for order in (1, 0):
            ctx.set_tuning(order=order)
            vol, steps = ctx.ftle_lattice(fld, (1, 1, 1), 60, f32(0.25))
            out.append((vol, steps))
    finally:
        ctx.set_tuning(**before)
    assert out[0][1] == out[1][1] and out[0][1] > 1000000
    assert np.array_equal(bits(out[0][0]), bits(out[1][0]))
