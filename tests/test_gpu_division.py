"""The verified reciprocal division.

On axes where field creation found it equal, bit for bit, to the correctly
rounded quotient over the whole axis range, contains/interpolate divide by a
spacing with its reciprocal and one exact-residual correction.  The checker
must find no mismatch for power-of-two spacings (the reciprocal is exact) and
must find mismatches for a reciprocal that is off; advection results must equal
the oracle's with the fast form on and off.
"""
import math
import sys

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_parity import compare_advect

pytestmark = pytest.mark.gpu
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_checker_on_exact_and_wrong_reciprocals(ctx):
    say("case division checker")
    for h in (1.0, 0.5, 4.0):
        assert ctx.verify_division(h, 1.0 / h, 600.0) == 0
    h = f32(0.75)
    r = f32(1.0) / h
    wrong = r * f32(1 + 2.0 ** -10)          # ~2^14 ulps high: one correction cannot recover
    assert ctx.verify_division(float(h), float(wrong), 60.0) > 1000000
    for h in (f32(0.75), f32(2 * math.pi / 511), f32(0.1)):
        n = ctx.verify_division(float(h), float(f32(1) / h), float(h) * 1022)
        say(f"  h={float(h):.9g}: {n} mismatches")


# every axis verified (reciprocal with one correction); power-of-two and verified
# axes mixed (the same form); every axis a power of two (exact reciprocals, fmod
# as one fmaf).  With fastdiv 0 every launch divides.
SPACINGS = [(0.75, float(f32(0.1)), float(f32(2 * math.pi / 511))), (0.5, 0.75, 1.0),
            (0.5, 2.0, 0.25)]


@pytest.mark.parametrize("spacing", SPACINGS)
@pytest.mark.parametrize("fastdiv", [0, 1])
def test_advect_matches_oracle_with_and_without(ctx, fastdiv, spacing):
    say(f"case fastdiv={fastdiv} spacing={spacing}")
    offset = (0.25, -0.5, 1.0)
    data = W.vortex_field(36, count=10)
    before = ctx.tuning()
    ctx.set_tuning(fastdiv=fastdiv)
    try:
        fld = ctx.field(data, offset, spacing)
        try:
            lo = np.array(offset, f32)
            hi = lo + np.array(spacing, f32) * 35
            seeds = W.uniform_seeds(4000, lo, hi, seed_offset=77)
            dt = f32(0.002)
            res = ctx.advect(fld, seeds, uniform_remaining=150, particles_per_round=4000,
                             step_size=dt)
            try:
                ora = po.run_rounds(po.Field(data, offset, spacing), seeds, 150, 4000, dt)
                compare_advect(res, ora, len(seeds))
                assert ora["total_steps"] > 100000
            finally:
                res.free()
        finally:
            fld.release()
    finally:
        ctx.set_tuning(**before)
