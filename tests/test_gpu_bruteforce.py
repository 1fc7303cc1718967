"""Brute force: millions of start positions, one and two steps, bit for bit.

contains / interpolate / step (regular_grid.hpp:41-75, particle_advector.cpp:
338-375) on 2^22 particles per case: positions drawn uniformly over a box a
little wider than the grid, and positions aimed at the special cases --
every node plane, the last valid cell (n-2) and the face n-1 with their float
neighbours, signed zeros, denormals and tiny offsets from the origin, NaN and
infinities.  Each case runs under one division mode (power-of-two spacings:
exact reciprocals; verified reciprocals; with fastdiv off every axis divides),
and the final positions, remaining counts and statuses must equal the
oracle's (which divides and calls fmodf as dpa does) bit for bit.

A second test pins the sign of a zero remainder: for a position coordinate of
-0.0, dpa's fmodf gives -0.0 and so a weight of -0.0, which turns a lerp of
corner values -0.0 into +0.0 (1 * -0 + -0 * -0).  A field storing -0.0 in the
component that moves that axis carries this into the position's sign bit.
"""
import math
import sys

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_parity import bits, compare_advect

pytestmark = pytest.mark.gpu
f32 = np.float32
N = 1 << 22


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def special_seeds(shape, offset, spacing, count, seed):
    """count x 3 float32 start positions: half uniform over the grid widened by
    5% per side, half with one to three coordinates on special values."""
    rng = np.random.default_rng(seed)
    n = np.asarray(shape, np.int64)
    o = np.asarray(offset, f32)
    h = np.asarray(spacing, f32)
    ext = (n - 1).astype(f32) * h
    pos = (o + rng.uniform(-0.05, 1.05, (count, 3)).astype(f32) * ext).astype(f32)
    half = count // 2
    for ax in range(3):
        k = rng.integers(0, n[ax], count)
        faces = (o[ax] + h[ax] * k.astype(f32)).astype(f32)
        last = (o[ax] + h[ax] * f32(n[ax] - 2)).astype(f32)
        top = (o[ax] + h[ax] * f32(n[ax] - 1)).astype(f32)
        tiny = np.array([0.0, -0.0, 1e-45, -1e-45, 1e-38, -1e-38, 1e-30, -1e-30, 1e-7, -1e-7],
                        f32)
        # per particle: a node plane or one of its float neighbours
        plane = np.stack([faces, np.nextafter(faces, f32(-np.inf)),
                          np.nextafter(faces, f32(np.inf))])[rng.integers(0, 3, count),
                                                             np.arange(count)]
        # a fixed list: the last cell, the face n-1, zeros, tiny, non-finite
        fixed = np.concatenate([
            np.array([last, np.nextafter(last, f32(np.inf)), np.nextafter(top, f32(-np.inf)),
                      top, np.nextafter(top, f32(np.inf)), last + f32(0.5) * h[ax]], f32),
            (o[ax] + tiny).astype(f32), tiny,
            np.array([np.nan, np.inf, -np.inf, f32(1e30), f32(-1e30)], f32),
        ]).astype(f32)
        pick = rng.random(count) < (0.5 if ax == 0 else 0.35)
        pick[half:] = False                    # the second half stays uniform
        from_fixed = pick & (rng.random(count) < 0.5)
        from_plane = pick & ~from_fixed
        pos[from_plane, ax] = plane[from_plane]
        pos[from_fixed, ax] = rng.choice(fixed, int(from_fixed.sum()))
    return np.ascontiguousarray(pos)


def smooth_field(shape, offset, spacing):
    """A smooth analytic field, no component zero anywhere."""
    n0, n1, n2 = shape
    o = np.asarray(offset, np.float64)
    h = np.asarray(spacing, np.float64)
    x0 = o[0] + h[0] * np.arange(n0)[:, None, None]
    x1 = o[1] + h[1] * np.arange(n1)[None, :, None]
    x2 = o[2] + h[2] * np.arange(n2)[None, None, :]
    f = np.empty((n0, n1, n2, 3), np.float64)
    f[..., 0] = 0.6 + 0.4 * np.sin(0.7 * x1) * np.cos(0.3 * x0) + 0 * x2
    f[..., 1] = -0.5 + 0.3 * np.cos(0.5 * x2 + 0.2 * x0) + 0 * x1
    f[..., 2] = 0.45 + 0.35 * np.sin(0.4 * x0 - 0.6 * x2) + 0 * x1
    return f.astype(f32)


ABC_H = float(f32(2 * math.pi / 63))
CASES = {
    # name: (field maker, offset, spacing, fastdiv, dt)
    "vortex64-pow2": (lambda: W.vortex_field(64, count=24), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                      1, 0.25),
    "abc64-reciprocal": (lambda: W.abc_field((64, 64, 64))[0], (0.0, 0.0, 0.0),
                         (ABC_H, ABC_H, ABC_H), 1, 0.05 * ABC_H),
    "aniso-offset-reciprocal": (lambda: smooth_field((40, 36, 44), (-3.5, 1.25, 0.1),
                                                     (0.75, float(f32(0.1)), 0.5)),
                                (-3.5, 1.25, 0.1), (0.75, float(f32(0.1)), 0.5), 1, 0.02),
    "aniso-offset-divide": (lambda: smooth_field((40, 36, 44), (-3.5, 1.25, 0.1),
                                                 (0.75, float(f32(0.1)), 0.5)),
                            (-3.5, 1.25, 0.1), (0.75, float(f32(0.1)), 0.5), 0, 0.02),
}


@pytest.mark.parametrize("name", list(CASES))
def test_special_positions_one_and_two_steps(ctx, name):
    make, offset, spacing, fastdiv, dt = CASES[name]
    say(f"case brute force {name}: {N} seeds")
    data = make()
    seeds = special_seeds(data.shape[:3], offset, spacing, N, seed=len(name) * 7919)
    ofld = po.Field(data, offset, spacing)
    before = ctx.tuning()
    ctx.set_tuning(fastdiv=fastdiv)
    d_s = d_p = d_r = d_t = None
    try:
        fld = ctx.field(data, offset, spacing)
        try:
            d_s = ctx.to_device(seeds)
            d_p, d_r, d_t = ctx.alloc(12 * N), ctx.alloc(8 * N), ctx.alloc(N)
            for iters in (1, 2):
                steps = ctx.advect_endpoints_device(fld, d_s, N, d_p, uniform_remaining=iters,
                                                    step_size=f32(dt), d_rem_out=d_r,
                                                    d_status_out=d_t)
                pos = ctx.to_host(d_p, np.empty((N, 3), f32))
                rem = ctx.to_host(d_r, np.empty(N, np.uint64))
                st = ctx.to_host(d_t, np.empty(N, np.uint8))
                ora = po.advect_endpoints(ofld, seeds, iters, f32(dt))
                bad = np.nonzero(np.any(bits(pos) != bits(ora["position"]), axis=1))[0]
                assert bad.size == 0, (iters, bad[:5], seeds[bad[:5]], pos[bad[:5]],
                                       ora["position"][bad[:5]])
                assert np.array_equal(rem, ora["remaining"])
                assert np.array_equal(st, ora["status"])
                assert steps == ora["total_steps"]
                say(f"  {iters} step(s): {steps} particle-steps, "
                    f"{int(np.sum(st == dpa_amd.STATUS_EXITED))} exits, bit-exact")
        finally:
            for p in (d_s, d_p, d_r, d_t):
                if p is not None:
                    ctx.free(p)
            fld.release()
    finally:
        ctx.set_tuning(**before)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("spacing", [(1.0, 1.0, 1.0), (0.75, float(f32(0.1)), 0.5)])
def test_negative_zero_coordinate_weight(ctx, axis, spacing):
    say(f"case -0.0 coordinate on axis {axis}, spacing {spacing}")
    shape = (12, 14, 10)
    data = smooth_field(shape, (0, 0, 0), spacing)
    data[..., 2 - axis] = f32(-0.0)         # the component that moves `axis`
    rng = np.random.default_rng(11 + axis)
    hi = (np.asarray(shape) - 1) * np.asarray(spacing, f32)
    seeds = (rng.uniform(0.2, 0.8, (512, 3)) * hi).astype(f32)
    seeds[:, axis] = f32(-0.0)
    seeds[::4, axis] = f32(0.0)             # +0 beside -0
    dt = f32(0.05)
    fld = ctx.field(data, (0, 0, 0), spacing)
    ofld = po.Field(data, (0, 0, 0), spacing)
    try:
        ora = po.advect_endpoints(ofld, seeds, 3, dt)
        # dpa's arithmetic makes the coordinate +0.0 after a step from -0.0
        moved = ora["status"] != dpa_amd.STATUS_EXITED
        assert np.all(bits(ora["position"][moved, axis]) == 0)
        res = ctx.advect(fld, seeds, uniform_remaining=3, particles_per_round=512, step_size=dt)
        try:
            compare_advect(res, po.run_rounds(ofld, seeds, 3, 512, dt), len(seeds))
            pos, rem, st = res.final_state()
            assert np.array_equal(bits(pos), bits(ora["position"]))
        finally:
            res.free()
        d_s = ctx.to_device(seeds)
        d_p = ctx.alloc(seeds.nbytes)
        try:
            ctx.advect_endpoints_device(fld, d_s, len(seeds), d_p, uniform_remaining=3,
                                        step_size=dt)
            pos = ctx.to_host(d_p, np.empty_like(seeds))
            assert np.array_equal(bits(pos), bits(ora["position"]))
        finally:
            ctx.free(d_s)
            ctx.free(d_p)
    finally:
        fld.release()
