"""GPU parity: libdpa_hip.so (through the C ABI) against the CPU oracle.

Bar: vertex arrays, index arrays, final positions / remaining / status and the
step count bit-exact; colours within 4 ulp (the GPU's atan2f is not glibc's,
every input to it is bit-identical).
"""
import sys

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_colors_close(got, want):
    tol = 4 * np.spacing(np.abs(want).astype(np.float32)) + np.float32(1e-30)
    bad = np.abs(got.astype(np.float64) - want.astype(np.float64)) > tol
    assert not bad.any(), (np.flatnonzero(bad)[:10], got[bad][:5], want[bad][:5])


def compare_advect(res, ora, n):
    assert res.round_count == len(ora["rounds"])
    for r, (v, c, i) in enumerate(ora["rounds"]):
        gv, gc, gi = res.round(r)
        assert gv.shape == v.shape, (r, gv.shape, v.shape)
        assert np.array_equal(bits(gv), bits(v)), f"round {r} vertices differ"
        assert gi.dtype == i.dtype and np.array_equal(gi, i), f"round {r} indices differ"
        assert_colors_close(gc, c)
    pos, rem, st = res.final_state()
    assert np.array_equal(bits(pos), bits(ora["position"]))
    assert np.array_equal(rem, ora["remaining"])
    assert np.array_equal(st, ora["status"])
    assert res.total_steps == ora["total_steps"]


def run_both(ctx, field_np, offset, spacing, seeds, remaining, ppr, dt, integrator=po.DPA_RK4,
             record=True, use64=False, t0=0.0, time_offset=0.0, time_spacing=1.0):
    if field_np.ndim == 4:
        fld = ctx.field(field_np, offset, spacing)
        ofld = po.Field(field_np, offset, spacing)
    else:
        fld = ctx.field_unsteady(field_np, offset, spacing, time_offset, time_spacing)
        ofld = po.Field(field_np, offset, spacing, time_offset, time_spacing)
    try:
        n = len(seeds)
        rem_arr = (np.full(n, remaining, np.uint64) if np.isscalar(remaining)
                   else np.asarray(remaining, np.uint64))
        res = ctx.advect(fld, seeds, remaining=rem_arr, particles_per_round=ppr, step_size=dt,
                         start_time=t0, integrator=integrator, record=record, use_64_bit=use64)
        ora = po.run_rounds(ofld, seeds, rem_arr, ppr, dt, t0, integrator, record, use64)
        compare_advect(res, ora, n)
        res.free()
        return ora
    finally:
        fld.release()


def test_abc64_streamlines_bit_exact(ctx):
    say("case ABC 64^3, 4096 seeds x 256 steps, record, u32")
    c = W.abc64_streamlines()
    ora = run_both(ctx, c["field"], c["offset"], c["spacing"], c["seeds"], c["iterations"],
                   c["particles_per_round"], c["step_size"])
    assert ora["total_steps"] > 100000


def test_rounds_tail_first_uneven_remaining(ctx):
    say("case rounds: 1000 seeds, ppr 300, remaining 0..299")
    c = W.abc64_streamlines()
    rng = np.random.default_rng(11)
    seeds = c["seeds"][:1000]
    rem = rng.integers(0, 300, 1000)
    ora = run_both(ctx, c["field"], c["offset"], c["spacing"], seeds, rem, 300, c["step_size"])
    assert len(ora["rounds"]) == 4


def test_u64_indices(ctx):
    say("case u64 indices")
    c = W.abc64_streamlines()
    run_both(ctx, c["field"], c["offset"], c["spacing"], c["seeds"][:700], 64, 256,
             c["step_size"], use64=True)


def test_sentinel_valued_vertices_and_edges(ctx):
    say("case sentinels: seeds equal to (-1,-1,-1) / (-2,-2,-2), outside, zero region")
    n = 24
    o = np.array([-4.0, -4.0, -4.0], f32)
    h = np.array([0.5, 0.5, 0.5], f32)
    rng = np.random.default_rng(12)
    field = rng.uniform(-1, 1, (n, n, n, 3)).astype(f32)
    field[:, :, 16:] = 0.0                        # a zero-vector region: isZero retires
    seeds = rng.uniform(-4, 7, (64, 3)).astype(f32)
    seeds[0] = (-1, -1, -1)                       # round position 0: pruned seed + g==1 quirk
    seeds[5] = (-2, -2, -2)                       # treated as terminal by indices / colours
    seeds[9] = (-1, -1, -1)
    seeds[10] = (100, 0, 0)                       # outside
    seeds[11] = (7.4, 7.4, 7.4)                   # in the zero region
    run_both(ctx, field, o, h, seeds, 40, 64, f32(0.1))


def test_empty_and_zero_remaining(ctx):
    say("case empty input / zero remaining")
    c = W.abc64_streamlines()
    run_both(ctx, c["field"], c["offset"], c["spacing"], np.zeros((0, 3), f32), 10, 5,
             c["step_size"])
    run_both(ctx, c["field"], c["offset"], c["spacing"], c["seeds"][:50], 0, 50, c["step_size"])


def test_vortex_endpoints_bit_exact(ctx):
    say("case vortex 64^3 endpoints, 20000 seeds x 1000 steps")
    field = W.vortex_field(64)
    seeds = W.uniform_seeds(20000, (0, 0, 0), (63, 63, 63), seed_offset=21)
    run_both(ctx, field, np.zeros(3, f32), np.ones(3, f32), seeds, 1000, 20000, f32(0.25),
             record=False)


def test_rk4_mode(ctx):
    say("case classic rk4, record")
    c = W.abc64_streamlines()
    run_both(ctx, c["field"], c["offset"], c["spacing"], c["seeds"][:512], 200, 512,
             c["step_size"], integrator=po.RK4)


def test_unsteady_pathlines(ctx):
    say("case unsteady 32^3 x 4 slices, pathlines dpa + rk4")
    field = W.unsteady_vortex_field(32, time_steps=4, count=8)
    seeds = W.uniform_seeds(300, (0, 0, 0), (31, 31, 31), seed_offset=22)
    for integ in (po.DPA_RK4, po.RK4):
        run_both(ctx, field, np.zeros(3, f32), np.ones(3, f32), seeds, 120, 300, f32(0.02),
                 integrator=integ, t0=0.3, time_offset=0.0, time_spacing=1.0)


def test_endpoints_device_path(ctx):
    say("case device-resident endpoints")
    c = W.abc64_streamlines()
    fld = ctx.field(c["field"], c["offset"], c["spacing"])
    seeds = c["seeds"]
    n = len(seeds)
    d_seeds = ctx.to_device(seeds)
    d_pos = ctx.alloc(seeds.nbytes)
    d_rem = ctx.alloc(8 * n)
    d_st = ctx.alloc(n)
    try:
        steps = ctx.advect_endpoints_device(fld, d_seeds, n, d_pos, uniform_remaining=256,
                                            step_size=c["step_size"], d_rem_out=d_rem,
                                            d_status_out=d_st)
        pos = ctx.to_host(d_pos, np.empty((n, 3), f32))
        rem = ctx.to_host(d_rem, np.empty(n, np.uint64))
        st = ctx.to_host(d_st, np.empty(n, np.uint8))
        ora = po.advect_endpoints(po.Field(c["field"], c["offset"], c["spacing"]), seeds, 256,
                                  c["step_size"])
        assert steps == ora["total_steps"]
        assert np.array_equal(bits(pos), bits(ora["position"]))
        assert np.array_equal(rem, ora["remaining"]) and np.array_equal(st, ora["status"])
    finally:
        for p in (d_seeds, d_pos, d_rem, d_st):
            ctx.free(p)
        fld.release()


def test_vortex256_sample_and_full_run_properties(ctx):
    say("case vortex 256^3; 16384-seed sample vs oracle; full 2^20 run properties")
    field = W.vortex_field(256)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    ofld = po.Field(field, (0, 0, 0), (1, 1, 1))
    seeds = W.vortex_seeds()
    n = len(seeds)
    d_seeds = ctx.to_device(seeds)
    d_pos = ctx.alloc(seeds.nbytes)
    d_rem = ctx.alloc(8 * n)
    d_st = ctx.alloc(n)
    try:
        s1 = ctx.advect_endpoints_device(fld, d_seeds, n, d_pos, uniform_remaining=1000,
                                         step_size=0.25, d_rem_out=d_rem, d_status_out=d_st)
        pos = ctx.to_host(d_pos, np.empty((n, 3), f32))
        rem = ctx.to_host(d_rem, np.empty(n, np.uint64))
        st = ctx.to_host(d_st, np.empty(n, np.uint8))
        # size-independent properties of the full run
        assert s1 == int(np.sum(1000 - rem.astype(np.int64)))
        assert np.all((st == dpa_amd.STATUS_COMPLETED) == (rem == 0))
        assert np.all(st != dpa_amd.STATUS_ACTIVE)
        s2 = ctx.advect_endpoints_device(fld, d_seeds, n, d_pos, uniform_remaining=1000,
                                         step_size=0.25)
        pos2 = ctx.to_host(d_pos, np.empty((n, 3), f32))
        assert s2 == s1 and np.array_equal(bits(pos2), bits(pos))   # deterministic
        # bit-exact against the oracle on a strided sample of the same run
        idx = np.arange(0, n, 64)
        ora = po.advect_endpoints(ofld, seeds[idx], 1000, f32(0.25))
        assert np.array_equal(bits(pos[idx]), bits(ora["position"]))
        assert np.array_equal(rem[idx], ora["remaining"])
        assert np.array_equal(st[idx], ora["status"])
    finally:
        for p in (d_seeds, d_pos, d_rem, d_st):
            ctx.free(p)
        fld.release()
