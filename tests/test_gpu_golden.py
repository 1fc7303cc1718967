"""libdpa_hip.so against the committed regression fixtures (tests/golden/),
which the CPU oracle wrote (self-generated; dpa holds no golden vectors).
Same bar as the parity tests: positions, vertices, indices, remaining,
status and step counts bit-exact; colours within 4 ulp; FTLE within
1e-3 |ref| + 1e-5 / T of the oracle's float64 estimate.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from test_gpu_ftle import assert_ftle_close
from test_gpu_parity import assert_colors_close, bits

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def check_result(res, g, prefix=""):
    assert res.round_count == int(g[prefix + "round_count"])
    for r in range(res.round_count):
        v, c, i = res.round(r)
        assert np.array_equal(bits(v), bits(g[f"{prefix}vertices_{r}"])), f"round {r} vertices"
        assert np.array_equal(i, g[f"{prefix}indices_{r}"]), f"round {r} indices"
        assert_colors_close(c, g[f"{prefix}colors_{r}"])
    pos, rem, st = res.final_state()
    assert np.array_equal(bits(pos), bits(g[prefix + "position"]))
    assert np.array_equal(rem, g[prefix + "remaining"])
    assert np.array_equal(st, g[prefix + "status"])
    assert res.total_steps == int(g[prefix + "total_steps"])


def test_abc16_rounds(ctx):
    say("case golden abc16 rounds")
    g = load("abc16_rounds")
    fld = ctx.field(g["field"], g["offset"], g["spacing"])
    try:
        res = ctx.advect(fld, g["seeds"], remaining=g["remaining_in"],
                         particles_per_round=int(g["particles_per_round"]),
                         step_size=f32(g["step_size"]))
        try:
            check_result(res, g)
        finally:
            res.free()
    finally:
        fld.release()


def test_vortex24_endpoints(ctx):
    say("case golden vortex24 endpoints")
    g = load("vortex24_endpoints")
    n = len(g["seeds"])
    fld = ctx.field(g["field"], g["offset"], g["spacing"])
    d_s = ctx.to_device(g["seeds"])
    d_p, d_r, d_t = ctx.alloc(12 * n), ctx.alloc(8 * n), ctx.alloc(n)
    try:
        steps = ctx.advect_endpoints_device(fld, d_s, n, d_p, uniform_remaining=int(g["iterations"]),
                                            step_size=f32(g["step_size"]), d_rem_out=d_r,
                                            d_status_out=d_t)
        pos = ctx.to_host(d_p, np.empty((n, 3), f32))
        rem = ctx.to_host(d_r, np.empty(n, np.uint64))
        st = ctx.to_host(d_t, np.empty(n, np.uint8))
    finally:
        for p in (d_s, d_p, d_r, d_t):
            ctx.free(p)
        fld.release()
    assert np.array_equal(bits(pos), bits(g["position"]))
    assert np.array_equal(rem, g["remaining"]) and np.array_equal(st, g["status"])
    assert steps == int(g["total_steps"])


@pytest.mark.parametrize("name,integ", [("dpa", 0), ("rk4", 1)])
def test_unsteady12_rounds(ctx, name, integ):
    say(f"case golden unsteady12 {name}")
    g = load("unsteady12_rounds")
    fld = ctx.field_unsteady(g["field"], g["offset"], g["spacing"], float(g["time_offset"]),
                             float(g["time_spacing"]))
    try:
        res = ctx.advect(fld, g["seeds"], uniform_remaining=int(g["iterations"]),
                         particles_per_round=len(g["seeds"]), step_size=f32(g["step_size"]),
                         start_time=f32(g["start_time"]), integrator=integ)
        try:
            check_result(res, g, name + "_")
        finally:
            res.free()
    finally:
        fld.release()


def test_ftle_aux_abc(ctx):
    say("case golden ftle aux abc")
    g = load("ftle_aux_abc")
    shape = tuple(int(x) for x in g["lattice_shape"])
    n = int(np.prod(shape))
    fld = ctx.field(g["field"], g["offset"], g["spacing"])
    d_out = ctx.alloc(4 * n)
    try:
        steps = ctx.ftle_aux_device(fld, g["offset"], g["spacing"], shape, g["delta"],
                                    int(g["iterations"]), f32(g["step_size"]), d_out)
        got = ctx.to_host(d_out, np.empty(n, f32))
    finally:
        ctx.free(d_out)
        fld.release()
    # delta is half the spacing: each particle shared by two points once
    assert steps == int(g["shared_steps"])
    T = float(f32(g["step_size"] * f32(g["iterations"])))
    assert_ftle_close(got, g["ftle"].reshape(-1), T)
