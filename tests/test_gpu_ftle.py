"""GPU FTLE against the CPU oracle.

The advection feeding the flow map is bit-exact (step counts compared
exactly); the gradient, Cauchy-Green tensor and eigenvalue run in float32 on
the GPU (float64 near repeated eigenvalues) and in float64 in the oracle, on
the same float32 flow map.  Tolerance: |gpu - oracle| <= 1e-3 |oracle| + 1e-5 / T.
"""
import math
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
    return np.ascontiguousarray(a).view(np.uint32)


def assert_ftle_close(got, want, T):
    assert got.shape == want.shape
    fin = np.isfinite(want)
    assert np.array_equal(np.isfinite(got), fin)
    assert np.array_equal(got[~fin], want[~fin])          # -inf where lambda = 0
    err = np.abs(got[fin].astype(np.float64) - want[fin].astype(np.float64))
    tol = 1e-3 * np.abs(want[fin].astype(np.float64)) + 1e-5 / T
    assert np.all(err <= tol), (err.max(), np.argmax(err - tol))


@pytest.mark.parametrize("stride", [(1.0, 1.0, 1.0), (2.0, 2.0, 3.0)])
def test_ftle_dpa_lattice_mode(ctx, stride):
    say(f"case ftle dpa lattice, stride {stride}")
    field, h = W.abc_field((40, 36, 32))
    off = np.zeros(3, f32)
    iters, dt = 100, f32(0.02)
    fld = ctx.field(field, off, h)
    try:
        got, steps = ctx.ftle_lattice(fld, stride, iters, dt)
    finally:
        fld.release()
    shape = np.array(field.shape[:3])
    size = (h * shape.astype(f32)).astype(f32)
    st = (h * np.asarray(stride, f32)).astype(f32)
    seeds, per = po.lattice_seeds(off, size, st)
    res = po.advect_endpoints(po.Field(field, off, h), seeds, iters, dt)
    want, coll = po.ftle_dpa(shape, off, h, stride, iters, dt, seeds, res["position"])
    assert steps == res["total_steps"]
    # with stride 1, float rounding in dpa's lattice/cell arithmetic sends some
    # seeds to a neighbouring cell (coll > 0); both sides let the highest
    # particle index win, so the volumes still agree
    assert coll == 0 or stride == (1.0, 1.0, 1.0)
    assert_ftle_close(got, want, float(f32(dt * f32(iters))))


@pytest.mark.parametrize("frac", [0.5, 0.3])
def test_ftle_aux_mode(ctx, frac):
    # delta = h / 2: each particle between two neighbouring points is advected
    # once (three per point plus the low faces); 0.3 h: six per point
    say(f"case ftle auxiliary neighbours, ABC 48x48x24, 200 steps, delta {frac} h")
    field, h = W.abc_field((48, 48, 24), spacing=(2 * math.pi / 47,) * 3)
    off = np.zeros(3, f32)
    iters, dt = 200, f32(0.02)
    shape = field.shape[:3]
    delta = (f32(frac) * h).astype(f32)
    assert po.aux_shares(h, delta, shape) == (frac == 0.5)
    fld = ctx.field(field, off, h)
    n = int(np.prod(shape))
    d_out = ctx.alloc(4 * n)
    try:
        steps = ctx.ftle_aux_device(fld, off, h, shape, delta, iters, dt, d_out)
        got = ctx.to_host(d_out, np.empty(n, f32))
    finally:
        ctx.free(d_out)
        fld.release()
    seeds = po.aux_seeds(off, h, shape, delta)
    res = po.advect_endpoints(po.Field(field, off, h), seeds, iters, dt)
    T = float(f32(dt * f32(iters)))
    want = po.ftle_aux(res["position"], delta, T)
    assert steps == po.aux_particle_steps(res["remaining"], shape, iters,
                                          po.aux_shares(h, delta, shape))
    assert_ftle_close(got, want, T)


@pytest.mark.parametrize("frac", [0.5, 0.3])
def test_ftle_aux_empty_lattice(ctx, frac):
    # a lattice with no points along one axis advects nothing
    say(f"case ftle auxiliary, empty lattice, delta {frac} h")
    field, h = W.abc_field((12, 10, 8), spacing=(2 * math.pi / 11,) * 3)
    off = np.zeros(3, f32)
    fld = ctx.field(field, off, h)
    d = ctx.alloc(64)
    try:
        for shape in ((0, 10, 8), (12, 0, 8), (12, 10, 0)):
            assert dpa_amd.aux_particles(h, (f32(frac) * h).astype(f32), shape) == 0
            steps = ctx.ftle_aux_device(fld, off, h, shape, (f32(frac) * h).astype(f32), 50,
                                        f32(0.02), d)
            assert steps == 0
    finally:
        ctx.free(d)
        fld.release()


@pytest.mark.parametrize("frac", [0.5, 0.3])
def test_ftle_aux_chunks(frac):
    # a context with a small memory budget processes the lattice in chunks
    # (whole axis-0 slabs when particles are shared, each chunk advecting its
    # own low face); the volume equals the one-launch volume bit for bit
    say(f"case ftle auxiliary chunks, delta {frac} h")
    field, h = W.abc_field((23, 17, 13), spacing=(2 * math.pi / 22,) * 3)
    off = np.zeros(3, f32)
    iters, dt = 80, f32(0.02)
    shape = field.shape[:3]
    n = int(np.prod(shape))
    delta = (f32(frac) * h).astype(f32)
    out = []
    for budget in (0, 1 << 18):
        with dpa_amd.Context(0, memory_budget=budget) as c:
            fld = c.field(field, off, h)
            d = c.alloc(4 * n)
            try:
                steps = c.ftle_aux_device(fld, off, h, shape, delta, iters, dt, d)
                out.append((c.to_host(d, np.empty(n, f32)), steps))
            finally:
                c.free(d)
                fld.release()
    (one, s1), (chunked, s2) = out
    assert np.array_equal(bits(one), bits(chunked))
    if frac == 0.5:
        assert s2 > s1   # each chunk after the first repeats a slab of particles
    else:
        assert s2 == s1
    seeds = po.aux_seeds(off, h, shape, delta)
    res = po.advect_endpoints(po.Field(field, off, h), seeds, iters, dt)
    assert s1 == po.aux_particle_steps(res["remaining"], shape, iters, frac == 0.5)


def test_ftle_estimate_from_given_positions(ctx):
    say("case ftle_estimate_device with caller positions")
    rng = np.random.default_rng(31)
    fshape, off, h, stride = (12, 10, 8), np.array([1.0, -2.0, 0.5], f32), np.array([0.5, 0.25, 1.0], f32), (1.0, 1.0, 1.0)
    seeds, per = po.lattice_seeds(off, (h * np.array(fshape, f32)).astype(f32), h)
    finals = (seeds + rng.normal(0, 0.1, seeds.shape)).astype(f32)
    want, coll = po.ftle_dpa(fshape, off, h, stride, 50, f32(0.01), seeds, finals)
    d_o, d_f = ctx.to_device(seeds), ctx.to_device(finals)
    shape = ctx.ftle_estimate_device(fshape, off, h, stride, 50, f32(0.01), d_o, d_f, len(seeds), None)
    d_out = ctx.alloc(4 * int(np.prod(shape)))
    try:
        ctx.ftle_estimate_device(fshape, off, h, stride, 50, f32(0.01), d_o, d_f, len(seeds), d_out)
        got = ctx.to_host(d_out, np.empty(shape, f32))
    finally:
        for p in (d_o, d_f, d_out):
            ctx.free(p)
    assert coll == 0
    assert_ftle_close(got, want, 0.5)


@pytest.mark.parametrize("kind", ["lattice", "aux"])
def test_brick_walk_matches_c_order(ctx, kind):
    # whole-lattice launches walk their points in 4x4x4 bricks cut at the far
    # faces (odd sides here), or in groups of bricks above them (groups cut at
    # the far faces too); the volume and the step count must equal the C-order
    # walk bit for bit
    say(f"case brick walk vs C order, {kind}")
    field, h = W.abc_field((23, 17, 13), spacing=(2 * math.pi / 22,) * 3)
    off = np.zeros(3, f32)
    iters, dt = 80, f32(0.02)
    shape = field.shape[:3]
    n = int(np.prod(shape))
    fld = ctx.field(field, off, h)
    before = ctx.tuning()
    out = []
    try:
        walks = (dict(order=1), dict(order=1, lattice_group=1), dict(order=1, lattice_group=2),
                 dict(order=1, lattice_brick=2, lattice_group=4), dict(order=0))
        for walk in walks:
            ctx.set_tuning(**{**before, **walk})
            if kind == "lattice":
                vol, steps = ctx.ftle_lattice(fld, (1.0, 1.0, 1.0), iters, dt)
            else:
                d_out = ctx.alloc(4 * n)
                try:
                    steps = ctx.ftle_aux_device(fld, off, h, shape, (f32(0.5) * h).astype(f32),
                                                iters, dt, d_out)
                    vol = ctx.to_host(d_out, np.empty(n, f32))
                finally:
                    ctx.free(d_out)
            out.append((vol, steps))
    finally:
        ctx.set_tuning(**before)
        fld.release()
    assert out[-1][1] > 1000
    for vol, steps in out[:-1]:
        assert steps == out[-1][1]
        assert np.array_equal(vol.view(np.uint32), out[-1][0].view(np.uint32))
