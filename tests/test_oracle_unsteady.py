"""Pins of the oracle's time-variant (4-D) sampling, the arithmetic every
pathline test holds the GPU to.

dpa defines a time-variant vector field as regular_grid<vector3, 4>
(include/dpa/types/regular_fields.hpp:25): the N-dimensional contains and
interpolate of regular_grid.hpp:41-75 with time as axis 3, so time is the
innermost corner bit of permute_for's list and the first axis the cascade
lerps.  The time of iteration i is iteration_index * step_size
(particle_advector.cpp:375), offset here by the request's start time.  dpa's
own stages never sample a time-variant field and its tests hold nothing for
it, so each pin below comes from what the mathematics fixes:

  pin_multilinear      a field multilinear in (x0, x1, x2, t) is reproduced
                       to rounding at random (p, t);
  pin_slice_index      a field whose value is its slice index returns
                       (t - t0) / th, and exactly k at slice k's time;
  pin_decomposition    the 4-D sample equals, bit for bit, the (separately
                       pinned) 3-D sample of the node field blended in time
                       first -- time is lerped before any spatial axis;
  pin_time_contains    the time axis's contains: [t0, t0 + (nt-1) th);
  pin_exit_time        a particle leaves at the first iteration whose time
                       falls outside, counted by brute force; a start time
                       outside retires it before any step;
  pin_ramp             a spatially uniform field ramping linearly in time
                       moves a particle by the closed-form sum (dpa's
                       one-sample step) or integral (classic RK4).

test_pins_catch_planted_slips builds the oracle with each planted mistake
(dpa_oracle.c, DPO_PLANT) and shows the pins that must catch it fail.
"""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import py_oracle as po

f32 = np.float32
ROOT = Path(__file__).resolve().parents[1]

T0, TH, NT = -1.25, 0.5, 5          # slices at -1.25 .. 0.75; t valid in [-1.25, 0.75)


def _coords(shape, h, o):
    i = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return [o[k] + h[k] * i[k] for k in range(3)]


def _t_inside(rng):
    return f32(T0 + rng.uniform(0, 1) * (NT - 1) * TH * 0.999)


def _p_inside(rng, shape, h, o):
    return (o + rng.uniform(0, 1, 3) * h * (np.array(shape) - 1) * 0.999).astype(f32)


# ---------------------------------------------------------------- the pins

def pin_multilinear():
    rng = np.random.default_rng(11)
    shape, h, o = (5, 6, 7), np.array([0.5, 0.25, 1.0]), np.array([-1.0, 2.0, 0.5])
    x = _coords(shape, h, o)
    coef = rng.normal(size=(3, 16))

    def F(c, a0, a1, a2, t):
        v = 0.0
        for m in range(16):          # every product of a subset of (x0, x1, x2, t)
            term = c[m]
            for bit, var in zip(range(4), (a0, a1, a2, t)):
                if (m >> bit) & 1:
                    term = term * var
            v = v + term
        return v

    data = np.stack([np.stack([F(coef[e], *x, T0 + TH * k) for e in range(3)], -1)
                     for k in range(NT)]).astype(f32)
    fld = po.Field(data, o, h, T0, TH)
    for _ in range(300):
        p, t = _p_inside(rng, shape, h, o), _t_inside(rng)
        got = po.interpolate(fld, p, t).astype(np.float64)
        want = np.array([F(coef[e], *p.astype(np.float64), float(t)) for e in range(3)])
        scale = 1.0 + np.abs(coef).sum(1) * 8.0
        assert np.all(np.abs(got - want) <= 3e-6 * scale), (p, t, got, want)


def pin_slice_index():
    shape = (4, 4, 4)
    k = np.arange(NT, dtype=np.float64)[:, None, None, None, None]
    data = (k * np.array([1.0, 2.0, -1.0]) * np.ones(shape + (3,))).astype(f32)
    fld = po.Field(data, (0, 0, 0), (1, 1, 1), T0, TH)
    p = np.array([1.25, 0.5, 2.75], f32)
    for s in range(NT - 1):                  # at slice s's time: exactly slice s
        t = f32(T0 + TH * s)
        assert np.array_equal(po.interpolate(fld, p, t), np.array([s, 2 * s, -s], f32))
    rng = np.random.default_rng(12)
    for _ in range(400):
        t = _t_inside(rng)
        want = (float(t) - T0) / TH
        v = po.interpolate(fld, p, t)
        assert abs(float(v[0]) - want) <= 4e-6 * NT, (t, v, want)
        assert abs(float(v[1]) - 2 * want) <= 8e-6 * NT and abs(float(v[2]) + want) <= 4e-6 * NT


def pin_decomposition():
    rng = np.random.default_rng(13)
    shape, h, o = (4, 5, 6), np.array([0.5, 0.75, 1.0], f32), np.array([0.25, -0.5, 1.0], f32)
    data = (rng.normal(size=(NT,) + shape + (3,))
            * 10.0 ** rng.uniform(-2, 2, (NT,) + shape + (3,))).astype(f32)
    fld = po.Field(data, o, h, T0, TH)
    th, t0 = f32(TH), f32(T0)
    for _ in range(400):
        p, t = _p_inside(rng, shape, h, o), _t_inside(rng)
        d = f32(t - t0)
        ct = int(np.floor(f32(d / th)))
        wt = f32(f32(np.fmod(d, th)) / th)
        blended = (f32(f32(1) - wt) * data[ct] + wt * data[ct + 1]).astype(f32)
        want = po.interpolate(po.Field(blended, o, h), p)
        got = po.interpolate(fld, p, t)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (p, t, got, want)


def pin_time_contains():
    fld = po.Field(np.ones((NT, 4, 4, 4, 3), f32), (0, 0, 0), (1, 1, 1), T0, TH)
    p = np.array([1, 1, 1], f32)
    last = f32(T0 + (NT - 1) * TH)
    below = f32(last - f32(2.0 ** -22))          # fl(below - t0) = 2 - 2^-22: slice nt-2
    for t in (f32(T0), f32(T0 + 0.3), below):
        assert po.contains(fld, p, t), t
    # the float just below the last slice's time is outside too: fl(t - t0)
    # rounds up to (nt - 1) th, so its slice is nt - 1
    assert f32(np.nextafter(last, f32(-np.inf)) - f32(T0)) == f32((NT - 1) * TH)
    for t in (np.nextafter(f32(T0), f32(-np.inf)), np.nextafter(last, f32(-np.inf)), last,
              f32(last + 1), f32(np.nan), f32(np.inf), f32(-np.inf)):
        assert not po.contains(fld, p, t), t


def pin_exit_time():
    # field slices at t = 0, 1, 2, 3: times in [0, 3) are inside
    fld = po.Field(np.tile(np.array([0.1, 0.2, 0.3], f32), (4, 16, 16, 16, 1)),
                   (0, 0, 0), (1, 1, 1), 0.0, 1.0)
    p0 = np.array([5, 5, 5], f32)
    for ts, dt in ((0.3, 0.25), (0.0, 0.1), (2.2, 0.05), (1.0, 0.5)):
        ts, dt = f32(ts), f32(dt)
        k = 0
        while True:                      # brute force: t_i = ts + fl(i * dt)
            t = f32(ts + f32(f32(k) * dt))
            if not (0 <= math.floor(float(t)) <= 2):
                break
            k += 1
        out = po.advect_particle(fld, p0, 100, dt, t0=ts)
        assert out["steps"] == k and out["status"] == po.EXITED, (ts, dt, out["steps"], k)
    for ts in (-0.01, 3.0, 7.5):         # outside before the first step
        out = po.advect_particle(fld, p0, 10, f32(0.25), t0=f32(ts))
        assert out["steps"] == 0 and out["status"] == po.EXITED, ts
        assert np.array_equal(out["vertices"], np.array([p0, [-2, -2, -2]], f32))


def pin_ramp():
    nt, th = 9, 0.5                      # slices at t = 0 .. 4
    a = np.array([0.2, -0.1, 0.3])       # stored (v0, v1, v2) = a + b t
    b = np.array([0.05, 0.1, -0.08])
    data = np.stack([np.tile((a + b * (th * k)).astype(f32), (24, 24, 24, 1))
                     for k in range(nt)]).astype(f32)
    fld = po.Field(data, (0, 0, 0), (1, 1, 1), 0.0, th)
    p0 = np.array([10.0, 11.0, 12.0], f32)
    ts, dt, N = 0.1, 0.05, 60
    ka, kb = a[::-1], b[::-1]            # axis i moves with stored component 2 - i
    # dpa's step samples once at t_i = ts + i dt: sum of dt (a + b t_i)
    want = p0 + dt * (N * ka + kb * (N * ts + dt * N * (N - 1) / 2))
    out = po.advect_particle(fld, p0, N, f32(dt), t0=f32(ts))
    assert out["steps"] == N and out["status"] == po.COMPLETED
    assert np.allclose(out["position"], want, rtol=0, atol=2e-5), (out["position"], want)
    # classic RK4 integrates a field linear in t exactly
    want = p0 + N * dt * ka + kb * ((ts + N * dt) ** 2 - ts ** 2) / 2
    out = po.advect_particle(fld, p0, N, f32(dt), t0=f32(ts), integrator=po.RK4)
    assert out["steps"] == N and out["status"] == po.COMPLETED
    assert np.allclose(out["position"], want, rtol=0, atol=2e-5), (out["position"], want)


PINS = [pin_multilinear, pin_slice_index, pin_decomposition, pin_time_contains, pin_exit_time,
        pin_ramp]


@pytest.mark.parametrize("pin", PINS, ids=lambda f: f.__name__)
def test_pin(pin):
    pin()


# ---------------------------------------------------- planted mistakes

# DPO_PLANT value -> the pins that must fail on that mistake
PLANTED = {
    1: [pin_decomposition],                                       # time lerped last
    2: [pin_multilinear, pin_slice_index, pin_decomposition, pin_ramp],   # slice + 1
    3: [pin_multilinear, pin_slice_index, pin_decomposition, pin_ramp],   # 1 - w
    4: [pin_exit_time, pin_ramp],                                 # time of i + 1
    5: [pin_time_contains],                                       # last node inside
}


@pytest.fixture(scope="module")
def planted_builds(tmp_path_factory):
    out = tmp_path_factory.mktemp("planted")
    paths = {}
    for k in PLANTED:
        so = out / f"libdpa_oracle_plant{k}.so"
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-fopenmp",
                        "-std=c11", f"-DDPO_PLANT={k}", str(ROOT / "oracle" / "dpa_oracle.c"),
                        "-shared", "-o", str(so), "-lm"], check=True)
        paths[k] = so
    return paths


@pytest.mark.parametrize("plant", sorted(PLANTED))
def test_pins_catch_planted_slips(planted_builds, plant):
    with po.using(planted_builds[plant]):
        for pin in PLANTED[plant]:
            with pytest.raises(AssertionError):
                pin()
