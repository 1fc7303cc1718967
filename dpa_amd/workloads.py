"""Synthetic inputs of the benchmark configurations (no datasets travel).

Storage follows dpa's convention (regular_grid_loader.cpp:65-88,
particle_advector.cpp:369): arrays are [i0][i1][i2][c] with axis 0 = z,
axis 1 = y, axis 2 = x, and components c = (vx, vy, vz); dpa's swizzle then
moves position component 0 (z) by stored component 2 (vz).  Fields are
computed in float64 and rounded once to float32.

Master seed 20260101; each configuration adds its own offset.
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

MASTER_SEED = 20260101
ABC = (math.sqrt(3.0), math.sqrt(2.0), 1.0)


def abc_field(shape=(64, 64, 64), spacing=None, coeffs=ABC) -> tuple[np.ndarray, np.ndarray]:
    """ABC flow u = A sin z + C cos y, v = B sin x + A cos z, w = C sin y + B cos x
    sampled node-inclusively with origin 0.  Default spacing 2*pi/(n-1) per axis
    (the [0, 2*pi]^3 cube for a cubic grid).  Returns (field, spacing[3] f32)."""
    A, B, Cc = coeffs
    n0, n1, n2 = shape
    if spacing is None:
        spacing = (2 * math.pi / (n0 - 1), 2 * math.pi / (n1 - 1), 2 * math.pi / (n2 - 1))
    h = np.asarray(spacing, np.float32)
    z = np.arange(n0, dtype=np.float64) * float(h[0])
    y = np.arange(n1, dtype=np.float64) * float(h[1])
    x = np.arange(n2, dtype=np.float64) * float(h[2])
    out = np.empty((n0, n1, n2, 3), np.float32)
    sz, cz = np.sin(z)[:, None, None], np.cos(z)[:, None, None]
    sy, cy = np.sin(y)[None, :, None], np.cos(y)[None, :, None]
    sx, cx = np.sin(x)[None, None, :], np.cos(x)[None, None, :]
    out[..., 0] = A * sz + Cc * cy
    out[..., 1] = B * sx + A * cz
    out[..., 2] = Cc * sy + B * cx
    return out, h


def _unit(rng, k):
    v = rng.normal(size=(k, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def vortex_params(count=64, n=256, seed_offset=2, radii=(8.0, 32.0)):
    """Random vortex tubes: centres uniform in [0, n)^3, radii uniform in
    `radii`, axes uniform on the sphere, strengths uniform in [0.5, 1.5] with a
    random sign, plus a unit drift direction.  Vectors are (x, y, z)."""
    rng = np.random.default_rng(MASTER_SEED + seed_offset)
    centres = rng.uniform(0.0, n, size=(count, 3))
    rho = rng.uniform(radii[0], radii[1], size=count)
    axes = _unit(rng, count)
    strength = rng.uniform(0.5, 1.5, size=count) * rng.choice([-1.0, 1.0], size=count)
    drift = _unit(rng, 1)[0]
    return centres, rho, axes, strength, drift


def _vortex_slab(z0, z1, n, params, drift_speed, centre_shift=None, strength_scale=None):
    centres, rho, axes, strength, drift = params
    zz = np.arange(z0, z1, dtype=np.float64)[:, None, None]
    yy = np.arange(n, dtype=np.float64)[None, :, None]
    xx = np.arange(n, dtype=np.float64)[None, None, :]
    vx = np.full((z1 - z0, n, n), drift_speed * drift[0])
    vy = np.full((z1 - z0, n, n), drift_speed * drift[1])
    vz = np.full((z1 - z0, n, n), drift_speed * drift[2])
    for m in range(len(rho)):
        c = centres[m] if centre_shift is None else centres[m] + centre_shift[m]
        s = strength[m] if strength_scale is None else strength[m] * strength_scale[m]
        a = axes[m]
        rx, ry, rz = xx - c[0], yy - c[1], zz - c[2]
        along = rx * a[0] + ry * a[1] + rz * a[2]
        px, py, pz = rx - along * a[0], ry - along * a[1], rz - along * a[2]
        g = (s / rho[m]) * np.exp(-(px * px + py * py + pz * pz) / (2.0 * rho[m] * rho[m]))
        # (a x r) scaled: peak tangential speed |s| e^-1/2 at distance rho
        vx += g * (a[1] * rz - a[2] * ry)
        vy += g * (a[2] * rx - a[0] * rz)
        vz += g * (a[0] * ry - a[1] * rx)
    return np.stack([vx, vy, vz], axis=-1).astype(np.float32)


def vortex_field(n=256, count=64, seed_offset=2, drift_speed=0.1, threads=None,
                 centre_shift=None, strength_scale=None) -> np.ndarray:
    """Sum of `count` Gaussian vortex tubes plus a uniform drift of
    `drift_speed`, on n^3 nodes with spacing 1 and origin 0:
    v(x) = drift_speed d + sum_m (s_m / rho_m) (a_m x r) exp(-|r_perp|^2 / (2 rho_m^2)),
    r = x - c_m, r_perp its part orthogonal to a_m."""
    params = vortex_params(count, n, seed_offset)
    out = np.empty((n, n, n, 3), np.float32)
    threads = threads or min(16, os.cpu_count() or 1)
    slab = max(1, n // (4 * threads))

    def work(z0):
        z1 = min(n, z0 + slab)
        out[z0:z1] = _vortex_slab(z0, z1, n, params, drift_speed, centre_shift, strength_scale)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(0, n, slab)))
    return out


def unsteady_vortex_field(n=256, time_steps=16, count=64, seed_offset=5, drift_speed=0.1,
                          threads=None) -> np.ndarray:
    """Time-variant vortex field [nt][n][n][n][3]: centres move linearly
    (velocity uniform in [-1, 1]^3 cells per time unit) and strengths vary as
    1 + 0.5 sin(omega t + phase), sampled at t = 0, 1, ..., time_steps-1."""
    rng = np.random.default_rng(MASTER_SEED + seed_offset + 1000)
    vel = rng.uniform(-1.0, 1.0, size=(count, 3))
    omega = rng.uniform(0.1, 0.5, size=count)
    phase = rng.uniform(0.0, 2 * math.pi, size=count)
    out = np.empty((time_steps, n, n, n, 3), np.float32)
    for t in range(time_steps):
        out[t] = vortex_field(n, count, seed_offset, drift_speed, threads,
                              centre_shift=vel * t, strength_scale=1.0 + 0.5 * np.sin(omega * t + phase))
    return out


def uniform_seeds(count: int, lo, hi, seed_offset: int) -> np.ndarray:
    """count x 3 float32 seeds uniform in the box [lo, hi) per component."""
    rng = np.random.default_rng(MASTER_SEED + seed_offset)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    s = rng.uniform(0.0, 1.0, size=(count, 3)) * (hi - lo) + lo
    s = s.astype(np.float32)
    # keep float32 rounding from landing exactly on the open upper bound
    return np.minimum(s, np.nextafter(hi.astype(np.float32), np.float32(-np.inf)))


# ------------------------------------------------------------- configs ----

def abc64_streamlines():
    """Plain steady streamlines: ABC on 64^3 over [0, 2pi]^3, 4096 seeds,
    dt = 0.05 cell (0.05 h in time units), 256 iterations, record, uint32."""
    field, h = abc_field((64, 64, 64))
    seeds = uniform_seeds(4096, (0, 0, 0), np.float32(63) * h, seed_offset=1)
    dt = np.float32(0.05) * h[0]
    return dict(field=field, spacing=h, offset=np.zeros(3, np.float32), seeds=seeds,
                iterations=256, step_size=np.float32(dt), record=True, use_64_bit=False,
                particles_per_round=4096)


def vortex_seeds(count=1 << 20):
    return uniform_seeds(count, (0, 0, 0), (255, 255, 255), seed_offset=2)


def vortex256_endpoints(count=1 << 20, n=256):
    """Large steady endpoint advection: vortex field 256^3 (h = 1), 2^20 seeds
    uniform in [0, 255)^3, dt = 0.25, 1000 iterations, endpoints only."""
    return dict(field=vortex_field(n), spacing=np.ones(3, np.float32),
                offset=np.zeros(3, np.float32), seeds=vortex_seeds(count), iterations=1000,
                step_size=np.float32(0.25), record=False)


def ftle_abc512(shape=(128, 512, 512)):
    """FTLE: ABC on 512 x 512 x 128 nodes (z = axis 0 has 128, extent 127 h)
    with isotropic spacing h = 2 pi / 511 (origin 0); every node is a lattice
    point with auxiliary neighbours at +-0.5 h; dt 0.02 for 200 steps (T = 4)."""
    h = np.float32(2 * math.pi / 511)
    field, sp = abc_field(shape, spacing=(h, h, h))
    return dict(field=field, spacing=sp, offset=np.zeros(3, np.float32),
                lattice_origin=np.zeros(3, np.float32), lattice_spacing=sp,
                lattice_shape=shape, delta=(np.float32(0.5) * sp).astype(np.float32),
                iterations=200, step_size=np.float32(0.02))


def batch_jobs(count=64, n=256, seed_offset=4):
    """64 jobs on the 256^3 vortex field: seed counts round(exp(U[ln 256,
    ln 262144])), steps uniform integer in [64, 4096], seeds uniform in a random
    axis-aligned box (extent U[16, 128] per axis) inside [0, 255)^3; full curves
    recorded with 32-bit indices (dpa's rule with ppr = the job's seed count)."""
    rng = np.random.default_rng(MASTER_SEED + seed_offset)
    jobs = []
    for j in range(count):
        seeds = int(round(math.exp(rng.uniform(math.log(256), math.log(262144)))))
        steps = int(rng.integers(64, 4097))
        ext = rng.uniform(16, 128, 3)
        lo = rng.uniform(0, (n - 1) - ext)
        s = uniform_seeds(seeds, lo, lo + ext, seed_offset=1000 + j)
        jobs.append(dict(seeds=s, uniform_remaining=steps, step_size=np.float32(0.25),
                         record=True, use_64_bit=seeds * steps > 0xFFFFFFFF,
                         particles_per_round=seeds))
    return jobs


def moving_field(base: np.ndarray, time_steps: int, shift=(0, 0, 1), phase=0.0,
                 offset=(0, 0, 0), threads=None) -> np.ndarray:
    """Time-variant field from a steady one: slice t is the base pattern moved by
    offset + t * shift nodes (periodic roll) with its amplitude pulsing as
    1 + 0.25 sin(0.7 t + phase).  [nt][n0][n1][n2][3] float32.  Slabs of
    axis 0 of every slice are filled in parallel."""
    n0 = base.shape[0]
    out = np.empty((time_steps,) + base.shape, np.float32)
    threads = threads or min(16, os.cpu_count() or 1)
    slab = max(1, n0 // (2 * threads))

    def work(job):
        t, z0 = job
        z1 = min(n0, z0 + slab)
        s0, s1, s2 = (offset[a] + t * shift[a] for a in range(3))
        amp = np.float32(1.0 + 0.25 * math.sin(0.7 * t + phase))
        rows = (np.arange(z0, z1) - s0) % n0          # out row z holds base row z - s0
        np.multiply(np.roll(base[rows], shift=(s1, s2), axis=(1, 2)), amp, out=out[t, z0:z1])

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, [(t, z0) for t in range(time_steps) for z0 in range(0, n0, slab)]))
    return out


def serving_fields(base: np.ndarray, count=4, time_steps=16):
    """The served field set: `count` time-variant fields of `time_steps`
    slices, field k moving the base pattern from a roll of (37k, 71k, 13k)
    nodes with its own pulse phase."""
    return [moving_field(base, time_steps, phase=k, offset=(37 * k, 71 * k, 13 * k))
            for k in range(count)]


def serving_trace(rate, count, seed_offset=5, fields=4, ftle_fraction=0.2,
                  seeds_range=(1000, 262144), steps_range=(128, 2048), n=256,
                  ftle_shape=(16, 256, 256), time_steps=16, dt=np.float32(0.005)):
    """Poisson arrivals at `rate` requests/s: a mix of pathline queries (seeds
    log-uniform in seeds_range, steps uniform in steps_range, start time so the
    pathline stays inside the time span, record) and FTLE slabs
    (ftle_shape lattice with +-0.5 cell neighbours, T = 4 at dt 0.02, start
    time uniform).  Returns a list of dicts with an 'at' arrival offset."""
    rng = np.random.default_rng(MASTER_SEED + seed_offset + int(rate))
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        f = int(rng.integers(0, fields))
        if rng.uniform() < ftle_fraction:
            z0 = int(rng.integers(0, n - ftle_shape[0]))
            t0 = float(rng.uniform(0, time_steps - 1 - 4.5))
            out.append(dict(at=t, kind="ftle", field=f, origin=(float(z0), 0.0, 0.0),
                            spacing=(1.0, 1.0, 1.0), shape=ftle_shape, delta=(0.5, 0.5, 0.5),
                            iterations=200, step_size=np.float32(0.02), start_time=np.float32(t0)))
        else:
            k = int(round(math.exp(rng.uniform(math.log(seeds_range[0]), math.log(seeds_range[1])))))
            steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
            span = float(dt) * steps
            t0 = float(rng.uniform(0, max(time_steps - 1 - span - 0.5, 0.0)))
            ext = rng.uniform(16, 128, 3)
            lo = rng.uniform(0, (n - 1) - ext)
            out.append(dict(at=t, kind="advect", field=f,
                            seeds=uniform_seeds(k, lo, lo + ext, seed_offset=5000 + i),
                            uniform_remaining=steps, step_size=dt, start_time=np.float32(t0),
                            record=True))
    return out
