"""Writes the regression fixtures of this directory.

dpa holds no golden vectors, known-answer tests or fixtures for the advection
path (its tests/ is a Catch2 main only), and it cannot be built here, so these
fixtures are SELF-GENERATED by the CPU oracle (oracle/dpa_oracle.c), not held
by the reference.  They do not pin the oracle to dpa -- tests/test_oracle.py
does that against closed forms, invariants and brute force -- they freeze its
outputs on small inputs, so that a later change to the oracle or to the
library that alters a single bit on these cases is seen, and give the GPU
tests a fixed input/output set that needs no oracle run.

Each case is one .npz (numpy arrays only, no pickled objects) holding its
inputs (the field too, so the fixture does not depend on the generators) and
the oracle's outputs.  Run from the repository root after building the
oracle:  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from dpa_amd import workloads as W  # noqa: E402
from oracle import py_oracle as po  # noqa: E402

f32 = np.float32


def rounds_arrays(ora):
    out = {}
    for r, (v, c, i) in enumerate(ora["rounds"]):
        out[f"vertices_{r}"] = v
        out[f"colors_{r}"] = c
        out[f"indices_{r}"] = i
    out["round_count"] = np.array(len(ora["rounds"]), np.int64)
    out["position"] = ora["position"]
    out["remaining"] = ora["remaining"]
    out["status"] = ora["status"]
    out["total_steps"] = np.array(ora["total_steps"], np.uint64)
    return out


def abc16_rounds():
    """ABC on 16^3 nodes over [0, 2 pi]^3, 96 seeds (some with fewer
    iterations, two outside the grid), dt = 0.05 h, 48 iterations, 40
    particles per round (3 rounds), recorded, u32 indices."""
    field, h = W.abc_field((16, 16, 16))
    off = np.zeros(3, f32)
    rng = np.random.default_rng(W.MASTER_SEED + 101)
    seeds = (rng.uniform(0, 15, (96, 3)) * h).astype(f32)
    seeds[5] = (-0.5 * h).astype(f32)
    seeds[77] = (16.5 * h).astype(f32)
    rem = np.full(len(seeds), 48, np.uint64)
    rem[::9] = 7
    dt = f32(f32(0.05) * h[0])
    ora = po.run_rounds(po.Field(field, off, h), seeds, rem, 40, dt)
    return dict(field=field, offset=off, spacing=h, seeds=seeds, remaining_in=rem,
                particles_per_round=np.array(40, np.uint64), step_size=np.array(dt, f32),
                **rounds_arrays(ora))


def vortex24_endpoints():
    """Gaussian vortices on 24^3 nodes (h = 1), 256 seeds, dt 0.25, 120
    iterations, endpoints only."""
    field = W.vortex_field(24, count=6)
    off, h = np.zeros(3, f32), np.ones(3, f32)
    seeds = W.uniform_seeds(256, (0, 0, 0), (23, 23, 23), seed_offset=102)
    dt = f32(0.25)
    ora = po.advect_endpoints(po.Field(field, off, h), seeds, 120, dt)
    return dict(field=field, offset=off, spacing=h, seeds=seeds, step_size=np.array(dt, f32),
                iterations=np.array(120, np.uint64), position=ora["position"],
                remaining=ora["remaining"], status=ora["status"],
                total_steps=np.array(ora["total_steps"], np.uint64))


def unsteady12_rounds():
    """Time-variant vortices, 12^3 nodes x 3 slices (origin 0, h = 1, slices
    0.5 apart from t = 0.25), 64 seeds from t0 = 0.4, dt 0.01, 60 iterations,
    recorded, one round; dpa's runge_kutta_4 and classic RK4."""
    field = W.unsteady_vortex_field(12, time_steps=3, count=4)
    off, h = np.zeros(3, f32), np.ones(3, f32)
    seeds = W.uniform_seeds(64, (0, 0, 0), (11, 11, 11), seed_offset=103)
    dt, t0 = f32(0.01), f32(0.4)
    out = dict(field=field, offset=off, spacing=h, seeds=seeds, step_size=np.array(dt, f32),
               start_time=np.array(t0, f32), time_offset=np.array(0.25, f32),
               time_spacing=np.array(0.5, f32), iterations=np.array(60, np.uint64))
    for name, integ in (("dpa", po.DPA_RK4), ("rk4", po.RK4)):
        ora = po.run_rounds(po.Field(field, off, h, f32(0.25), f32(0.5)), seeds, 60, 64, dt, t0,
                            integ)
        for k, v in rounds_arrays(ora).items():
            out[f"{name}_{k}"] = v
    return out


def ftle_aux_abc():
    """Auxiliary-neighbour FTLE on ABC 20 x 18 x 12 nodes (h = 2 pi / 19),
    +-0.5 h neighbours of every node, dt 0.02, 100 iterations."""
    field, h = W.abc_field((20, 18, 12), spacing=(2 * math.pi / 19,) * 3)
    off = np.zeros(3, f32)
    shape = field.shape[:3]
    delta = (f32(0.5) * h).astype(f32)
    seeds = po.aux_seeds(off, h, shape, delta)
    dt, iters = f32(0.02), 100
    res = po.advect_endpoints(po.Field(field, off, h), seeds, iters, dt)
    T = float(f32(dt * f32(iters)))
    ftle = po.ftle_aux(res["position"], delta, T)
    # delta is half the spacing: the library advects each particle shared by
    # two neighbouring points once and reports those particle-steps
    assert po.aux_shares(h, delta, shape)
    shared = po.aux_particle_steps(res["remaining"], shape, iters, True)
    return dict(field=field, offset=off, spacing=h, delta=delta, step_size=np.array(dt, f32),
                iterations=np.array(iters, np.uint64), lattice_shape=np.array(shape, np.uint64),
                ftle=ftle.astype(f32), total_steps=np.array(res["total_steps"], np.uint64),
                shared_steps=np.array(shared, np.uint64))


CASES = {"abc16_rounds": abc16_rounds, "vortex24_endpoints": vortex24_endpoints,
         "unsteady12_rounds": unsteady12_rounds, "ftle_aux_abc": ftle_aux_abc}


def main() -> None:
    for name, make in CASES.items():
        arrays = make()
        np.savez_compressed(HERE / f"{name}.npz", **arrays)
        size = (HERE / f"{name}.npz").stat().st_size
        print(f"{name}.npz: {len(arrays)} arrays, {size} bytes")


if __name__ == "__main__":
    main()
