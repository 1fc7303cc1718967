"""The forms of a recorded launch.

Persistent launches of at least 2^15 recorded particles (a batch's, by
default; a single request's with slot_stage 2) keep each lane's current
block of eight slots in LDS and write whole 96-byte blocks: as one run per
wave when the refill threshold is 64 (lanes in lockstep, lane l on row
64t + l), lane by lane otherwise.  A single request's rounds run one thread
per row by default (slot_stage 0, or 3 explicitly) and store every slot
directly, as smaller persistent launches do.  Every form must give the
oracle's rounds bit for bit, with particles retiring at every position
within a block (uneven remaining counts, exits, NaN seeds).
"""
import sys

import numpy as np
import pytest

from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_batch import result_arrays, same
from test_gpu_parity import compare_advect

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


def uneven(n, seed):
    rng = np.random.default_rng(seed)
    seeds = W.uniform_seeds(n, (0, 0, 0), (63, 63, 63), seed_offset=seed)
    seeds[3::1777] = (np.nan, 5.0, 5.0)           # retires before its first step
    seeds[9::2111] = (63.0, 20.0, 20.0)           # on the face n-1: outside
    rem = rng.integers(0, 150, n).astype(np.uint64)   # 0: completed before a step
    return seeds, rem


@pytest.mark.parametrize("refill,bpc,stage", [(0, 0, 2), (16, 0, 2), (1, 2, 2), (0, 0, 0),
                                              (0, 0, 3), (0, 0, 1)])
def test_staged_round_matches_oracle(ctx, vortex, refill, bpc, stage):
    field, fld = vortex
    say(f"case recorded round, refill={refill} blocks_per_cu={bpc} slot_stage={stage}")
    seeds, rem = uneven(40000, 501)
    dt = f32(0.25)
    ora = po.run_rounds(po.Field(field, (0, 0, 0), (1, 1, 1)), seeds, rem, 40000, dt)
    before = ctx.tuning()
    ctx.set_tuning(refill=refill, blocks_per_cu=bpc, slot_stage=stage)
    try:
        res = ctx.advect(fld, seeds, remaining=rem, particles_per_round=40000, step_size=dt)
        try:
            compare_advect(res, ora, len(seeds))
        finally:
            res.free()
    finally:
        ctx.set_tuning(**before)


def test_staged_batch_equals_direct_jobs(ctx, vortex):
    """Three jobs of 15000 rows: alone each runs one thread per row and stores
    slots directly; in one batch (45000 rows) the persistent launch stages
    them.  Same bits either way."""
    field, fld = vortex
    say("case staged batch vs direct jobs")
    jobs = []
    for k in range(3):
        seeds, rem = uneven(15000, 610 + k)
        jobs.append(dict(field=fld, seeds=seeds, remaining=rem, uniform_remaining=0,
                         particles_per_round=None if k else 6000, step_size=f32(0.25),
                         record=True, use_64_bit=k == 2))
    batch = ctx.advect_batch(jobs)
    try:
        for k, (j, rb) in enumerate(zip(jobs, batch)):
            alone = ctx.advect(fld, j["seeds"], remaining=j["remaining"],
                               particles_per_round=j["particles_per_round"] or len(j["seeds"]),
                               step_size=j["step_size"], use_64_bit=j["use_64_bit"])
            try:
                same(result_arrays(rb), result_arrays(alone))
            finally:
                alone.free()
        ora = po.run_rounds(po.Field(field, (0, 0, 0), (1, 1, 1)), jobs[1]["seeds"],
                            jobs[1]["remaining"], 15000, f32(0.25))
        compare_advect(batch[1], ora, 15000)
    finally:
        for r in batch:
            r.free()


def test_staged_batch_pass_replays_equal_eager(ctx, vortex):
    """The staged kernel instance inside a captured batch pass (45000 rows):
    every replay equals the eager batch bit for bit."""
    field, fld = vortex
    say("case staged batch pass, 2 replays")
    jobs = []
    for k in range(3):
        seeds, rem = uneven(15000, 640 + k)
        jobs.append(dict(field=fld, seeds=seeds, remaining=rem, uniform_remaining=0,
                         particles_per_round=None, step_size=f32(0.25), record=True,
                         use_64_bit=False))
    with ctx.batch_pass(jobs) as bp:
        for rep in range(2):
            eager = ctx.advect_batch(jobs)
            bp.launch()
            got = bp.results()
            try:
                for g, e in zip(got, eager):
                    same(result_arrays(g), result_arrays(e))
            finally:
                for r in got + eager:
                    r.free()
            say(f"  replay {rep} matches")


@pytest.mark.parametrize("integrator,stage", [(po.DPA_RK4, 2), (po.RK4, 2), (po.DPA_RK4, 0),
                                              (po.RK4, 0)])
def test_staged_pathlines_match_oracle(ctx, integrator, stage):
    """Staged slot blocks (slot_stage 2) and one thread per row (0) on a
    time-variant field (pathlines; the sample lerps time first), one round
    of 36000 rows, both integrators."""
    say(f"case pathlines, integrator {integrator}, slot_stage {stage}")
    before = ctx.tuning()
    ctx.set_tuning(slot_stage=stage)
    field = W.unsteady_vortex_field(32, time_steps=4, count=8)
    fld = ctx.field_unsteady(field, (0, 0, 0), (1, 1, 1), 0.0, 1.0)
    try:
        rng = np.random.default_rng(720 + integrator)
        seeds = W.uniform_seeds(36000, (0, 0, 0), (31, 31, 31), seed_offset=720 + integrator)
        rem = rng.integers(0, 45, len(seeds)).astype(np.uint64)
        dt, t0 = f32(0.02), 0.25
        ora = po.run_rounds(po.Field(field, (0, 0, 0), (1, 1, 1), 0.0, 1.0), seeds, rem, 36000,
                            dt, t0, integrator)
        res = ctx.advect(fld, seeds, remaining=rem, particles_per_round=36000, step_size=dt,
                         start_time=t0, integrator=integrator)
        try:
            compare_advect(res, ora, len(seeds))
        finally:
            res.free()
    finally:
        fld.release()
        ctx.set_tuning(**before)


@pytest.mark.parametrize("fastdiv,stage", [(1, 2), (0, 2), (1, 0), (0, 0)])
def test_staged_round_other_division_modes(ctx, fastdiv, stage):
    """Staged slot blocks (slot_stage 2) and one thread per row (0) in the
    verified-reciprocal (fastdiv 1) and the dividing (fastdiv 0) kernel
    instances: an anisotropic, offset grid whose spacings are not all powers
    of two."""
    say(f"case round, spacing (0.75, 0.1, 0.5), fastdiv={fastdiv}, slot_stage={stage}")
    offset, spacing = (0.25, -0.5, 1.0), (0.75, float(f32(0.1)), 0.5)
    data = W.vortex_field(36, count=10)
    before = ctx.tuning()
    ctx.set_tuning(fastdiv=fastdiv, slot_stage=stage)
    try:
        fld = ctx.field(data, offset, spacing)
        try:
            lo = np.array(offset, f32)
            hi = lo + np.array(spacing, f32) * 35
            seeds = W.uniform_seeds(36000, lo, hi, seed_offset=730 + fastdiv)
            rem = np.random.default_rng(731).integers(0, 80, len(seeds)).astype(np.uint64)
            dt = f32(0.002)
            ora = po.run_rounds(po.Field(data, offset, spacing), seeds, rem, 36000, dt)
            res = ctx.advect(fld, seeds, remaining=rem, particles_per_round=36000, step_size=dt)
            try:
                compare_advect(res, ora, len(seeds))
            finally:
                res.free()
        finally:
            fld.release()
    finally:
        ctx.set_tuning(**before)
