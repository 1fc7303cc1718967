"""The cell-load counting instance behind the roofline's byte count.

A counting launch must give the same results as a normal one, and the count
must be what the trajectories imply: a particle samples the field at x_0 ..
x_{k-1} for its k steps (and once more at x_k when it retires on a zero
vector; a failed contains test or a finished count samples nothing), and
every sample whose cell floor((x - o) / h) differs from the previous
sample's (the first sample always) reloads the register cell cache.  The
oracle's recorded vertices give the cells.
"""
import numpy as np
import pytest

from dpa_amd import workloads as W
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu
f32 = np.float32


def expected_loads(ora, offset, spacing):
    v, _, _ = ora["rounds"][0]
    steps, status = ora["steps"], ora["status"]
    o, h = np.asarray(offset, f32), np.asarray(spacing, f32)
    total, base = 0, 0
    for k, st in zip(steps.astype(np.int64), status):
        samples = k + (1 if st == po.ZERO else 0)
        x = v[base:base + samples]
        base += k + 2
        if samples == 0:
            continue
        cell = np.floor((x - o).astype(f32) / h).astype(f32)
        total += 1 + int(np.count_nonzero(np.any(cell[1:] != cell[:-1], axis=1)))
    return total


@pytest.mark.parametrize("cells", [1, 0])
def test_cell_loads_match_trajectories(ctx, cells):
    field = W.vortex_field(40, count=10)
    offset, spacing = (0.5, -1.0, 0.25), (0.75, 1.0, 0.5)
    prev = ctx.tuning()
    ctx.set_tuning(cells=cells)
    try:
        fld = ctx.field(field, offset, spacing)
        lo = np.array(offset, f32)
        seeds = W.uniform_seeds(3000, lo, lo + np.array(spacing, f32) * 39, seed_offset=41)
        seeds[::101] = (-5.0, 0.0, 0.0)                     # outside: no sample, no load
        n, iters, dt = len(seeds), 150, f32(0.2)
        d_s = ctx.to_device(seeds)
        d_p, d_q = ctx.alloc(12 * n), ctx.alloc(12 * n)
        try:
            s0 = ctx.advect_endpoints_device(fld, d_s, n, d_p, uniform_remaining=iters,
                                             step_size=dt)
            ctx.set_count_cell_loads(True)
            s1 = ctx.advect_endpoints_device(fld, d_s, n, d_q, uniform_remaining=iters,
                                             step_size=dt)
            rec, node = ctx.cell_loads()
            ctx.set_count_cell_loads(False)
            a = ctx.to_host(d_p, np.empty((n, 3), f32))
            b = ctx.to_host(d_q, np.empty((n, 3), f32))
        finally:
            ctx.set_count_cell_loads(False)
            for p in (d_s, d_p, d_q):
                ctx.free(p)
        fld.release()
    finally:
        ctx.set_tuning(**{k: prev[k] for k in ("cells",)})
    assert s0 == s1 and np.array_equal(a.view(np.uint32), b.view(np.uint32))
    ora = po.run_rounds(po.Field(field, offset, spacing), seeds, iters, n, dt, record=True)
    assert s1 == ora["total_steps"]
    want = expected_loads(ora, offset, spacing)
    got = rec if cells else node
    assert (rec if not cells else node) == 0
    assert got == want, (got, want)
    assert 0.05 * s1 < got < s1


def test_counting_off_counts_nothing(ctx):
    field = W.vortex_field(24, count=4)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    seeds = W.uniform_seeds(500, (0, 0, 0), (23, 23, 23), seed_offset=42)
    d_s, d_p = ctx.to_device(seeds), ctx.alloc(12 * 500)
    try:
        ctx.set_count_cell_loads(True)
        ctx.set_count_cell_loads(False)
        ctx.advect_endpoints_device(fld, d_s, 500, d_p, uniform_remaining=50, step_size=0.25)
        assert ctx.cell_loads() == (0, 0)
    finally:
        ctx.free(d_s)
        ctx.free(d_p)
        fld.release()
