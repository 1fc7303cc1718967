"""Both field layouts against the oracle.

A field keeps its nodes and, when the context's tuning asks for it and it
fits, a cell layout (each cell's eight corners in one 128-B record).  The
kernels read corners from whichever the launch is given; the values are the
same copies, so every result must equal the oracle's bit for bit either way,
on steady and time-variant fields, with both integrators, for recorded
curves and for the auxiliary-neighbour FTLE.
"""
import sys
from contextlib import contextmanager

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_parity import bits, compare_advect

pytestmark = pytest.mark.gpu
f32 = np.float32
OFFSET, SPACING = (0.5, -1.0, 0.25), (0.75, 1.0, 0.5)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@contextmanager
def tuned(ctx, **kw):
    before = ctx.tuning()
    ctx.set_tuning(**kw)
    try:
        yield
    finally:
        ctx.set_tuning(**before)


def make_field(ctx, data, unsteady):
    if unsteady:
        return ctx.field_unsteady(data, OFFSET, SPACING, 0.5, 0.5)
    return ctx.field(data, OFFSET, SPACING)


@pytest.mark.parametrize("node_brick", [0, 2])
@pytest.mark.parametrize("cells", [0, 1, 2])
@pytest.mark.parametrize("unsteady", [False, True])
@pytest.mark.parametrize("integrator", [po.DPA_RK4, po.RK4])
def test_layout_matches_oracle(ctx, cells, node_brick, unsteady, integrator):
    # cells 0: the launches read the nodes, in C order (node_brick 0) or in
    # 4^3-node bricks (node_brick 2: steady and time-variant fields; the 40^3
    # and 24^3 grids pad their last bricks)
    say(f"case layout cells={cells} node_brick={node_brick} unsteady={unsteady} "
        f"integrator={integrator}")
    if unsteady:
        data = W.unsteady_vortex_field(24, time_steps=5, count=6)
        ofld = po.Field(data, OFFSET, SPACING, 0.5, 0.5)
    else:
        data = W.vortex_field(40, count=12)
        ofld = po.Field(data, OFFSET, SPACING)
    n = data.shape[-2]
    lo = np.array(OFFSET, f32)
    hi = lo + np.array(SPACING, f32) * (n - 1)
    seeds = W.uniform_seeds(3000, lo, hi, seed_offset=900 + cells)
    rem = np.full(len(seeds), 120, np.uint64)
    dt, t0 = (f32(0.01) if unsteady else f32(0.2)), f32(0.75)
    with tuned(ctx, cells=cells, node_brick=node_brick):
        fld = make_field(ctx, data, unsteady)
        try:
            res = ctx.advect(fld, seeds, remaining=rem, particles_per_round=1000, step_size=dt,
                             start_time=t0, integrator=integrator)
            try:
                ora = po.run_rounds(ofld, seeds, rem, 1000, dt, t0, integrator)
                compare_advect(res, ora, len(seeds))
                assert ora["total_steps"] > 50000
            finally:
                res.free()
        finally:
            fld.release()


@pytest.mark.parametrize("unsteady", [False, True])
def test_ftle_aux_same_on_both_layouts(ctx, unsteady):
    say(f"case ftle aux layouts unsteady={unsteady}")
    data = (W.unsteady_vortex_field(24, time_steps=5, count=6) if unsteady
            else W.vortex_field(40, count=12))
    shape = (6, 9, 11)
    out = []
    for cells, node_brick in ((0, 0), (0, 2), (2, 0)):   # C-order nodes, node bricks, cells
        with tuned(ctx, cells=cells, node_brick=node_brick):
            fld = make_field(ctx, data, unsteady)
            d = ctx.alloc(4 * int(np.prod(shape)))
            try:
                steps = ctx.ftle_aux_device(fld, (3.0, 2.0, 4.0), (1.5, 1.0, 0.75), shape,
                                            (0.25, 0.25, 0.25), 60, f32(0.02 if unsteady else 0.2),
                                            d, start_time=f32(0.6))
                out.append((ctx.to_host(d, np.empty(int(np.prod(shape)), f32)), steps))
            finally:
                ctx.free(d)
                fld.release()
    assert out[0][1] == out[1][1] == out[2][1] and out[0][1] > 0
    assert np.array_equal(bits(out[0][0]), bits(out[1][0]))
    assert np.array_equal(bits(out[0][0]), bits(out[2][0]))


def test_ftle_lattice_same_on_node_layouts(ctx):
    # dpa's lattice FTLE (every node a seed) through C-order nodes and node
    # bricks: the same volume bit for bit on a 37 x 42 x 29 grid, whose sides
    # leave partial bricks on every axis
    say("case ftle lattice node layouts")
    data = W.vortex_field(42, count=9)[:37, :, :29].copy()
    vols = []
    for node_brick in (0, 1):
        with tuned(ctx, cells=1, node_brick=node_brick):
            fld = ctx.field(data, OFFSET, SPACING)
            try:
                vols.append(ctx.ftle_lattice(fld, (1, 1, 1), 80, f32(0.2)))
            finally:
                fld.release()
    assert vols[0][1] == vols[1][1] and vols[0][1] > 0
    assert np.array_equal(bits(vols[0][0]), bits(vols[1][0]))


@pytest.mark.parametrize("cell_brick", [2, 3, 5])
@pytest.mark.parametrize("unsteady", [False, True])
def test_brick_ordered_cells_match_oracle(ctx, cell_brick, unsteady):
    # cell records stored brick by brick (2^k cells per axis, the cell grid
    # padded to whole bricks: 39 and 23 cells per axis here); every launch
    # reads the cells
    say(f"case brick-ordered cells k={cell_brick} unsteady={unsteady}")
    if unsteady:
        data = W.unsteady_vortex_field(24, time_steps=5, count=6)
        ofld = po.Field(data, OFFSET, SPACING, 0.5, 0.5)
    else:
        data = W.vortex_field(40, count=12)
        ofld = po.Field(data, OFFSET, SPACING)
    n = data.shape[-2]
    lo = np.array(OFFSET, f32)
    hi = lo + np.array(SPACING, f32) * (n - 1)
    seeds = W.uniform_seeds(3000, lo, hi, seed_offset=950 + cell_brick)
    rem = np.full(len(seeds), 120, np.uint64)
    dt, t0 = (f32(0.01) if unsteady else f32(0.2)), f32(0.75)
    with tuned(ctx, cells=2, cell_brick=cell_brick):
        fld = make_field(ctx, data, unsteady)
        try:
            res = ctx.advect(fld, seeds, remaining=rem, particles_per_round=1000, step_size=dt,
                             start_time=t0)
            try:
                ora = po.run_rounds(ofld, seeds, rem, 1000, dt, t0)
                compare_advect(res, ora, len(seeds))
                assert ora["total_steps"] > 50000
            finally:
                res.free()
        finally:
            fld.release()
