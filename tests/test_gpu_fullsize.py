"""The FTLE, batch and serving workloads at their full sizes.

  FTLE lattice   ABC on 512 x 512 x 128 nodes, T = 4 at dt 0.02 (200 steps),
                 in dpa's lattice mode (stride 1) and in the +-0.5-cell
                 auxiliary mode.  The whole flow map is also advected as a
                 seed array: its per-particle steps must add up to the lattice
                 launch's total (in the auxiliary mode each particle shared by
                 two neighbouring points counted once), a strided sample of its
                 final positions must
                 equal the oracle's bit for bit, and FTLE at a strided sample
                 of points must be within 1e-3 |ref| + 1e-5 / T of the float64
                 reduction of the oracle on the same flow map.
  batch          64 recorded jobs (256 .. 262144 seeds, 64 .. 4096 steps) on
                 the 256^3 Gaussian-vortex field in one persistent launch:
                 every job equals the same job run alone, bit for bit, and the
                 six smallest equal the oracle.
  serving        4 resident 256^3 x 16-slice time-variant fields, a 500 req/s
                 Poisson trace of recorded pathlines (1k .. 256k seeds,
                 128 .. 2048 steps) and 256 x 256 x 16 FTLE slabs through the
                 continuous-batching server: every result equals the request
                 run alone, bit for bit, and the smallest pathline request
                 equals the oracle.
"""
import math
import sys

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_ftle import assert_ftle_close
from test_gpu_parity import bits, compare_advect

pytestmark = pytest.mark.gpu
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def endpoints(ctx, fld, seeds, iters, dt):
    """Device endpoints of a host seed array: (positions, remaining, status, total)."""
    n = len(seeds)
    d_s = ctx.to_device(seeds)
    d_p, d_r, d_t = ctx.alloc(12 * n), ctx.alloc(8 * n), ctx.alloc(n)
    try:
        total = ctx.advect_endpoints_device(fld, d_s, n, d_p, uniform_remaining=iters,
                                            step_size=dt, d_rem_out=d_r, d_status_out=d_t)
        return (ctx.to_host(d_p, np.empty((n, 3), f32)), ctx.to_host(d_r, np.empty(n, np.uint64)),
                ctx.to_host(d_t, np.empty(n, np.uint8)), total)
    finally:
        for p in (d_s, d_p, d_r, d_t):
            ctx.free(p)


@pytest.fixture(scope="module")
def abc512(ctx):
    c = W.ftle_abc512()
    fld = ctx.field(c["field"], c["offset"], c["spacing"])
    yield c, fld
    fld.release()


def test_ftle_lattice_mode_abc512(ctx, abc512):
    c, fld = abc512
    h, off, iters, dt = c["spacing"], c["offset"], c["iterations"], c["step_size"]
    shape = np.array(c["field"].shape[:3])
    T = float(f32(dt * f32(iters)))
    say("case full-size FTLE, dpa lattice mode, ABC 512x512x128")
    vol, steps = ctx.ftle_lattice(fld, (1.0, 1.0, 1.0), iters, dt)
    assert vol.shape == tuple(shape)
    # the same lattice seeds as a seed array: per-particle steps and the flow map
    seeds, per = po.lattice_seeds(off, (h * shape.astype(f32)).astype(f32), h)
    n = len(seeds)
    F, rem, st, total = endpoints(ctx, fld, seeds, iters, dt)
    assert total == steps == int(np.sum(iters - rem.astype(np.int64)))
    sample = np.arange(0, n, 1009)
    ora = po.advect_endpoints(po.Field(c["field"], off, h), seeds[sample], iters, dt)
    assert np.array_equal(bits(F[sample]), bits(ora["position"]))
    assert np.array_equal(rem[sample], ora["remaining"]) and np.array_equal(st[sample], ora["status"])
    # dpa's scatter (ftle_estimator.cpp:39-43): the cell of each seed, the
    # highest particle index winning a shared cell; unscattered cells stay 0
    cell = np.floor((seeds - off).astype(f32) / h).astype(np.int64)
    ok = np.all((cell >= 0) & (cell < shape), axis=1)
    lin = (cell[:, 0] * shape[1] + cell[:, 1]) * shape[2] + cell[:, 2]
    idx = np.flatnonzero(ok)
    rev = lin[idx][::-1]
    uniq, first_in_rev = np.unique(rev, return_index=True)
    owner = idx[::-1][first_in_rev]
    flow = np.zeros((int(np.prod(shape)), 3), f32)
    flow[uniq] = F[owner]
    flow = flow.reshape(tuple(shape) + (3,))
    # clamped central differences divided by 2h (regular_grid.hpp:95-105),
    # reduced in float64 by the oracle, at a strided sample of points
    pts = np.arange(0, int(np.prod(shape)), 997)
    p3 = np.stack(np.unravel_index(pts, tuple(shape)), 1)
    fin = np.empty((len(pts), 6, 3), f32)
    for d in range(3):
        lo, hi = p3.copy(), p3.copy()
        lo[:, d] = np.maximum(lo[:, d] - 1, 0)
        hi[:, d] = np.minimum(hi[:, d] + 1, shape[d] - 1)
        fin[:, 2 * d] = flow[lo[:, 0], lo[:, 1], lo[:, 2]]
        fin[:, 2 * d + 1] = flow[hi[:, 0], hi[:, 1], hi[:, 2]]
    want = po.ftle_aux(fin, h, T)
    assert_ftle_close(vol.reshape(-1)[pts], want, T)


def test_ftle_aux_mode_abc512(ctx, abc512):
    c, fld = abc512
    h, off, iters, dt = c["spacing"], c["offset"], c["iterations"], c["step_size"]
    shape, delta = c["lattice_shape"], c["delta"]
    P = int(np.prod(shape))
    T = float(f32(dt * f32(iters)))
    say("case full-size FTLE, +-0.5-cell auxiliary mode, ABC 512x512x128")
    d_out = ctx.alloc(4 * P)
    try:
        steps = ctx.ftle_aux_device(fld, c["lattice_origin"], c["lattice_spacing"], shape, delta,
                                    iters, dt, d_out)
        vol = ctx.to_host(d_out, np.empty(P, f32))
    finally:
        ctx.free(d_out)
    # the six particles of every point as a seed array: the steps of the
    # particles the launch advected (each shared by two points once) add up
    seeds = po.aux_seeds(c["lattice_origin"], c["lattice_spacing"], shape, delta)
    _, rem, _, total = endpoints(ctx, fld, seeds, iters, dt)
    assert po.aux_shares(c["lattice_spacing"], delta, shape)
    assert steps == po.aux_particle_steps(rem, shape, iters, True)
    assert total == int(np.sum(iters - rem.astype(np.int64)))
    del rem
    # a strided sample of points against the oracle
    pts = np.arange(0, P, 499)
    s6 = seeds.reshape(P, 6, 3)[pts].copy()
    del seeds
    ora = po.advect_endpoints(po.Field(c["field"], off, h), s6.reshape(-1, 3), iters, dt)
    want = po.ftle_aux(ora["position"], delta, T)
    assert_ftle_close(vol[pts], want, T)


def test_batch64_vortex256(ctx):
    say("case full-size batch: 64 recorded jobs on the 256^3 vortex field")
    field = W.vortex_field(256)
    jobs = W.batch_jobs()
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    try:
        for j in jobs:
            j["field"] = fld
        batch = ctx.advect_batch(jobs)
        smallest = sorted(range(len(jobs)), key=lambda i: len(jobs[i]["seeds"]))[:6]
        for i, (j, res) in enumerate(zip(jobs, batch)):
            alone = ctx.advect(fld, j["seeds"], uniform_remaining=j["uniform_remaining"],
                               particles_per_round=j["particles_per_round"],
                               step_size=j["step_size"], record=True,
                               use_64_bit=j["use_64_bit"])
            assert res.total_steps == alone.total_steps and res.round_count == alone.round_count
            for r in range(res.round_count):
                (va, ca, ia), (vb, cb, ib) = res.round(r), alone.round(r)
                assert np.array_equal(bits(va), bits(vb)) and np.array_equal(ia, ib), (i, r)
                assert np.array_equal(bits(ca), bits(cb)), (i, r)
                del va, ca, ia, vb, cb, ib
            for x, y in zip(res.final_state(), alone.final_state()):
                assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), i
            alone.free()
            if i in smallest:
                ora = po.run_rounds(po.Field(field, (0, 0, 0), (1, 1, 1)), j["seeds"],
                                    j["uniform_remaining"], j["particles_per_round"],
                                    j["step_size"], record=True)
                compare_advect(res, ora, len(j["seeds"]))
            res.free()
    finally:
        fld.release()


def test_serve_four_fields_500rps(ctx):
    say("case full-size serving: 4 x 16-slice 256^3 fields, 500 req/s, 256x256x16 FTLE slabs")
    base = W.vortex_field(256)
    hosts, fields = [], []
    try:
        hosts = W.serving_fields(base)
        for h in hosts:
            fields.append(ctx.field_unsteady(h, (0, 0, 0), (1, 1, 1), 0.0, 1.0))
        trace = W.serving_trace(500.0, 40)
        assert sum(r["kind"] == "ftle" for r in trace) >= 4
        with dpa_amd.Server(ctx, mode=dpa_amd.SERVE_BATCHED) as server:
            tickets = []
            for r in trace:
                f = fields[r["field"]]
                if r["kind"] == "ftle":
                    tickets.append(server.submit_ftle(f, r["origin"], r["spacing"], r["shape"],
                                                      r["delta"], r["iterations"],
                                                      r["step_size"], start_time=r["start_time"]))
                else:
                    tickets.append(server.submit_advect(
                        f, r["seeds"], uniform_remaining=r["uniform_remaining"],
                        step_size=r["step_size"], start_time=r["start_time"], record=True))
            for t, r in zip(tickets, trace):
                res = server.wait(t, 600000, n=len(r.get("seeds", [])))
                f = fields[r["field"]]
                if r["kind"] == "ftle":
                    P = int(np.prod(r["shape"]))
                    d = ctx.alloc(4 * P)
                    try:
                        s = ctx.ftle_aux_device(f, r["origin"], r["spacing"], r["shape"],
                                                r["delta"], r["iterations"], r["step_size"], d,
                                                start_time=r["start_time"])
                        want = ctx.to_host(d, np.empty(P, f32))
                    finally:
                        ctx.free(d)
                    assert res.total_steps == s
                    assert np.array_equal(bits(dpa_amd.result_ftle(res)), bits(want))
                else:
                    alone = ctx.advect(f, r["seeds"], uniform_remaining=r["uniform_remaining"],
                                       particles_per_round=len(r["seeds"]),
                                       step_size=r["step_size"], start_time=r["start_time"],
                                       record=True)
                    assert res.total_steps == alone.total_steps
                    (va, ca, ia), (vb, cb, ib) = res.round(0), alone.round(0)
                    assert np.array_equal(bits(va), bits(vb)) and np.array_equal(ia, ib)
                    assert np.array_equal(bits(ca), bits(cb))
                    del va, ca, ia, vb, cb, ib
                    alone.free()
                res.free()
                server.forget(t)
        # the smallest pathline request against the oracle
        paths = [r for r in trace if r["kind"] == "advect"]
        r = min(paths, key=lambda q: len(q["seeds"]) * q["uniform_remaining"])
        res = ctx.advect(fields[r["field"]], r["seeds"], uniform_remaining=r["uniform_remaining"],
                         particles_per_round=len(r["seeds"]), step_size=r["step_size"],
                         start_time=r["start_time"], record=True)
        ora = po.run_rounds(po.Field(hosts[r["field"]], (0, 0, 0), (1, 1, 1), 0.0, 1.0),
                            r["seeds"], r["uniform_remaining"], len(r["seeds"]), r["step_size"],
                            r["start_time"], record=True)
        compare_advect(res, ora, len(r["seeds"]))
        res.free()
    finally:
        for f in fields:
            f.release()
