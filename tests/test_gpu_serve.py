"""Serving: continuous batching gives each request the result it gets alone.

Requests of different kinds (recorded pathlines on a time-variant field,
endpoint streamlines on a steady field, auxiliary-neighbour FTLE) are admitted
together and advanced tick by tick; every result must equal, bit for bit, the
same request run alone through dpah_advect / dpah_ftle_aux_device (and the
oracle for the advection).
"""
import sys
import time

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_parity import bits, compare_advect

pytestmark = pytest.mark.gpu
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def arrays(res):
    return [res.round(r) for r in range(res.round_count)], res.final_state(), res.total_steps


def same(a, b):
    (ra, sa, ta), (rb, sb, tb) = a, b
    assert ta == tb and len(ra) == len(rb)
    for (va, ca, ia), (vb, cb, ib) in zip(ra, rb):
        assert np.array_equal(bits(va), bits(vb)) and np.array_equal(ia, ib)
        assert np.array_equal(bits(ca), bits(cb))
    for x, y in zip(sa, sb):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


@pytest.fixture(scope="module")
def fields(ctx):
    steady = W.vortex_field(48, count=16)
    unsteady = W.unsteady_vortex_field(32, time_steps=4, count=8)
    fs = ctx.field(steady, (0, 0, 0), (1, 1, 1))
    fu = ctx.field_unsteady(unsteady, (0, 0, 0), (1, 1, 1), 0.0, 1.0)
    yield dict(steady=(fs, steady), unsteady=(fu, unsteady))
    fs.release()
    fu.release()


def make_requests(fields):
    fs, _ = fields["steady"]
    fu, _ = fields["unsteady"]
    reqs = []
    for k in range(8):
        n = [40, 300, 1200, 77, 500, 9, 640, 2000][k]
        if k % 3 == 0:
            seeds = W.uniform_seeds(n, (0, 0, 0), (31, 31, 31), seed_offset=300 + k)
            reqs.append(dict(kind="advect", field=fu, seeds=seeds, uniform_remaining=60 + 37 * k,
                             step_size=f32(0.02), start_time=f32(0.1 * k), record=True,
                             particles_per_round=max(1, n // 2)))
        else:
            seeds = W.uniform_seeds(n, (0, 0, 0), (47, 47, 47), seed_offset=300 + k)
            reqs.append(dict(kind="advect", field=fs, seeds=seeds, uniform_remaining=50 + 90 * k,
                             step_size=f32(0.25), start_time=f32(0.0), record=(k % 2 == 0),
                             particles_per_round=None))
    # delta half the spacing (shared particles) and not (six per point)
    reqs.append(dict(kind="ftle", field=fs, origin=(2, 2, 2), spacing=(1, 1, 1), shape=(8, 10, 6),
                     delta=(0.5, 0.5, 0.5), iterations=150, step_size=f32(0.25)))
    reqs.append(dict(kind="ftle", field=fs, origin=(3, 1, 2), spacing=(1, 1, 1), shape=(7, 9, 5),
                     delta=(0.3, 0.3, 0.3), iterations=120, step_size=f32(0.25)))
    return reqs


def submit(server, r):
    if r["kind"] == "ftle":
        return server.submit_ftle(r["field"], r["origin"], r["spacing"], r["shape"], r["delta"],
                                  r["iterations"], r["step_size"])
    return server.submit_advect(r["field"], r["seeds"], uniform_remaining=r["uniform_remaining"],
                                particles_per_round=r["particles_per_round"],
                                step_size=r["step_size"], start_time=r["start_time"],
                                record=r["record"])


def alone(ctx, r):
    if r["kind"] == "ftle":
        n = int(np.prod(r["shape"]))
        d = ctx.alloc(4 * n)
        try:
            steps = ctx.ftle_aux_device(r["field"], r["origin"], r["spacing"], r["shape"],
                                        r["delta"], r["iterations"], r["step_size"], d)
            return ctx.to_host(d, np.empty(n, f32)), steps
        finally:
            ctx.free(d)
    return ctx.advect(r["field"], r["seeds"], uniform_remaining=r["uniform_remaining"],
                      particles_per_round=r["particles_per_round"] or len(r["seeds"]),
                      step_size=r["step_size"], start_time=r["start_time"], record=r["record"])


@pytest.mark.parametrize("mode,tick,host", [(dpa_amd.SERVE_BATCHED, 16, False),
                                            (dpa_amd.SERVE_BATCHED, 1000, False),
                                            (dpa_amd.SERVE_BATCHED, 0, False),
                                            (dpa_amd.SERVE_BATCHED, 0, True),
                                            (dpa_amd.SERVE_PER_REQUEST, 0, False),
                                            (dpa_amd.SERVE_PER_REQUEST, 0, True),
                                            (dpa_amd.SERVE_SERIAL, 0, False)])
def test_mixed_requests_match_alone(ctx, fields, mode, tick, host):
    say(f"case serve mode={mode} steps_per_tick={tick} results_on_host={host}")
    reqs = make_requests(fields)
    with dpa_amd.Server(ctx, mode=mode, steps_per_tick=tick, results_on_host=host) as server:
        tickets = [submit(server, r) for r in reqs]
        results = [server.wait(t, 120000, n=len(r.get("seeds", []))) for t, r in zip(tickets, reqs)]
        infos = [server.info(t) for t in tickets]
        st = server.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for r, res, info in zip(reqs, results, infos):
        assert info["state"] == dpa_amd.TICKET_DONE and info["latency_seconds"] > 0
        ref = alone(ctx, r)
        if r["kind"] == "ftle":
            got = dpa_amd.result_ftle(res)
            assert np.array_equal(bits(got), bits(ref[0]))
            assert res.total_steps == ref[1]
        else:
            same(arrays(res), arrays(ref))
            ref.free()
        res.free()
    # one pathline request against the oracle as well
    fu, unsteady = fields["unsteady"]
    r0 = reqs[0]
    ora = po.run_rounds(po.Field(unsteady, (0, 0, 0), (1, 1, 1), 0.0, 1.0), r0["seeds"],
                        r0["uniform_remaining"], r0["particles_per_round"], r0["step_size"],
                        r0["start_time"], record=True)
    res = alone(ctx, r0)
    compare_advect(res, ora, len(r0["seeds"]))
    res.free()


def test_admission_rejects_and_serialises(ctx, fields):
    say("case serve admission: budget fits one request at a time; oversize rejected")
    fs, _ = fields["steady"]
    seeds = W.uniform_seeds(1000, (0, 0, 0), (47, 47, 47), seed_offset=400)
    with dpa_amd.Server(ctx, memory_budget=200_000, steps_per_tick=64) as server:
        with pytest.raises(dpa_amd.DpahError) as e:
            server.submit_advect(fs, W.uniform_seeds(20000, (0, 0, 0), (47, 47, 47), 401),
                                 uniform_remaining=100, step_size=0.25, record=False)
        assert e.value.code == dpa_amd.ERR_ADMISSION
        ts = [server.submit_advect(fs, seeds, uniform_remaining=200, step_size=0.25, record=False)
              for _ in range(3)]
        res = [server.wait(t, 120000, n=1000) for t in ts]
        st = server.stats()
    assert st["rejected"] == 1 and st["completed"] == 3
    for r in res[1:]:
        same(arrays(r), arrays(res[0]))
    for r in res:
        r.free()


def test_cancel_queued_request(ctx, fields):
    say("case serve cancel: second request cancelled while the first runs")
    fs, _ = fields["steady"]
    big = W.uniform_seeds(50000, (0, 0, 0), (47, 47, 47), seed_offset=402)
    small = W.uniform_seeds(10, (0, 0, 0), (47, 47, 47), seed_offset=403)
    with dpa_amd.Server(ctx, max_active=1, steps_per_tick=8) as server:
        t1 = server.submit_advect(fs, big, uniform_remaining=2000, step_size=0.25, record=False)
        t2 = server.submit_advect(fs, small, uniform_remaining=10, step_size=0.25, record=False)
        server.cancel(t2)
        with pytest.raises(dpa_amd.DpahError) as e:
            server.wait(t2, 120000)
        assert e.value.code == dpa_amd.ERR_CANCELLED
        server.wait(t1, 120000, n=len(big)).free()
        assert server.stats()["cancelled"] == 1
