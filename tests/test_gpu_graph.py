"""Endpoint passes captured into a hipGraph.

Every replay of the captured pass must reproduce the eager call's outputs bit
for bit (and the oracle's), with the program's own eager work and
synchronisations between replays, for a pass that orders its queue (>= 2^15
seeds) and one that does not.
"""
import sys

import numpy as np
import pytest

import dpa_amd
from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_parity import bits

pytestmark = pytest.mark.gpu
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [50000, 3000])
def test_graph_replays_match_eager_and_oracle(ctx, n):
    say(f"case graph pass n={n}")
    field = W.vortex_field(64, count=24)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    seeds = W.uniform_seeds(n, (0, 0, 0), (63, 63, 63), seed_offset=321)
    dt, iters = f32(0.25), 120
    ora = po.advect_endpoints(po.Field(field, (0, 0, 0), (1, 1, 1)), seeds, iters, dt)
    d_s = ctx.to_device(seeds)
    d_p, d_r, d_t = ctx.alloc(12 * n), ctx.alloc(8 * n), ctx.alloc(n)
    e_p, e_r, e_t = ctx.alloc(12 * n), ctx.alloc(8 * n), ctx.alloc(n)
    try:
        with ctx.pass_endpoints(fld, d_s, n, d_p, uniform_remaining=iters, step_size=dt,
                                d_rem_out=d_r, d_status_out=d_t) as gp:
            for rep in range(4):
                # eager work between replays, on other buffers
                steps_eager = ctx.advect_endpoints_device(fld, d_s, n, e_p, uniform_remaining=iters,
                                                          step_size=dt, d_rem_out=e_r,
                                                          d_status_out=e_t)
                eager = (ctx.to_host(e_p, np.empty((n, 3), f32)),
                         ctx.to_host(e_r, np.empty(n, np.uint64)), ctx.to_host(e_t, np.empty(n, np.uint8)))
                ctx.memset_zero(d_p, 12 * n)     # the replay must rewrite every output
                gp.launch()
                steps = gp.total_steps()
                got = (ctx.to_host(d_p, np.empty((n, 3), f32)),
                       ctx.to_host(d_r, np.empty(n, np.uint64)), ctx.to_host(d_t, np.empty(n, np.uint8)))
                assert steps == steps_eager == ora["total_steps"], (rep, steps, steps_eager)
                assert np.array_equal(bits(got[0]), bits(eager[0]))
                assert np.array_equal(bits(got[0]), bits(ora["position"]))
                assert np.array_equal(got[1], ora["remaining"]) and np.array_equal(got[2], ora["status"])
                assert np.array_equal(got[1], eager[1]) and np.array_equal(got[2], eager[2])
    finally:
        for p in (d_s, d_p, d_r, d_t, e_p, e_r, e_t):
            ctx.free(p)
        fld.release()


def test_batch_pass_replays_match_eager_batch_and_oracle(ctx):
    """A captured batch (three launch groups: recorded at two step sizes and an
    endpoint group with both integrators' jobs split by integrator) replayed
    three times, with an eager batch of the same jobs before every replay."""
    from test_gpu_batch import result_arrays, same
    from test_gpu_parity import compare_advect
    say("case batch pass: 9 jobs, recorded and endpoint groups, 3 replays")
    field = W.vortex_field(64, count=24)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    ofld = po.Field(field, (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(77)
    jobs = []
    for k in range(9):
        n = int(np.exp(rng.uniform(np.log(16), np.log(1500))))
        lo = rng.uniform(0, 40, 3)
        seeds = W.uniform_seeds(n, lo, lo + rng.uniform(4, 20, 3), seed_offset=700 + k)
        rem = rng.integers(8, 200, n) if k % 3 == 0 else None
        jobs.append(dict(field=fld, seeds=seeds, remaining=rem,
                         uniform_remaining=int(rng.integers(8, 200)),
                         particles_per_round=None if k % 2 else max(1, n // 3),
                         step_size=f32(0.25 if k < 6 else 0.2), record=k not in (2, 5),
                         integrator=po.RK4 if k == 5 else po.DPA_RK4, use_64_bit=(k == 4)))
    try:
        with pytest.raises(dpa_amd.DpahError):
            ctx.batch_pass([])
        with ctx.batch_pass(jobs) as bp:
            with pytest.raises(dpa_amd.DpahError):
                bp.results()   # nothing replayed yet
            for rep in range(3):
                eager = ctx.advect_batch(jobs)
                bp.launch()
                got = bp.results()
                for k, (g, e) in enumerate(zip(got, eager)):
                    same(result_arrays(g), result_arrays(e))
                    if rep == 0:
                        j = jobs[k]
                        n = len(j["seeds"])
                        rem = j["remaining"] if j["remaining"] is not None else j["uniform_remaining"]
                        ora = po.run_rounds(ofld, j["seeds"], rem, j["particles_per_round"] or n,
                                            j["step_size"], integrator=j["integrator"],
                                            record=j["record"], use_64_bit=j["use_64_bit"])
                        compare_advect(g, ora, n)
                for r in got + eager:
                    r.free()
                say(f"  replay {rep} matches")
    finally:
        fld.release()


def test_batch_pass_time_variant_both_integrators(ctx):
    """A captured batch over a time-variant field (pathlines): recorded and
    endpoint jobs, dpa runge_kutta_4 and classic RK4, start times apart, so
    the graph holds several groups of unsteady launches.  Two replays, each
    equal to the eager batch, and the oracle's rounds."""
    from test_gpu_batch import result_arrays, same
    from test_gpu_parity import compare_advect
    say("case batch pass, time-variant field, 6 jobs")
    field = W.unsteady_vortex_field(32, time_steps=4, count=8)
    fld = ctx.field_unsteady(field, (0, 0, 0), (1, 1, 1), 0.0, 1.0)
    ofld = po.Field(field, (0, 0, 0), (1, 1, 1), 0.0, 1.0)
    jobs = []
    for k in range(6):
        seeds = W.uniform_seeds(150 + 70 * k, (0, 0, 0), (31, 31, 31), seed_offset=900 + k)
        jobs.append(dict(field=fld, seeds=seeds, remaining=None, uniform_remaining=60 + 15 * k,
                         particles_per_round=None if k % 2 else 100, step_size=f32(0.02),
                         start_time=float(f32(0.1 * k)), record=k != 3,
                         integrator=po.RK4 if k >= 4 else po.DPA_RK4, use_64_bit=False))
    try:
        with ctx.batch_pass(jobs) as bp:
            for rep in range(2):
                eager = ctx.advect_batch(jobs)
                bp.launch()
                got = bp.results()
                try:
                    for k, (g, e) in enumerate(zip(got, eager)):
                        same(result_arrays(g), result_arrays(e))
                        if rep == 0:
                            j = jobs[k]
                            n = len(j["seeds"])
                            ora = po.run_rounds(ofld, j["seeds"], j["uniform_remaining"],
                                                j["particles_per_round"] or n, j["step_size"],
                                                t0=j["start_time"], integrator=j["integrator"],
                                                record=j["record"])
                            compare_advect(g, ora, n)
                finally:
                    for r in got + eager:
                        r.free()
                say(f"  replay {rep} matches")
    finally:
        fld.release()
