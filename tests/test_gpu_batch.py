"""Batched jobs in one persistent launch: every job's output equals the same
job run alone (and the oracle), bit for bit."""
import sys

import numpy as np
import pytest

from dpa_amd import workloads as W
from oracle import py_oracle as po
from test_gpu_parity import compare_advect, bits

pytestmark = pytest.mark.gpu
f32 = np.float32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def result_arrays(res):
    rounds = [res.round(r) for r in range(res.round_count)]
    return rounds, res.final_state(), res.total_steps


def same(a, b):
    (ra, sa, ta), (rb, sb, tb) = a, b
    assert ta == tb and len(ra) == len(rb)
    for (va, ca, ia), (vb, cb, ib) in zip(ra, rb):
        assert np.array_equal(bits(va), bits(vb)) and np.array_equal(ia, ib)
        assert np.array_equal(bits(ca), bits(cb))
    for x, y in zip(sa, sb):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_batch_uneven_jobs_match_alone_and_oracle(ctx):
    say("case batch: 12 uneven record jobs + 3 endpoint jobs on a vortex 64^3 field")
    field = W.vortex_field(64)
    fld = ctx.field(field, (0, 0, 0), (1, 1, 1))
    ofld = po.Field(field, (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(41)
    jobs = []
    try:
        for k in range(15):
            n = int(np.exp(rng.uniform(np.log(16), np.log(3000))))
            lo = rng.uniform(0, 40, 3)
            seeds = W.uniform_seeds(n, lo, lo + rng.uniform(4, 20, 3), seed_offset=100 + k)
            rem = rng.integers(8, 400, n) if k % 4 == 0 else None
            jobs.append(dict(field=fld, seeds=seeds, remaining=rem,
                             uniform_remaining=int(rng.integers(8, 400)),
                             particles_per_round=None if k % 3 else max(1, n // 3),
                             step_size=f32(0.25), record=k < 12, use_64_bit=(k == 5)))
        batch = ctx.advect_batch(jobs)
        for k, (j, rb) in enumerate(zip(jobs, batch)):
            alone = ctx.advect(fld, j["seeds"], remaining=j["remaining"],
                               uniform_remaining=j["uniform_remaining"],
                               particles_per_round=j["particles_per_round"] or len(j["seeds"]),
                               step_size=j["step_size"], record=j["record"],
                               use_64_bit=j["use_64_bit"])
            same(result_arrays(rb), result_arrays(alone))
            n = len(j["seeds"])
            rem = j["remaining"] if j["remaining"] is not None else j["uniform_remaining"]
            ora = po.run_rounds(ofld, j["seeds"], rem, j["particles_per_round"] or n,
                                j["step_size"], record=j["record"], use_64_bit=j["use_64_bit"])
            compare_advect(rb, ora, n)
            rb.free()
            alone.free()
    finally:
        fld.release()
