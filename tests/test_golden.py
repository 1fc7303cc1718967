"""The CPU oracle against the committed regression fixtures (tests/golden/).

The fixtures are self-generated by the oracle (tests/golden/make_fixtures.py),
not held by dpa, which ships none: they freeze the oracle's outputs on small
inputs, so any change that alters a bit of them fails here.  Every array is
compared bit for bit.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import py_oracle as po

GOLDEN = Path(__file__).resolve().parent / "golden"
f32 = np.float32


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float32:
        return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return a.shape == b.shape and np.array_equal(a, b)


def check_rounds(ora, g, prefix=""):
    assert len(ora["rounds"]) == int(g[prefix + "round_count"])
    for r, (v, c, i) in enumerate(ora["rounds"]):
        assert same(v, g[f"{prefix}vertices_{r}"]), f"round {r} vertices"
        assert same(c, g[f"{prefix}colors_{r}"]), f"round {r} colours"
        assert same(i, g[f"{prefix}indices_{r}"]), f"round {r} indices"
    assert same(ora["position"], g[prefix + "position"])
    assert same(ora["remaining"], g[prefix + "remaining"])
    assert same(ora["status"], g[prefix + "status"])
    assert ora["total_steps"] == int(g[prefix + "total_steps"])


def test_abc16_rounds():
    g = load("abc16_rounds")
    ora = po.run_rounds(po.Field(g["field"], g["offset"], g["spacing"]), g["seeds"],
                        g["remaining_in"], int(g["particles_per_round"]), g["step_size"])
    check_rounds(ora, g)
    assert int(g["round_count"]) == 3 and set(np.unique(g["status"])) >= {1, 2}


def test_vortex24_endpoints():
    g = load("vortex24_endpoints")
    ora = po.advect_endpoints(po.Field(g["field"], g["offset"], g["spacing"]), g["seeds"],
                              int(g["iterations"]), g["step_size"])
    for k in ("position", "remaining", "status"):
        assert same(ora[k], g[k]), k
    assert ora["total_steps"] == int(g["total_steps"])


@pytest.mark.parametrize("name,integ", [("dpa", po.DPA_RK4), ("rk4", po.RK4)])
def test_unsteady12_rounds(name, integ):
    g = load("unsteady12_rounds")
    fld = po.Field(g["field"], g["offset"], g["spacing"], g["time_offset"], g["time_spacing"])
    ora = po.run_rounds(fld, g["seeds"], int(g["iterations"]), len(g["seeds"]), g["step_size"],
                        g["start_time"], integ)
    check_rounds(ora, g, name + "_")


def test_ftle_aux_abc():
    g = load("ftle_aux_abc")
    seeds = po.aux_seeds(g["offset"], g["spacing"], tuple(int(x) for x in g["lattice_shape"]),
                         g["delta"])
    res = po.advect_endpoints(po.Field(g["field"], g["offset"], g["spacing"]), seeds,
                              int(g["iterations"]), g["step_size"])
    T = float(f32(g["step_size"] * f32(g["iterations"])))
    assert res["total_steps"] == int(g["total_steps"])
    assert same(po.ftle_aux(res["position"], g["delta"], T).astype(f32), g["ftle"])
    shape = tuple(int(x) for x in g["lattice_shape"])
    assert po.aux_shares(g["spacing"], g["delta"], shape)
    assert po.aux_particle_steps(res["remaining"], shape, int(g["iterations"]),
                                 True) == int(g["shared_steps"])
