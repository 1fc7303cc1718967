"""The dpa_hip command line on the CPU: configuration parsing with dpa's keys
(argument_parser.cpp:17-57) and field files in dpa's HDF5 layout
(regular_grid_loader.cpp:65-88).  No GPU is touched."""
import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

import h5util
from dpa_amd import workloads as W

CLI = Path(__file__).resolve().parent.parent / "dpa_amd" / "bin" / "dpa_hip"

CONFIG = {
    "input_dataset_filepath": "in.h5", "input_dataset_name": "Data3",
    "input_dataset_spacing_name": "spacing", "seed_generation_stride": [1, 2, 4],
    "seed_generation_iterations": "1000",
    "seed_generation_boundaries": {"minimum": [0.0, 0.25, 0.5], "maximum": [1.0, 1.0, 1.0]},
    "particle_advector_particles_per_round": "10000000",
    "particle_advector_load_balancer": "diffuse_lesser_average",
    "particle_advector_integrator": "runge_kutta_4", "particle_advector_step_size": 0.001,
    "particle_advector_gather_particles": False, "particle_advector_record": True,
    "estimate_ftle": False, "output_dataset_filepath": "out.h5"}


def run(*args, **kw):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, **kw)


def test_check_config_reads_dpa_keys(tmp_path):
    p = tmp_path / "c.json"
    p.write_text(json.dumps(CONFIG))
    r = run("--check-config", p)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    assert got["seed_generation_iterations"] == 1000          # 64-bit ints come as strings
    assert got["particle_advector_particles_per_round"] == 10000000
    assert got["seed_generation_stride"] == [1, 2, 4]
    assert got["seed_generation_boundaries"]["minimum"] == [0.0, 0.25, 0.5]
    assert abs(got["particle_advector_step_size"] - 0.001) < 1e-9


def test_check_config_rejects_number_for_string_integer(tmp_path):
    bad = dict(CONFIG, seed_generation_iterations=1000)   # dpa requires a string here
    p = tmp_path / "c.json"
    p.write_text(json.dumps(bad))
    r = run("--check-config", p)
    assert r.returncode != 0 and "not a string" in r.stderr


def test_check_config_missing_key(tmp_path):
    bad = dict(CONFIG)
    del bad["particle_advector_step_size"]
    p = tmp_path / "c.json"
    p.write_text(json.dumps(bad))
    r = run("--check-config", p)
    assert r.returncode != 0 and "particle_advector_step_size" in r.stderr


@pytest.mark.skipif(not h5util.available(), reason="h5dump not in this image")
def test_make_abc_field_layout(tmp_path):
    out = tmp_path / "abc.h5"
    r = run("--make-abc-field", out, 6, 5, 4)
    assert r.returncode == 0, r.stderr
    data = h5util.read(out, "/Data3", np.float32).reshape(6, 5, 4, 3)
    want, h = W.abc_field((6, 5, 4))
    assert np.allclose(data, want, rtol=1e-6, atol=1e-6)
    assert "Data3" in h5util.datasets(out)
