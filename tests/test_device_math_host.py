"""Device arithmetic of dpa_amd/csrc/device_math.hpp, built for the host.

* tests/native/cell_cache_check.cpp: the register cell cache keeps the corners
  of a particle's last cell; every sample through a warm cache must equal the
  sample through an empty one, bit for bit, over random walks on steady and
  time-variant grids, with the node layout.
* tests/native/brick_order_check.cpp: the closed-form brick order of lattice
  launches (queue position -> lattice index) is a bijection that visits
  bricks in C order, one contiguous run per brick, points in C order inside,
  also on lattices whose sides are not multiples of the brick.
* tests/native/slot_layout_check.cpp: the record-slot layout for slot shifts
  0 and 3 -- shift 0 is the one-slot-run formula, every slot of a region is
  distinct and inside it, staged blocks start 16-byte aligned, and the
  compaction's chunk offsets land on the same floats as slot_float.
* tests/native/div_check.cpp: the reciprocal division used on verified axes
  must equal the correctly rounded x / h bit for bit (all patterns around the
  switch to plain division and over [h/2, 2h], plus a sample of the whole
  axis range) for power-of-two, decimal and 2*pi/(n-1) spacings, and the plain
  reciprocal multiply of power-of-two spacings against x / h.  The device
  repeats the first check exhaustively per field before it uses the fast form.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = ["-O2", "-ffp-contract=off", "--offload-arch=gfx950", "-Xarch_host", "-mfma",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build_and_run(tmp_path, name):
    exe = tmp_path / name
    src = os.path.join(HERE, "native", name + ".cpp")
    subprocess.run([HIPCC, *FLAGS, "-o", str(exe), src], check=True, capture_output=True,
                   timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_warm_cache_equals_full_reload(tmp_path):
    words = build_and_run(tmp_path, "cell_cache_check").split()
    samples, shifts, roundups = int(words[1]), int(words[3]), int(words[5])
    assert samples > 100000 and shifts > 10000 and roundups >= 100


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_reciprocal_division_is_correctly_rounded(tmp_path):
    lines = build_and_run(tmp_path, "div_check").split("\n")
    rows = [l.split() for l in lines if l.strip()]
    assert len(rows) == 17 and all(r[1] == "0" for r in rows), lines


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_lattice_brick_order_is_a_bijection(tmp_path):
    words = build_and_run(tmp_path, "brick_order_check").split()
    assert int(words[1]) == 35 and int(words[3]) > 160000000 and int(words[5]) == 0, words


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_slot_layout_is_consistent(tmp_path):
    words = build_and_run(tmp_path, "slot_layout_check").split()
    assert int(words[1]) == 80 and int(words[3]) > 400000 and int(words[5]) == 0, words
