"""Pins of the CPU oracle against closed forms, invariants and brute force.

dpa ships no tests, fixtures or golden vectors for this path (its tests/ holds
only the Catch2 main), and the reference cannot be built here, so every check
below comes from mathematics the reference's algorithm must satisfy.  Each is
chosen so that a plausible slip (swapped axes, wrong lerp weight, missing
swizzle, Euler instead of the four-term sum, off-by-one exit) fails.
"""
import math

import numpy as np
import pytest

from oracle import py_oracle as po

f32 = np.float32


def grid_coords(shape, h, o):
    i0, i1, i2 = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return [o[k] + h[k] * g for k, g in enumerate((i0, i1, i2))]


# ------------------------------------------------------ contains / interpolate

def test_trilinear_reproduces_multilinear_fields():
    # Any function multilinear in (x0, x1, x2) is reproduced exactly by
    # trilinear interpolation; a swapped weight / corner order breaks it.
    rng = np.random.default_rng(1)
    shape, h, o = (7, 9, 11), np.array([0.5, 0.25, 1.0]), np.array([-1.0, 2.0, 0.5])
    x0, x1, x2 = grid_coords(shape, h, o)
    coef = rng.normal(size=(3, 8))

    def F(c, a, b, d):
        return (c[0] + c[1] * a + c[2] * b + c[3] * d + c[4] * a * b + c[5] * a * d
                + c[6] * b * d + c[7] * a * b * d)

    data = np.stack([F(coef[e], x0, x1, x2) for e in range(3)], -1).astype(f32)
    fld = po.Field(data, o, h)
    for _ in range(300):
        p = o + rng.uniform(0, 1, 3) * h * (np.array(shape) - 1) * 0.999
        p = p.astype(f32)
        got = po.interpolate(fld, p).astype(np.float64)
        pd = p.astype(np.float64)
        want = np.array([F(coef[e], *pd) for e in range(3)])
        assert np.allclose(got, want, rtol=2e-5, atol=2e-5), (p, got, want)


def test_corner_layout_axis0_slowest():
    # field value encodes the node index; at a node the sample is that node.
    shape = (4, 5, 6)
    i0, i1, i2 = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    data = np.stack([i0 * 100 + i1 * 10 + i2, i0, i2], -1).astype(f32)
    fld = po.Field(data)
    for p in [(1, 2, 3), (2, 0, 4), (0, 3, 0)]:
        v = po.interpolate(fld, np.array(p, f32))
        assert v[0] == p[0] * 100 + p[1] * 10 + p[2] and v[1] == p[0] and v[2] == p[2]


def test_contains_special_cases():
    n = 8
    fld = po.Field(np.zeros((n, n, n, 3), f32), (0, 0, 0), (0.5, 0.5, 0.5))
    last = (n - 2) * 0.5
    assert po.contains(fld, [0, 0, 0])
    assert po.contains(fld, [-0.0, 0, 0])                    # floor(-0) = -0, inside
    assert po.contains(fld, [last, last, last])              # last valid cell n-2
    assert po.contains(fld, np.nextafter(f32(last + 0.5), f32(0)) * np.ones(3, f32))
    assert not po.contains(fld, [last + 0.5, 0, 0])          # cell n-1 is outside
    assert not po.contains(fld, [0, np.nextafter(f32(0), f32(-1)), 0])
    assert not po.contains(fld, [0, 0, -1])
    assert not po.contains(fld, [np.nan, 0, 0])
    assert not po.contains(fld, [np.inf, 0, 0])


def test_weights_follow_fmod_identity():
    # w = fmod(d, h) / h with d = k h + r exactly; check against float64 math.
    rng = np.random.default_rng(2)
    h = f32(2 * math.pi / 63)
    n = 64
    i0 = np.arange(n)[:, None, None] * np.ones((1, n, n))
    data = np.stack([i0, 0 * i0, 0 * i0], -1).astype(f32)   # v0 = node index along axis 0
    fld = po.Field(data, (0, 0, 0), (h, h, h))
    for _ in range(200):
        d = f32(rng.uniform(0, 62 * float(h)))
        v = po.interpolate(fld, np.array([d, 0, 0], f32))
        k = math.floor(float(f32(d / h)))
        r = math.fmod(float(d), float(h))
        w = float(f32(f32(r) / h))
        # linear lerp between k and k+1, rounded like dpa: (1-w)k + w(k+1)
        want = float(f32(f32(f32(1 - f32(w)) * f32(k)) + f32(f32(w) * f32(k + 1))))
        assert v[0] == f32(want)


# ---------------------------------------------------------------- stepping

def test_swizzle_axis0_moves_with_stored_component2():
    fld = po.Field(np.tile(np.array([0, 0, 1], f32), (8, 8, 8, 1)))
    out = po.advect_particle(fld, [1, 1, 1], 4, 0.5)
    assert out["position"][1] == 1 and out["position"][2] == 1
    assert out["position"][0] > 1


def test_uniform_field_straight_line():
    v = np.array([0.3, -0.2, 0.7], f32)                     # stored (v0, v1, v2)
    fld = po.Field(np.tile(v, (64, 64, 64, 1)))
    p0 = np.array([30.25, 30.5, 30.75], f32)
    out = po.advect_particle(fld, p0, 40, 0.25)
    assert out["steps"] == 40 and out["status"] == po.COMPLETED
    expect = p0.astype(np.float64) + 40 * 0.25 * v[::-1].astype(np.float64)
    assert np.allclose(out["position"], expect, atol=40 * 4e-6)


def test_four_term_sum_and_sensitivity():
    # dpa's runge_kutta_4 applies one sample four times: x + a k + b k + b k + a k,
    # rounded after every term.  It equals a float32 emulation of that order and
    # differs from forward Euler x + dt k in the last bits for many inputs.
    rng = np.random.default_rng(3)
    dt = f32(0.0137)
    a, b = f32(f32(1) / f32(6)) * dt, f32(f32(1) / f32(3)) * dt
    differ = 0
    for _ in range(2000):
        x = rng.uniform(-10, 10, 3).astype(f32)
        v = rng.uniform(-3, 3, 3).astype(f32)
        got = po.step_dpa(x, v, dt)
        k = v[::-1]
        y = x.copy()
        for c in (a, b, b, a):
            y = (y + (c * k).astype(f32)).astype(f32)
        assert np.array_equal(got, y)
        differ += int(not np.array_equal(got, (x + (dt * k).astype(f32)).astype(f32)))
    assert differ > 100


def _rotation_field(omega, L=16, h=0.25):
    # k = (0, -w x2, w x1): rotation in the (x1, x2) plane; stored v = (k2, k1, k0).
    n = int(2 * L / h) + 1
    o = np.array([-L, -L, -L], np.float64)
    x0, x1, x2 = grid_coords((n, n, n), np.array([h, h, h]), o)
    data = np.stack([omega * x1, -omega * x2, 0 * x0], -1).astype(f32)
    return po.Field(data, o, (h, h, h))


def test_rotation_discriminates_dpa_and_rk4():
    w, dt, steps = 1.0, 0.05, 60
    fld = _rotation_field(w)
    p0 = np.array([0.0, 5.0, 0.0], f32)
    dpa = po.advect_particle(fld, p0, steps, dt, integrator=po.DPA_RK4)
    rk4 = po.advect_particle(fld, p0, steps, dt, integrator=po.RK4)
    r0 = 5.0
    r_dpa = float(np.hypot(dpa["position"][1], dpa["position"][2]))
    r_rk4 = float(np.hypot(rk4["position"][1], rk4["position"][2]))
    # one-sample (Euler) growth sqrt(1 + (w dt)^2) per step; RK4 drift O((w dt)^6)
    assert abs(r_dpa / r0 - (1 + (w * dt) ** 2) ** (steps / 2)) < 2e-5
    assert abs(r_rk4 / r0 - 1.0) < 2e-5
    # angle advanced: atan(w dt) per step for dpa mode
    ang = math.atan2(dpa["position"][2], dpa["position"][1])
    assert abs(ang - steps * math.atan(w * dt)) < 1e-4


# ------------------------------------------------------------ termination

def test_zero_field_gives_seed_and_terminal():
    fld = po.Field(np.zeros((8, 8, 8, 3), f32))
    out = po.advect_particle(fld, [2, 2, 2], 10, 0.1)
    assert out["steps"] == 0 and out["status"] == po.ZERO and out["remaining"] == 10
    assert np.array_equal(out["vertices"], np.array([[2, 2, 2], [-2, -2, -2]], f32))


def test_is_zero_threshold():
    assert po.is_zero([0.9e-5, -1e-5, 0])
    assert not po.is_zero([1.1e-5, 0, 0])


def test_outside_seed():
    fld = po.Field(np.ones((8, 8, 8, 3), f32))
    out = po.advect_particle(fld, [9, 2, 2], 10, 0.1)
    assert out["steps"] == 0 and out["status"] == po.EXITED
    assert np.array_equal(out["vertices"], np.array([[9, 2, 2], [-2, -2, -2]], f32))


def test_exit_step_count_brute_force():
    # uniform field along axis 0: brute-force float32 march gives the exit step.
    n, dt, u = 16, f32(0.3), f32(0.7)
    fld = po.Field(np.tile(np.array([0, 0, u], f32), (n, n, n, 1)))
    a, b = f32(f32(1) / f32(6)) * dt, f32(f32(1) / f32(3)) * dt
    for x0 in [0.0, 3.3, 13.9, 14.0]:
        x = f32(x0)
        k = 0
        while math.floor(float(x)) <= n - 2 and k < 100:
            for c in (a, b, b, a):
                x = f32(x + f32(c * u))
            k += 1
        out = po.advect_particle(fld, [x0, 1, 1], 100, dt)
        assert out["steps"] == k and out["status"] == po.EXITED
        assert out["position"][0] == x


# ------------------------------------------------------------------ rounds

def test_rounds_tail_first_and_closed_forms():
    rng = np.random.default_rng(4)
    n = 37
    fld = po.Field(np.tile(np.array([0.1, 0.2, 0.3], f32), (12, 12, 12, 1)))
    seeds = rng.uniform(0, 10, (n, 3)).astype(f32)
    rem = rng.integers(0, 30, n).astype(np.uint64)
    res = po.run_rounds(fld, seeds, rem, 10, 0.5)
    assert len(res["rounds"]) == 4                           # ceil(37 / 10)
    blocks = [(27, 37), (17, 27), (7, 17), (0, 7)]          # tail of the vector first
    for (lo, hi), (v, c, idx) in zip(blocks, res["rounds"]):
        k = res["steps"][lo:hi].astype(np.int64)
        assert v.shape[0] == int(np.sum(k + 2))
        base = np.concatenate([[0], np.cumsum(k + 2)[:-1]])
        want = np.concatenate([np.stack([b + np.arange(kk), b + np.arange(kk) + 1], 1).ravel()
                               for b, kk in zip(base, k)]).astype(np.uint32)
        assert np.array_equal(idx, want)
        assert np.array_equal(v[base], seeds[lo:hi])          # curve order = slice order
        assert np.all(v[base + k + 1] == -2)
    assert res["total_steps"] == int(res["steps"].sum())
    assert np.array_equal(res["remaining"], rem - res["steps"])


def test_empty_input_gives_one_empty_round():
    fld = po.Field(np.ones((4, 4, 4, 3), f32))
    res = po.run_rounds(fld, np.zeros((0, 3), f32), 5, 10, 0.1)
    assert len(res["rounds"]) == 1 and res["rounds"][0][0].shape == (0, 3)


def test_record_off_has_no_curves():
    fld = po.Field(np.ones((4, 4, 4, 3), f32))
    res = po.run_rounds(fld, np.ones((3, 3), f32), 5, 2, 0.1, record=False)
    assert res["rounds"] == [] and res["total_steps"] > 0


# ------------------------------------------------- prune / indices / colors

def test_prune_removes_only_invalid():
    v = np.array([[1, 2, 3], [-1, -1, -1], [-1, -1, 0], [-2, -2, -2], [-1, -1, -1]], f32)
    assert np.array_equal(po.prune(v), v[[0, 2, 3]])


def test_indices_skip_terminals():
    T = [-2, -2, -2]
    v = np.array([[0, 0, 0], [1, 0, 0], T, [5, 5, 5], T, T, [1, 1, 1], [2, 2, 2]], f32)
    assert po.indices(v).tolist() == [0, 1, 6, 7]
    assert po.indices(v, True).dtype == np.uint64
    assert po.indices(np.zeros((0, 3), f32)).size == 0


def test_colors_straight_line_and_zigzag_quirk():
    line = np.array([[0, 0, i] for i in range(6)] + [[-2, -2, -2]], f32)
    assert np.all(po.colors(line) == 0)
    zig = np.array([[0, 0, 0], [0, 0, 1], [0, 1, 1], [0, 1, 2], [0, 2, 2], [-2, -2, -2],
                    [0, 0, 0], [0, 0, 1], [0, 1, 1], [-2, -2, -2]], f32)
    c = po.colors(zig)
    # vertex 1 has no "prev" (vertex_index > 1 test): atan2(0, 0) = 0
    assert c[0] == 0 and c[1] == 0
    assert np.allclose(c[2:4], math.pi / 2, atol=1e-6)
    assert c[4] == 0 and c[5] == 0                    # last before terminal, terminal
    assert c[6] == 0 and abs(c[7] - math.pi / 2) < 1e-6 and c[8] == 0 and c[9] == 0


# ------------------------------------------------------------------ seeds

def test_lattice_seeds_closed_form():
    seeds, per = po.lattice_seeds([0.5, 1.0, 0.0], [4.0, 3.0, 2.5], [1.0, 1.5, 0.5])
    assert per == (4, 2, 5)
    i0, i1, i2 = np.unravel_index(np.arange(40), per)
    want = np.stack([0.5 + 1.0 * i0, 1.0 + 1.5 * i1, 0.0 + 0.5 * i2], 1).astype(f32)
    assert np.array_equal(seeds, want)


# -------------------------------------------------------------------- ftle

@pytest.mark.parametrize("m", [np.diag([3.0, 2.0, 1.0]), 4 * np.eye(3), np.diag([2.0, 2.0, 0.5]),
                               np.diag([1.0, 0.0, 0.0])])
def test_eigen_special_cases(m):
    assert abs(po.sym3_max_eigenvalue(m) - np.linalg.eigvalsh(m).max()) < 1e-12


def test_eigen_random_symmetric():
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = rng.normal(size=(3, 3)) * 10 ** rng.uniform(-3, 3)
        m = a @ a.T
        ref = np.linalg.eigvalsh(m).max()
        assert abs(po.sym3_max_eigenvalue(m) - ref) <= 1e-12 * max(ref, 1e-300) + 1e-300


def _linear_field(K, shape, h, o, b=(0.0, 0.0, 0.0)):
    x0, x1, x2 = grid_coords(shape, h, o)
    X = np.stack([x0, x1, x2], -1)
    k = X @ K.T + np.asarray(b)                             # dx/dt = K x + b
    return k[..., ::-1].astype(f32)                         # stored (k2, k1, k0)


def test_ftle_dpa_lattice_linear_field():
    K = np.array([[0.2, 0.05, 0.0], [0.0, -0.1, 0.03], [0.02, 0.0, 0.05]])
    shape, h = (20, 20, 20), np.array([0.5, 0.5, 0.5])
    o = np.array([-5.0, -5.0, -5.0])
    # affine field: the drift keeps the stagnation point -K^-1 b off the grid
    fld = po.Field(_linear_field(K, shape, h, o, b=(0.5, -0.4, 0.3)), o, h)
    stride = np.array([2.0, 2.0, 2.0], f32)
    size = (np.array(shape) * h).astype(f32)
    seeds, per = po.lattice_seeds(o.astype(f32), size, (h * stride).astype(f32))
    N, dt = 8, f32(0.05)
    res = po.advect_endpoints(fld, seeds, N, dt)
    ftle, coll = po.ftle_dpa(shape, o, h, stride, N, dt, seeds, res["position"])
    assert coll == 0 and ftle.shape == per
    M = np.linalg.matrix_power(np.eye(3) + float(dt) * K, N)
    want = math.log(np.linalg.svd(M, compute_uv=False).max()) / abs(float(f32(dt * f32(N))))
    inner = ftle[2:-2, 2:-2, 2:-2]
    ok = res["status"].reshape(per)[1:-1, 1:-1, 1:-1] == po.COMPLETED
    assert ok.all()
    assert np.allclose(inner, want, rtol=2e-3), (inner.min(), inner.max(), want)
    # border: one-sided difference still divided by 2h (regular_grid.hpp:100-104)
    assert ftle[0, 5, 5] < want


def test_ftle_aux_diagonal_closed_form():
    a = np.array([0.3, -0.2, 0.1])
    dt, N = 0.02, 50
    delta = np.array([0.1, 0.1, 0.1], f32)
    growth = (1 + dt * a) ** N
    pts = 5
    fin = np.zeros((pts, 6, 3), f32)
    for d in range(3):
        fin[:, 2 * d, d] = -delta[d] * growth[d]
        fin[:, 2 * d + 1, d] = delta[d] * growth[d]
    T = dt * N
    out = po.ftle_aux(fin, delta, T)
    want = math.log(np.abs(growth).max()) / T
    assert np.allclose(out, want, rtol=1e-5)


def test_dpa_structured_cpu_variant_equals_oracle():
    """The CPU baseline with dpa's per-interpolation heap list and recursive
    corner walk computes the same endpoints as the oracle, bit for bit."""
    from dpa_amd import workloads as W
    field = W.vortex_field(32, count=6)
    f = po.Field(field, (0.5, -1.0, 0.25), (0.75, 1.0, 0.5))
    lo = np.array([0.5, -1.0, 0.25], np.float32)
    seeds = W.uniform_seeds(500, lo, lo + np.array([0.75, 1.0, 0.5], np.float32) * 31, seed_offset=4)
    a = po.advect_endpoints(f, seeds, 200, np.float32(0.2))
    b = po.advect_endpoints_dpa_structured(f, seeds, 200, np.float32(0.2))
    assert a["total_steps"] == b["total_steps"] > 10000
    assert np.array_equal(a["position"].view(np.uint32), b["position"].view(np.uint32))
    assert np.array_equal(a["remaining"], b["remaining"]) and np.array_equal(a["status"], b["status"])


def _round_f32(fr):
    """A rational rounded to float32, to nearest with ties to even."""
    from fractions import Fraction
    c = np.float32(float(fr))
    best = None
    for cand in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
        err = abs(Fraction(float(cand)) - fr)
        even = (int(np.float32(cand).view(np.uint32)) & 1) == 0
        if best is None or err < best[0] or (err == best[0] and even):
            best = (err, cand)
    return np.float32(best[1])


def test_aux_seeds_offset_rounded_once():
    # the moved coordinate is origin + round(spacing * i +- delta), the offset
    # position rounded once, computed here in exact rational arithmetic; the
    # other coordinates are the lattice point's origin + round(spacing * i)
    from fractions import Fraction
    rng = np.random.default_rng(7)
    shape = (5, 4, 3)
    idx = np.indices(shape).reshape(3, -1).T
    for _ in range(30):
        sp = rng.uniform(0.01, 2.0, 3).astype(f32)
        dl = (rng.uniform(0.1, 0.9, 3) * sp).astype(f32)
        o = rng.uniform(-5.0, 5.0, 3).astype(f32)
        s = po.aux_seeds(o, sp, shape, dl).reshape(-1, 6, 3)
        for p in rng.choice(len(idx), 8, replace=False):
            for a in range(6):
                ax, sign = a // 2, (1 if a & 1 else -1)
                m = _round_f32(Fraction(float(sp[ax])) * int(idx[p][ax]) +
                               sign * Fraction(float(dl[ax])))
                assert s[p, a, ax].view(np.uint32) == f32(o[ax] + m).view(np.uint32)
                for e in range(3):
                    if e != ax:
                        want = f32(o[e] + f32(sp[e] * f32(idx[p][e])))
                        assert s[p, a, e].view(np.uint32) == want.view(np.uint32)


def test_aux_seeds_half_spacing_particles_coincide():
    # with delta = spacing / 2, point k's +delta particle is point k + 1's
    # -delta particle, bit for bit, on every axis: the library advects it once.
    # Two roundings (lattice point, then +- delta) would not give that.
    sp = np.array([0.1, 2 * np.pi / 511, 1 / 3], f32)
    dl = (f32(0.5) * sp).astype(f32)
    o = np.array([0.0, -1.25, 3.0], f32)
    shape = (40, 33, 29)
    assert po.aux_shares(sp, dl, shape)
    assert not po.aux_shares(sp, (dl * f32(0.9)).astype(f32), shape)
    s = po.aux_seeds(o, sp, shape, dl).reshape(shape + (6, 3))
    twice = 0
    for d in range(3):
        plus = np.take(s[..., 2 * d + 1, :], np.arange(shape[d] - 1), axis=d)
        minus = np.take(s[..., 2 * d, :], np.arange(1, shape[d]), axis=d)
        assert np.array_equal(plus.view(np.uint32), minus.view(np.uint32))
        k = np.arange(shape[d], dtype=np.float32)
        base = (o[d] + (sp[d] * k).astype(f32)).astype(f32)
        twice += int(np.sum((base[:-1] + dl[d]).astype(f32) != (base[1:] - dl[d]).astype(f32)))
    assert twice > 0


def test_aux_particle_steps_counts_shared_particles_once():
    # a 2 x 2 x 1 lattice: 24 particles, of which the +d particles of the
    # lower points equal the -d particles of the upper ones (4 + 4 + 0 pairs
    # shared along axes 0 and 1; axis 2 has one point)
    shape = (2, 2, 1)
    rem = np.arange(24, dtype=np.uint64)           # steps = 100 - rem
    steps = 100 - rem.astype(np.int64)
    assert po.aux_particle_steps(rem, shape, 100, False) == int(steps.sum())
    st = steps.reshape(4, 6)
    idx = np.indices(shape).reshape(3, -1).T
    want = int(st[:, 1::2].sum())
    want += sum(int(st[p, 2 * d]) for p in range(4) for d in range(3) if idx[p][d] == 0)
    assert po.aux_particle_steps(rem, shape, 100, True) == want
    assert want == int(steps.sum()) - sum(int(st[p, 2 * d]) for p in range(4) for d in range(3)
                                          if idx[p][d] > 0)
