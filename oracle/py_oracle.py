"""ctypes binding of the CPU oracle (oracle/libdpa_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libdpa_oracle.so"

DPA_RK4, RK4 = 0, 1
ACTIVE, COMPLETED, EXITED, ZERO = range(4)


class Grid(C.Structure):
    _fields_ = [("data", C.c_void_p), ("n", C.c_int64 * 3), ("nt", C.c_int64),
                ("offset", C.c_float * 3), ("spacing", C.c_float * 3),
                ("t_offset", C.c_float), ("t_spacing", C.c_float)]


_lib = None
_vp = C.c_void_p
_u64 = C.c_uint64


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    _lib = bind(LIB_PATH)
    return _lib


@contextlib.contextmanager
def using(path):
    """Routes every function of this module through another build of the
    oracle (a planted-mistake variant of a pin test) while the block runs."""
    global _lib
    saved = load()
    _lib = bind(path)
    try:
        yield
    finally:
        _lib = saved


def bind(path) -> C.CDLL:
    lib = C.CDLL(str(path))
    G = C.POINTER(Grid)
    sig = {
        "dpo_contains": (C.c_int, [G, _vp, C.c_float]),
        "dpo_interpolate": (None, [G, _vp, C.c_float, _vp]),
        "dpo_is_zero": (C.c_int, [_vp]),
        "dpo_step_dpa": (None, [_vp, _vp, C.c_float]),
        "dpo_advect_particle": (_u64, [G, _vp, _vp, C.c_float, C.c_float, C.c_int, _vp, _vp]),
        "dpo_run_rounds": (_vp, [G, _vp, _vp, _u64, _u64, C.c_float, C.c_float, C.c_int, C.c_int,
                                 C.c_int, C.c_int]),
        "dpo_rounds_count": (_u64, [_vp]),
        "dpo_rounds_vertex_count": (_u64, [_vp, _u64]),
        "dpo_rounds_index_count": (_u64, [_vp, _u64]),
        "dpo_rounds_vertices": (_vp, [_vp, _u64]),
        "dpo_rounds_colors": (_vp, [_vp, _u64]),
        "dpo_rounds_indices": (_vp, [_vp, _u64]),
        "dpo_rounds_final_positions": (_vp, [_vp]),
        "dpo_rounds_final_remaining": (_vp, [_vp]),
        "dpo_rounds_final_status": (_vp, [_vp]),
        "dpo_rounds_steps": (_vp, [_vp]),
        "dpo_rounds_total_steps": (_u64, [_vp]),
        "dpo_rounds_free": (None, [_vp]),
        "dpo_advect_endpoints": (_u64, [G, _vp, _vp, _vp, _u64, C.c_float, C.c_float, C.c_int,
                                        C.c_int]),
        "dpo_advect_endpoints_dpa_structured": (_u64, [G, _vp, _vp, _vp, _u64, C.c_float,
                                                        C.c_float, C.c_int]),
        "dpo_prune": (_u64, [_vp, _u64]),
        "dpo_indices": (_u64, [_vp, _u64, _vp, C.c_int]),
        "dpo_colors": (None, [_vp, _u64, _vp]),
        "dpo_lattice_seeds": (_u64, [_vp, _vp, _vp, _vp, _vp]),
        "dpo_ftle_dpa": (None, [_vp, _vp, _vp, _vp, _u64, C.c_float, _vp, _vp, _u64, _vp, _vp,
                                _vp]),
        "dpo_ftle_aux": (None, [_vp, _u64, _vp, C.c_double, _vp]),
        "dpo_aux_seeds": (None, [_vp, _vp, _vp, _vp, _vp]),
        "dpo_sym3_max_eigenvalue": (C.c_double, [_vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _p(a):
    return a.ctypes.data_as(_vp)


def _f32(v, n=3):
    return np.ascontiguousarray(np.asarray(v, np.float32).reshape(n))


class Field:
    """Keeps the numpy data alive next to the dpo_grid that points at it."""

    def __init__(self, data, offset=(0, 0, 0), spacing=(1, 1, 1), time_offset=0.0,
                 time_spacing=1.0):
        self.data = np.ascontiguousarray(data, np.float32)
        g = Grid()
        g.data = self.data.ctypes.data
        if self.data.ndim == 4:
            g.n[:] = self.data.shape[:3]
            g.nt = 1
        else:
            g.n[:] = self.data.shape[1:4]
            g.nt = self.data.shape[0]
        g.offset[:] = [float(x) for x in np.asarray(offset, np.float32)]
        g.spacing[:] = [float(x) for x in np.asarray(spacing, np.float32)]
        g.t_offset = float(np.float32(time_offset))
        g.t_spacing = float(np.float32(time_spacing))
        self.grid = g
        self.ref = C.byref(g)


def contains(f: Field, p, t=0.0) -> bool:
    return bool(load().dpo_contains(f.ref, _p(_f32(p)), float(np.float32(t))))


def interpolate(f: Field, p, t=0.0) -> np.ndarray:
    out = np.empty(3, np.float32)
    load().dpo_interpolate(f.ref, _p(_f32(p)), float(np.float32(t)), _p(out))
    return out


def is_zero(v) -> bool:
    return bool(load().dpo_is_zero(_p(_f32(v))))


def step_dpa(x, v, dt) -> np.ndarray:
    x = _f32(x).copy()
    load().dpo_step_dpa(_p(x), _p(_f32(v)), float(np.float32(dt)))
    return x


def advect_particle(f: Field, p, remaining, dt, t0=0.0, integrator=DPA_RK4, record=True):
    p = _f32(p).copy()
    rem = np.array([remaining], np.uint64)
    st = np.zeros(1, np.uint8)
    rec = np.empty((int(remaining) + 2, 3), np.float32) if record else None
    steps = load().dpo_advect_particle(f.ref, _p(p), _p(rem), float(np.float32(dt)),
                                       float(np.float32(t0)), integrator,
                                       _p(rec) if rec is not None else None, _p(st))
    return dict(position=p, remaining=int(rem[0]), status=int(st[0]), steps=int(steps),
                vertices=None if rec is None else rec[:steps + 2])


def run_rounds(f: Field, seeds, remaining, particles_per_round, dt, t0=0.0, integrator=DPA_RK4,
               record=True, use_64_bit=False, threads=0):
    """Returns dict(rounds=[(vertices, colors, indices)], position, remaining, status,
    steps, total_steps)."""
    lib = load()
    seeds = np.ascontiguousarray(seeds, np.float32).reshape(-1, 3)
    n = seeds.shape[0]
    rem = (np.full(n, remaining, np.uint64) if np.isscalar(remaining)
           else np.ascontiguousarray(remaining, np.uint64))
    h = lib.dpo_run_rounds(f.ref, _p(seeds), _p(rem), n, int(particles_per_round),
                           float(np.float32(dt)), float(np.float32(t0)), integrator,
                           1 if record else 0, 1 if use_64_bit else 0, threads)
    if not h:
        raise ValueError("oracle rejected the configuration")
    try:
        rounds = []
        for r in range(lib.dpo_rounds_count(h)):
            V = lib.dpo_rounds_vertex_count(h, r)
            I = lib.dpo_rounds_index_count(h, r)
            v = np.ctypeslib.as_array(C.cast(lib.dpo_rounds_vertices(h, r), C.POINTER(C.c_float)),
                                      shape=(max(V, 1) * 3,))[:3 * V].reshape(V, 3).copy()
            c = np.ctypeslib.as_array(C.cast(lib.dpo_rounds_colors(h, r), C.POINTER(C.c_float)),
                                      shape=(max(V, 1),))[:V].copy()
            it = C.c_uint64 if use_64_bit else C.c_uint32
            i = np.ctypeslib.as_array(C.cast(lib.dpo_rounds_indices(h, r), C.POINTER(it)),
                                      shape=(max(I, 1),))[:I].copy()
            rounds.append((v, c, i))

        def arr(ptr, ct, shape):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=shape).copy()

        m = max(n, 1)
        out = dict(rounds=rounds,
                   position=arr(lib.dpo_rounds_final_positions(h), C.c_float, (m, 3))[:n],
                   remaining=arr(lib.dpo_rounds_final_remaining(h), C.c_uint64, (m,))[:n],
                   status=arr(lib.dpo_rounds_final_status(h), C.c_uint8, (m,))[:n],
                   steps=arr(lib.dpo_rounds_steps(h), C.c_uint64, (m,))[:n],
                   total_steps=int(lib.dpo_rounds_total_steps(h)))
    finally:
        lib.dpo_rounds_free(h)
    return out


def advect_endpoints(f: Field, seeds, remaining, dt, t0=0.0, integrator=DPA_RK4, threads=0):
    seeds = np.ascontiguousarray(seeds, np.float32).reshape(-1, 3)
    n = seeds.shape[0]
    pos = seeds.copy()
    rem = (np.full(n, remaining, np.uint64) if np.isscalar(remaining)
           else np.ascontiguousarray(remaining, np.uint64).copy())
    st = np.zeros(n, np.uint8)
    total = load().dpo_advect_endpoints(f.ref, _p(pos), _p(rem), _p(st), n,
                                        float(np.float32(dt)), float(np.float32(t0)), integrator,
                                        threads)
    return dict(position=pos, remaining=rem, status=st, total_steps=int(total))


def advect_endpoints_dpa_structured(f: Field, seeds, remaining, dt, t0=0.0, threads=0):
    """dpa-mode endpoints through the dpa-structured interpolation (heap corner
    list, recursive corner walk): the CPU baseline with dpa's per-step costs."""
    seeds = np.ascontiguousarray(seeds, np.float32).reshape(-1, 3)
    n = seeds.shape[0]
    pos = seeds.copy()
    rem = (np.full(n, remaining, np.uint64) if np.isscalar(remaining)
           else np.ascontiguousarray(remaining, np.uint64).copy())
    st = np.zeros(n, np.uint8)
    total = load().dpo_advect_endpoints_dpa_structured(f.ref, _p(pos), _p(rem), _p(st), n,
                                                       float(np.float32(dt)),
                                                       float(np.float32(t0)), threads)
    return dict(position=pos, remaining=rem, status=st, total_steps=int(total))


def prune(vertices) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3).copy()
    n = load().dpo_prune(_p(v), v.shape[0])
    return v[:n]


def indices(vertices, use_64_bit=False) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    out = np.empty(max(2 * v.shape[0], 1), np.uint64 if use_64_bit else np.uint32)
    n = load().dpo_indices(_p(v), v.shape[0], _p(out), 1 if use_64_bit else 0)
    return out[:n]


def colors(vertices) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    out = np.empty(max(v.shape[0], 1), np.float32)
    load().dpo_colors(_p(v), v.shape[0], _p(out))
    return out[:v.shape[0]]


def lattice_seeds(offset, size, stride):
    per = np.zeros(3, np.uint64)
    o, s, st = _f32(offset), _f32(size), _f32(stride)
    n = load().dpo_lattice_seeds(_p(o), _p(s), _p(st), _p(per), None)
    out = np.empty((max(n, 1), 3), np.float32)
    load().dpo_lattice_seeds(_p(o), _p(s), _p(st), _p(per), _p(out))
    return out[:n], tuple(int(x) for x in per)


def ftle_dpa(field_shape, offset, spacing, stride, iterations, dt, original, final):
    lib = load()
    fn = np.ascontiguousarray(field_shape, np.int64)
    shape = np.zeros(3, np.uint64)
    o, h, st = _f32(offset), _f32(spacing), _f32(stride)
    orig = np.ascontiguousarray(original, np.float32).reshape(-1, 3)
    fin = np.ascontiguousarray(final, np.float32).reshape(-1, 3)
    lib.dpo_ftle_dpa(_p(fn), _p(o), _p(h), _p(st), int(iterations), float(np.float32(dt)),
                     _p(orig), _p(fin), orig.shape[0], None, _p(shape), None)
    out = np.empty(tuple(int(x) for x in shape), np.float32)
    coll = np.zeros(1, np.uint64)
    lib.dpo_ftle_dpa(_p(fn), _p(o), _p(h), _p(st), int(iterations), float(np.float32(dt)),
                     _p(orig), _p(fin), orig.shape[0], _p(out), _p(shape), _p(coll))
    return out, int(coll[0])


def ftle_aux(finals, delta, T) -> np.ndarray:
    f = np.ascontiguousarray(finals, np.float32).reshape(-1, 6, 3)
    out = np.empty(f.shape[0], np.float32)
    load().dpo_ftle_aux(_p(f), f.shape[0], _p(_f32(delta)), float(T), _p(out))
    return out


def aux_shares(spacing, delta, shape) -> bool:
    """Whether an auxiliary lattice's offsets are half its spacing on every
    axis (exactly, in float32; every axis below 2^24 points), so that each
    particle between two neighbouring points is one particle."""
    sp, dl = _f32(spacing), _f32(delta)
    return all(dl[i] > 0 and np.float32(2.0) * dl[i] == sp[i] and int(shape[i]) < (1 << 24)
               for i in range(3))


def aux_particle_steps(remaining, shape, iterations, shared) -> int:
    """Particle-steps of an auxiliary lattice's advection from the remaining
    counts of its six-per-point seeds (aux_seeds order): all six per point, or,
    when shared, each particle between two points once -- the +delta particles
    of every point and the -delta particles of the points on the low faces."""
    shape = tuple(int(x) for x in shape)
    steps = (int(iterations) - np.asarray(remaining, np.int64)).reshape(-1, 6)
    if not shared:
        return int(steps.sum())
    idx = np.indices(shape).reshape(3, -1)
    total = int(steps[:, 1::2].sum())
    for d in range(3):
        total += int(steps[idx[d] == 0, 2 * d].sum())
    return total


def sym3_max_eigenvalue(m) -> float:
    a = np.ascontiguousarray(m, np.float64).reshape(9)
    return float(load().dpo_sym3_max_eigenvalue(_p(a)))


def aux_seeds(origin, spacing, shape, delta) -> np.ndarray:
    """Seeds of the auxiliary-neighbour FTLE (dpo_aux_seeds): for each lattice
    point i (C order) the six points with coordinate d moved to
    origin_d + fl(spacing_d * i_d -/+ delta_d), axes 0, 1, 2 in that order;
    shape [points * 6, 3]."""
    shp = np.ascontiguousarray([int(x) for x in shape], np.uint64)
    out = np.empty((int(np.prod(shp)) * 6, 3), np.float32)
    load().dpo_aux_seeds(_p(_f32(origin)), _p(_f32(spacing)), _p(shp), _p(_f32(delta)), _p(out))
    return out
