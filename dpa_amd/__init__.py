"""Python handle on libdpa_hip.so (the C ABI in include/dpa_hip.h).

This module is plumbing for tests and the benchmark: it binds the C ABI with
ctypes and moves numpy arrays across it.  All computation happens in the HIP
kernels behind the ABI.  There is no fallback: if the shared library is missing
or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libdpa_hip.so"
HEADER = HERE.parent / "include" / "dpa_hip.h"

OK = 0
ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_ADMISSION, ERR_CANCELLED, ERR_UNSUPPORTED, ERR_TIMEOUT, ERR_PENDING = range(1, 9)
INTEGRATOR_DPA_RK4 = 0
INTEGRATOR_RK4 = 1
STATUS_ACTIVE, STATUS_COMPLETED, STATUS_EXITED, STATUS_ZERO = range(4)

_u64p = C.POINTER(C.c_uint64)
_f32p = C.POINTER(C.c_float)


class DpahError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dpah error {code}: {msg}")
        self.code = code


class _Options(C.Structure):
    _fields_ = [("memory_budget_bytes", C.c_uint64), ("slot_budget_bytes", C.c_uint64),
                ("kernel_timing", C.c_int), ("autotune_cache_path", C.c_char_p)]


class _AdvectParams(C.Structure):
    _fields_ = [("particles_per_round", C.c_uint64), ("step_size", C.c_float),
                ("start_time", C.c_float), ("integrator", C.c_int), ("record", C.c_int),
                ("use_64_bit_indices", C.c_int)]


class _Job(C.Structure):
    _fields_ = [("field", C.c_void_p), ("seeds", C.c_void_p), ("remaining", C.c_void_p),
                ("uniform_remaining", C.c_uint64), ("count", C.c_uint64),
                ("params", _AdvectParams)]


class _ServerOptions(C.Structure):
    _fields_ = [("memory_budget_bytes", C.c_uint64), ("max_active", C.c_uint32),
                ("steps_per_tick", C.c_uint32), ("mode", C.c_int),
                ("results_on_host", C.c_int)]


class _Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("field", C.c_void_p), ("seeds", C.c_void_p),
                ("remaining", C.c_void_p), ("uniform_remaining", C.c_uint64),
                ("count", C.c_uint64), ("params", _AdvectParams), ("origin", C.c_float * 3),
                ("lattice_spacing", C.c_float * 3), ("lattice_shape", C.c_uint64 * 3),
                ("delta", C.c_float * 3), ("iterations", C.c_uint64)]


class _Tuning(C.Structure):
    _fields_ = [("block", C.c_int), ("blocks_per_cu", C.c_int), ("order", C.c_int),
                ("parts", C.c_int), ("refill", C.c_int), ("cells", C.c_int),
                ("fastdiv", C.c_int), ("cell_brick", C.c_int), ("order_brick", C.c_int),
                ("lattice_brick", C.c_int), ("slot_stage", C.c_int),
                ("lattice_group", C.c_int), ("node_brick", C.c_int)]


def _tuning_dict(t) -> dict:
    return {name: getattr(t, name) for name, _ in _Tuning._fields_}


class _RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int), ("queue_seconds", C.c_double),
                ("latency_seconds", C.c_double), ("particle_steps", C.c_uint64),
                ("footprint_bytes", C.c_uint64), ("launch_seconds", C.c_double),
                ("device_seconds", C.c_double)]


class _ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("cancelled", C.c_uint64), ("failed", C.c_uint64), ("pending", C.c_uint64),
                ("ticks", C.c_uint64), ("particle_steps", C.c_uint64),
                ("bytes_in_use", C.c_uint64), ("busy_seconds", C.c_double)]


REQUEST_ADVECT, REQUEST_FTLE_AUX = 0, 1
SERVE_BATCHED, SERVE_PER_REQUEST, SERVE_SERIAL = 0, 1, 2
MAX_NOALLOC_PARTICLES = 1 << 21     # DPAH_NOALLOC_MAX_PARTICLES
TICKET_PENDING, TICKET_RUNNING, TICKET_DONE, TICKET_FAILED, TICKET_CANCELLED = range(5)

_lib = None


def load_library() -> C.CDLL:
    """Loads libdpa_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dpa_amd.build` "
                           "(or __graft_entry__.build())")
    lib = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    sig = {
        "dpah_last_error": (C.c_char_p, []),
        "dpah_version": (C.c_int, []),
        "dpah_create": (C.c_int, [C.c_int, C.POINTER(_Options), C.POINTER(vp)]),
        "dpah_destroy": (C.c_int, [vp]),
        "dpah_synchronize": (C.c_int, [vp]),
        "dpah_device_alloc": (C.c_int, [vp, C.c_uint64, C.POINTER(vp)]),
        "dpah_device_free": (C.c_int, [vp, vp]),
        "dpah_copy_to_device": (C.c_int, [vp, vp, vp, C.c_uint64]),
        "dpah_copy_to_host": (C.c_int, [vp, vp, vp, C.c_uint64]),
        "dpah_field_create": (C.c_int, [vp, vp, _u64p, _f32p, _f32p, C.POINTER(vp)]),
        "dpah_field_create_unsteady": (C.c_int, [vp, vp, _u64p, C.c_uint64, _f32p, _f32p,
                                                 C.c_float, C.c_float, C.POINTER(vp)]),
        "dpah_field_release": (C.c_int, [vp]),
        "dpah_advect": (C.c_int, [vp, vp, vp, vp, C.c_uint64, C.c_uint64,
                                  C.POINTER(_AdvectParams), C.POINTER(vp)]),
        "dpah_advect_batch": (C.c_int, [vp, C.POINTER(_Job), C.c_uint64, C.POINTER(vp)]),
        "dpah_result_round_count": (C.c_int, [vp, _u64p]),
        "dpah_result_round_sizes": (C.c_int, [vp, C.c_uint64, _u64p, _u64p, C.POINTER(C.c_int)]),
        "dpah_result_copy_round": (C.c_int, [vp, C.c_uint64, vp, vp, vp]),
        "dpah_result_final_state": (C.c_int, [vp, vp, vp, vp]),
        "dpah_result_total_steps": (C.c_int, [vp, _u64p]),
        "dpah_result_free": (C.c_int, [vp]),
        "dpah_advect_endpoints_device": (C.c_int, [vp, vp, vp, vp, C.c_uint64, C.c_uint64,
                                                   C.c_float, C.c_float, C.c_int, vp, vp, vp,
                                                   _u64p]),
        "dpah_advect_endpoints_device_noalloc": (C.c_int, [vp, vp, vp, vp, C.c_uint64,
                                                           C.c_uint64, C.c_float, C.c_float,
                                                           C.c_int, vp, vp, vp, _u64p]),
        "dpah_ftle_estimate_device": (C.c_int, [vp, _u64p, _f32p, _f32p, _f32p, C.c_uint64,
                                                C.c_float, vp, vp, C.c_uint64, vp, _u64p]),
        "dpah_ftle_lattice": (C.c_int, [vp, vp, _f32p, C.c_uint64, C.c_float, C.c_int, vp,
                                        _u64p, _u64p]),
        "dpah_ftle_aux_device": (C.c_int, [vp, vp, _f32p, _f32p, _u64p, _f32p, C.c_uint64,
                                           C.c_float, C.c_float, C.c_int, vp, _u64p]),
        "dpah_server_create": (C.c_int, [vp, C.POINTER(_ServerOptions), C.POINTER(vp)]),
        "dpah_server_submit": (C.c_int, [vp, C.POINTER(_Request), _u64p]),
        "dpah_server_poll": (C.c_int, [vp, C.c_uint64, C.POINTER(C.c_int)]),
        "dpah_server_wait": (C.c_int, [vp, C.c_uint64, C.c_int64, C.POINTER(vp)]),
        "dpah_server_request_info": (C.c_int, [vp, C.c_uint64, C.POINTER(_RequestInfo)]),
        "dpah_server_cancel": (C.c_int, [vp, C.c_uint64]),
        "dpah_server_forget": (C.c_int, [vp, C.c_uint64]),
        "dpah_server_get_stats": (C.c_int, [vp, C.POINTER(_ServerStats)]),
        "dpah_server_destroy": (C.c_int, [vp]),
        "dpah_result_ftle": (C.c_int, [vp, vp, _u64p]),
        "dpah_kernel_timing": (C.c_int, [vp, C.c_char_p, _u64p, C.POINTER(C.c_double)]),
        "dpah_kernel_timing_reset": (C.c_int, [vp]),
        "dpah_set_count_cell_loads": (C.c_int, [vp, C.c_int]),
        "dpah_cell_loads": (C.c_int, [vp, _u64p, _u64p]),
        "dpah_get_tuning": (C.c_int, [vp, C.POINTER(_Tuning)]),
        "dpah_set_tuning": (C.c_int, [vp, C.POINTER(_Tuning)]),
        "dpah_verify_division": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, _u64p]),
        "dpah_pass_create_endpoints": (C.c_int, [vp, vp, vp, vp, C.c_uint64, C.c_uint64,
                                                 C.c_float, C.c_float, C.c_int, vp, vp, vp,
                                                 C.POINTER(vp)]),
        "dpah_pass_launch": (C.c_int, [vp]),
        "dpah_pass_total_steps": (C.c_int, [vp, _u64p]),
        "dpah_pass_destroy": (C.c_int, [vp]),
        "dpah_batch_pass_create": (C.c_int, [vp, C.POINTER(_Job), C.c_uint64, C.POINTER(vp)]),
        "dpah_batch_pass_launch": (C.c_int, [vp]),
        "dpah_batch_pass_results": (C.c_int, [vp, C.POINTER(vp)]),
        "dpah_batch_pass_destroy": (C.c_int, [vp]),
        "dpah_autotune_endpoints": (C.c_int, [vp, vp, vp, C.c_uint64, C.c_uint64, C.c_float,
                                              C.c_float, C.c_int, C.POINTER(_Tuning),
                                              C.POINTER(C.c_double)]),
        "dpah_autotune_batch": (C.c_int, [vp, C.POINTER(_Job), C.c_uint64, C.POINTER(_Tuning),
                                          C.POINTER(C.c_double)]),
        "dpah_autotune_lattice": (C.c_int, [vp, vp, _f32p, _f32p, _u64p, _f32p, C.c_uint64,
                                            C.c_float, C.c_float, C.c_int, C.POINTER(_Tuning),
                                            C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols() -> list[str]:
    """Function names declared in include/dpa_hip.h."""
    import re
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dpah_\w+)\s*\(", text, re.M)))


def _check(rc: int) -> None:
    if rc != OK:
        msg = load_library().dpah_last_error().decode(errors="replace")
        raise DpahError(rc, msg)


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f3(v) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in v])


def _u3(v) -> C.Array:
    return (C.c_uint64 * 3)(*[int(x) for x in v])


class Context:
    """A dpah_context on one GPU."""

    def __init__(self, device: int = 0, memory_budget: int = 0, slot_budget: int = 0,
                 kernel_timing: bool = False, autotune_cache: str | None = None):
        self.lib = load_library()
        self._cache_path = None if autotune_cache is None else str(autotune_cache).encode()
        opt = _Options(memory_budget, slot_budget, 1 if kernel_timing else 0, self._cache_path)
        h = C.c_void_p()
        _check(self.lib.dpah_create(device, C.byref(opt), C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            _check(self.lib.dpah_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self) -> None:
        _check(self.lib.dpah_synchronize(self.h))

    # device memory
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        _check(self.lib.dpah_device_alloc(self.h, nbytes, C.byref(p)))
        return p.value

    def free(self, ptr: int) -> None:
        _check(self.lib.dpah_device_free(self.h, C.c_void_p(ptr)))

    def to_device(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.alloc(max(arr.nbytes, 1))
        _check(self.lib.dpah_copy_to_device(self.h, C.c_void_p(p), _ptr(arr), arr.nbytes))
        return p

    def memset_zero(self, ptr: int, nbytes: int) -> None:
        """Zeroes device memory (through a host copy: test and setup use only)."""
        z = np.zeros(nbytes, np.uint8)
        _check(self.lib.dpah_copy_to_device(self.h, C.c_void_p(ptr), _ptr(z), nbytes))

    def to_host(self, ptr: int, out: np.ndarray) -> np.ndarray:
        _check(self.lib.dpah_copy_to_host(self.h, _ptr(out), C.c_void_p(ptr), out.nbytes))
        return out

    def field(self, data: np.ndarray, offset=(0, 0, 0), spacing=(1, 1, 1)) -> "Field":
        data = np.ascontiguousarray(data, dtype=np.float32)
        assert data.ndim == 4 and data.shape[3] == 3, "field must be [n0][n1][n2][3]"
        h = C.c_void_p()
        _check(self.lib.dpah_field_create(self.h, _ptr(data), _u3(data.shape[:3]), _f3(offset),
                                          _f3(spacing), C.byref(h)))
        return Field(self, h, data.shape[:3], 1)

    def field_unsteady(self, data: np.ndarray, offset=(0, 0, 0), spacing=(1, 1, 1),
                       time_offset=0.0, time_spacing=1.0) -> "Field":
        data = np.ascontiguousarray(data, dtype=np.float32)
        assert data.ndim == 5 and data.shape[4] == 3, "field must be [nt][n0][n1][n2][3]"
        h = C.c_void_p()
        _check(self.lib.dpah_field_create_unsteady(
            self.h, _ptr(data), _u3(data.shape[1:4]), data.shape[0], _f3(offset), _f3(spacing),
            float(time_offset), float(time_spacing), C.byref(h)))
        return Field(self, h, data.shape[1:4], data.shape[0])

    def advect(self, field: "Field", seeds: np.ndarray, remaining=None, uniform_remaining=0,
               particles_per_round=None, step_size=0.1, start_time=0.0,
               integrator=INTEGRATOR_DPA_RK4, record=True, use_64_bit=False) -> "Result":
        seeds = np.ascontiguousarray(seeds, dtype=np.float32).reshape(-1, 3)
        n = seeds.shape[0]
        rem = None if remaining is None else np.ascontiguousarray(remaining, dtype=np.uint64)
        ppr = n if particles_per_round is None else particles_per_round
        p = _AdvectParams(int(ppr), float(step_size), float(start_time), int(integrator),
                          1 if record else 0, 1 if use_64_bit else 0)
        h = C.c_void_p()
        _check(self.lib.dpah_advect(self.h, field.h, _ptr(seeds), _ptr(rem),
                                    int(uniform_remaining), n, C.byref(p), C.byref(h)))
        return Result(self, h, n)

    @staticmethod
    def _job_array(jobs: list[dict]):
        """The dpah_job array of `jobs` and the numpy arrays it points into."""
        keep = []
        arr = (_Job * max(len(jobs), 1))()
        for i, j in enumerate(jobs):
            seeds = np.ascontiguousarray(j["seeds"], dtype=np.float32).reshape(-1, 3)
            rem = j.get("remaining")
            rem = None if rem is None else np.ascontiguousarray(rem, dtype=np.uint64)
            keep += [seeds, rem]
            n = seeds.shape[0]
            ppr = j.get("particles_per_round") or max(n, 1)
            arr[i].field = j["field"].h
            arr[i].seeds = seeds.ctypes.data if n else None
            arr[i].remaining = None if rem is None else rem.ctypes.data
            arr[i].uniform_remaining = int(j.get("uniform_remaining", 0))
            arr[i].count = n
            arr[i].params = _AdvectParams(int(ppr), float(j.get("step_size", 0.1)),
                                          float(j.get("start_time", 0.0)),
                                          int(j.get("integrator", INTEGRATOR_DPA_RK4)),
                                          1 if j.get("record", True) else 0,
                                          1 if j.get("use_64_bit", False) else 0)
        return arr, keep

    def advect_batch(self, jobs: list[dict]) -> list["Result"]:
        """jobs: dicts with field, seeds and the keyword arguments of advect()."""
        arr, _keep = self._job_array(jobs)
        hs = (C.c_void_p * max(len(jobs), 1))()
        _check(self.lib.dpah_advect_batch(self.h, arr, len(jobs), hs))
        return [Result(self, C.c_void_p(hs[i]), arr[i].count) for i in range(len(jobs))]

    def batch_pass(self, jobs: list[dict]) -> "BatchPass":
        """advect_batch's jobs captured once into a hipGraph (dpah_batch_pass_*)."""
        arr, _keep = self._job_array(jobs)
        h = C.c_void_p()
        _check(self.lib.dpah_batch_pass_create(self.h, arr, len(jobs), C.byref(h)))
        return BatchPass(self, h, [arr[i].count for i in range(len(jobs))])

    def advect_endpoints_device(self, field: "Field", d_seeds: int, count: int, d_pos_out: int,
                                d_remaining: int | None = None, uniform_remaining: int = 0,
                                step_size: float = 0.1, start_time: float = 0.0,
                                integrator: int = INTEGRATOR_DPA_RK4,
                                d_rem_out: int | None = None, d_status_out: int | None = None,
                                wait: bool = True) -> int | None:
        steps = C.c_uint64()
        _check(self.lib.dpah_advect_endpoints_device(
            self.h, field.h, C.c_void_p(d_seeds), C.c_void_p(d_remaining) if d_remaining else None,
            int(uniform_remaining), int(count), float(step_size), float(start_time), int(integrator),
            C.c_void_p(d_pos_out), C.c_void_p(d_rem_out) if d_rem_out else None,
            C.c_void_p(d_status_out) if d_status_out else None,
            C.byref(steps) if wait else None))
        return steps.value if wait else None

    def advect_endpoints_device_noalloc(self, field: "Field", d_seeds: int, count: int,
                                        d_pos_out: int, d_remaining: int | None = None,
                                        uniform_remaining: int = 0, step_size: float = 0.1,
                                        start_time: float = 0.0,
                                        integrator: int = INTEGRATOR_DPA_RK4,
                                        d_rem_out: int | None = None,
                                        d_status_out: int | None = None,
                                        wait: bool = True) -> int | None:
        """advect_endpoints_device with no device allocation: its scratch is the
        library's fixed per-GPU array (at most MAX_NOALLOC_PARTICLES particles)."""
        steps = C.c_uint64()
        _check(self.lib.dpah_advect_endpoints_device_noalloc(
            self.h, field.h, C.c_void_p(d_seeds), C.c_void_p(d_remaining) if d_remaining else None,
            int(uniform_remaining), int(count), float(step_size), float(start_time), int(integrator),
            C.c_void_p(d_pos_out), C.c_void_p(d_rem_out) if d_rem_out else None,
            C.c_void_p(d_status_out) if d_status_out else None,
            C.byref(steps) if wait else None))
        return steps.value if wait else None

    def ftle_lattice(self, field: "Field", stride, iterations: int, step_size: float,
                     integrator: int = INTEGRATOR_DPA_RK4, out: np.ndarray | None = None):
        """dpa's FTLE stage (lattice seeds, advection, estimate); host volume.

        `out`, if given, is a C-contiguous float32 array of the lattice shape
        that receives the volume (reusing one array across calls spares the
        page faults of a fresh 134-MB allocation per call)."""
        shape = (C.c_uint64 * 3)()
        _check(self.lib.dpah_ftle_lattice(self.h, field.h, _f3(stride), int(iterations),
                                          float(step_size), int(integrator), None, shape, None))
        if out is None:
            out = np.empty(tuple(shape), np.float32)
        elif (out.dtype != np.float32 or out.shape != tuple(shape)
              or not out.flags.c_contiguous or not out.flags.writeable):
            raise ValueError(f"out must be a writeable C-contiguous float32 array of shape "
                             f"{tuple(shape)}")
        steps = C.c_uint64()
        _check(self.lib.dpah_ftle_lattice(self.h, field.h, _f3(stride), int(iterations),
                                          float(step_size), int(integrator), _ptr(out), shape,
                                          C.byref(steps)))
        return out, steps.value

    def ftle_estimate_device(self, field_shape, offset, spacing, stride, iterations, step_size,
                             d_original: int, d_final: int, count: int, d_out: int | None):
        shape = (C.c_uint64 * 3)()
        _check(self.lib.dpah_ftle_estimate_device(
            self.h, _u3(field_shape), _f3(offset), _f3(spacing), _f3(stride), int(iterations),
            float(step_size), C.c_void_p(d_original) if d_original else None,
            C.c_void_p(d_final) if d_final else None, int(count),
            C.c_void_p(d_out) if d_out else None, shape))
        return tuple(int(x) for x in shape)

    def ftle_aux_device(self, field: "Field", origin, spacing, shape, delta, iterations: int,
                        step_size: float, d_out: int, start_time: float = 0.0,
                        integrator: int = INTEGRATOR_DPA_RK4) -> int:
        steps = C.c_uint64()
        _check(self.lib.dpah_ftle_aux_device(self.h, field.h, _f3(origin), _f3(spacing),
                                             _u3(shape), _f3(delta), int(iterations),
                                             float(step_size), float(start_time), int(integrator),
                                             C.c_void_p(d_out), C.byref(steps)))
        return steps.value

    def kernel_timing(self, name: str) -> tuple[int, float]:
        n = C.c_uint64()
        ms = C.c_double()
        _check(self.lib.dpah_kernel_timing(self.h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def kernel_timing_reset(self) -> None:
        _check(self.lib.dpah_kernel_timing_reset(self.h))

    def set_count_cell_loads(self, on: bool) -> None:
        """Counting of the cell cache's reloads by steady dpa-mode endpoint
        launches (roofline instrumentation); turning it on zeroes the counts."""
        _check(self.lib.dpah_set_count_cell_loads(self.h, 1 if on else 0))

    def cell_loads(self) -> tuple[int, int]:
        """(cell-record loads, node-group loads) since counting was turned on."""
        a, b = C.c_uint64(), C.c_uint64()
        _check(self.lib.dpah_cell_loads(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def verify_division(self, h: float, r: float, bound: float) -> int:
        """Floats in [-bound, bound] where the reciprocal division differs from x / h."""
        n = C.c_uint64()
        _check(self.lib.dpah_verify_division(self.h, float(h), float(r), float(bound),
                                             C.byref(n)))
        return n.value

    def pass_endpoints(self, field: "Field", d_seeds: int, count: int, d_pos_out: int,
                       d_remaining: int | None = None, uniform_remaining: int = 0,
                       step_size: float = 0.1, start_time: float = 0.0,
                       integrator: int = INTEGRATOR_DPA_RK4, d_rem_out: int | None = None,
                       d_status_out: int | None = None) -> "Pass":
        """An endpoint pass captured into a hipGraph (see Pass)."""
        h = C.c_void_p()
        _check(self.lib.dpah_pass_create_endpoints(
            self.h, field.h, C.c_void_p(d_seeds), C.c_void_p(d_remaining) if d_remaining else None,
            int(uniform_remaining), int(count), float(step_size), float(start_time),
            int(integrator), C.c_void_p(d_pos_out), C.c_void_p(d_rem_out) if d_rem_out else None,
            C.c_void_p(d_status_out) if d_status_out else None, C.byref(h)))
        return Pass(self, h)

    def autotune_endpoints(self, field: "Field", d_seeds: int, count: int, uniform_remaining: int,
                           step_size: float, start_time: float = 0.0,
                           integrator: int = INTEGRATOR_DPA_RK4) -> tuple[dict, float]:
        """Times candidate tunings on a device-resident seed sample; the fastest
        applies to later launches of the same kind.  Returns (tuning, ms)."""
        t = _Tuning()
        ms = C.c_double()
        _check(self.lib.dpah_autotune_endpoints(self.h, field.h, C.c_void_p(d_seeds), int(count),
                                                int(uniform_remaining), float(step_size),
                                                float(start_time), int(integrator), C.byref(t),
                                                C.byref(ms)))
        return _tuning_dict(t), ms.value

    def autotune_batch(self, jobs: list[dict]) -> tuple[dict, float]:
        """Times candidate launch shapes of each group of advect_batch's jobs on
        its own persistent launch; returns the tuning of the largest group and
        its milliseconds."""
        arr, _keep = self._job_array(jobs)
        t = _Tuning()
        ms = C.c_double()
        _check(self.lib.dpah_autotune_batch(self.h, arr, len(jobs), C.byref(t), C.byref(ms)))
        return _tuning_dict(t), ms.value

    def autotune_lattice(self, field: "Field", origin, spacing, shape, iterations: int,
                         step_size: float, delta=None, start_time: float = 0.0,
                         integrator: int = INTEGRATOR_DPA_RK4) -> tuple[dict, float]:
        """Times candidate launch shapes of a whole-lattice launch (auxiliary
        FTLE particles when delta is given, else the lattice points)."""
        t = _Tuning()
        ms = C.c_double()
        _check(self.lib.dpah_autotune_lattice(self.h, field.h, _f3(origin), _f3(spacing),
                                              _u3(shape), _f3(delta) if delta is not None else None,
                                              int(iterations), float(step_size), float(start_time),
                                              int(integrator), C.byref(t), C.byref(ms)))
        return _tuning_dict(t), ms.value

    def tuning(self) -> dict:
        t = _Tuning()
        _check(self.lib.dpah_get_tuning(self.h, C.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, **kw) -> None:
        cur = self.tuning()
        cur.update(kw)
        t = _Tuning(*[cur[name] for name, _ in _Tuning._fields_])
        _check(self.lib.dpah_set_tuning(self.h, C.byref(t)))


class Pass:
    """A captured endpoint pass: launch() replays it on the context's stream."""

    def __init__(self, ctx: "Context", h):
        self.ctx, self.h = ctx, h

    def launch(self) -> None:
        _check(self.ctx.lib.dpah_pass_launch(self.h))

    def total_steps(self) -> int:
        s = C.c_uint64()
        _check(self.ctx.lib.dpah_pass_total_steps(self.h, C.byref(s)))
        return s.value

    def destroy(self) -> None:
        if self.h:
            _check(self.ctx.lib.dpah_pass_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


class BatchPass:
    """A captured batch: launch() replays it; results() returns one Result per job."""

    def __init__(self, ctx: "Context", h, counts: list[int]):
        self.ctx, self.h, self.counts = ctx, h, counts

    def launch(self) -> None:
        _check(self.ctx.lib.dpah_batch_pass_launch(self.h))

    def results(self) -> list["Result"]:
        k = len(self.counts)
        hs = (C.c_void_p * max(k, 1))()
        _check(self.ctx.lib.dpah_batch_pass_results(self.h, hs))
        return [Result(self.ctx, C.c_void_p(hs[i]), self.counts[i]) for i in range(k)]

    def destroy(self) -> None:
        if self.h:
            _check(self.ctx.lib.dpah_batch_pass_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


class Field:
    def __init__(self, ctx: Context, h, shape, nt):
        self.ctx, self.h, self.shape, self.nt = ctx, h, tuple(int(s) for s in shape), int(nt)

    def release(self) -> None:
        if self.h:
            _check(self.ctx.lib.dpah_field_release(self.h))
            self.h = None


class Result:
    def __init__(self, ctx: Context, h, n: int):
        self.ctx, self.h, self.n = ctx, h, n

    def free(self) -> None:
        if self.h:
            _check(self.ctx.lib.dpah_result_free(self.h))
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def round_count(self) -> int:
        v = C.c_uint64()
        _check(self.ctx.lib.dpah_result_round_count(self.h, C.byref(v)))
        return v.value

    def round(self, r: int):
        """(vertices [V,3] f32, colors [V] f32, indices [I] u32|u64) of round r."""
        V, I, ib = C.c_uint64(), C.c_uint64(), C.c_int()
        _check(self.ctx.lib.dpah_result_round_sizes(self.h, r, C.byref(V), C.byref(I), C.byref(ib)))
        verts = np.empty((V.value, 3), np.float32)
        cols = np.empty(V.value, np.float32)
        idx = np.empty(I.value, np.uint64 if ib.value == 8 else np.uint32)
        _check(self.ctx.lib.dpah_result_copy_round(self.h, r, _ptr(verts), _ptr(cols), _ptr(idx)))
        return verts, cols, idx

    def final_state(self):
        pos = np.empty((self.n, 3), np.float32)
        rem = np.empty(self.n, np.uint64)
        st = np.empty(self.n, np.uint8)
        _check(self.ctx.lib.dpah_result_final_state(self.h, _ptr(pos), _ptr(rem), _ptr(st)))
        return pos, rem, st

    @property
    def total_steps(self) -> int:
        v = C.c_uint64()
        _check(self.ctx.lib.dpah_result_total_steps(self.h, C.byref(v)))
        return v.value


class Server:
    """A dpah_server: continuous batching (mode SERVE_BATCHED) or one of the
    comparators that run requests one at a time (SERVE_PER_REQUEST: allocate,
    launch and free per request, one thread per particle; SERVE_SERIAL: the
    library's single-request path).  results_on_host: every mode copies each
    result to host memory before its ticket is ready."""

    def __init__(self, ctx: Context, mode: int = SERVE_BATCHED, memory_budget: int = 0,
                 max_active: int = 0, steps_per_tick: int = 0, results_on_host: bool = False):
        self.ctx, self.lib = ctx, ctx.lib
        o = _ServerOptions(memory_budget, max_active, steps_per_tick, mode,
                           1 if results_on_host else 0)
        h = C.c_void_p()
        _check(self.lib.dpah_server_create(ctx.h, C.byref(o), C.byref(h)))
        self.h = h
        self._keep = {}

    def close(self) -> None:
        if self.h:
            _check(self.lib.dpah_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit_advect(self, field: Field, seeds, remaining=None, uniform_remaining=0,
                      particles_per_round=None, step_size=0.1, start_time=0.0,
                      integrator=INTEGRATOR_DPA_RK4, record=True, use_64_bit=False) -> int:
        seeds = np.ascontiguousarray(seeds, dtype=np.float32).reshape(-1, 3)
        rem = None if remaining is None else np.ascontiguousarray(remaining, dtype=np.uint64)
        n = seeds.shape[0]
        r = _Request()
        r.kind = REQUEST_ADVECT
        r.field = field.h
        r.seeds = seeds.ctypes.data if n else None
        r.remaining = None if rem is None else rem.ctypes.data
        r.uniform_remaining = int(uniform_remaining)
        r.count = n
        r.params = _AdvectParams(int(particles_per_round or max(n, 1)), float(step_size),
                                 float(start_time), int(integrator), 1 if record else 0,
                                 1 if use_64_bit else 0)
        t = C.c_uint64()
        _check(self.lib.dpah_server_submit(self.h, C.byref(r), C.byref(t)))
        return t.value

    def submit_ftle(self, field: Field, origin, spacing, shape, delta, iterations, step_size,
                    start_time=0.0, integrator=INTEGRATOR_DPA_RK4) -> int:
        r = _Request()
        r.kind = REQUEST_FTLE_AUX
        r.field = field.h
        r.params = _AdvectParams(1, float(step_size), float(start_time), int(integrator), 0, 0)
        r.origin[:] = [float(x) for x in np.asarray(origin, np.float32)]
        r.lattice_spacing[:] = [float(x) for x in np.asarray(spacing, np.float32)]
        r.lattice_shape[:] = [int(x) for x in shape]
        r.delta[:] = [float(x) for x in np.asarray(delta, np.float32)]
        r.iterations = int(iterations)
        t = C.c_uint64()
        _check(self.lib.dpah_server_submit(self.h, C.byref(r), C.byref(t)))
        return t.value

    def poll(self, ticket: int) -> int:
        s = C.c_int()
        _check(self.lib.dpah_server_poll(self.h, ticket, C.byref(s)))
        return s.value

    def wait(self, ticket: int, timeout_ms: int = -1, n: int = 0) -> Result:
        h = C.c_void_p()
        _check(self.lib.dpah_server_wait(self.h, ticket, timeout_ms, C.byref(h)))
        return Result(self.ctx, h, n)

    def info(self, ticket: int) -> dict:
        i = _RequestInfo()
        _check(self.lib.dpah_server_request_info(self.h, ticket, C.byref(i)))
        return dict(state=i.state, queue_seconds=i.queue_seconds,
                    latency_seconds=i.latency_seconds, particle_steps=i.particle_steps,
                    footprint_bytes=i.footprint_bytes, launch_seconds=i.launch_seconds,
                    device_seconds=i.device_seconds)

    def cancel(self, ticket: int) -> None:
        _check(self.lib.dpah_server_cancel(self.h, ticket))

    def forget(self, ticket: int) -> None:
        _check(self.lib.dpah_server_forget(self.h, ticket))

    def stats(self) -> dict:
        s = _ServerStats()
        _check(self.lib.dpah_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _ServerStats._fields_}


def aux_particles(spacing, delta, shape) -> int:
    """Particles the library advects for an auxiliary-neighbour FTLE lattice of
    `shape` points: with delta half the spacing on every axis (float32, every
    axis below 2^24 points) a particle between two neighbouring points is one
    particle -- three per point plus the -delta particles of the low faces --
    else six per point."""
    sp = np.asarray(spacing, np.float32).reshape(3)
    dl = np.asarray(delta, np.float32).reshape(3)
    d0, d1, d2 = (int(x) for x in shape)
    if d0 * d1 * d2 == 0:
        return 0
    if all(dl[i] > 0 and np.float32(2.0) * dl[i] == sp[i] and int(shape[i]) < (1 << 24)
           for i in range(3)):
        return 3 * d0 * d1 * d2 + d1 * d2 + d0 * d2 + d0 * d1
    return 6 * d0 * d1 * d2


def result_ftle(res: Result) -> np.ndarray:
    n = C.c_uint64()
    _check(res.ctx.lib.dpah_result_ftle(res.h, None, C.byref(n)))
    out = np.empty(n.value, np.float32)
    _check(res.ctx.lib.dpah_result_ftle(res.h, _ptr(out), C.byref(n)))
    return out
