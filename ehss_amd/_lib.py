"""ctypes binding of libehss_hip.so (the C-ABI declared in include/ehss_hip.h).

The library is loaded from the package directory.  There is no fallback: if the
shared object is missing or no HIP device is present, every call raises.
"""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libehss_hip.so"

EHSS_OK = 0
EHSS_ERR_INVALID = 1
EHSS_ERR_HIP = 2
EHSS_ERR_NOMEM = 3
EHSS_ERR_ADMISSION = 4
EHSS_ERR_BUSY = 5

PRECISION_FP64 = 0
PRECISION_FP32_FIXUP = 1
PRECISIONS = {"fp64": PRECISION_FP64, "fp32-fixup": PRECISION_FP32_FIXUP}

FLAG_PA = 1
FLAG_KEEP_RAYS = 2
FLAG_GRAPH = 4
FLAG_TIMESTAMPS = 8
FLAG_LDS_PACKETS = 16
FLAG_PER_REPLICATE = 32
FLAG_CONTINUABLE = 64
FLAG_UNSORTED = 128
VARIANT_SHIFT = 8
COOP_SHIFT = 12
GROUP_SHIFT = 19
STAGE1_SHIFT = 22
PACKETS_SHIFT = 26

RAY_ESCAPED = 0x1
RAY_UNRESOLVED = 0x2
RAY_PA_HIT = 0x4
RAY_EHSS_HIT = 0x8
RAY_GRAZING = 0x10

_P = ct.c_void_p
_I64 = ct.c_int64
_INT = ct.c_int


class EHSSError(RuntimeError):
    """A libehss_hip call returned a nonzero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class AdmissionError(EHSSError):
    """The request does not fit the configured device memory budget."""


class Sobol(ct.Structure):
    _fields_ = [("sv", (ct.c_uint32 * 32) * 4), ("shift", ct.c_uint32 * 4)]


class Sobol5(ct.Structure):
    _fields_ = [("sv", (ct.c_uint32 * 32) * 5), ("shift", ct.c_uint32 * 5)]


SOURCE_BOX, SOURCE_ELLIPSOID, SOURCE_REVERSE_PA = 1, 2, 3


class Source(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("pad", ct.c_int32), ("center", ct.c_double * 3),
                ("rotation", ct.c_double * 9), ("axes", ct.c_double * 3), ("area", ct.c_double),
                ("cdf", ct.c_void_p)]


class SceneInfo(ct.Structure):
    _fields_ = [("atoms", ct.c_int64), ("groups", ct.c_int64), ("nodes", ct.c_int64), ("pa_nodes", ct.c_int64),
                ("leaf_size", ct.c_int32), ("max_depth", ct.c_int32), ("center", ct.c_double * 3),
                ("radius", ct.c_double), ("scale", ct.c_double), ("device_bytes", ct.c_int64),
                ("build_path", ct.c_int32), ("packet_nodes", ct.c_int32)]


BUILD_GENERAL, BUILD_SHAPED = 0, 1


class Totals(ct.Structure):
    _fields_ = [("omega", ct.c_double), ("omega_err", ct.c_double), ("pa", ct.c_double),
                ("pa_from_ehss", ct.c_double), ("residual_mass", ct.c_double), ("tail_bound", ct.c_double),
                ("weight", ct.c_double), ("rays", ct.c_int64), ("escaped", ct.c_int64),
                ("unresolved", ct.c_int64), ("pa_hits", ct.c_int64), ("ehss_hits", ct.c_int64),
                ("box_tests", ct.c_int64), ("sphere_tests", ct.c_int64), ("pa_box_tests", ct.c_int64),
                ("pa_sphere_tests", ct.c_int64), ("fp64_retests", ct.c_int64), ("grazing_rays", ct.c_int64),
                ("max_bounces", ct.c_int64), ("kernel_ms", ct.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# NumPy view of an array of Totals (every field is 8 bytes: double or int64).
TOTALS_DTYPE = np.dtype([(name, "<f8" if typ is ct.c_double else "<i8") for name, typ in Totals._fields_])
assert TOTALS_DTYPE.itemsize == ct.sizeof(Totals)


def totals_list(arr, count: int) -> list[dict]:
    """Dicts of the first `count` entries of a ctypes Totals array."""
    view = np.frombuffer(arr, dtype=TOTALS_DTYPE, count=count)
    names = TOTALS_DTYPE.names
    return [dict(zip(names, row)) for row in view.tolist()]


# Signatures: name -> (restype, argtypes)
_SIGS = {
    "ehss_abi_version": (_INT, []),
    "ehss_last_error": (_INT, [ct.c_char_p, ct.c_size_t]),
    "ehss_init": (_INT, [_INT]),
    "ehss_device_info": (_INT, [ct.c_char_p, ct.c_size_t, _P, _P, _P, _P]),
    "ehss_selftest_fp64": (_INT, [_I64, _P, _P, _P, _P]),
    "ehss_trace": (None, [_I64, _I64, _INT, _INT, _INT] + [_P] * 17),
    "ehss_trace_status": (_INT, [_I64, _I64, _INT, _INT, _INT] + [_P] * 16),
    "ehss_scene_create": (_INT, [_I64, _P, _P, _INT, ct.POINTER(_P)]),
    "ehss_scene_create_grouped": (_INT, [_I64, _P, _P, _P, _INT, ct.POINTER(_P)]),
    "ehss_scene_destroy": (_INT, [_P]),
    "ehss_scramble": (_INT, [ct.c_uint64, ct.POINTER(Sobol)]),
    "ehss_scene_rebuild": (_INT, [_P, ct.c_int64, _P, _P, _INT]),
    "ehss_scene_get_info": (_INT, [_P, ct.POINTER(SceneInfo)]),
    "ehss_tree_counts": (_INT, [_I64, _INT, _P]),
    "ehss_scene_export_tree": (_INT, [_P] + [_P] * 7),
    "ehss_scene_export_pa_tree": (_INT, [_P] + [_P] * 6),
    "ehss_scene_trace": (_INT, [_P, _I64, _INT, _INT, _INT] + [_P] * 8),
    "ehss_scene_pa_query": (_INT, [_P, _I64, _INT] + [_P] * 5),
    "ehss_solve": (_INT, [_P, ct.POINTER(Sobol), _INT, _INT, _INT, ct.c_uint, ct.POINTER(Totals), _P, _P]),
    "ehss_solve_async": (_INT, [_P, ct.POINTER(Sobol), _INT, _INT, _INT, ct.c_uint]),
    "ehss_solve_many": (_INT, [_P, _INT, _P, _INT, _INT, _INT, ct.c_uint, _P, _P, _P]),
    "ehss_solve_many_async": (_INT, [_P, _INT, _P, _INT, _INT, _INT, ct.c_uint]),
    "ehss_solve_query": (_INT, [_P, ct.POINTER(_INT)]),
    "ehss_solve_wait": (_INT, [_P, ct.POINTER(Totals), _P, _P]),
    "ehss_scene_last_rays": (_INT, [_P, _I64, _P, _P, _P]),
    "ehss_scene_last_ray_times": (_INT, [_P, _I64, _P, _P]),
    "ehss_ray_map": (_INT, [ct.POINTER(Sobol), _INT, _P]),
    "ehss_cross_map": (_INT, [_P, _INT, _INT, _INT, _P]),
    "ehss_source_rays": (_INT, [_P, ct.POINTER(Sobol), _INT, _P, _P, _P]),
    "ehss_solve_source": (_INT, [_P, ct.POINTER(Source), ct.POINTER(Sobol5), _INT, _INT, _P, _P, _P, _P]),
    "ehss_source_states": (_INT, [_P, ct.POINTER(Source), ct.POINTER(Sobol5), _INT] + [_P] * 7),
    "ehss_trig_compat_info": (_INT, [_P, _P, ct.c_size_t, _P, ct.c_size_t]),
    "ehss_scramble_many": (_INT, [_I64, _P, _P]),
    "ehss_solve_extend": (_INT, [_P, _INT, _P, _P, _P]),
    "ehss_solve_noalloc": (_INT, [_P, ct.POINTER(Sobol), _INT, _INT, _INT, ct.c_uint, _P, _P, _P, _P]),
    "ehss_sobol_sincos": (_INT, [_I64, _P, _P, _P]),
    "ehss_pa_grid": (_INT, [_I64, _P, _P, _INT, _P, _INT, _P, _P]),
    "ehss_pa_grid_count": (_INT, [_I64, _P, _P, _P, ct.c_double, _INT, _P]),
    "ehss_plan_create": (_INT, [_INT, _P, _P, _P, _P, _INT, ct.c_uint, ct.POINTER(_P)]),
    "ehss_plan_set_sobol": (_INT, [_P, _INT, ct.POINTER(Sobol)]),
    "ehss_plan_retarget": (_INT, [_P, _INT, _P, _P, _P, _P]),
    "ehss_plan_launch": (_INT, [_P]),
    "ehss_plan_query": (_INT, [_P, ct.POINTER(_INT)]),
    "ehss_plan_wait": (_INT, [_P, _P, _P, _P]),
    "ehss_plan_destroy": (_INT, [_P]),
}

_lib = None


def lib() -> ct.CDLL:
    """Load libehss_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ehss_amd.build`")
        handle = ct.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def fn(name: str):
    f = getattr(lib(), name, None)
    if f is None:
        raise RuntimeError(f"libehss_hip.so does not export {name}")
    return f


def last_error() -> str:
    buf = ct.create_string_buffer(2048)
    lib().ehss_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    if status != EHSS_OK:
        msg = last_error()
        if status == EHSS_ERR_ADMISSION:
            raise AdmissionError(status, msg)
        raise EHSSError(status, msg)


def call(name: str, *args) -> None:
    check(fn(name)(*args))


def ptr(a: np.ndarray | None) -> ct.c_void_p:
    return ct.c_void_p(0 if a is None else a.ctypes.data)


def device_info() -> dict:
    name = ct.create_string_buffer(256)
    cu, arch_major, arch_minor = ct.c_int(), ct.c_int(), ct.c_int()
    mem = ct.c_int64()
    call("ehss_device_info", name, len(name), ct.byref(cu), ct.byref(mem), ct.byref(arch_major),
         ct.byref(arch_minor))
    return dict(name=name.value.decode(), compute_units=cu.value, total_mem=mem.value,
                arch=f"{arch_major.value}.{arch_minor.value}")


def trig_compat_info() -> dict:
    """The host-libm compatibility table of the source angles: entries listed
    (-1: no table, cos/sin stay correctly rounded), its libm and file."""
    n = ct.c_int64()
    libm = ct.create_string_buffer(64)
    path = ct.create_string_buffer(4096)
    call("ehss_trig_compat_info", ct.byref(n), libm, len(libm), path, len(path))
    return dict(entries=n.value, libm=libm.value.decode(), path=path.value.decode())


def sobol_sincos(q: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Device sin and cos of the Sobol angles 2 pi q 2^-30 (q: 30-bit words)."""
    q = np.ascontiguousarray(q, dtype=np.uint32)
    s, c = np.empty(len(q)), np.empty(len(q))
    call("ehss_sobol_sincos", len(q), ptr(q), ptr(s), ptr(c))
    return s, c
