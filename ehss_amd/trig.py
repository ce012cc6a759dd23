"""Host side of the source-angle trig (ehss_amd/csrc/ehss_trig.h).

The device source computes cos/sin of the Sobol angles x = fl(2pi q 2^-30)
correctly rounded and applies a compatibility table that lists every 30-bit word
q whose host-libm result (what NumPy returns, and so what the reference's
external_source uses: ehss_state.py:45-49) rounds the other way.  This module
loads that table and the host build of the same C++ source
(libehss_trighost.so), so tests and tools can evaluate exactly what the device
evaluates.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
HOST_LIB = PKG / "libehss_trighost.so"
TABLE = PKG / "trig_compat.bin"
TWO_PI = 2 * np.pi
HEADER_BYTES = 64

_lib = None


def host_lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not HOST_LIB.exists():
            raise RuntimeError(f"{HOST_LIB} is missing: build it with `python -m ehss_amd.build`")
        L = ct.CDLL(str(HOST_LIB))
        L.ehss_trig_host_sincos.argtypes = [ct.c_int64] + [ct.c_void_p] * 4
        L.ehss_trig_host_sobol.argtypes = [ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_double,
                                           ct.c_void_p, ct.c_void_p]
        L.ehss_trig_libm_version.argtypes = [ct.c_char_p, ct.c_size_t]
        _lib = L
    return _lib


def _p(a):
    return ct.c_void_p(0 if a is None else a.ctypes.data)


@dataclass
class CompatTable:
    offsets: np.ndarray  # uint32, buckets + 1
    entries: np.ndarray  # uint16: (q & (2^shift - 1)) << 4 | flags
    shift: int
    libm: str
    not_sure: int
    lookup_from: float  # look up only when |lo| >= lookup_from * ulp(hi) for cos or sin

    def words(self) -> tuple[np.ndarray, np.ndarray]:
        """Every listed word q and its flags (1 cos differs, 2 cos up, 4 sin differs, 8 sin up)."""
        counts = np.diff(self.offsets.astype(np.int64))
        bucket = np.repeat(np.arange(len(counts), dtype=np.uint32), counts)
        q = (bucket << np.uint32(self.shift)) | (self.entries >> np.uint16(4)).astype(np.uint32)
        return q, (self.entries & np.uint16(15)).astype(np.uint8)


def load_table(path: Path | str = TABLE) -> CompatTable:
    raw = Path(path).read_bytes()
    if raw[:8] != b"EHSSTRG2":
        raise ValueError(f"{path}: not a trig compatibility table")
    key_bits, shift, buckets, entries, not_sure, far = np.frombuffer(raw[8:32], np.uint32)
    lookup_from = float(np.frombuffer(raw[32:40], np.float64)[0])
    if key_bits != 30 or far != 0 or not 0 < lookup_from <= 0.5:
        raise ValueError(f"{path}: unsupported table")
    libm = raw[40:64].split(b"\0")[0].decode()
    off_end = HEADER_BYTES + 4 * (int(buckets) + 1)
    offsets = np.frombuffer(raw[HEADER_BYTES:off_end], np.uint32).copy()
    ent = np.frombuffer(raw[off_end:], np.uint16).copy()
    if len(ent) != entries or offsets[-1] != entries:
        raise ValueError(f"{path}: truncated table")
    return CompatTable(offsets, ent, int(shift), libm, int(not_sure), lookup_from)


def host_libm() -> str:
    buf = ct.create_string_buffer(64)
    host_lib().ehss_trig_libm_version(buf, len(buf))
    return buf.value.decode()


def angles(q: np.ndarray) -> np.ndarray:
    """x = fl(2pi * q 2^-30), as NumPy's 2 * np.pi * u forms it."""
    return TWO_PI * (np.asarray(q, dtype=np.uint32).astype(np.float64) * 2.0 ** -30)


def sincos_cr(x: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Host build of the correctly rounded sin/cos: (sin, cos, rounding test passed)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c, sure = np.empty_like(x), np.empty_like(x), np.empty(len(x), np.int8)
    host_lib().ehss_trig_host_sincos(len(x), _p(x), _p(s), _p(c), _p(sure))
    return s, c, sure.astype(bool)


def sobol_sincos(q: np.ndarray, table: CompatTable | None) -> tuple[np.ndarray, np.ndarray]:
    """Host build of the device's source-angle sin/cos of words q (table may be None)."""
    q = np.ascontiguousarray(q, dtype=np.uint32)
    s, c = np.empty(len(q)), np.empty(len(q))
    off = None if table is None else table.offsets
    ent = None if table is None else table.entries
    start = 0.5 if table is None else table.lookup_from
    host_lib().ehss_trig_host_sobol(len(q), _p(q), _p(off), _p(ent), start, _p(s), _p(c))
    return s, c
