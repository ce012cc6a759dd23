"""Thin helpers that call libehss_hip.so through its C-ABI for the GPU tests."""
from __future__ import annotations

import ctypes as ct
import sys

import numpy as np

from ehss_amd import _lib as L


def say(*args):
    print("[gpu-test]", *args, flush=True)
    sys.stdout.flush()


def compat_trace(tree, centers, radii, source, cap, record=True, precision=0, void_abi=False, threads=4):
    """ehss_trace array contract through the C-ABI (host arrays in, host arrays out)."""
    p = np.ascontiguousarray(source["origins"], dtype=np.float64).copy()
    d = np.ascontiguousarray(source["directions"], dtype=np.float64).copy()
    n = len(p)
    lasts = np.ascontiguousarray(source.get("last_atom", np.full(n, -1)), dtype=np.int64)
    counts = np.ascontiguousarray(source.get("initial_bounces", np.zeros(n)), dtype=np.int64).copy()
    escaped = np.zeros(n, np.uint8)
    unresolved = np.zeros(n, np.uint8)
    history = np.full((n, cap if record else 0), -1, np.int32)
    counters = np.zeros(3, np.int64)
    world = np.ascontiguousarray(centers, dtype=np.float64)
    rad = np.ascontiguousarray(radii, dtype=np.float64)
    arrays = [p, d, lasts, counts, escaped, unresolved, history, tree["lo"], tree["hi"], tree["end"],
              tree["leaf_group"], tree["offsets"], tree["ids"], world, rad, counters]
    observed = None
    if void_abi:
        obs = np.zeros(1, np.int32)
        L.fn("ehss_trace")(n, len(tree["end"]), cap, threads, int(record), *(L.ptr(a) for a in arrays + [obs]))
        observed = int(obs[0])
        if observed < 0:
            raise L.EHSSError(-1, L.last_error())
    else:
        L.call("ehss_trace_status", n, len(tree["end"]), cap, precision, int(record), *(L.ptr(a) for a in arrays))
    return dict(positions=p, outgoing=d, escaped=escaped.astype(bool), unresolved=unresolved.astype(bool),
                bounces=counts, collider_ids=history, box_tests=int(counters[0]), sphere_tests=int(counters[1]),
                observed_threads=observed)


FIELDS = ("positions", "outgoing", "escaped", "unresolved", "bounces", "collider_ids")


def assert_same_paths(actual, expected, label, counters=True):
    for k in FIELDS:
        a, e = actual[k], expected[k]
        if not np.array_equal(a, e):
            bad = np.argwhere(a != e)
            raise AssertionError(f"{label}: field {k} differs at {len(bad)} entries, first {bad[:3].tolist()}")
    if counters:
        assert actual["box_tests"] == expected["box_tests"], (label, actual["box_tests"], expected["box_tests"])
        assert actual["sphere_tests"] == expected["sphere_tests"], label
