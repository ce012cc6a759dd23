"""CPU oracle for the EHSS/PA hot path: NumPy restatements plus ctypes access to
the C restatement in ehss_oracle.c.

TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check in __graft_entry__.py and
the cpu_baseline leg of bench.py may import this module, and only as the checker
or the timed CPU baseline.  The product package ehss_amd never imports it.

Each function cites the reference file:line whose arithmetic it restates.  The
NumPy functions use the same NumPy expressions in the same order as the
reference so that, on one machine, their bits agree with the reference's.

Pinning summary (see tests/test_oracle_*.py):
  sobol_scramble / sobol_points   == SciPy 1.15.3 qmc.Sobol golden vectors (bitwise).
  external_source                 per-ray closed forms that fix the u -> ray assignment
                                  (d_z = 1 - 2 u0, azimuths 2 pi u1 and 2 pi u3, impact
                                  radius R sqrt(u2)), planted swaps failing them, and
                                  the single-sphere pi r^2 (tests/test_oracle_source.py).
  readout                         math.fsum of the same contributions; clip bounds.
  compact_source (box, ellipsoid) per-ray closed forms: box area = sum |d_k| x face area,
                                  entry on the face facing the ray, inside its rectangle;
                                  ellipsoid area = pi abc |diag(1/axes) R^T d|; every ray
                                  enters the body; mean weight = surface / 4 (Cauchy)
                                  (tests/test_oracle_compact_source.py).  Which of u3/u4
                                  spans which in-face axis follows the reference text:
                                  parity unpinned beyond it (no reference fixture).
  sample_source / reverse_pa_source  the reference's single-sphere test restated (all
                                  rays escape after one bounce, omega = PA = pi r^2).
  trace / static_tree / pa_*      through ehss_oracle.c (see its header).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libehss_ref.so"

_P = ct.c_void_p
_I64 = ct.c_int64
_INT = ct.c_int


def build(force: bool = False) -> None:
    """Compile liboracle.so (and the reference kernel when its source exists)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE), "liboracle.so"], check=True, capture_output=True)
    if Path("/root/reference/src/native/ehss_cpu.cpp").exists():
        subprocess.run(["make", "-C", str(HERE), "ref"], check=True, capture_output=True)


_lib = None


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ct.CDLL(str(LIB_PATH))
        L.oracle_entry.restype = ct.c_double
        L.oracle_entry.argtypes = [_P, _P, _P, ct.c_double]
        L.oracle_box.restype = _INT
        L.oracle_box.argtypes = [_P, _P, _P, _P, ct.c_double]
        L.oracle_reflect.restype = None
        L.oracle_reflect.argtypes = [_P, _P, _P, ct.c_double, ct.c_double]
        L.oracle_sphere_hit.restype = _INT
        L.oracle_sphere_hit.argtypes = [_P, _P, _P, ct.c_double]
        L.oracle_trace.restype = None
        L.oracle_trace.argtypes = [_I64, _I64, _INT, _INT, _INT] + [_P] * 16
        L.oracle_trace_flat.restype = None
        L.oracle_trace_flat.argtypes = [_I64, _I64, _INT, _INT, _INT] + [_P] * 10
        L.oracle_trace_allatom.restype = None
        L.oracle_trace_allatom.argtypes = [_I64, _I64, _INT] + [_P] * 9
        L.oracle_group_keys.restype = None
        L.oracle_group_keys.argtypes = [_I64, _P, _P, _P]
        L.oracle_groups.restype = _I64
        L.oracle_groups.argtypes = [_I64, _P, _I64, _P]
        L.oracle_topology.restype = None
        L.oracle_topology.argtypes = [_I64] + [_P] * 7
        L.oracle_static_tree.restype = None
        L.oracle_static_tree.argtypes = [_I64, _I64] + [_P] * 9
        L.oracle_pa_build.restype = _I64
        L.oracle_pa_build.argtypes = [_I64, _P, _P, _I64] + [_P] * 6
        L.oracle_pa_query.restype = None
        L.oracle_pa_query.argtypes = [_I64] + [_P] * 11 + [_INT, _P, _P]
        L.oracle_pa_brute.restype = None
        L.oracle_pa_brute.argtypes = [_I64, _P, _P, _P, _I64, _P, _P, _P]
        L.oracle_solve_cpu.restype = None
        L.oracle_solve_cpu.argtypes = ([_I64, _P, _P, _P, ct.c_double, ct.c_double, _INT, _INT, _INT, _I64]
                                       + [_P] * 8 + [_I64] + [_P] * 6 + [_P])
        L.oracle_max_threads.restype = _INT
        L.oracle_row_span.restype = None
        L.oracle_row_span.argtypes = [ct.c_double, ct.c_double, ct.c_double, _P, ct.c_double, _I64, _I64, _P]
        L.oracle_grid_count.restype = _I64
        L.oracle_grid_count.argtypes = [_I64, _P, _P, _P, ct.c_double, _I64, _P]
        L.oracle_grid_brute.restype = _I64
        L.oracle_grid_brute.argtypes = [_I64, _P, _P, _P, ct.c_double, _I64]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> ct.c_void_p:
    return ct.c_void_p(a.ctypes.data)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


# ---------------------------------------------------------------------------
# Scrambled Sobol (third-party: scipy.stats.qmc.Sobol, SciPy 1.15.3 here; the
# reference recorded 1.18.0).  Called at ehss_state.py:44 as
# qmc.Sobol(4, scramble=True, seed=seed).random_base2(power).
# ---------------------------------------------------------------------------

# Joe & Kuo (2008) new-joe-kuo-6.21201 parameters (s, a, m_1..m_s) for the
# first dimensions after the van der Corput dimension.
_JOE_KUO = {1: (1, 0, (1,)), 2: (2, 1, (1, 3)), 3: (3, 1, (1, 3, 1)), 4: (3, 2, (1, 1, 1))}


def sobol_direction_numbers(dim: int, bits: int = 30) -> np.ndarray:
    """Unscrambled direction numbers v[d, j] = m_j << (bits-1-j) (Bratley-Fox recurrence)."""
    if not 1 <= dim <= 5:
        raise ValueError("oracle restates dimensions 1..5 only")
    v = np.zeros((dim, bits), dtype=np.uint64)
    for j in range(bits):
        v[0, j] = 1 << (bits - 1 - j)
    for d in range(1, dim):
        s, a, m0 = _JOE_KUO[d]
        m = list(m0)
        for j in range(s, bits):
            new = m[j - s] ^ (m[j - s] << s)
            for k in range(1, s):
                if (a >> (s - 1 - k)) & 1:
                    new ^= m[j - k] << k
            m.append(new)
        for j in range(bits):
            v[d, j] = m[j] << (bits - 1 - j)
    return v.astype(np.uint32)


def sobol_scramble(seed: int, dim: int = 4, bits: int = 30):
    """LMS + digital shift as in scipy/stats/_qmc.py:1808-1823 (Sobol._scramble).

    shift = rng.integers(2, (d, bits)) . 2^arange(bits); ltm = tril(rng.integers(2, (d, bits, bits)))
    with the unit diagonal forced; output bit p (MSB-first) of each scrambled
    direction number = parity(row p of ltm & input bits, MSB-first).
    """
    rng = np.random.default_rng(seed)
    shift = np.dot(rng.integers(0, 2, size=(dim, bits), dtype=np.uint32),
                   2 ** np.arange(bits, dtype=np.uint32)).astype(np.uint32)
    ltm = np.tril(rng.integers(0, 2, size=(dim, bits, bits), dtype=np.uint32))
    v = sobol_direction_numbers(dim, bits)
    sv = np.zeros_like(v)
    weights = np.array([1 << (bits - 1 - k) for k in range(bits)], dtype=np.uint64)
    for d in range(dim):
        L = ltm[d].astype(np.uint64)
        np.fill_diagonal(L, 1)
        rows = (L * weights[None, :]).sum(axis=1)          # row p read MSB-first
        for j in range(bits):
            vdj = int(v[d, j])
            out = 0
            for p in range(bits):
                if bin(int(rows[p]) & vdj).count("1") & 1:
                    out |= 1 << (bits - 1 - p)
            sv[d, j] = out
    return sv, shift


def sobol_points(sv: np.ndarray, shift: np.ndarray, m: int, bits: int = 30) -> np.ndarray:
    """First 2^m points in SciPy's Gray-code order: q_i = shift ^ XOR_{j in gray(i)} sv[:, j]."""
    n = 1 << m
    i = np.arange(n, dtype=np.uint64)
    g = i ^ (i >> np.uint64(1))
    q = np.tile(shift.astype(np.uint64), (n, 1))
    for j in range(bits):
        sel = ((g >> np.uint64(j)) & np.uint64(1)).astype(bool)
        q[sel] ^= sv[:, j].astype(np.uint64)
    return q.astype(np.float64) * (1.0 / 2 ** bits)


# ---------------------------------------------------------------------------
# Source mapping: ehss_state.py:42-52, inside_out_pa.py:45-58.
# ---------------------------------------------------------------------------

def unit_sphere(uv):
    z = 1 - 2 * uv[:, 0]
    phi = 2 * np.pi * uv[:, 1]
    rho = np.sqrt(np.maximum(0, 1 - z * z))
    return np.column_stack((rho * np.cos(phi), rho * np.sin(phi), z))


def tangent_frames(normals):
    axes = np.zeros_like(normals)
    axes[:, 2] = 1
    axes[np.abs(normals[:, 2]) > .9] = [1, 0, 0]
    t = np.cross(axes, normals)
    t /= np.linalg.norm(t, axis=1)[:, None]
    return t, np.cross(normals, t)


def enclosing(centers, radii):
    """Centre and radius of the impact disk (ehss_state.py:46-47)."""
    center = centers.mean(axis=0)
    radius = float(np.max(np.linalg.norm(centers - center, axis=1) + radii)) * (1. + 1e-9)
    return center, radius


def source_from_u(u, centers, radii):
    """ehss_state.py:45-52 given the Sobol matrix u (n, 4)."""
    directions = unit_sphere(u[:, :2])
    t, b = tangent_frames(directions)
    center, radius = enclosing(centers, radii)
    rho = radius * np.sqrt(u[:, 2])
    phi = 2 * np.pi * u[:, 3]
    origins = center - radius * directions + rho[:, None] * (np.cos(phi)[:, None] * t + np.sin(phi)[:, None] * b)
    n = len(u)
    return dict(origins=np.ascontiguousarray(origins), directions=np.ascontiguousarray(directions),
                incoming=directions.copy(), weights=np.full(n, np.pi * radius * radius / n),
                last_atom=np.full(n, -1, dtype=np.int64), initial_bounces=np.zeros(n, dtype=np.int64))


def external_source(centers, radii, power, seed):
    sv, shift = sobol_scramble(seed)
    return source_from_u(sobol_points(sv, shift, power), centers, radii)


def sample_source(centers, radii, power, seed):
    """inside_out_pa.py:72-90: area-selected surface points, cosine hemisphere."""
    sv, shift = sobol_scramble(seed, dim=5)
    u = sobol_points(sv, shift, power)
    weights = radii ** 2
    ids = np.searchsorted(np.cumsum(weights) / weights.sum(), u[:, 0], side="right")
    normals = unit_sphere(u[:, 1:3])
    t, b = tangent_frames(normals)
    phi = 2 * np.pi * u[:, 4]
    rho = np.sqrt(u[:, 3])
    directions = (np.sqrt(1 - u[:, 3])[:, None] * normals
                  + (rho * np.cos(phi))[:, None] * t
                  + (rho * np.sin(phi))[:, None] * b)
    origins = centers[ids] + radii[ids, None] * normals
    area = float(4 * np.pi * np.dot(radii, radii))
    return dict(origins=np.ascontiguousarray(origins), directions=np.ascontiguousarray(directions),
                owners=ids.astype(np.int64), area=area)


def compact_source(kind, center, rotation, axes, power, seed):
    """ehss_compact_source.py:44-87, kinds 'box' and 'ellipsoid': directions
    uniform on S2, a uniform point of the body's projection (box: a face chosen
    in proportion to its projected area), weight = projected area / n."""
    center, rotation, axes = _f64(center), _f64(rotation).reshape(3, 3), _f64(axes)
    sv, shift = sobol_scramble(seed, dim=5 if kind == "box" else 4)
    u = sobol_points(sv, shift, power)
    d = unit_sphere(u[:, :2])
    n = len(d)
    if kind == "box":
        local_d = d @ rotation
        face_areas = 4 * np.prod(axes) / axes
        projected = np.abs(local_d) * face_areas
        area = projected.sum(axis=1)
        face = (u[:, 2, None] * area[:, None] >= np.cumsum(projected, axis=1)).sum(axis=1)
        p = np.empty((n, 3))
        for axis in range(3):
            selected = face == axis
            p[selected, axis] = -np.sign(local_d[selected, axis]) * axes[axis]
            other = [k for k in range(3) if k != axis]
            p[np.ix_(selected, other)] = (2 * u[selected, 3:5] - 1) * axes[other]
        origins = center + p @ rotation.T - 1e-10 * max(1., axes.max()) * d
    elif kind == "ellipsoid":
        t, b = tangent_frames(d)
        lt, lb = (t @ rotation) * axes, (b @ rotation) * axes
        g00 = np.sum(lt * lt, axis=1)
        g01 = np.sum(lt * lb, axis=1)
        g11 = np.sum(lb * lb, axis=1)
        l00 = np.sqrt(g00)
        l10 = g01 / l00
        l11 = np.sqrt(np.maximum(0., g11 - l10 * l10))
        rho, phi = np.sqrt(u[:, 2]), 2 * np.pi * u[:, 3]
        x, y = rho * np.cos(phi), rho * np.sin(phi)
        offset = (l00 * x)[:, None] * t + (l10 * x + l11 * y)[:, None] * b
        extent = np.linalg.norm((d @ rotation) * axes, axis=1)
        origins = center + offset - (extent * (1 + 1e-10))[:, None] * d
        area = np.pi * l00 * l11
    else:
        raise ValueError("kind must be box or ellipsoid")
    return dict(origins=np.ascontiguousarray(origins), directions=np.ascontiguousarray(d), incoming=d.copy(),
                weights=area / n, last_atom=np.full(n, -1, np.int64), initial_bounces=np.zeros(n, np.int64))


def reverse_pa_source(centers, radii, power, seed):
    """ehss_state.py:55-69: unblocked surface points, first reflection applied."""
    c, r = _f64(centers), _f64(radii)
    s = sample_source(c, r, power, seed)
    blocked = pa_brute(c, r, s["origins"], s["directions"], s["owners"])
    keep = ~blocked
    owners = s["owners"][keep]
    origins = s["origins"][keep]
    incoming = -s["directions"][keep]
    normals = (origins - c[owners]) / r[owners, None]
    normals /= np.linalg.norm(normals, axis=1)[:, None]
    outgoing = incoming - 2 * np.sum(incoming * normals, axis=1)[:, None] * normals
    outgoing /= np.linalg.norm(outgoing, axis=1)[:, None]
    n = len(owners)
    return dict(origins=np.ascontiguousarray(origins), directions=np.ascontiguousarray(outgoing),
                incoming=np.ascontiguousarray(incoming),
                weights=np.full(n, s["area"] / (4 * len(s["owners"]))),
                last_atom=owners.astype(np.int64), initial_bounces=np.ones(n, dtype=np.int64))


# ---------------------------------------------------------------------------
# Trees and traces (through the C restatement).
# ---------------------------------------------------------------------------

def group_keys(centers, radii):
    c, r = _f64(centers), _f64(radii)
    keys = np.empty_like(c)
    lib().oracle_group_keys(len(r), _ptr(c), _ptr(r), _ptr(keys))
    return keys


def static_groups(centers, radii, leaf_size=4):
    """ehss_static.py:53-59 -> (groups (N,), group count)."""
    keys = group_keys(centers, radii)
    groups = np.empty(len(keys), dtype=np.int64)
    ng = lib().oracle_groups(len(keys), _ptr(keys), leaf_size, _ptr(groups))
    if ng < 0:
        raise RuntimeError("oracle grouping failed")
    return groups, int(ng)


def static_tree(centers, radii, leaf_size=4, groups=None):
    """PreparedStaticCPUEHSS tree arrays (ehss_static.py:62-91)."""
    c, r = _f64(centers), _f64(radii)
    if groups is None:
        groups, ng = static_groups(c, r, leaf_size)
    else:
        groups = _i64(groups)
        ng = int(groups.max()) + 1
    nn = 2 * ng - 1
    lo, hi = np.empty((nn, 3)), np.empty((nn, 3))
    end, leaf = np.empty(nn, np.int64), np.empty(nn, np.int64)
    offsets, ids = np.empty(ng + 1, np.int64), np.empty(len(r), np.int64)
    lib().oracle_static_tree(len(r), ng, _ptr(groups), _ptr(c), _ptr(r), _ptr(lo), _ptr(hi), _ptr(end),
                             _ptr(leaf), _ptr(offsets), _ptr(ids))
    return dict(lo=lo, hi=hi, end=end, leaf_group=leaf, offsets=offsets, ids=ids, groups=groups)


def topology(ng):
    nn = 2 * ng - 1
    arrs = [np.empty(nn, np.int64) for _ in range(5)]
    leaf_node = np.empty(ng, np.int64)
    depth = np.empty(nn, np.int64)
    left, right, end, parent, leaf = arrs
    lib().oracle_topology(ng, _ptr(left), _ptr(right), _ptr(end), _ptr(parent), _ptr(leaf), _ptr(leaf_node),
                          _ptr(depth))
    return dict(left=left, right=right, end=end, parent=parent, leaf=leaf, leaf_node=leaf_node, depth=depth)


def pa_tree(centers, radii, leaf_size=4):
    """pa_spatial.py:38-75 build_tree(c - r, c + r, 'bvh', leaf_size)."""
    c, r = _f64(centers), _f64(radii)
    n = len(r)
    cap = 2 * n + 1
    lo, hi = np.empty((cap, 3)), np.empty((cap, 3))
    end, first, count = (np.empty(cap, np.int64) for _ in range(3))
    ids = np.empty(n, np.int64)
    nodes = lib().oracle_pa_build(n, _ptr(c), _ptr(r), leaf_size, _ptr(lo), _ptr(hi), _ptr(end), _ptr(first),
                                  _ptr(count), _ptr(ids))
    return dict(lo=lo[:nodes].copy(), hi=hi[:nodes].copy(), end=end[:nodes].copy(),
                first=first[:nodes].copy(), count=count[:nodes].copy(), ids=ids)


def pa_groups(tree, n):
    """ccs_cli.py:129-139 spatial_groups: group ids in PA leaf order."""
    groups = np.empty(n, dtype=np.int64)
    g = 0
    for first, count in zip(tree["first"], tree["count"]):
        if count:
            groups[tree["ids"][first:first + count]] = g
            g += 1
    return groups


def trace(tree, centers, radii, origins, directions, last_atom=None, initial_bounces=None, cap=64,
          record=True, threads=1, lib_handle=None, fn="oracle_trace"):
    """ehss_cpu.py:66-100 / ehss_cpu.cpp:42-82 on copies of the rays."""
    p, d = _f64(origins).copy(), _f64(directions).copy()
    n = len(p)
    lasts = _i64(np.full(n, -1) if last_atom is None else last_atom)
    counts = _i64(np.zeros(n) if initial_bounces is None else initial_bounces).copy()
    escaped = np.zeros(n, np.uint8)
    unresolved = np.zeros(n, np.uint8)
    history = np.full((n, cap if record else 0), -1, np.int32)
    counters = np.zeros(2, np.int64)
    world, rad = _f64(centers), _f64(radii)
    arrays = [p, d, lasts, counts, escaped, unresolved, history, tree["lo"], tree["hi"], tree["end"],
              tree["leaf_group"], tree["offsets"], tree["ids"], world, rad, counters]
    if lib_handle is None:
        getattr(lib(), fn)(n, len(tree["end"]), cap, threads, int(record), *(_ptr(a) for a in arrays))
    else:  # the reference's own ehss_trace (oracle/_ref), which adds observed_threads
        observed = np.zeros(1, np.int32)
        lib_handle.ehss_trace(n, len(tree["end"]), cap, threads, int(record),
                              *(_ptr(a) for a in arrays + [observed]))
    return dict(positions=p, outgoing=d, escaped=escaped.astype(bool), unresolved=unresolved.astype(bool),
                bounces=counts, collider_ids=history, box_tests=int(counters[0]), sphere_tests=int(counters[1]))


def trace_flat(centers, radii, origins, directions, last_atom=None, initial_bounces=None, cap=64, threads=1):
    p, d = _f64(origins).copy(), _f64(directions).copy()
    n = len(p)
    lasts = _i64(np.full(n, -1) if last_atom is None else last_atom)
    counts = _i64(np.zeros(n) if initial_bounces is None else initial_bounces).copy()
    escaped, unresolved = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    history = np.full((n, cap), -1, np.int32)
    counters = np.zeros(2, np.int64)
    world, rad = _f64(centers), _f64(radii)
    lib().oracle_trace_flat(n, len(rad), cap, threads, 1, _ptr(p), _ptr(d), _ptr(lasts), _ptr(counts),
                            _ptr(escaped), _ptr(unresolved), _ptr(history), _ptr(world), _ptr(rad), _ptr(counters))
    return dict(positions=p, outgoing=d, escaped=escaped.astype(bool), unresolved=unresolved.astype(bool),
                bounces=counts, collider_ids=history, sphere_tests=int(counters[1]))


def trace_allatom(centers, radii, origins, directions, last_atom=None, initial_bounces=None, cap=64):
    p, d = _f64(origins).copy(), _f64(directions).copy()
    n = len(p)
    lasts = _i64(np.full(n, -1) if last_atom is None else last_atom)
    counts = _i64(np.zeros(n) if initial_bounces is None else initial_bounces).copy()
    escaped, unresolved = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    history = np.full((n, cap), -1, np.int32)
    world, rad = _f64(centers), _f64(radii)
    lib().oracle_trace_allatom(n, len(rad), cap, _ptr(p), _ptr(d), _ptr(lasts), _ptr(counts), _ptr(escaped),
                               _ptr(unresolved), _ptr(history), _ptr(world), _ptr(rad))
    return dict(positions=p, outgoing=d, escaped=escaped.astype(bool), unresolved=unresolved.astype(bool),
                bounces=counts, collider_ids=history)


def pa_query(tree, centers, radii, origins, directions, owners=None, threads=1):
    o, d = _f64(origins), _f64(directions)
    n = len(o)
    own = _i64(np.full(n, -1) if owners is None else owners)
    hits = np.zeros(n, np.uint8)
    counters = np.zeros(2, np.int64)
    c, r = _f64(centers), _f64(radii)
    lib().oracle_pa_query(n, _ptr(o), _ptr(d), _ptr(own), _ptr(c), _ptr(r), _ptr(tree["lo"]), _ptr(tree["hi"]),
                          _ptr(tree["end"]), _ptr(tree["first"]), _ptr(tree["count"]), _ptr(tree["ids"]),
                          threads, _ptr(hits), _ptr(counters))
    return hits.astype(bool), int(counters[0]), int(counters[1])


def pa_brute(centers, radii, origins, directions, owners=None):
    o, d = _f64(origins), _f64(directions)
    n = len(o)
    own = _i64(np.full(n, -1) if owners is None else owners)
    hits = np.zeros(n, np.uint8)
    c, r = _f64(centers), _f64(radii)
    lib().oracle_pa_brute(n, _ptr(o), _ptr(d), _ptr(own), len(r), _ptr(c), _ptr(r), _ptr(hits))
    return hits.astype(bool)


# ---------------------------------------------------------------------------
# Readout: ehss_state.py:84-103, ccs_cli.py:211-219.
# ---------------------------------------------------------------------------

def readout(source, result, cap):
    angle = np.clip(1 - np.sum(source["incoming"] * result["outgoing"], axis=1), 0., 2.)
    contributions = source["weights"] * angle * result["escaped"]
    unresolved = result["unresolved"]
    bounces = result["bounces"]
    return dict(contributions=contributions, omega=float(contributions.sum()),
                residual_mass=float(source["weights"][unresolved].sum()),
                tail_bound=2 * float(source["weights"][unresolved].sum()),
                pa=float(source["weights"][(bounces > 0) | unresolved].sum()),
                histogram=np.bincount(bounces, minlength=cap + 1),
                ledger=np.bincount(bounces, weights=contributions, minlength=cap + 1))


def solve_cpu(centers, radii, seed, power, cap=64, threads=None, pa=True, leaf_size=4):
    """Whole path in C (Sobol -> source -> PA -> trace -> readout), for timing."""
    c, r = _f64(centers), _f64(radii)
    tree = static_tree(c, r, leaf_size)
    ptree = pa_tree(c, r, leaf_size)
    sv, shift = sobol_scramble(seed)
    sv = np.ascontiguousarray(sv, dtype=np.uint32)
    shift = np.ascontiguousarray(shift, dtype=np.uint32)
    center, radius = enclosing(c, r)
    center = _f64(center)
    n = 1 << power
    w = np.pi * radius * radius / n
    out = np.zeros(7)
    threads = threads or os.cpu_count() or 1
    lib().oracle_solve_cpu(n, _ptr(sv), _ptr(shift), _ptr(center), radius, w, cap, threads, int(pa),
                           len(tree["end"]), _ptr(tree["lo"]), _ptr(tree["hi"]), _ptr(tree["end"]),
                           _ptr(tree["leaf_group"]), _ptr(tree["offsets"]), _ptr(tree["ids"]), _ptr(c), _ptr(r),
                           len(ptree["end"]), _ptr(ptree["lo"]), _ptr(ptree["hi"]), _ptr(ptree["end"]),
                           _ptr(ptree["first"]), _ptr(ptree["count"]), _ptr(ptree["ids"]), _ptr(out))
    return dict(omega=out[0], pa=out[1], pa_from_ehss=out[2], residual_mass=out[3], unresolved=int(out[4]),
                box_tests=int(out[5]), sphere_tests=int(out[6]), rays=n, threads=threads)


def ref_lib():
    """The reference's own ehss_trace compiled into oracle/_ref (None if absent)."""
    if not REF_PATH.exists():
        return None
    L = ct.CDLL(str(REF_PATH))
    L.ehss_trace.restype = None
    L.ehss_trace.argtypes = [_I64, _I64, _INT, _INT, _INT] + [_P] * 17
    return L


# ---------------------------------------------------------------------------
# Finite-grid ("adaptive") PA of the CLI (ccs_cli.py:142-160)
# ---------------------------------------------------------------------------
def project_grid(centers, radii, direction, n):
    """pa_walsh_geometry.py:171-180: projection frame, centred disk centres,
    enclosing square (lower corner, pixel spacing)."""
    direction = np.asarray(direction, float)
    direction = direction / np.linalg.norm(direction)
    axis = np.array([0., 0., 1.]) if abs(direction[2]) < .9 else np.array([1., 0., 0.])
    u = np.cross(axis, direction)
    u /= np.linalg.norm(u)
    v = np.cross(direction, u)
    centered = np.asarray(centers, float) - np.asarray(centers, float).mean(axis=0)
    xy = np.column_stack((centered @ u, centered @ v))
    extent = float(np.max(np.abs(xy) + np.asarray(radii, float)[:, None])) * (1. + 1e-10)
    lower = np.array([-extent, -extent])
    spacing = 2 * extent / n
    return xy, lower, spacing, u, v, direction


def grid_count(xy, radii, lower, spacing, n):
    """Covered-pixel count of the disk union (pa_dyadic_bitmask.py:86-97)."""
    xy = _f64(xy)
    r = _f64(radii)
    lo = _f64(lower)
    rows = ct.c_int64(0)
    c = lib().oracle_grid_count(len(r), _ptr(xy), _ptr(r), _ptr(lo), float(spacing), int(n), ct.byref(rows))
    return int(c), int(rows.value)


def grid_brute(xy, radii, lower, spacing, n):
    xy, r, lo = _f64(xy), _f64(radii), _f64(lower)
    return int(lib().oracle_grid_brute(len(r), _ptr(xy), _ptr(r), _ptr(lo), float(spacing), int(n)))


def row_span(cx, cy, r, lower, spacing, y, n):
    out = np.zeros(2, np.int64)
    lo = _f64(lower)
    lib().oracle_row_span(float(cx), float(cy), float(r), _ptr(lo), float(spacing), int(y), int(n), _ptr(out))
    return int(out[0]), int(out[1])


def adaptive_directions(directions, seed):
    """ccs_cli.py:148: unit_sphere of the first `directions` points of a 2-D
    scrambled Sobol net."""
    m = int(directions).bit_length() - 1
    sv, shift = sobol_scramble(seed, dim=2)
    return unit_sphere(sobol_points(sv, shift, m))


def adaptive_pa(centers, radii, directions, grid_size, seed):
    """ccs_cli.py:142-160 with the union count of oracle_grid_count."""
    values = []
    for direction in adaptive_directions(directions, seed):
        xy, lower, spacing, *_ = project_grid(centers, radii, direction, grid_size)
        values.append(float(grid_count(xy, radii, lower, spacing, grid_size)[0]) * spacing ** 2)
    return float(np.mean(values)), values
