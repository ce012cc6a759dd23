"""Python host layer over libehss_hip.so: Sobol scrambles and device scenes.

Everything numerical runs on the GPU through the C-ABI; this module only
validates arguments, owns handles and converts between NumPy arrays and the
C structures.
"""
from __future__ import annotations

import ctypes as ct
import functools
import time

import numpy as np

from . import _lib as L

BITS = 30

# Joe & Kuo (2008), new-joe-kuo-6.21201: (s, a, m_1..m_s) for Sobol dimensions
# 2..5 (dimension 1 is van der Corput).  The same published table SciPy uses.
_JOE_KUO = ((1, 0, (1,)), (2, 1, (1, 3)), (3, 1, (1, 3, 1)), (3, 2, (1, 1, 1)))


@functools.lru_cache(maxsize=None)
def _direction_numbers(dim: int, bits: int) -> np.ndarray:
    v = _direction_numbers_uncached(dim, bits)
    v.setflags(write=False)
    return v


def direction_numbers(dim: int = 4, bits: int = BITS) -> np.ndarray:
    """v[d, j] = m_j << (bits - 1 - j), m_j from the Bratley-Fox recurrence."""
    return _direction_numbers(int(dim), int(bits)).copy()


def _direction_numbers_uncached(dim: int, bits: int) -> np.ndarray:
    if not 1 <= dim <= len(_JOE_KUO) + 1:
        raise ValueError(f"dimension must be 1..{len(_JOE_KUO) + 1}")
    v = np.zeros((dim, bits), dtype=np.int64)
    v[0] = [1 << (bits - 1 - j) for j in range(bits)]
    for d in range(1, dim):
        s, a, init = _JOE_KUO[d - 1]
        m = list(init)
        for j in range(s, bits):
            x = m[j - s] ^ (m[j - s] << s)
            for k in range(1, s):
                if (a >> (s - 1 - k)) & 1:
                    x ^= m[j - k] << k
            m.append(x)
        v[d] = [m[j] << (bits - 1 - j) for j in range(bits)]
    return v


def scramble_matrices(seed: int, dim: int = 4, bits: int = BITS):
    """Scrambled direction numbers and digital shift of scipy.stats.qmc.Sobol(dim,
    scramble=True, seed=seed): LMS + shift, _qmc.py:1808-1823.  The generator is
    numpy.random.default_rng(seed) (PCG64); the shift bits are drawn first, then
    the lower-triangular matrices (unit diagonal)."""
    if not 0 <= int(seed) < 2 ** 64:
        raise ValueError("seed must be a nonnegative integer")
    rng = np.random.default_rng(int(seed))
    shift_bits = rng.integers(0, 2, size=(dim, bits), dtype=np.uint32).astype(np.int64)
    shift = shift_bits @ (np.int64(1) << np.arange(bits, dtype=np.int64))
    ltm = np.tril(rng.integers(0, 2, size=(dim, bits, bits), dtype=np.uint32)).astype(np.int64)
    idx = np.arange(bits)
    ltm[:, idx, idx] = 1                                    # unit diagonal
    v = _direction_numbers(int(dim), int(bits))
    msb = np.int64(1) << (bits - 1 - np.arange(bits, dtype=np.int64))  # weight of MSB-first bit p
    vb = ((v[:, :, None] & msb[None, None, :]) != 0).astype(np.float64)  # vb[d, j, k]: MSB-first bit k of v[d, j]
    # out[d, j, p] = parity(L[d, p, :] . vb[d, j, :]); the dot products are
    # small integers (<= bits), exact in FP64.
    out = np.matmul(vb, ltm.transpose(0, 2, 1).astype(np.float64)).astype(np.int64) & 1
    sv = out @ msb
    return sv.astype(np.uint32), shift.astype(np.uint32)


def scramble(seed: int) -> L.Sobol:
    """ehss_sobol structure for a 4-D scrambled net with the given seed, built
    by the library (ehss_scramble: the same bits as scramble_py, ~100x faster)."""
    if not 0 <= int(seed) < 2 ** 64:
        raise ValueError("seed must be a nonnegative integer")
    s = L.Sobol()
    L.call("ehss_scramble", int(seed), ct.byref(s))
    return s


def scramble_many(seeds) -> ct.Array:
    """scramble() of every seed, on the host's cores (ehss_scramble_many); the
    result is a launch table (sobol_table) as it is."""
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    out = (L.Sobol * len(seeds))()
    L.call("ehss_scramble_many", len(seeds), L.ptr(seeds), ct.byref(out))
    return out


def scramble_py(seed: int) -> L.Sobol:
    """scramble() from the NumPy restatement (scramble_matrices)."""
    sv, shift = scramble_matrices(seed, 4)
    words = np.zeros(4 * 32 + 4, dtype=np.uint32)
    words[:4 * 32].reshape(4, 32)[:, :BITS] = sv
    words[4 * 32:] = shift
    s = L.Sobol()
    ct.memmove(ct.addressof(s), words.ctypes.data, words.nbytes)
    return s


def sobol_points(seed: int, dim: int, m: int) -> np.ndarray:
    """First 2^m points of scipy.stats.qmc.Sobol(dim, scramble=True,
    seed).random_base2(m) (Gray-code order, _qmc.py:1866-1879), on the host."""
    sv, shift = scramble_matrices(seed, dim)
    n = 1 << int(m)
    i = np.arange(n, dtype=np.uint64)
    g = i ^ (i >> np.uint64(1))
    q = np.tile(np.asarray(shift, dtype=np.uint64), (n, 1))
    for j in range(int(m)):
        sel = ((g >> np.uint64(j)) & np.uint64(1)).astype(bool)
        q[sel] ^= np.asarray(sv[:, j], dtype=np.uint64)
    return q.astype(np.float64) * 2.0 ** -BITS


def sobol_table(sobols):
    """ctypes array of ehss_sobol for a list of scrambles (returned unchanged when
    it already is one), so launches that reuse a set of scrambles convert it once."""
    if isinstance(sobols, ct.Array) and sobols._type_ is L.Sobol:
        return sobols
    return (L.Sobol * len(sobols))(*sobols)


def replicate_seed(seed: int, repeat: int) -> int:
    """ccs_cli.py:190: (seed + 104729 * repeat) mod 2^32."""
    return (int(seed) + 104729 * int(repeat)) % (2 ** 32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


def precision_code(precision) -> int:
    if isinstance(precision, str):
        if precision not in L.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(L.PRECISIONS)}")
        return L.PRECISIONS[precision]
    if precision not in (0, 1):
        raise ValueError("precision must be 0 (fp64) or 1 (fp32-fixup)")
    return int(precision)


def coop_flags(coop: int | None, group: int | None = None, stage1: int | None = None,
               packets: int | None = None) -> int:
    """Flag bits of a tail threshold, of the drain's lanes per ray, of the
    queries a ray gets in the first drain and of the packet queries before the
    drain (None: library defaults; stage1 0: one drain)."""
    flags = 0
    if packets is not None:
        if not 1 <= int(packets) <= 15:
            raise ValueError("packets must be in 1..15")
        flags |= int(packets) << L.PACKETS_SHIFT
    if stage1 is not None:
        if not 0 <= int(stage1) <= 14:
            raise ValueError("stage1 must be in 0..14")
        flags |= (int(stage1) + 1) << L.STAGE1_SHIFT
    if coop is not None:
        if not 0 <= int(coop) <= 64:
            raise ValueError("coop must be in 0..64")
        flags |= (int(coop) + 1) << L.COOP_SHIFT
    if group is not None:
        g = int(group)
        if g < 1 or g > 64 or g & (g - 1):
            raise ValueError("group must be a power of two in 1..64")
        flags |= (g.bit_length()) << L.GROUP_SHIFT
    return flags


class GPUScene:
    """Device-resident spheres with the static EHSS tree and the PA tree built
    on the GPU (ehss_scene_create)."""

    def __init__(self, centers, radii, leaf_size: int = 4, groups=None):
        c, r = _f64(centers), _f64(radii)
        if c.ndim != 2 or c.shape[1] != 3 or r.shape != (len(c),) or len(c) == 0:
            raise ValueError("Expected nonempty centers (N,3) and radii (N,).")
        if type(leaf_size) is not int or leaf_size < 1:
            raise ValueError("Positive integer leaf size required")
        self.centers, self.radii = c, r
        handle = ct.c_void_p()
        t0 = time.perf_counter()
        if groups is None:
            L.call("ehss_scene_create", len(r), L.ptr(c), L.ptr(r), leaf_size, ct.byref(handle))
        else:
            g = _i64(groups)
            if g.shape != (len(r),):
                raise ValueError("One integer group per atom required")
            L.call("ehss_scene_create_grouped", len(r), L.ptr(c), L.ptr(r), L.ptr(g), leaf_size, ct.byref(handle))
        self.build_s = time.perf_counter() - t0
        self._h = handle
        info = L.SceneInfo()
        L.call("ehss_scene_get_info", self._h, ct.byref(info))
        self.info = {name: (list(getattr(info, name)) if name == "center" else getattr(info, name))
                     for name, _ in L.SceneInfo._fields_}

    def rebuild(self, centers, radii, leaf_size: int | None = None) -> None:
        """Replace the structure (a new conformer) in place: the device arrays
        are reused and only grow (ehss_scene_rebuild)."""
        c, r = _f64(centers), _f64(radii)
        if c.ndim != 2 or c.shape[1] != 3 or r.shape != (len(c),) or len(c) == 0:
            raise ValueError("Expected nonempty centers (N,3) and radii (N,).")
        leaf = self.info["leaf_size"] if leaf_size is None else leaf_size
        if type(leaf) is not int or leaf < 1:
            raise ValueError("Positive integer leaf size required")
        t0 = time.perf_counter()
        L.call("ehss_scene_rebuild", self.handle, len(r), L.ptr(c), L.ptr(r), leaf)
        self.build_s = time.perf_counter() - t0
        self.centers, self.radii = c, r
        info = L.SceneInfo()
        L.call("ehss_scene_get_info", self._h, ct.byref(info))
        self.info = {name: (list(getattr(info, name)) if name == "center" else getattr(info, name))
                     for name, _ in L.SceneInfo._fields_}

    def device_bytes(self) -> int:
        """Device memory the scene holds now (trees + solve workspace)."""
        info = L.SceneInfo()
        L.call("ehss_scene_get_info", self.handle, ct.byref(info))
        return int(info.device_bytes)

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            L.check(L.fn("ehss_scene_destroy")(self._h))  # refused while a plan holds the scene
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("scene is closed")
        return self._h

    # -- tree export -------------------------------------------------------------
    def export_tree(self) -> dict:
        i = self.info
        nn, ng, n = i["nodes"], i["groups"], i["atoms"]
        lo, hi = np.empty((nn, 3)), np.empty((nn, 3))
        end, leaf = np.empty(nn, np.int64), np.empty(nn, np.int64)
        offsets, ids, groups = np.empty(ng + 1, np.int64), np.empty(n, np.int64), np.empty(n, np.int64)
        L.call("ehss_scene_export_tree", self.handle, L.ptr(lo), L.ptr(hi), L.ptr(end), L.ptr(leaf), L.ptr(offsets),
               L.ptr(ids), L.ptr(groups))
        return dict(lo=lo, hi=hi, end=end, leaf_group=leaf, offsets=offsets, ids=ids, groups=groups)

    def export_pa_tree(self) -> dict:
        i = self.info
        nn, n = i["pa_nodes"], i["atoms"]
        lo, hi = np.empty((nn, 3)), np.empty((nn, 3))
        end, first, count = (np.empty(nn, np.int64) for _ in range(3))
        ids = np.empty(n, np.int64)
        L.call("ehss_scene_export_pa_tree", self.handle, L.ptr(lo), L.ptr(hi), L.ptr(end), L.ptr(first),
               L.ptr(count), L.ptr(ids))
        return dict(lo=lo, hi=hi, end=end, first=first, count=count, ids=ids)

    # -- caller rays ----------------------------------------------------------------
    def trace(self, origins, directions, last_atom=None, initial_bounces=None, cap=64, record=False,
              precision="fp32-fixup") -> dict:
        p, d = _f64(origins).copy(), _f64(directions).copy()
        n = len(p)
        if p.shape != (n, 3) or d.shape != (n, 3):
            raise ValueError("origins and directions must be (N,3)")
        lasts = _i64(np.full(n, -1) if last_atom is None else last_atom)
        counts = _i64(np.zeros(n) if initial_bounces is None else initial_bounces).copy()
        escaped, unresolved = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        history = np.full((n, cap if record else 0), -1, np.int32)
        counters = np.zeros(3, np.int64)
        L.call("ehss_scene_trace", self.handle, n, int(cap), precision_code(precision), int(bool(record)),
               L.ptr(p), L.ptr(d), L.ptr(lasts), L.ptr(counts), L.ptr(escaped), L.ptr(unresolved), L.ptr(history),
               L.ptr(counters))
        return dict(positions=p, outgoing=d, escaped=escaped.astype(bool), unresolved=unresolved.astype(bool),
                    bounces=counts, collider_ids=history, box_tests=int(counters[0]), sphere_tests=int(counters[1]),
                    fp64_retests=int(counters[2]))

    def pa_query(self, origins, directions, owners=None, precision="fp64"):
        o, d = _f64(origins), _f64(directions)
        n = len(o)
        own = None if owners is None else _i64(owners)
        hits = np.zeros(n, np.uint8)
        counters = np.zeros(2, np.int64)
        L.call("ehss_scene_pa_query", self.handle, n, precision_code(precision), L.ptr(o), L.ptr(d), L.ptr(own),
               L.ptr(hits), L.ptr(counters))
        return hits.astype(bool), int(counters[0]), int(counters[1])

    # -- fused device solve ---------------------------------------------------------
    def solve(self, sobol: L.Sobol, power: int, cap: int = 64, precision="fp32-fixup", pa: bool = True,
              keep_rays: bool = False, histogram: bool = False, variant: int = 0, coop: int | None = None,
              timestamps: bool = False, group: int | None = None, stage1: int | None = None,
              lds: bool = False, packets: int | None = None, continuable: bool = False) -> dict:
        """One fused device solve of 2^power rays.  `variant` selects the FP32
        filter kernel (0 default, 1 per-lane, 2 packet first query, 3 packet
        throughout, 4 packet first queries then queue + persistent drain);
        `coop` is the tail threshold (a wave with at most that many live rays
        finishes them with wave-cooperative queries; 0 off, None default);
        `group` is the lanes per ray of variant 4's drain, `stage1` the queries a
        ray gets there before the second drain takes it (None: the library's choice; from 2^18 rays the
        drains pick their lanes per ray from their queue lengths), `packets` the packet queries
        before the drain.  All settings return identical per-ray results."""
        totals = L.Totals()
        hist = np.zeros(cap + 1, np.int64) if histogram else None
        ledger = np.zeros(cap + 1, np.float64) if histogram else None
        flags = (L.FLAG_PA if pa else 0) | (L.FLAG_KEEP_RAYS if keep_rays else 0) | (int(variant) << L.VARIANT_SHIFT)
        flags |= coop_flags(coop, group, stage1, packets) | (L.FLAG_TIMESTAMPS if timestamps else 0)
        flags |= (L.FLAG_LDS_PACKETS if lds else 0) | (L.FLAG_CONTINUABLE if continuable else 0)
        L.call("ehss_solve", self.handle, ct.byref(sobol), int(power), int(cap), precision_code(precision),
               flags, ct.byref(totals), L.ptr(hist), L.ptr(ledger))
        out = totals.as_dict()
        if histogram:
            out["histogram"] = hist
            out["ledger"] = ledger
        return out

    def solve_many(self, sobols, power: int, cap: int = 64, precision="fp32-fixup", pa: bool = True,
                   histogram: bool = False, variant: int = 0, coop: int | None = None, group: int | None = None,
                   keep_rays: bool = False, stage1: int | None = None, lds: bool = False,
                   per_replicate: bool = False, packets: int | None = None, as_array: bool = False,
                   continuable: bool = False, unsorted: bool = False):
        """len(sobols) independent solves in one launch; one result dict per
        scramble, each equal to solve() with that scramble.  per_replicate:
        never schedule the replicates across the lanes of a wave.  unsorted: with
        256 or more replicates of small nets, keep the shared-partition schedule
        instead of ordering the launch's rays by 4-D cell.  as_array: the
        totals as one NumPy structured array (fields of ehss_totals) instead of
        dicts, for launches of thousands of replicates.  `sobols` may be a
        prebuilt sobol_table()."""
        if as_array and histogram:
            raise ValueError("as_array returns totals only")
        table = sobol_table(sobols)
        reps = len(table)
        totals = (L.Totals * reps)()
        hist = np.zeros(reps * (cap + 1), np.int64) if histogram else None
        ledger = np.zeros(reps * (cap + 1), np.float64) if histogram else None
        flags = (L.FLAG_PA if pa else 0) | (L.FLAG_KEEP_RAYS if keep_rays else 0) | \
            (int(variant) << L.VARIANT_SHIFT) | coop_flags(coop, group, stage1, packets) | \
            (L.FLAG_LDS_PACKETS if lds else 0)
        flags |= (L.FLAG_PER_REPLICATE if per_replicate else 0) | (L.FLAG_CONTINUABLE if continuable else 0)
        flags |= L.FLAG_UNSORTED if unsorted else 0
        L.call("ehss_solve_many", self.handle, reps, table, int(power), int(cap), precision_code(precision), flags,
               totals, L.ptr(hist), L.ptr(ledger))
        if as_array:
            return np.frombuffer(totals, dtype=L.TOTALS_DTYPE, count=reps).copy()
        return self._split(totals, reps, cap, hist, ledger)

    def extend(self, new_cap: int, replicates: int = 1, histogram: bool = False) -> list[dict]:
        """Continues the last solve (made with continuable=True, or a previous
        extend) at a larger cap: only the rays that ended unresolved are traced
        further.  One result per replicate of that solve, bit-identical to
        solving the same rays again at new_cap (work counters aside)."""
        totals = (L.Totals * replicates)()
        hist = np.zeros(replicates * (new_cap + 1), np.int64) if histogram else None
        ledger = np.zeros(replicates * (new_cap + 1), np.float64) if histogram else None
        L.call("ehss_solve_extend", self.handle, int(new_cap), totals, L.ptr(hist), L.ptr(ledger))
        return self._split(totals, replicates, new_cap, hist, ledger)

    @staticmethod
    def _split(totals, reps, cap, hist, ledger) -> list[dict]:
        out = L.totals_list(totals, reps)
        if hist is not None:
            for r, d in enumerate(out):
                d["histogram"] = hist[r * (cap + 1):(r + 1) * (cap + 1)].copy()
                d["ledger"] = ledger[r * (cap + 1):(r + 1) * (cap + 1)].copy()
        return out

    def solve_async(self, sobol, power: int, cap: int = 64, precision="fp32-fixup", pa: bool = True,
                    variant: int = 0, coop: int | None = None, group: int | None = None,
                    stage1: int | None = None, packets: int | None = None) -> None:
        """Enqueue solve() -- or, given a list of scrambles, solve_many() -- on
        the scene's stream and return at once; collect with solve_wait().
        Launches on different scenes overlap on the device."""
        flags = (L.FLAG_PA if pa else 0) | (int(variant) << L.VARIANT_SHIFT) | \
            coop_flags(coop, group, stage1, packets)
        if isinstance(sobol, (list, tuple, ct.Array)):
            table = sobol_table(sobol)
            reps = len(table)
            L.call("ehss_solve_many_async", self.handle, reps, table, int(power), int(cap),
                   precision_code(precision), flags)
        else:
            reps = 0
            L.call("ehss_solve_async", self.handle, ct.byref(sobol), int(power), int(cap),
                   precision_code(precision), flags)
        self._pending = (int(cap), reps)

    def solve_done(self) -> bool:
        flag = ct.c_int(0)
        L.call("ehss_solve_query", self.handle, ct.byref(flag))
        return bool(flag.value)

    def solve_wait(self, histogram: bool = False, as_array: bool = False):
        """Result of the pending launch: a dict for solve_async(sobol), a list of
        dicts for solve_async([sobols]) (as_array: one structured array)."""
        if as_array and histogram:
            raise ValueError("as_array returns totals only")
        cap, reps = getattr(self, "_pending", (0, 0))
        n = max(reps, 1)
        totals = (L.Totals * n)()
        hist = np.zeros(n * (cap + 1), np.int64) if histogram else None
        ledger = np.zeros(n * (cap + 1), np.float64) if histogram else None
        L.call("ehss_solve_wait", self.handle, totals, L.ptr(hist), L.ptr(ledger))
        if as_array:
            return np.frombuffer(totals, dtype=L.TOTALS_DTYPE, count=n).copy()
        out = self._split(totals, n, cap, hist, ledger)
        return out if reps else out[0]

    def last_rays(self, n: int) -> dict:
        flags = np.empty(n, np.uint32)
        contrib = np.empty(n, np.float64)
        bounces = np.empty(n, np.int32)
        L.call("ehss_scene_last_rays", self.handle, n, L.ptr(flags), L.ptr(contrib), L.ptr(bounces))
        return dict(flags=flags, contributions=contrib, bounces=bounces.astype(np.int64),
                    escaped=(flags & L.RAY_ESCAPED) != 0, unresolved=(flags & L.RAY_UNRESOLVED) != 0,
                    pa_hit=(flags & L.RAY_PA_HIT) != 0, ehss_hit=(flags & L.RAY_EHSS_HIT) != 0,
                    grazing=(flags & L.RAY_GRAZING) != 0)

    def last_ray_times(self, n: int) -> np.ndarray:
        """Completion time of every ray of the last solve run with
        timestamps=True, in microseconds after the first ray finished."""
        ticks = np.empty(n, np.int64)
        hz = ct.c_double(0.0)
        L.call("ehss_scene_last_ray_times", self.handle, n, L.ptr(ticks), ct.byref(hz))
        return (ticks - ticks.min()) * (1e6 / hz.value)

    def source_rays(self, sobol: L.Sobol, power: int) -> dict:
        n = 1 << power
        u, o, d = np.empty((n, 4)), np.empty((n, 3)), np.empty((n, 3))
        L.call("ehss_source_rays", self.handle, ct.byref(sobol), int(power), L.ptr(u), L.ptr(o), L.ptr(d))
        return dict(u=u, origins=o, directions=d)

    # -- alternative incidence sources, generated and traced on the device --------
    def solve_source(self, source: L.Source, sobol: L.Sobol5, power: int, cap: int = 64,
                     histogram: bool = False, keep_rays: bool = False) -> dict:
        """One device solve of 2^power states of an alternative source
        (ehss_solve_source); see source_params()."""
        totals = L.Totals()
        n = 1 << int(power)
        hist = np.zeros(cap + 1, np.int64) if histogram else None
        info = np.empty(n, np.uint32) if keep_rays else None
        contrib = np.empty(n, np.float64) if keep_rays else None
        L.call("ehss_solve_source", self.handle, ct.byref(source), ct.byref(sobol), int(power), int(cap),
               ct.byref(totals), L.ptr(hist), L.ptr(info), L.ptr(contrib))
        out = totals.as_dict()
        if histogram:
            out["histogram"] = hist
        if keep_rays:  # per ray, in ray order
            flags = info >> np.uint32(20)
            out["rays_out"] = dict(bounces=(info & np.uint32(0xFFFFF)).astype(np.int64),
                                   escaped=(flags & L.RAY_ESCAPED) != 0,
                                   unresolved=(flags & L.RAY_UNRESOLVED) != 0,
                                   grazing=(flags & L.RAY_GRAZING) != 0, contributions=contrib)
        return out

    def source_states(self, source: L.Source, sobol: L.Sobol5, power: int) -> dict:
        """The states the device generates for an alternative source (ehss_source_states)."""
        n = 1 << int(power)
        o, d, inc = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3))
        w, own, cnt = np.empty(n), np.empty(n, np.int64), np.empty(n, np.int64)
        blocked = np.empty(n, np.uint8)
        L.call("ehss_source_states", self.handle, ct.byref(source), ct.byref(sobol), int(power), L.ptr(o),
               L.ptr(d), L.ptr(inc), L.ptr(w), L.ptr(own), L.ptr(cnt), L.ptr(blocked))
        return dict(origins=o, directions=d, incoming=inc, weights=w, owners=own, bounces=cnt,
                    blocked=blocked.astype(bool))


def scramble5(seed: int, dim: int = 5) -> L.Sobol5:
    """ehss_sobol5 of scipy.stats.qmc.Sobol(dim, scramble=True, seed) for dim 4 or 5
    (dimensions past `dim` are zero)."""
    if dim not in (4, 5):
        raise ValueError("dim must be 4 or 5")
    sv, shift = scramble_matrices(seed, dim)
    s = L.Sobol5()
    for k in range(dim):
        for j in range(BITS):
            s.sv[k][j] = int(sv[k, j])
        s.shift[k] = int(shift[k])
    return s


def source_params(kind: str, center=None, rotation=None, axes=None, area: float = 0.0, cdf=None) -> L.Source:
    """ehss_source for kind 'box' or 'ellipsoid' (a CompactBoundary's center,
    rotation, axes) or 'reverse-pa' (Spheres.area and the searchsorted table
    cumsum(r^2) / sum(r^2), inside_out_pa.py:79).  The table array must stay
    alive while the source is used."""
    kinds = {"box": L.SOURCE_BOX, "ellipsoid": L.SOURCE_ELLIPSOID, "reverse-pa": L.SOURCE_REVERSE_PA}
    if kind not in kinds:
        raise ValueError("source kind must be box, ellipsoid or reverse-pa")
    s = L.Source()
    s.kind = kinds[kind]
    if kind in ("box", "ellipsoid"):
        c = np.asarray(center, np.float64).reshape(3)
        R = np.ascontiguousarray(np.asarray(rotation, np.float64).reshape(3, 3))
        a = np.asarray(axes, np.float64).reshape(3)
        for k in range(3):
            s.center[k] = float(c[k])
            s.axes[k] = float(a[k])
        for k in range(9):
            s.rotation[k] = float(R.reshape(9)[k])
    else:
        table = np.ascontiguousarray(cdf, np.float64)
        s.area = float(area)
        s.cdf = table.ctypes.data
        s._keep = table  # noqa: SLF001 -- holds the table for the call
    return s


class _MultiSolve:
    """Shared wrapper of the multi-structure entry points (ehss_<prefix>_*):
    launch() is asynchronous, done() polls, wait() collects per-job results."""

    _prefix = ""

    def _create(self, scenes, sobols, powers, caps, flags: int, precision, *extra):
        n = len(scenes)
        if n == 0 or not (len(sobols) == len(powers) == len(caps) == n):
            raise ValueError("one scene, scramble, power and cap per job required")
        self.scenes = list(scenes)  # keep the scenes alive while the handle exists
        self.caps = [int(c) for c in caps]
        self.powers = [int(p) for p in powers]
        handles = (ct.c_void_p * n)(*[s.handle for s in self.scenes])
        sob = (L.Sobol * n)(*sobols)
        pw = np.asarray(self.powers, np.int32)
        cp = np.asarray(self.caps, np.int32)
        h = ct.c_void_p()
        L.call(f"ehss_{self._prefix}_create", n, handles, sob, L.ptr(pw), L.ptr(cp), precision_code(precision),
               flags, *extra, ct.byref(h))
        self._h = h
        self.n_jobs = n

    @property
    def handle(self):
        if not getattr(self, "_h", None):
            raise RuntimeError(f"{self._prefix} is closed")
        return self._h

    def set_sobol(self, job: int, sobol: L.Sobol) -> None:
        L.call(f"ehss_{self._prefix}_set_sobol", self.handle, int(job), ct.byref(sobol))

    def retarget(self, scenes, sobols, powers, caps) -> None:
        """Point the plan at another job list (ehss_plan_retarget): its buffers
        and stream are reused; the previous scenes are released.  An empty list
        only releases them."""
        n = len(scenes)
        if not (len(sobols) == len(powers) == len(caps) == n):
            raise ValueError("one scene, scramble, power and cap per job required")
        handles = (ct.c_void_p * max(n, 1))(*[s.handle for s in scenes])
        sob = sobol_table(sobols) if n else (L.Sobol * 1)()
        pw = np.asarray(powers, np.int32)
        cp = np.asarray(caps, np.int32)
        L.call(f"ehss_{self._prefix}_retarget", self.handle, n, handles, sob, L.ptr(pw), L.ptr(cp))
        self.scenes = list(scenes)
        self.caps = [int(c) for c in caps]
        self.powers = [int(p) for p in powers]
        self.n_jobs = n

    def release(self) -> None:
        """Let the scenes go (rebuild and destroy are refused while held)."""
        self.retarget([], [], [], [])

    def launch(self) -> None:
        L.call(f"ehss_{self._prefix}_launch", self.handle)

    def done(self) -> bool:
        flag = ct.c_int(0)
        L.call(f"ehss_{self._prefix}_query", self.handle, ct.byref(flag))
        return bool(flag.value)

    def wait(self, histogram: bool = False) -> list[dict]:
        totals = (L.Totals * self.n_jobs)()
        bins = sum(c + 1 for c in self.caps)
        hist = np.zeros(bins, np.int64) if histogram else None
        ledger = np.zeros(bins, np.float64) if histogram else None
        L.call(f"ehss_{self._prefix}_wait", self.handle, totals, L.ptr(hist), L.ptr(ledger))
        out, off = [], 0
        for j in range(self.n_jobs):
            d = totals[j].as_dict()
            if histogram:
                d["histogram"] = hist[off:off + self.caps[j] + 1].copy()
                d["ledger"] = ledger[off:off + self.caps[j] + 1].copy()
            off += self.caps[j] + 1
            out.append(d)
        return out

    def run(self, histogram: bool = False) -> list[dict]:
        self.launch()
        return self.wait(histogram)

    def close(self):
        if getattr(self, "_h", None):
            L.fn(f"ehss_{self._prefix}_destroy")(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GPUPlan(_MultiSolve):
    """Many fused solves -- different scenes, scrambles, ray counts and caps -- as
    one prepared launch sequence (ehss_plan_*): a packet grid over (ray block,
    job), one drain queue shared by every job, per-job fixed-order reductions.
    Each job's totals are bit-identical to the same solve run alone
    (GPUScene.solve with the same variant/coop/group/stage1 settings).

    With graph=True the sequence is captured once into a hipGraph and replayed
    by later launches; set_sobol() changes a job's scramble between launches
    (graph or not)."""

    _prefix = "plan"

    def __init__(self, scenes, sobols, powers, caps, precision="fp32-fixup", pa: bool = True, graph: bool = False,
                 variant: int = 0, coop: int | None = None, group: int | None = None, stage1: int | None = None,
                 lds: bool = False, packets: int | None = None):
        flags = (L.FLAG_PA if pa else 0) | (L.FLAG_GRAPH if graph else 0) | (int(variant) << L.VARIANT_SHIFT)
        flags |= coop_flags(coop, group, stage1, packets) | (L.FLAG_LDS_PACKETS if lds else 0)
        self._create(scenes, sobols, powers, caps, flags, precision)
