"""GPU: ehss_solve_noalloc, the allocation-free solve, called through the C-ABI.

Its totals, histogram and ledger must equal ehss_solve's on the same scene,
scramble and flags (bit for bit: same kernels, same fixed-order reduction), on
two calls in a row, the second with a smaller input; its limits are refused with
a clear error.  Output buffers are device memory (torch allocates them here, as a
caller would beforehand); the call itself runs on a caller stream.
"""
import ctypes as ct

import numpy as np
import pytest

from ehss_amd import _lib as L
from ehss_amd.engine import GPUScene, scramble
from tests import scenes
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _noalloc(sc, sob, power, cap, precision="fp32-fixup", flags=L.FLAG_PA):
    tot = torch.zeros(ct.sizeof(L.Totals), dtype=torch.uint8, device="cuda")
    hist = torch.full((cap + 1,), -1, dtype=torch.int64, device="cuda")
    led = torch.full((cap + 1,), -1.0, dtype=torch.float64, device="cuda")
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    L.call("ehss_solve_noalloc", sc.handle, ct.byref(sob), int(power), int(cap), L.PRECISIONS[precision], int(flags),
           ct.c_void_p(stream.cuda_stream), ct.c_void_p(tot.data_ptr()), ct.c_void_p(hist.data_ptr()),
           ct.c_void_p(led.data_ptr()))
    stream.synchronize()
    t = L.Totals.from_buffer_copy(tot.cpu().numpy().tobytes()).as_dict()
    t["histogram"] = hist.cpu().numpy()
    t["ledger"] = led.cpu().numpy()
    return t


def _same(a, b, what):
    for k in a:
        if k == "kernel_ms":
            continue
        if isinstance(a[k], np.ndarray):
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")
        else:
            assert a[k] == b[k], (what, k, a[k], b[k])


def test_noalloc_solve_equals_solve_twice_then_smaller():
    runs = [("globule650", scenes.globule(650, 1), 11, 16, 64, "fp32-fixup"),
            ("c60", scenes.c60(), 12, 12, 16, "fp32-fixup"),
            ("c60", scenes.c60(), 13, 10, 8, "fp64")]
    for name, (c, r), seed, power, cap, prec in runs:
        say("noalloc", name, seed, power, cap, prec)
        with GPUScene(c, r) as sc:
            sob = scramble(seed)
            got = _noalloc(sc, sob, power, cap, prec)
            ref = sc.solve(sob, power, cap, precision=prec, pa=True, histogram=True)
        assert got["rays"] == 1 << power and got["histogram"].sum() == 1 << power
        _same(ref, got, name)


def test_noalloc_refuses_its_limits():
    c, r = scenes.c60()
    with GPUScene(c, r) as sc:
        sob = scramble(1)
        for kw, msg in ((dict(power=17, cap=64), "power 17 out of range"),
                        (dict(power=10, cap=5000), "cap 5000 out of range"),
                        (dict(power=10, cap=64, flags=L.FLAG_KEEP_RAYS), "no per-ray records")):
            with pytest.raises(L.EHSSError, match=msg):
                _noalloc(sc, sob, **kw)


def test_noalloc_on_a_caller_stream_right_after_rebuild():
    """The caller's stream waits for the scene's last (re)build, which returns
    before its kernels have run."""
    c, r = scenes.globule(3000, 12)
    with GPUScene(c, r) as fresh:
        want = fresh.solve(scramble(4), 12, 64, pa=True, histogram=True)
    with GPUScene(*scenes.globule(150, 13)) as sc:
        tot = torch.zeros(ct.sizeof(L.Totals), dtype=torch.uint8, device="cuda")
        hist = torch.zeros(65, dtype=torch.int64, device="cuda")
        led = torch.zeros(65, dtype=torch.float64, device="cuda")
        stream = torch.cuda.Stream()
        torch.cuda.synchronize()
        say("noalloc after rebuild")
        sc.rebuild(c, r)
        L.call("ehss_solve_noalloc", sc.handle, ct.byref(scramble(4)), 12, 64, L.PRECISIONS["fp32-fixup"],
               int(L.FLAG_PA), ct.c_void_p(stream.cuda_stream), ct.c_void_p(tot.data_ptr()),
               ct.c_void_p(hist.data_ptr()), ct.c_void_p(led.data_ptr()))
        stream.synchronize()
        got = L.Totals.from_buffer_copy(tot.cpu().numpy().tobytes()).as_dict()
        got["histogram"] = hist.cpu().numpy()
    for k in ("omega", "pa", "escaped", "unresolved", "max_bounces"):
        assert got[k] == want[k], k
    np.testing.assert_array_equal(got["histogram"], want["histogram"])
