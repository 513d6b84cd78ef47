"""Allocation-free residual norms (cfdg_residual_norms_noalloc) against
cfdg_residual_norms, the entry point whose kernels it shares.

The residual is evaluated by a context, copied out in mesh order and handed
back to the new entry point in caller-owned device memory (hipMalloc of
the HIP runtime the library itself links, on a stream of the caller's),
padded to a leading dimension whose tail holds NaN, which must never be
read.  The two entry points add the same values in different cell orders
(the context sums in its device order), so the tolerance is the round-off of
a sum of squares: 1e-13 relative.  Two calls in a row, the second on a much
smaller input, check that nothing from the first call's scratch contents
leaks into the second.
"""
import ctypes as C

import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from helpers import perturbed_state

pytestmark = pytest.mark.gpu

RTOL = 1e-13


def _context_norms(case):
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    s.set_state(perturbed_state(s, case, seed=11))
    s.eval_residual()
    R, _ = s.residual()
    l2 = s.residual_norms()
    s.close()
    return R, l2


def _library_hip_runtime():
    """Path of the HIP runtime that libcfdg.so runs on.  Another copy can be
    mapped into the process as well (torch ships its own, loaded by an
    earlier test that imports torch); device pointers of one runtime are not
    valid in the other, so the test must allocate through the library's."""
    import os

    rocm = os.path.realpath("/opt/rocm/lib/libamdhip64.so.7")
    mapped = set()
    with open("/proc/self/maps") as fh:
        for line in fh:
            path = line.split()[-1]
            if "libamdhip64.so" in os.path.basename(path):
                mapped.add(os.path.realpath(path))
    if rocm in mapped:
        return rocm
    assert len(mapped) == 1, f"cannot tell which HIP runtime libcfdg.so uses: {sorted(mapped)}"
    return mapped.pop()


class _Hip:
    """The HIP runtime already loaded by libcfdg.so (the same mapped file)."""

    def __init__(self):
        self.h = C.CDLL(_library_hip_runtime())
        for n in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipStreamCreate",
                  "hipStreamSynchronize", "hipStreamDestroy"):
            getattr(self.h, n).restype = C.c_int
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def ok(self, rc, what):
        assert rc == 0, f"{what} failed: hipError {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))), "hipMalloc")
        if a.nbytes:
            self.ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, n):
        out = np.empty(n)
        self.ok(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out


def _noalloc(R, pad):
    hip = _Hip()
    n = R.shape[1]
    ld = n + pad
    host = np.full((5, ld), np.nan)
    host[:, :n] = R
    buf = hip.upload(host)
    out = hip.upload(np.full(5, np.nan))
    stream = C.c_void_p()
    hip.ok(hip.h.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    rc = cfdg.lib.cfdg_residual_norms_noalloc(buf, ld, n, out, stream)
    hip.ok(hip.h.hipStreamSynchronize(stream), "hipStreamSynchronize")
    res = hip.download(out, 5)
    hip.ok(hip.h.hipStreamDestroy(stream), "hipStreamDestroy")
    hip.h.hipFree(buf)
    hip.h.hipFree(out)
    return rc, res


def _check(R, l2_ctx, l2_na, what):
    assert np.all(np.isfinite(l2_na)), f"{what}: non-finite norms {l2_na}"
    err = np.abs(l2_na - l2_ctx) / np.maximum(np.abs(l2_ctx), 1e-300)
    assert err.max() <= RTOL, f"{what}: vs cfdg_residual_norms rel err {err.max():.3e}"
    # independent restatement: sqrt(sum R^2 / n) with numpy's pairwise sum
    n = R.shape[1]
    ref = np.array([np.sqrt(np.sum(R[v].astype(np.float64) ** 2) / n) for v in range(5)])
    err2 = np.abs(l2_na - ref) / np.maximum(np.abs(ref), 1e-300)
    assert err2.max() <= RTOL, f"{what}: vs numpy rel err {err2.max():.3e}"


def test_noalloc_norms_match_context_two_calls():
    # first call: 294,912 cells -> 1024 reduction blocks with grid-stride loops
    big = cases.plate(ni=96, nj=64, nk=48)
    R1, l1 = _context_norms(big)
    # second call: 3,072 cells -> 12 reduction blocks (a smaller input)
    small = cases.ramp()
    R2, l2 = _context_norms(small)
    assert R1.shape[1] > 262144 and R2.shape[1] < R1.shape[1]

    print(f"noalloc call 1: {R1.shape[1]} cells", flush=True)
    rc, n1 = _noalloc(R1, pad=13)
    assert rc == cfdg.OK
    _check(R1, l1, n1, "plate 96x64x48")
    print(f"noalloc call 2: {R2.shape[1]} cells", flush=True)
    rc, n2 = _noalloc(R2, pad=0)
    assert rc == cfdg.OK
    _check(R2, l2, n2, "ramp 48x16x4")
    # the same input again gives the same bits (fixed reduction order)
    rc, n2b = _noalloc(R2, pad=5)
    assert rc == cfdg.OK and np.array_equal(n2, n2b)


def test_noalloc_norms_edge_cases_and_rejections():
    rc, z = _noalloc(np.zeros((5, 0)), pad=3)
    assert rc == cfdg.OK and np.array_equal(z, np.zeros(5))
    rc, one = _noalloc(np.full((5, 1), -2.5), pad=0)
    assert rc == cfdg.OK and np.array_equal(one, np.full(5, 2.5))

    hip = _Hip()
    p = hip.upload(np.zeros(8))
    o = hip.upload(np.full(5, 7.0))
    f = cfdg.lib.cfdg_residual_norms_noalloc
    big = cfdg.NOALLOC_MAX_CELLS + 1
    assert f(p, big, big, o, None) == cfdg.ERR_CAPACITY
    assert f(None, 1, 1, o, None) == cfdg.ERR_ARG
    assert f(p, 1, 1, None, None) == cfdg.ERR_ARG
    assert f(p, 1, -1, o, None) == cfdg.ERR_ARG
    assert f(p, 1, 2, o, None) == cfdg.ERR_ARG          # ld < n_cells
    assert f(p, 1 << 31, 1, o, None) == cfdg.ERR_ARG    # ld beyond int32
    hip.ok(hip.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
    # rejected calls launch nothing: the output is untouched
    assert np.array_equal(hip.download(o, 5), np.full(5, 7.0))
    hip.h.hipFree(p)
    hip.h.hipFree(o)
