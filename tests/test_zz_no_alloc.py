"""GPU: the allocation-free solve sedi_solve_noalloc against sedi_solve.

sedi_solve_noalloc runs the plain solvers' kernels with every temporary in a
fixed device array of the library (no hipMalloc); on the same device arrays
and grid it must return what sedi_solver_create + sedi_solve return: the same
result code, final iteration and final error, and the same x bitwise (same
kernels, same launch plan, same reduction order).  Each case runs two calls
in a row, the second on a smaller grid, so stale contents of the shared array
from the first call must not leak into the second.  Device arrays come from
hipMalloc of the HIP runtime the library itself is linked to (plumbing
only), so they are device memory of that runtime whatever else the process
has loaded."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

sedi = pytest.importorskip("sedi_amd")

pytestmark = pytest.mark.gpu

vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)


def _lib():
    L = sedi.lib()
    L.sedi_solve_noalloc.argtypes = [ctypes.POINTER(sedi._Grid), ctypes.POINTER(sedi._Stencil), ctypes.c_int,
                                     vp, vp, vp, ctypes.c_int, ctypes.c_double, dp, ip]
    return L


_hip = None


def _hiprt():
    """The HIP runtime libsedi_amd is linked to (already loaded with it)."""
    global _hip
    if _hip is None:
        _lib()
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevArray:
    """A device copy of a float64 array (hipMalloc + hipMemcpy)."""

    def __init__(self, a):
        self.host = np.ascontiguousarray(a, dtype=np.float64)
        H = _hiprt()
        self.ptr = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.ptr), max(8, self.host.nbytes)) == 0
        assert H.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.host)
        assert _hiprt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def any(self):
        return bool(np.any(self.get()))

    def free(self):
        if self.ptr:
            _hiprt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _dev(a):
    return DevArray(a)


def solve_existing(kind, shape, abc, b, x0, iM, maxiter, tol):
    """sedi_solver_create + sedi_solve on device arrays."""
    L = _lib()
    g = sedi._Grid(*shape)
    st = sedi._Stencil(*abc)
    h = ctypes.c_void_p()
    assert L.sedi_solver_create(ctypes.byref(g), ctypes.byref(st), kind, 0, ctypes.byref(h)) == 0
    bd, xd = _dev(b), _dev(x0)
    md = _dev(iM) if iM is not None else None
    fe, fi = ctypes.c_double(-1.0), ctypes.c_int(-1)
    code = L.sedi_solve(h, bd.ptr, xd.ptr, md.ptr if md is not None else None,
                        maxiter, tol, ctypes.byref(fe), ctypes.byref(fi))
    assert _hiprt().hipDeviceSynchronize() == 0
    L.sedi_solver_destroy(h)
    x = xd.get()
    for a in (bd, xd, md):
        if a is not None:
            a.free()
    return code, x, fe.value, fi.value


def solve_noalloc(kind, shape, abc, b, x0, iM, maxiter, tol):
    L = _lib()
    g = sedi._Grid(*shape)
    st = sedi._Stencil(*abc)
    bd, xd = _dev(b), _dev(x0)
    md = _dev(iM) if iM is not None else None
    fe, fi = ctypes.c_double(-1.0), ctypes.c_int(-1)
    code = L.sedi_solve_noalloc(ctypes.byref(g), ctypes.byref(st), kind, bd.ptr, xd.ptr,
                                md.ptr if md is not None else None, maxiter, tol, ctypes.byref(fe),
                                ctypes.byref(fi))
    assert code >= 0, (code, L.sedi_last_error())
    assert _hiprt().hipDeviceSynchronize() == 0
    x = xd.get()
    for a in (bd, xd, md):
        if a is not None:
            a.free()
    return code, x, fe.value, fi.value


def check_same(kind, shape, abc=(6.0, 1.0, 1.0), seed=1, maxiter=500, tol=1e-9, x0=None, iM=None):
    n = int(np.prod(shape))
    b = O.rademacher(seed, 0, n)
    x0 = np.zeros(n) if x0 is None else x0
    print("noalloc kind", kind, "shape", shape, "seed", seed, flush=True)
    ref = solve_existing(kind, shape, abc, b, x0, iM, maxiter, tol)
    got = solve_noalloc(kind, shape, abc, b, x0, iM, maxiter, tol)
    assert got[0] == ref[0] and got[3] == ref[3], (got[0], got[3], ref[0], ref[3])
    assert got[2] == ref[2], (got[2], ref[2])
    assert np.array_equal(got[1], ref[1])
    return got


@pytest.mark.parametrize("kind", [O.CG, O.PCG, O.BICGSTAB])
def test_noalloc_equals_solve_two_calls(kind):
    # first call on a larger grid, then a smaller one reusing the same array
    c1 = check_same(kind, (48, 40, 36), seed=3)
    c2 = check_same(kind, (20, 12, 9), seed=4)
    assert c1[0] == 0 and c2[0] == 0


def test_noalloc_matches_the_oracle():
    # the bar of the plain parity tests, on the reference's 32^3 probe
    shape = (32, 32, 32)
    b = O.rademacher(1234, 0, 32768)
    code, x, fe, fi = solve_noalloc(O.CG, shape, (6.0, 1.0, 1.0), b, np.zeros(32768), None, 500, 1e-9)
    ocode, ox, ofe, ofi, _ = O.solve(O.CG, shape, b, maxiter=500, tol=1e-9)
    assert code == ocode == 0 and abs(fi - ofi) <= 1
    assert np.max(np.abs(x - ox)) / np.max(np.abs(ox)) < 1e-9


def test_noalloc_vector_preconditioner_guess_and_maxiter():
    shape = (20, 12, 9)
    n = int(np.prod(shape))
    rng = np.random.default_rng(3)
    x0 = rng.standard_normal(n) * 0.1
    iM = 1.0 / (6.0 + np.abs(rng.standard_normal(n)))
    for kind in (O.CG, O.PCG, O.BICGSTAB):
        got = check_same(kind, shape, seed=8, maxiter=12, tol=1e-12, x0=x0, iM=iM)
        assert got[0] == 1 and got[3] == 12
        # an odd final iteration leaves CG a pending update to flush
        got = check_same(kind, shape, seed=8, maxiter=7, tol=1e-12, x0=x0, iM=iM)
        assert got[3] == 7


def test_noalloc_nonsymmetric_bicgstab_and_zero_rhs():
    check_same(O.BICGSTAB, (16, 16, 8), abc=(6.0, 1.4, 0.6), seed=21, maxiter=25, tol=1e-14)
    L = _lib()
    g = sedi._Grid(8, 8, 8)
    st = sedi._Stencil(6.0, 1.0, 1.0)
    bd = DevArray(np.zeros(512))
    xd = DevArray(np.zeros(512))
    fe, fi = ctypes.c_double(-1.0), ctypes.c_int(-1)
    for kind in (O.CG, O.PCG, O.BICGSTAB):
        code = L.sedi_solve_noalloc(ctypes.byref(g), ctypes.byref(st), kind, bd.ptr, xd.ptr, None,
                                    10, 1e-9, ctypes.byref(fe), ctypes.byref(fi))
        assert (code, fi.value, fe.value) == (0, 0, 0.0)
    assert not xd.any()
    bd.free()
    xd.free()


def test_noalloc_rejects_host_arrays_and_oversize():
    L = _lib()
    g = sedi._Grid(8, 8, 8)
    st = sedi._Stencil(6.0, 1.0, 1.0)
    b = np.ones(512)
    x = np.zeros(512)
    fe, fi = ctypes.c_double(), ctypes.c_int()
    rc = L.sedi_solve_noalloc(ctypes.byref(g), ctypes.byref(st), O.CG, b.ctypes.data, x.ctypes.data, None, 10,
                              1e-9, ctypes.byref(fe), ctypes.byref(fi))
    assert rc == -1 and b"device arrays" in L.sedi_last_error()
    big = sedi._Grid(256, 256, 256)  # 5 vectors of 128 MiB: above the 64 MiB array
    rc = L.sedi_solve_noalloc(ctypes.byref(big), ctypes.byref(st), O.CG, b.ctypes.data, x.ctypes.data, None, 10,
                              1e-9, ctypes.byref(fe), ctypes.byref(fi))
    assert rc == -1 and b"allocation-free limit" in L.sedi_last_error()
