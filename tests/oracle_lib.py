"""ctypes binding of the oracle (oracle/libsedi_oracle.so) for the tests.

The oracle is the CPU restatement of the reference used as the checker; it is
test infrastructure and never part of the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "libsedi_oracle.so")
REF_BIN = os.path.join(ORACLE_DIR, "_ref", "sedi_ref")

BICGSTAB, CG, PCG = 1, 2, 3

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "liboracle"])
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        L.oracle_mt_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.oracle_mt_next.argtypes = [ctypes.c_void_p]
        L.oracle_mt_next.restype = ctypes.c_uint32
        L.oracle_uniform01.argtypes = [ctypes.c_void_p]
        L.oracle_uniform01.restype = ctypes.c_double
        L.oracle_rademacher_fill.argtypes = [ctypes.c_uint32, ctypes.c_uint64, i64, dp]
        L.oracle_stencil_apply.argtypes = [i64, i64, i64, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, dp, dp]
        L.oracle_dots.argtypes = [i64, i64, i64, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, dp]
        L.oracle_solve.argtypes = [ctypes.c_int, i64, i64, i64, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, dp, dp, dp,
                                   ctypes.c_int, ctypes.c_double, dp,
                                   ctypes.POINTER(ctypes.c_int), dp]
        L.oracle_solve.restype = ctypes.c_int
        L.oracle_estimate.argtypes = [ctypes.c_int, i64, i64, i64, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_uint32,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int), dp,
                                      ctypes.POINTER(ctypes.c_int)]
        L.oracle_estimate.restype = ctypes.c_int
        L.oracle_exact_diag_inv.argtypes = [i64, i64, i64, ctypes.c_double, ctypes.c_double,
                                            i64, ctypes.POINTER(ctypes.c_int64), dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class MT19937:
    """The oracle's mt19937 stream."""

    def __init__(self, seed):
        self._buf = ctypes.create_string_buffer(624 * 4 + 16)
        lib().oracle_mt_seed(self._buf, seed)

    def next(self):
        return lib().oracle_mt_next(self._buf)

    def uniform01(self):
        return lib().oracle_uniform01(self._buf)


def rademacher(seed, first_draw, ncells):
    v = np.empty(ncells, dtype=np.float64)
    lib().oracle_rademacher_fill(seed, first_draw, ncells, _dp(v))
    return v


def stencil_apply(shape, x, a=6.0, b=1.0, c=1.0):
    n1, n2, n3 = shape
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    y = np.empty_like(x)
    lib().oracle_stencil_apply(n1, n2, n3, a, b, c, _dp(x), _dp(y))
    return y


def dots(shape, xs, ys):
    n1, n2, n3 = shape
    xs = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1) for v in xs]
    ys = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1) for v in ys]
    xp = (ctypes.c_void_p * len(xs))(*[v.ctypes.data for v in xs])
    yp = (ctypes.c_void_p * len(ys))(*[v.ctypes.data for v in ys])
    out = np.zeros(len(xs))
    lib().oracle_dots(n1, n2, n3, len(xs), xp, yp, _dp(out))
    return out


def solve(kind, shape, b, x0=None, iM=None, maxiter=500, tol=1e-9, abc=(6.0, 1.0, 1.0)):
    """Returns (code, x, final_err, final_iter, err_history)."""
    n1, n2, n3 = shape
    n = n1 * n2 * n3
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64).reshape(-1)
    hist = np.zeros(max(maxiter, 1))
    fe = ctypes.c_double(0)
    fi = ctypes.c_int(0)
    iMp = None if iM is None else _dp(np.ascontiguousarray(iM, dtype=np.float64))
    code = lib().oracle_solve(kind, n1, n2, n3, abc[0], abc[1], abc[2], _dp(b), _dp(x), iMp,
                              maxiter, tol, ctypes.byref(fe), ctypes.byref(fi), _dp(hist))
    return code, x, fe.value, fi.value, hist[: fi.value]


def estimate(kind, shape, seed, samples, maxiter, tol, ignore_errors=False, abc=(6.0, 1.0, 1.0)):
    """Returns (status, D, iters, errs, codes)."""
    n1, n2, n3 = shape
    D = np.zeros(n1 * n2 * n3)
    iters = np.zeros(samples, dtype=np.int32)
    codes = np.full(samples, -1, dtype=np.int32)
    errs = np.zeros(samples)
    ip = ctypes.POINTER(ctypes.c_int)
    st = lib().oracle_estimate(kind, n1, n2, n3, abc[0], abc[1], abc[2], seed, samples,
                               maxiter, tol, int(ignore_errors), _dp(D),
                               iters.ctypes.data_as(ip), _dp(errs), codes.ctypes.data_as(ip))
    return st, D, iters, errs, codes


def exact_diag_inv(shape, cells, a=6.0, bc=1.0):
    n1, n2, n3 = shape
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    out = np.zeros(len(cells))
    lib().oracle_exact_diag_inv(n1, n2, n3, a, bc, len(cells),
                                cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _dp(out))
    return out


def read_ref_output(path):
    """Reads the -o file written by oracle/_ref/sedi_ref."""
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(40), dtype=np.int64)
        n1, n2, n3, done, solver = (int(v) for v in hdr)
        iters = np.frombuffer(f.read(4 * done), dtype=np.int32).copy()
        codes = np.frombuffer(f.read(4 * done), dtype=np.int32).copy()
        errs = np.frombuffer(f.read(8 * done), dtype=np.float64).copy()
        D = np.frombuffer(f.read(8 * n1 * n2 * n3), dtype=np.float64).copy()
    return dict(shape=(n1, n2, n3), solver=solver, iters=iters, codes=codes, errs=errs, D=D)
