"""sedi_amd -- MI355X (gfx950) stochastic diag(A^-1) estimator, a drop-in for
intel/sedi's solver path.

Python mirror of the reference's operator/solver interface over the C ABI of
libsedi_amd.so (include/sedi_c.h):

    Stencil(a, b, c).apply(x)           Stencil<double>::apply  (Stencil.hpp:21)
    Cg / PipelinedCg / BiCgStab(grid)   solver constructors     (Cg.hpp:48, ...)
      .solve(x, b, iM, maxiter, tol)    solver_t::solve         (Cg.hpp:70-82)
        -> (result_code, final_err, final_iter)
    rademacher_fill(seed, first, n)     DistrRngStd::fill_plusminusone
    estimate(...)                       stochastic-estimator.cc:158-216
    Batch(grid, K, precision)           multi-RHS CG (no reference counterpart)
    probe_batch(seed, n, k0, count)     the estimator's probes k0.., by jump-ahead

Arrays are numpy float64, n1*n2*n3 cells with i1 fastest (reshape a
(n3, n2, n1) C-ordered array to 1-D).  Everything runs in the HIP library; there
is no CPU fallback: if libsedi_amd.so is missing or no GPU is present, the
calls raise.
"""
import ctypes
import os

import numpy as np

__all__ = ["SOLUTION_FOUND", "REACHED_MAXITER", "BREAKDOWN_RHO", "BREAKDOWN_OMEGA",
           "BICGSTAB", "CG", "PCG", "FP64", "FP32_MIXED", "SediError", "lib", "lib_path",
           "Stencil", "Cg", "PipelinedCg", "BiCgStab", "Solver", "rademacher_fill",
           "Estimator", "estimate", "Batch", "probe_batch", "Server", "h5_available", "h5_read", "h5_write",
           "HOOK_X", "HOOK_V", "HOOK_D", "DeviceArray", "graphs_enable", "batch_compact_enable", "profile_enable", "profile_get",
           "profile_reset"]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# SEDI_AMD_LIB selects another build of the library (e.g. a tuning variant)
lib_path = os.environ.get("SEDI_AMD_LIB") or os.path.join(PKG_DIR, "libsedi_amd.so")

SOLUTION_FOUND, REACHED_MAXITER, BREAKDOWN_RHO, BREAKDOWN_OMEGA = 0, 1, 2, 3
BICGSTAB, CG, PCG = 1, 2, 3
FP64, FP32_MIXED = 0, 1
SOLVER_NAMES = {BICGSTAB: "BiCgStab", CG: "Cg", PCG: "PipelinedCg"}


class SediError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sedi error %d: %s" % (code, msg))
        self.code = code


class _Grid(ctypes.Structure):
    _fields_ = [("n1", ctypes.c_int64), ("n2", ctypes.c_int64), ("n3", ctypes.c_int64)]


class _Stencil(ctypes.Structure):
    _fields_ = [("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double)]


class _EstArgs(ctypes.Structure):
    _fields_ = [("grid", _Grid), ("stencil", _Stencil), ("kind", ctypes.c_int),
                ("mode", ctypes.c_int), ("samples", ctypes.c_int), ("maxiter", ctypes.c_int),
                ("tol", ctypes.c_double), ("ignore_errors", ctypes.c_int),
                ("seed", ctypes.c_uint32), ("precision", ctypes.c_int), ("batch_k", ctypes.c_int),
                ("rep_every", ctypes.c_int)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_double), ("precision", ctypes.c_int), ("rep_every", ctypes.c_int),
                ("chunk", ctypes.c_int), ("max_slots", ctypes.c_int), ("max_iter", ctypes.c_int)]


class _Job(ctypes.Structure):
    _fields_ = [("grid", _Grid), ("stencil", _Stencil), ("kind", ctypes.c_int), ("samples", ctypes.c_int),
                ("tol", ctypes.c_double), ("maxiter", ctypes.c_int), ("seed", ctypes.c_uint32)]


class _ServerInfo(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("done", ctypes.c_int64), ("probes", ctypes.c_int64),
                ("bytes_in_use", ctypes.c_double), ("peak_bytes", ctypes.c_double), ("pools", ctypes.c_int)]


LOG_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_char_p)
_DP = ctypes.POINTER(ctypes.c_double)
SAMPLE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, _DP, _DP, _DP)
HOOK_X, HOOK_V, HOOK_D = 1, 2, 4

_lib = None


def lib():
    """Loads libsedi_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(lib_path):
            raise SediError(-5, "%s not built; run `python -m sedi_amd.build`" % lib_path)
        L = ctypes.CDLL(lib_path)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.sedi_last_error.restype = ctypes.c_char_p
        L.sedi_init.argtypes = [ctypes.c_int]
        L.sedi_stencil_apply.argtypes = [ctypes.POINTER(_Grid), ctypes.POINTER(_Stencil), vp, vp]
        L.sedi_solver_create.argtypes = [ctypes.POINTER(_Grid), ctypes.POINTER(_Stencil),
                                         ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
        L.sedi_solve.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_double, dp, ip]
        L.sedi_solver_err_history.argtypes = [vp, dp, ctypes.c_int]
        L.sedi_solver_destroy.argtypes = [vp]
        L.sedi_solver_destroy.restype = None
        L.sedi_rademacher_fill.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, vp]
        L.sedi_estimator_create.argtypes = [ctypes.POINTER(_EstArgs), ctypes.POINTER(vp)]
        L.sedi_estimator_run.argtypes = [vp, vp, ip, dp, ip, dp]
        L.sedi_estimator_set_log.argtypes = [vp, LOG_FN, vp]
        L.sedi_estimator_destroy.argtypes = [vp]
        L.sedi_estimator_destroy.restype = None
        L.sedi_estimate.argtypes = [ctypes.POINTER(_EstArgs), vp, ip, dp, ip, dp]
        L.sedi_batch_create.argtypes = [ctypes.POINTER(_Grid), ctypes.POINTER(_Stencil), ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
        L.sedi_batch_solve.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_double, ip,
                                       dp, ip]
        L.sedi_batch_err_history.argtypes = [vp, ctypes.c_int, dp, ctypes.c_int]
        L.sedi_batch_destroy.argtypes = [vp]
        L.sedi_batch_destroy.restype = None
        L.sedi_batch_lane_moves.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.sedi_batch_compact_enable.argtypes = [ctypes.c_int]
        L.sedi_probe_batch.argtypes = [ctypes.c_uint32, ctypes.c_int64, ctypes.c_int, ctypes.c_int, vp]
        L.sedi_estimator_set_rhs.argtypes = [vp, vp]
        L.sedi_estimator_set_sample_hook.argtypes = [vp, SAMPLE_FN, ctypes.c_int, vp]
        L.sedi_h5_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_Grid), vp]
        L.sedi_h5_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                    ctypes.POINTER(vp), ctypes.POINTER(_Grid)]
        L.sedi_graphs_enable.argtypes = [ctypes.c_int]
        L.sedi_server_create.argtypes = [ctypes.POINTER(_ServerConfig), ctypes.POINTER(vp)]
        L.sedi_server_submit.argtypes = [vp, ctypes.POINTER(_Job)]
        L.sedi_server_submit.restype = ctypes.c_int64
        L.sedi_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_double, vp, ip, dp, ip, dp, ip]
        L.sedi_server_stats.argtypes = [vp, ctypes.POINTER(_ServerInfo)]
        L.sedi_server_destroy.argtypes = [vp]
        L.sedi_server_destroy.restype = None
        L.sedi_profile_enable.argtypes = [ctypes.c_char_p]
        L.sedi_profile_get.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), dp]
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise SediError(rc, lib().sedi_last_error().decode())
    return rc


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if n is not None and a.size != n:
        raise ValueError("expected %d cells, got %d" % (n, a.size))
    return a


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _grid(grid):
    n1, n2, n3 = (int(v) for v in grid)
    return _Grid(n1, n2, n3), n1 * n2 * n3


_hip = None


def _hiprt():
    """The HIP runtime libsedi_amd.so is linked to (loaded with it): device
    memory for DeviceArray comes from the same runtime the kernels run on, so
    the library sees it as device memory whatever else the process loaded."""
    global _hip
    if _hip is None:
        lib()
        H = ctypes.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        H.hipFree.argtypes = [ctypes.c_void_p]
        H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        H.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip = H
    return _hip


class DeviceArray:
    """A float64 array in device memory (hipMalloc), for the entry points that
    take device pointers: keeps inputs and outputs resident on the GPU.
    `ptr` is the device address; get() copies it back."""

    def __init__(self, a=None, n=None):
        H = _hiprt()
        host = None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        self.n = host.size if host is not None else int(n)
        self.ptr = ctypes.c_void_p()
        if H.hipMalloc(ctypes.byref(self.ptr), max(8, 8 * self.n)) != 0:
            raise SediError(-3, "hipMalloc of %d doubles failed" % self.n)
        if host is not None and H.hipMemcpy(self.ptr, host.ctypes.data, 8 * self.n, 1) != 0:
            raise SediError(-2, "hipMemcpy to the device failed")

    def get(self):
        out = np.empty(self.n)
        if _hiprt().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2) != 0:
            raise SediError(-2, "hipMemcpy from the device failed")
        return out

    def fill_bytes(self, value=0):
        """hipMemset of the whole array (synchronous)."""
        if _hiprt().hipMemset(self.ptr, int(value), 8 * self.n) != 0:
            raise SediError(-2, "hipMemset failed")

    def free(self):
        if self.ptr:
            _hiprt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stencil:
    """Stencil<double>(a, b, c) (Stencil.hpp:18-19)."""

    def __init__(self, a=6.0, b=1.0, c=1.0):
        self.a, self.b, self.c = float(a), float(b), float(c)

    def _c(self):
        return _Stencil(self.a, self.b, self.c)

    def apply(self, grid, x, y=None):
        """y = A x on the whole grid with zero boundary (Stencil.hpp:21-54)."""
        g, n = _grid(grid)
        x = _f64(x, n)
        if y is None:
            y = np.empty(n)
        _check(lib().sedi_stencil_apply(ctypes.byref(g), ctypes.byref(self._c()), _ptr(x), _ptr(y)))
        return y


class Solver:
    """Base of Cg / PipelinedCg / BiCgStab: allocates device temporaries once."""
    kind = None

    def __init__(self, grid, stencil=None, mode=0):
        self.grid = tuple(int(v) for v in grid)
        self.stencil = stencil or Stencil()
        g, self.n = _grid(self.grid)
        h = ctypes.c_void_p()
        _check(lib().sedi_solver_create(ctypes.byref(g), ctypes.byref(self.stencil._c()),
                                        self.kind, int(mode), ctypes.byref(h)))
        self._h = h

    def solve(self, x, b, iM=None, maxiter=1000, tol=1e-4):
        """Solves A x = b from the guess in x (updated in place).

        Returns (result_code, final_err, final_iter) like solver_t::solve.
        """
        if not isinstance(x, np.ndarray) or x.dtype != np.float64 or not x.flags.c_contiguous:
            raise TypeError("x must be a C-contiguous float64 numpy array (updated in place)")
        if x.size != self.n:
            raise ValueError("x has %d cells, grid has %d" % (x.size, self.n))
        b = _f64(b, self.n)
        m = None if iM is None else _f64(iM, self.n)
        fe = ctypes.c_double()
        fi = ctypes.c_int()
        rc = lib().sedi_solve(self._h, _ptr(b), _ptr(x), None if m is None else _ptr(m),
                              int(maxiter), float(tol), ctypes.byref(fe), ctypes.byref(fi))
        _check(rc)
        return rc, fe.value, fi.value

    def err_history(self):
        n = _check(lib().sedi_solver_err_history(self._h, None, 0))
        out = np.zeros(n)
        lib().sedi_solver_err_history(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().sedi_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Cg(Solver):
    kind = CG


class PipelinedCg(Solver):
    kind = PCG


class BiCgStab(Solver):
    kind = BICGSTAB


SOLVERS = {CG: Cg, PCG: PipelinedCg, BICGSTAB: BiCgStab}


def rademacher_fill(seed, first_draw, ncells):
    """+-1 probe cells from std::mt19937(seed) after first_draw outputs."""
    v = np.empty(int(ncells))
    _check(lib().sedi_rademacher_fill(int(seed), int(first_draw), int(ncells), _ptr(v)))
    return v


def probe_batch(seed, ncells, k0, count):
    """Probes k0 .. k0+count-1 of the estimator's stream, shape (count, ncells)."""
    v = np.empty((int(count), int(ncells)))
    _check(lib().sedi_probe_batch(int(seed) & 0xffffffff, int(ncells), int(k0), int(count), _ptr(v)))
    return v


class Batch:
    """Up to K right-hand sides solved together by CG, pipelined CG or
    BiCGStab (kind; the recurrences of Cg.hpp / PipelinedCg.hpp / BiCgStab.hpp
    per right-hand side).

    precision FP64 keeps the plain solver's arithmetic; FP32_MIXED stores the
    vectors in FP32 with FP64 reductions and an FP64 residual replacement
    every rep_every iterations (0 = 50) and at every tenfold decrease.
    """

    def __init__(self, grid, K, precision=FP64, rep_every=0, stencil=None, kind=CG):
        self.grid = tuple(int(v) for v in grid)
        self.stencil = stencil or Stencil()
        g, self.n = _grid(self.grid)
        self.K = (int(K) + 3) // 4 * 4
        h = ctypes.c_void_p()
        _check(lib().sedi_batch_create(ctypes.byref(g), ctypes.byref(self.stencil._c()), int(kind),
                                       int(K), int(precision), int(rep_every), ctypes.byref(h)))
        self._h = h

    def solve(self, b, maxiter=1000, tol=1e-4):
        """b: (nrhs, n).  Returns (x, iters, errs, codes), x of shape (nrhs, n)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.ndim == 1:
            b = b.reshape(1, -1)
        nrhs = b.shape[0]
        if b.shape[1] != self.n:
            raise ValueError("expected %d cells per right-hand side, got %d" % (self.n, b.shape[1]))
        x = np.empty_like(b)
        iters = np.zeros(nrhs, dtype=np.int32)
        codes = np.full(nrhs, -1, dtype=np.int32)
        errs = np.zeros(nrhs)
        ip = ctypes.POINTER(ctypes.c_int)
        _check(lib().sedi_batch_solve(self._h, _ptr(b), nrhs, _ptr(x), int(maxiter), float(tol),
                                      iters.ctypes.data_as(ip),
                                      errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      codes.ctypes.data_as(ip)))
        return x, iters, errs, codes

    @property
    def lane_moves(self):
        """Lanes that lane compaction moved in the last solve."""
        m = ctypes.c_int64()
        _check(lib().sedi_batch_lane_moves(self._h, ctypes.byref(m)))
        return int(m.value)

    def err_history(self, k):
        n = _check(lib().sedi_batch_err_history(self._h, int(k), None, 0))
        out = np.zeros(n)
        lib().sedi_batch_err_history(self._h, int(k), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().sedi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Estimator:
    """The sample loop of stochastic-estimator.cc:158-216 (seeded probes)."""

    def __init__(self, grid, kind=CG, samples=8, maxiter=1000, tol=1e-4, seed=1234,
                 ignore_errors=False, stencil=None, mode=0, precision=FP64, batch_k=1,
                 rep_every=0):
        st = stencil or Stencil()
        g, self.n = _grid(grid)
        self.samples = int(samples)
        self.args = _EstArgs(g, st._c(), int(kind), int(mode), self.samples, int(maxiter),
                             float(tol), int(bool(ignore_errors)), int(seed) & 0xffffffff,
                             int(precision), int(batch_k), int(rep_every))
        h = ctypes.c_void_p()
        _check(lib().sedi_estimator_create(ctypes.byref(self.args), ctypes.byref(h)))
        self._h = h
        self._log = None

    def set_log(self, fn):
        """fn(line: str) receives the reference's stdout lines."""
        if fn is None:
            self._log = None
            lib().sedi_estimator_set_log(self._h, LOG_FN(), None)
            return
        self._log = LOG_FN(lambda user, line: fn(line.decode()))
        _check(lib().sedi_estimator_set_log(self._h, self._log, None))

    def set_rhs(self, v):
        """Validation mode: every sample solves with right-hand side v (None: probes)."""
        if v is None:
            _check(lib().sedi_estimator_set_rhs(self._h, None))
            return
        v = _f64(v, self.n)
        _check(lib().sedi_estimator_set_rhs(self._h, _ptr(v)))

    def set_sample_hook(self, fn, want=HOOK_X | HOOK_V | HOOK_D):
        """fn(k, x, v, D) after each sample, numpy copies (None when not requested)."""
        if fn is None:
            self._hook = None
            _check(lib().sedi_estimator_set_sample_hook(self._h, SAMPLE_FN(), 0, None))
            return
        n = self.n

        def cb(user, k, x, v, D):
            cp = lambda p: None if not p else np.ctypeslib.as_array(p, shape=(n,)).copy()
            fn(k, cp(x), cp(v), cp(D))
        self._hook = SAMPLE_FN(cb)
        _check(lib().sedi_estimator_set_sample_hook(self._h, self._hook, int(want), None))

    def run(self, want_D=True):
        """Returns dict(status, D, iters, errs, codes, seconds)."""
        D = np.zeros(self.n) if want_D else None
        iters = np.zeros(self.samples, dtype=np.int32)
        codes = np.full(self.samples, -1, dtype=np.int32)
        errs = np.zeros(self.samples)
        secs = ctypes.c_double()
        ip = ctypes.POINTER(ctypes.c_int)
        rc = lib().sedi_estimator_run(self._h, None if D is None else _ptr(D),
                                      iters.ctypes.data_as(ip),
                                      errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      codes.ctypes.data_as(ip), ctypes.byref(secs))
        _check(rc)
        return dict(status=rc, D=D, iters=iters, errs=errs, codes=codes, seconds=secs.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().sedi_estimator_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def estimate(grid, kind=CG, samples=8, maxiter=1000, tol=1e-4, seed=1234, ignore_errors=False,
             **kw):
    est = Estimator(grid, kind, samples, maxiter, tol, seed, ignore_errors, **kw)
    try:
        return est.run()
    finally:
        est.close()


class Server:
    """Continuous-batching estimation server (sedi_server_*): submit jobs at
    any time, wait for their results; see include/sedi_c.h."""

    def __init__(self, hbm_budget=200e9, precision=FP32_MIXED, rep_every=0, chunk=0, max_slots=0,
                 max_iter=0):
        cfg = _ServerConfig(float(hbm_budget), int(precision), int(rep_every), int(chunk), int(max_slots),
                            int(max_iter))
        h = ctypes.c_void_p()
        _check(lib().sedi_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._jobs = {}

    def submit(self, grid, kind=CG, samples=8, tol=1e-8, maxiter=2000, seed=0, stencil=None):
        """Returns the job id."""
        g, n = _grid(grid)
        st = stencil or Stencil()
        job = _Job(g, st._c(), int(kind), int(samples), float(tol), int(maxiter), int(seed) & 0xffffffff)
        jid = lib().sedi_server_submit(self._h, ctypes.byref(job))
        _check(jid if jid < 0 else 0)
        self._jobs[jid] = (n, int(samples))
        return jid

    def wait(self, jid, timeout=-1.0):
        """dict(status, D, iters, errs, codes, times) or None on timeout."""
        n, s = self._jobs[jid]
        D = np.zeros(n)
        iters = np.zeros(s, dtype=np.int32)
        codes = np.full(s, -1, dtype=np.int32)
        errs = np.zeros(s)
        times = np.zeros(4)
        status = ctypes.c_int()
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        rc = _check(lib().sedi_server_wait(self._h, int(jid), float(timeout), _ptr(D), iters.ctypes.data_as(ip),
                                           errs.ctypes.data_as(dp), codes.ctypes.data_as(ip),
                                           times.ctypes.data_as(dp), ctypes.byref(status)))
        if rc == 1:
            return None
        del self._jobs[jid]
        return dict(status=status.value, D=D, iters=iters, errs=errs, codes=codes,
                    times=dict(submit=times[0], admit=times[1], start=times[2], done=times[3]))

    def stats(self):
        info = _ServerInfo()
        _check(lib().sedi_server_stats(self._h, ctypes.byref(info)))
        return {k: getattr(info, k) for k, _ in _ServerInfo._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().sedi_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def h5_available():
    """True if libhdf5 can be loaded (validation archives)."""
    return lib().sedi_h5_available() == 1


def h5_read(path, dataset, grid):
    """A [n1][n2][n3] (i3 fastest) FP64 dataset as an i1-fastest vector."""
    g, n = _grid(grid)
    out = np.empty(n)
    _check(lib().sedi_h5_read(str(path).encode(), dataset.encode(), ctypes.byref(g), _ptr(out)))
    return out


def h5_write(path, grid, **datasets):
    """Writes i1-fastest vectors as the reference's archive datasets."""
    g, n = _grid(grid)
    names = list(datasets)
    arrs = [_f64(datasets[k], n) for k in names]
    cn = (ctypes.c_char_p * len(names))(*[k.encode() for k in names])
    ca = (ctypes.c_void_p * len(names))(*[a.ctypes.data for a in arrs])
    _check(lib().sedi_h5_write(str(path).encode(), len(names), cn, ca, ctypes.byref(g)))


def graphs_enable(on=True):
    """Capture chunks of solver iterations into HIP graphs (default) or not."""
    _check(lib().sedi_graphs_enable(1 if on else 0))


def batch_compact_enable(on=True):
    """Lane compaction of batch solves (default) or not; results are the same."""
    _check(lib().sedi_batch_compact_enable(1 if on else 0))


def profile_enable(kernel_name):
    """Times every launch of `kernel_name` ("*" = all, None = off) with HIP events."""
    _check(lib().sedi_profile_enable(None if kernel_name is None else kernel_name.encode()))


def profile_get(kernel_name):
    """(launches, total_ms) of a timed kernel."""
    n = ctypes.c_int64()
    ms = ctypes.c_double()
    _check(lib().sedi_profile_get(kernel_name.encode(), ctypes.byref(n), ctypes.byref(ms)))
    return n.value, ms.value


def profile_reset():
    _check(lib().sedi_profile_reset())


def tuning_lookup(key, grid, fallback=0):
    """The autotuner's value for `key` on `grid` (see sedi_tuning_lookup)."""
    g, _ = _grid(grid)
    L = lib()
    L.sedi_tuning_lookup.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Grid), ctypes.c_int]
    return L.sedi_tuning_lookup(key.encode(), ctypes.byref(g), int(fallback))
