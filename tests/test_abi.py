"""The C ABI library loads and exports every entry point include/sedi_c.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "sedi_c.h")
LIB = os.path.join(ROOT, "sedi_amd", "libsedi_amd.so")
CLI = os.path.join(ROOT, "sedi_amd", "bin", "sedi-estimator")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|char|double)\s*\*?\s*(sedi_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["sedi_stencil_apply", "sedi_solver_create", "sedi_solve", "sedi_solver_destroy",
                 "sedi_rademacher_fill", "sedi_estimate", "sedi_estimator_create",
                 "sedi_estimator_run", "sedi_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libsedi_amd.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_reports_errors_without_gpu_work():
    lib = ctypes.CDLL(LIB)
    lib.sedi_last_error.restype = ctypes.c_char_p
    lib.sedi_solver_create.argtypes = [ctypes.c_void_p] * 5
    # NULL arguments are rejected before any device call
    rc = lib.sedi_solver_create(None, None, 2, 0, None)
    assert rc == -1
    assert b"NULL" in lib.sedi_last_error()
    assert lib.sedi_version() == 1


class Grid(ctypes.Structure):
    _fields_ = [("n1", ctypes.c_int64), ("n2", ctypes.c_int64), ("n3", ctypes.c_int64)]


class Stencil(ctypes.Structure):
    _fields_ = [("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double)]


class EstimateArgs(ctypes.Structure):
    _fields_ = [("grid", Grid), ("stencil", Stencil), ("kind", ctypes.c_int), ("mode", ctypes.c_int),
                ("samples", ctypes.c_int), ("maxiter", ctypes.c_int), ("tol", ctypes.c_double),
                ("ignore_errors", ctypes.c_int), ("seed", ctypes.c_uint32), ("precision", ctypes.c_int),
                ("batch_k", ctypes.c_int), ("rep_every", ctypes.c_int)]


@pytest.mark.parametrize("dims", [(1024, 1024, 512), (1024, 1024, 1024), (2048, 2048, 128)])
def test_plain_path_rejects_vectors_of_4gib(dims):
    """The plain FP64 kernels use 32-bit byte offsets: a grid of 2^29 cells or
    more (an FP64 vector of 4 GiB) is refused before any allocation or device
    call, by the stencil, the solver and the plain estimator."""
    lib = ctypes.CDLL(LIB)
    lib.sedi_last_error.restype = ctypes.c_char_p
    g = Grid(*dims)
    st = Stencil(6.0, 1.0, 1.0)
    x = (ctypes.c_double * 1)()
    lib.sedi_stencil_apply.argtypes = [ctypes.c_void_p] * 4
    assert lib.sedi_stencil_apply(ctypes.byref(g), ctypes.byref(st), x, x) == -5
    assert b"4 GiB" in lib.sedi_last_error()
    out = ctypes.c_void_p()
    lib.sedi_solver_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p]
    assert lib.sedi_solver_create(ctypes.byref(g), ctypes.byref(st), 2, 0, ctypes.byref(out)) == -5
    assert not out.value
    args = EstimateArgs(g, st, 2, 0, 4, 100, 1e-9, 0, 1, 0, 1, 0)
    lib.sedi_estimator_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert lib.sedi_estimator_create(ctypes.byref(args), ctypes.byref(out)) == -5
    assert b"4 GiB" in lib.sedi_last_error()


def test_plain_path_size_limit_is_exact():
    """One cell below 2^29 passes the size check (and then needs a device)."""
    lib = ctypes.CDLL(LIB)
    lib.sedi_last_error.restype = ctypes.c_char_p
    lib.sedi_solver_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p]
    st = Stencil(6.0, 1.0, 1.0)
    out = ctypes.c_void_p()
    rc = lib.sedi_solver_create(ctypes.byref(Grid(1 << 29, 1, 1)), ctypes.byref(st), 2, 0, ctypes.byref(out))
    assert rc == -5
    rc = lib.sedi_solver_create(ctypes.byref(Grid((1 << 29) - 1, 1, 1)), ctypes.byref(st), 2, 0,
                                ctypes.byref(out))
    assert rc != -5 and b"4 GiB" not in (lib.sedi_last_error() or b"")
    if rc == 0:
        lib.sedi_solver_destroy.argtypes = [ctypes.c_void_p]
        lib.sedi_solver_destroy(out)


def test_cli_usage_errors():
    if not os.path.exists(CLI):
        pytest.fail("sedi-estimator is not built")
    # -g and -p are required (CmdLineParser.hpp:81-93): usage text, exit 1
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([CLI, "-g", "8x8x8"], capture_output=True, text=True)
    assert r.returncode == 1
    r = subprocess.run([CLI, "-g", "8x8", "-p", "1x1x1"], capture_output=True, text=True)
    assert r.returncode == 1


TUNING_PROBE = r"""
import sys
sys.path.insert(0, %r)
import sedi_amd
print(sedi_amd.tuning_lookup("zc:batch_upd_kernel", (256, 256, 256), -1),
      sedi_amd.tuning_lookup("zc:batch_upd_kernel", (128, 128, 128), -1),
      sedi_amd.tuning_lookup("zc:cg_update_kernel", (96, 96, 96), -1),
      sedi_amd.tuning_lookup("zc:pcg_iter_kernel", (256, 256, 256), -1),
      sedi_amd.tuning_lookup("k:2:1", (200, 200, 200), -1),
      sedi_amd.tuning_lookup("k:2:1", (40, 40, 40), -1),
      sedi_amd.tuning_lookup("k:3:1", (256, 256, 256), -1),
      sedi_amd.tuning_lookup("sched:cg_update_kernel", (96, 96, 96), -1))
"""


def test_tuning_table_lookup(tmp_path):
    """Exact grid first, then an entry for any grid (0 0 0), then the
    wildcard kernel; batch sizes also from the nearest grid by cell count."""
    table = tmp_path / "tuning.txt"
    table.write_text("# comment line\n"
                     "zc:batch_upd_kernel 256 256 256 32   # trailing comment\n"
                     "zc:batch_upd_kernel 0 0 0 8\n"
                     "zc:* 96 96 96 16\n"
                     "k:2:1 64 64 64 16\n"
                     "k:2:1 256 256 256 64\n"
                     "sched:* 0 0 0 1\n"
                     "malformed line\n")
    env = dict(os.environ, SEDI_TUNING_FILE=str(table))
    r = subprocess.run([sys.executable, "-c", TUNING_PROBE % ROOT], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["32", "8", "16", "-1", "64", "16", "-1", "1"]
    env["SEDI_TUNING_FILE"] = ""  # no table: every lookup falls back
    r = subprocess.run([sys.executable, "-c", TUNING_PROBE % ROOT], capture_output=True, text=True, env=env)
    assert r.stdout.split() == ["-1"] * 8


def test_noalloc_solve_refuses_what_its_array_cannot_hold():
    """sedi_solve_noalloc checks the working set against its fixed array
    before any device call: 256^3 CG needs 3 x 128 MiB > 64 MiB."""
    lib = ctypes.CDLL(LIB)
    lib.sedi_last_error.restype = ctypes.c_char_p
    lib.sedi_solve_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_void_p, ctypes.c_void_p]
    st = Stencil(6.0, 1.0, 1.0)
    b = (ctypes.c_double * 1)()
    fe, fi = ctypes.c_double(), ctypes.c_int()
    for dims, kind in [((256, 256, 256), 2), ((128, 128, 128), 3), ((128, 128, 128), 1)]:
        rc = lib.sedi_solve_noalloc(ctypes.byref(Grid(*dims)), ctypes.byref(st), kind, b, b, None, 100, 1e-9,
                                    ctypes.byref(fe), ctypes.byref(fi))
        assert rc == -1 and b"allocation-free limit" in lib.sedi_last_error(), (dims, kind)
    # maxiter counts too (8 bytes of err history per iteration)
    rc = lib.sedi_solve_noalloc(ctypes.byref(Grid(16, 16, 16)), ctypes.byref(st), 2, b, b, None, 10 ** 7, 1e-9,
                                ctypes.byref(fe), ctypes.byref(fi))
    assert rc == -1 and b"allocation-free limit" in lib.sedi_last_error()
