"""GPU: the validation modes through the CLI driver (-V, -E, -X) and the
estimator's validation hooks, against the oracle and the reference's own
stochastic-estimator built WITH_HDF5 (oracle/_ref).

Bars: -V on an archive made from the oracle's solution reports rel L2 / Linf
errors below 1e-10 for every sample with CG and pipelined CG (which follow
the reference's trajectory to rounding) and below 1e-7 with BiCGStab (whose
rounding differences grow along the iterations: the bar is tol 1e-10 times
the condition number ~440); -E on the oracle's D of the same
seeded run reports an L2 error below 1e-12 after the last sample; an archive
written by -X is read by the reference's -V with errors below 1e-8.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from conftest import normwise

sedi = pytest.importorskip("sedi_amd")

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(sedi.__file__), "bin", "sedi-estimator")
REF_MAIN = os.path.join(O.ORACLE_DIR, "_ref", "stochastic-estimator-S%d")
SHAPE = (12, 10, 8)
N = 12 * 10 * 8
G = "12x10x8"


def need_h5():
    if not sedi.h5_available():
        pytest.skip("libhdf5 not available")


def driver(args, cwd):
    out = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=120, cwd=cwd)
    return out.returncode, out.stdout


def oracle_archive(path, kind, seed=3, tol=1e-10):
    v = O.rademacher(seed, 0, N)
    code, x, fe, fi, hist = O.solve(kind, SHAPE, v, maxiter=500, tol=tol)
    assert code == 0
    sedi.h5_write(path, SHAPE, x_k=x, v_k=v)
    return x, v


@pytest.mark.parametrize("kind", [O.CG, O.PCG, O.BICGSTAB])
def test_driver_solver_validation(tmp_path, kind):
    need_h5()
    f = str(tmp_path / "s.h5")
    oracle_archive(f, kind)
    rc, out = driver(["-g", G, "-p", "1x1x1", "-s", "2", "-e", "1e-10", "-m", "500", "-S", str(kind),
                      "-V", f], str(tmp_path))
    assert rc == 0, out
    lines = out.splitlines()
    assert "VALIDATION: rank 0 read arrays from %s" % f in lines
    k = lines.index("Solving for sample 1/2...")
    assert lines[k + 1] == "VALIDATION: filling RHS from validation dataset"
    errs = re.findall(r"VALIDATION: rel L2 error=(\S+), rel Linf error=(\S+)", out)
    assert len(errs) == 2
    # CG / pipelined CG follow the reference's trajectory to rounding;
    # BiCGStab's rounding differences grow along its iterations, so its bar is
    # what the tolerance admits (tol x condition number ~ 4e-8 here)
    bar = 1e-7 if kind == O.BICGSTAB else 1e-10
    for l2, linf in errs:
        assert float(l2) < bar and float(linf) < bar
    # the same sequence of line kinds as the reference's own binary
    exe = REF_MAIN % kind
    if os.path.exists(exe):
        r = subprocess.run([exe, "-g", G, "-p", "1x1x1", "-s", "2", "-e", "1e-10", "-m", "500", "-V", f],
                           capture_output=True, text=True, timeout=120, cwd=str(tmp_path)).stdout
        kinds = lambda text: [re.sub(r"[-+0-9.e=]+", "#", l) for l in text.splitlines()
                              if l.startswith(("VALIDATION", "Solving"))]
        assert kinds(out) == kinds(r)


def test_driver_estimator_validation(tmp_path):
    need_h5()
    st, D, iters, errs, codes = O.estimate(O.CG, SHAPE, 9, 6, 500, 1e-10)
    f = str(tmp_path / "e.h5")
    sedi.h5_write(f, SHAPE, diag_inv=D)
    for extra in ([], ["-k", "4"]):
        rc, out = driver(["-g", G, "-p", "1x1x1", "-s", "6", "-e", "1e-10", "-m", "500", "-r", "9", "-E", f]
                         + extra, str(tmp_path))
        assert rc == 0, out
        lines = re.findall(r"VALIDATION_ESTIMATOR: L2 error=(\S+), Linf error=(\S+)", out)
        assert len(lines) == 6
        assert float(lines[-1][0]) < 1e-12 and float(lines[-1][1]) < 1e-11
        assert float(lines[0][0]) > 1e-3  # after one sample D is still far from it


def test_both_validation_modes_disable_each_other(tmp_path):
    need_h5()
    f = str(tmp_path / "s.h5")
    x, v = oracle_archive(f, O.CG)
    sedi.h5_write(str(tmp_path / "e.h5"), SHAPE, diag_inv=x)
    rc, out = driver(["-g", G, "-p", "1x1x1", "-s", "1", "-e", "1e-8", "-m", "500", "-V", f, "-E",
                      str(tmp_path / "e.h5")], str(tmp_path))
    assert rc == 0
    assert "WARNING: cannot enable both solver and estimator validation mode: disabling both!" in out
    assert "VALIDATION: rel L2" not in out and "VALIDATION_ESTIMATOR: L2" not in out


@pytest.mark.parametrize("extra", [[], ["-P", "mixed", "-k", "4"]])
def test_archive_written_here_is_read_by_the_reference(tmp_path, extra):
    need_h5()
    exe = REF_MAIN % O.CG
    f = str(tmp_path / "x.h5")
    rc, out = driver(["-g", G, "-p", "1x1x1", "-s", "3", "-e", "1e-11", "-m", "500", "-r", "4", "-X", f]
                     + extra, str(tmp_path))
    assert rc == 0, out
    v = sedi.h5_read(f, "v_k", SHAPE)
    assert np.array_equal(v, O.rademacher(4, 2 * 2 * N, N))  # the last (third) probe
    st, D, iters, errs, codes = O.estimate(O.CG, SHAPE, 4, 3, 500, 1e-11)
    assert normwise(sedi.h5_read(f, "diag_inv", SHAPE), D) < 1e-8
    if not os.path.exists(exe):
        pytest.skip("reference stochastic-estimator (HDF5) not built")
    r = subprocess.run([exe, "-g", G, "-p", "1x1x1", "-s", "1", "-e", "1e-11", "-m", "500", "-V", f],
                       capture_output=True, text=True, timeout=120, cwd=str(tmp_path)).stdout
    l2, linf = re.findall(r"VALIDATION: rel L2 error=(\S+), rel Linf error=(\S+)", r)[0]
    assert float(l2) < 1e-8 and float(linf) < 1e-8


@pytest.mark.parametrize("batch_k,prec", [(1, sedi.FP64), (4, sedi.FP64), (4, sedi.FP32_MIXED)])
def test_estimator_rhs_and_sample_hook(batch_k, prec):
    v = O.rademacher(12, 0, N)
    code, ox, ofe, ofi, oh = O.solve(O.CG, SHAPE, v, maxiter=500, tol=1e-10)
    est = sedi.Estimator(SHAPE, kind=sedi.CG, samples=3, maxiter=500, tol=1e-10, seed=1,
                         batch_k=batch_k, precision=prec)
    est.set_rhs(v)
    seen = []
    est.set_sample_hook(lambda k, x, vv, D: seen.append((k, x, vv, D)))
    r = est.run()
    est.close()
    assert [s[0] for s in seen] == [0, 1, 2]
    for k, x, vv, D in seen:
        assert np.array_equal(vv, v)
        assert normwise(x, ox) < (1e-9 if prec == sedi.FP64 else 1e-6)
        # q counts samples exactly and every sample has the same x and v
        assert normwise(D, x * v) < 1e-12
    assert normwise(r["D"], seen[-1][3]) == 0.0
