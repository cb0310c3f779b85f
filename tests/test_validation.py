"""Validation archives (HDF5) on the CPU: the library's archive writer and
reader against the reference's own ValidateSolver / ValidateEstimator, run
in the reference's stochastic-estimator built WITH_HDF5 (oracle/_ref).

The reference reads datasets [n1][n2][n3] with i3 fastest (Slice3d.hpp:42-66,
144-160); an archive written here from the oracle's solution must make the
reference report a zero error (the oracle reproduces the reference bitwise),
and the same archive with the axes left untransposed must not.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O

sedi = pytest.importorskip("sedi_amd")

REF_MAIN = os.path.join(O.ORACLE_DIR, "_ref", "stochastic-estimator-S%d")
SHAPE = (12, 10, 8)
N = 12 * 10 * 8


def need_h5():
    if not os.path.exists(sedi.lib_path):
        pytest.skip("libsedi_amd.so not built")
    if not sedi.h5_available():
        pytest.skip("libhdf5 not available")


def run_ref_main(kind, args, cwd):
    exe = REF_MAIN % kind
    if not os.path.exists(exe):
        pytest.skip("reference stochastic-estimator (HDF5) not built")
    out = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120, cwd=cwd)
    return out.returncode, out.stdout


def test_h5_roundtrip(tmp_path):
    need_h5()
    x = np.random.default_rng(1).standard_normal(N)
    f = str(tmp_path / "a.h5")
    sedi.h5_write(f, SHAPE, x_k=x, v_k=2 * x, diag_inv=-x)
    for name, want in (("x_k", x), ("v_k", 2 * x), ("diag_inv", -x)):
        assert np.array_equal(sedi.h5_read(f, name, SHAPE), want)
    with pytest.raises(sedi.SediError):
        sedi.h5_read(f, "x_k", (10, 12, 8))
    with pytest.raises(sedi.SediError):
        sedi.h5_read(f, "missing", SHAPE)


@pytest.mark.parametrize("kind", [O.CG, O.PCG, O.BICGSTAB])
def test_reference_validates_oracle_archive(tmp_path, kind):
    need_h5()
    v = O.rademacher(3, 0, N)
    code, x, fe, fi, hist = O.solve(kind, SHAPE, v, maxiter=500, tol=1e-10)
    assert code == 0
    f = str(tmp_path / "solver.h5")
    sedi.h5_write(f, SHAPE, x_k=x, v_k=v)
    rc, out = run_ref_main(kind, ["-g", "12x10x8", "-p", "1x1x1", "-s", "2", "-e", "1e-10", "-m", "500",
                                  "-V", f], str(tmp_path))
    assert rc == 0, out
    assert "VALIDATION: rank 0 read arrays from %s" % f in out
    errs = re.findall(r"VALIDATION: rel L2 error=(\S+), rel Linf error=(\S+)", out)
    assert len(errs) == 2
    for l2, linf in errs:
        assert float(l2) == 0.0 and float(linf) == 0.0
    # the same buffers stored untransposed (i1 fastest in the file) are
    # another problem: the reference's solution no longer matches
    untransposed = lambda a: a.reshape(12, 10, 8).transpose(2, 1, 0).reshape(-1)
    wrong = str(tmp_path / "wrong.h5")
    sedi.h5_write(wrong, SHAPE, x_k=untransposed(x), v_k=untransposed(v))
    rc, out = run_ref_main(kind, ["-g", "12x10x8", "-p", "1x1x1", "-s", "1", "-e", "1e-10", "-m", "500",
                                  "-V", wrong], str(tmp_path))
    l2 = float(re.findall(r"rel L2 error=(\S+),", out)[0])
    assert l2 > 1e-3


def test_reference_estimator_archive_orientation(tmp_path):
    """diag_inv = the exact diagonal of A^-1 (closed form): the reference reads
    it and prints one line per sample; its running D after 64 random probes
    is within sampling error of it."""
    need_h5()
    cells = np.arange(N)
    exact = O.exact_diag_inv(SHAPE, cells)
    f = str(tmp_path / "est.h5")
    sedi.h5_write(f, SHAPE, diag_inv=exact)
    rc, out = run_ref_main(O.CG, ["-g", "12x10x8", "-p", "1x1x1", "-s", "64", "-e", "1e-10", "-m", "500",
                                  "-E", f], str(tmp_path))
    assert rc == 0, out
    lines = re.findall(r"VALIDATION_ESTIMATOR: L2 error=(\S+), Linf error=(\S+)", out)
    assert len(lines) == 64
    # sampling error ~ ||offdiag row||_2 / sqrt(s), about 0.03 here
    assert float(lines[-1][0]) < 0.06
