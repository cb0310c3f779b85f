"""GPU: the drop-in from the reference's side.

oracle/_ref/stochastic-estimator-hip-S{1,2,3} is the reference's own
stochastic-estimator.cc, unmodified, compiled with include/SolverHip.hpp
taking the place of its Cg / PipelinedCg / BiCgStab and DistrRngStd
(`-include SolverHip.hpp -DSEDI_HIP_AS_REFERENCE_SOLVERS`, oracle/Makefile
target ref_hip): its sample loop, accumulation, output lines and exit codes
run as written while every solve and probe runs on the GPU through
libsedi_amd.so.  Run at the reference configuration (-g 32x32x32 -p 1x1x1
-s 8 -e 1e-9 -m 500, probes from mt19937 seed 1234 via SEDI_RNG_SEED), its
printed per-iteration lines and summary line are checked against the
reference's own run on the same probes (tests/golden/ref_32cube.npz) with the
parity bars; a run that cannot converge exits 100 + code as the reference's
does (stochastic-estimator.cc:181-184).
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, iteration_bar

pytestmark = pytest.mark.gpu

NAMES = {1: "BiCgStab", 2: "Cg", 3: "PipelinedCg"}


def binary(kind):
    path = os.path.join(ROOT, "oracle", "_ref", "stochastic-estimator-hip-S%d" % kind)
    if not os.path.exists(path):
        pytest.skip("%s not built (make -C oracle ref_hip)" % path)
    return path


def run(kind, *args, seed=1234):
    env = dict(os.environ, SEDI_RNG_SEED=str(seed))
    cmd = [binary(kind)] + list(args)
    print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=240, cwd=ROOT)


def parse(out, kind):
    samples = out.split("Solving for sample")[1:]
    hist = [[float(v) for v in re.findall(NAMES[kind] + r": after iter=\d+, err=(\S+)", s)] for s in samples]
    m = re.search(r"Total solve time: \S+ seconds, iteration count min=(\d+) max=(\d+)", out)
    return hist, (int(m.group(1)), int(m.group(2))) if m else None


@pytest.mark.parametrize("kind", [2, 3, 1])
def test_reference_main_over_the_adapter(ref32, ref_spread, kind):
    r = run(kind, "-g", "32x32x32", "-p", "1x1x1", "-s", "8", "-e", "1e-9", "-m", "500")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Initialized %s solver" % NAMES[kind] in r.stdout
    hist, summary = parse(r.stdout, kind)
    assert len(hist) == 8 and summary is not None
    iters = np.array([len(h) for h in hist])
    ref_it = ref32[f"s{kind}_iters"]
    print("iterations", iters.tolist(), "reference", ref_it.tolist(), "summary", summary, flush=True)
    # the summary line is the loop's own min/max of final_iter
    assert summary == (iters.min(), iters.max())
    assert np.all(np.abs(iters - ref_it) <= iteration_bar(ref_spread, "c32", kind))
    # the err lines: the first iterations agree with the reference's to the
    # printed 10 digits' rounding
    lens, flat = ref32[f"s{kind}_hist_len"], ref32[f"s{kind}_hist"]
    off = np.concatenate([[0], np.cumsum(lens)])
    for k in range(8):
        ref_h = flat[off[k]:off[k + 1]]
        m = min(15, len(ref_h))
        assert np.max(np.abs(np.array(hist[k][:m]) - ref_h[:m]) / ref_h[:m]) < 1e-9
    # each printed err is a solver's final_err at its stopping iteration
    for h in hist:
        assert h[-1] <= 1e-9
    if kind != 1:
        assert np.array_equal(iters, ref_it)
    assert "Resident set size across ranks" in r.stdout


def test_reference_main_exit_code_on_failed_solve():
    # maxiter too small without -i: MPI_Abort(100 + REACHED_MAXITER) after
    # the first sample (stochastic-estimator.cc:181-184)
    r = run(2, "-g", "16x16x16", "-p", "1x1x1", "-s", "4", "-e", "1e-12", "-m", "5")
    assert r.returncode == 101, (r.returncode, r.stderr[-1000:])
    assert "Incomplete convergence in solver, errcode=1" in r.stderr
    assert r.stdout.count("Solving for sample") == 1


def test_reference_main_ignore_errors_runs_every_sample():
    r = run(3, "-g", "16x16x16", "-p", "1x1x1", "-s", "3", "-e", "1e-12", "-m", "5", "-i")
    assert r.returncode == 0, r.stderr[-1000:]
    hist, summary = parse(r.stdout, 3)
    assert [len(h) for h in hist] == [5, 5, 5] and summary == (5, 5)
