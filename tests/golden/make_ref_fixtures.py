"""Writes the solver/estimator fixtures from the compiled reference.

Runs oracle/_ref/sedi_ref (intel/sedi's unmodified solver headers driven by
oracle/ref_driver.cc; build it with `make -C oracle ref`) and stores what it
prints and writes as .npz files next to this script:

  ref_32cube.npz    32^3, seed 1234, 8 samples, tol 1e-9, maxiter 500, -p 1x1x1,
                    solvers 1/2/3: iterations, codes, final errors, D, and the
                    per-iteration err lines; plus the fixed-iteration mode
                    (-m 20 -i) D for each solver.
  ref_odd.npz       13x7x5, seed 5, 3 samples, tol 1e-8, maxiter 200 (odd rows,
                    odd planes), solvers 1/2/3.
  ref_32cube_p8.npz 32^3 CG on -p 2x2x2 under mpirun -n 8 (the domain-decomposed
                    path; differs from -p 1x1x1 only by summation order).
  ref_256cube.npz   (--big) 256^3, seed 7, 16 samples, tol 1e-9, maxiter 2000,
                    solvers 1/2/3 on -p 2x2x2 (mpirun -n 8): iterations, codes,
                    final errors, and D's summary (sum, sum of squares, min, max)
                    with D at 4096 strided cells.

Usage: python tests/golden/make_ref_fixtures.py [--big]
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle_lib  # noqa: E402

REF = oracle_lib.REF_BIN
MPIRUN = "/opt/conda/bin/mpirun"
NAMES = {1: "BiCgStab", 2: "Cg", 3: "PipelinedCg"}


def run_ref(grid, procs, samples, tol, maxiter, solver, seed, ignore=False, nranks=1):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.bin")
        cmd = [REF, "-g", "x".join(map(str, grid)), "-p", "x".join(map(str, procs)),
               "-s", str(samples), "-e", repr(tol), "-m", str(maxiter), "-S", str(solver),
               "-r", str(seed), "-o", out]
        if ignore:
            cmd.append("-i")
        if nranks > 1:
            cmd = [MPIRUN, "-n", str(nranks)] + cmd
        log = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
        res = oracle_lib.read_ref_output(out)
    hist = []
    for chunk in log.split("Solving for sample")[1:]:
        hist.append([float(m) for m in re.findall(NAMES[solver] + r": after iter=\d+, err=([0-9.eE+-]+)", chunk)])
    res["hist"] = hist
    res["log_tail"] = [l for l in log.splitlines() if l.startswith("Total solve time")]
    return res


def pack_hist(hist):
    lens = np.array([len(h) for h in hist], dtype=np.int32)
    flat = np.array([v for h in hist for v in h], dtype=np.float64)
    return lens, flat


def small():
    d = {}
    for s in (1, 2, 3):
        r = run_ref((32, 32, 32), (1, 1, 1), 8, 1e-9, 500, s, 1234)
        lens, flat = pack_hist(r["hist"])
        d.update({f"s{s}_iters": r["iters"], f"s{s}_codes": r["codes"], f"s{s}_errs": r["errs"],
                  f"s{s}_D": r["D"], f"s{s}_hist_len": lens, f"s{s}_hist": flat})
        f = run_ref((32, 32, 32), (1, 1, 1), 8, 1e-9, 20, s, 1234, ignore=True)
        d.update({f"s{s}_fixed20_iters": f["iters"], f"s{s}_fixed20_codes": f["codes"],
                  f"s{s}_fixed20_errs": f["errs"], f"s{s}_fixed20_D": f["D"]})
    np.savez_compressed(os.path.join(HERE, "ref_32cube.npz"), **d)

    d = {}
    for s in (1, 2, 3):
        r = run_ref((13, 7, 5), (1, 1, 1), 3, 1e-8, 200, s, 5)
        lens, flat = pack_hist(r["hist"])
        d.update({f"s{s}_iters": r["iters"], f"s{s}_codes": r["codes"], f"s{s}_errs": r["errs"],
                  f"s{s}_D": r["D"], f"s{s}_hist_len": lens, f"s{s}_hist": flat})
    np.savez_compressed(os.path.join(HERE, "ref_odd.npz"), **d)

    r = run_ref((32, 32, 32), (2, 2, 2), 8, 1e-9, 500, 2, 1234, nranks=8)
    np.savez_compressed(os.path.join(HERE, "ref_32cube_p8.npz"), s2_iters=r["iters"],
                        s2_codes=r["codes"], s2_errs=r["errs"], s2_D=r["D"])


def big():
    d = {}
    n = 256 ** 3
    idx = np.arange(0, n, n // 4096, dtype=np.int64)[:4096]
    for s in (1, 2, 3):
        r = run_ref((256, 256, 256), (2, 2, 2), 16, 1e-9, 2000, s, 7, ignore=True, nranks=8)
        D = r["D"]
        d.update({f"s{s}_iters": r["iters"], f"s{s}_codes": r["codes"], f"s{s}_errs": r["errs"],
                  f"s{s}_Dsum": np.array(D.sum()), f"s{s}_Dsq": np.array((D * D).sum()),
                  f"s{s}_Dmin": np.array(D.min()), f"s{s}_Dmax": np.array(D.max()),
                  f"s{s}_Dsample": D[idx], f"s{s}_time": np.array(r["log_tail"], dtype=object).astype(str)})
        print("solver", s, "iters", r["iters"], r["log_tail"], flush=True)
    d["sample_idx"] = idx
    np.savez_compressed(os.path.join(HERE, "ref_256cube.npz"), **d)


if __name__ == "__main__":
    if "--big" in sys.argv:
        big()
    else:
        small()
