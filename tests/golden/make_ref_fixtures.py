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

  ref_spread.npz    (--spread) how far the reference's own results move when
                    only its summation order changes: the three solvers at 32^3
                    (seed 1234, 8 samples, tol 1e-9) and 13x7x5 (seed 5, 3
                    samples, tol 1e-8) under other decompositions (every
                    p1 x p2 x p3 on 2, 4 and 8 ranks, 32^3 only) and wider-SIMD builds of the same driver (4-lane AVX2, 8-lane
                    AVX-512 dot products; oracle/Makefile ref_variants):
                    per-variant iteration counts and D's normwise distance
                    from the -p 1x1x1 SSE2 run.
  ref_256cube_extra.npz (--big-extra) at 256^3, seed 7: BiCGStab in
                    fixed-iteration mode (-m 100 -i, 16 samples, -p 2x2x2),
                    with D's summary and 4096 strided cells; and the first 4
                    probes of each solver at -p 1x1x1 (SSE2, and AVX-512 for
                    BiCGStab): iterations and final errors, beside the -p 2x2x2
                    run of ref_256cube.npz.  (--big-fixed) adds BiCGStab
                    for -m 100 -i at -p 1x1x8 and -m 20 -i at both
                    decompositions.

  ref_breakdown.json (--breakdown) BiCGStab on small systems built so that
                    its breakdowns happen in exact arithmetic (oracle/_ref/
                    sedi_ref_cases: the reference's own BiCgStab on one rank):
                    omega = <t,s>/<t,t> = 0 (code 3) and rho = <r,r^> = 0 at
                    the second iteration (code 2), along each axis, and omega
                    = 0 with err <= tol (code 0: the outcome order of
                    BiCgStab.hpp:237-259); code, final iteration, final error,
                    the err lines and x.

  ref_spread_256cube.npz (--spread256) BiCGStab at 256^3 (seed 7, 16
                    samples, tol 1e-9) under seven other summation orders
                    (-p 1x1x8, 1x8x1, 8x1x1, 1x2x4, 4x2x1 with 2-lane SIMD;
                    -p 2x2x2 with 4- and 8-lane SIMD): counts and D at the
                    4096 strided cells, beside ref_256cube.npz's -p 2x2x2.

Usage: python tests/golden/make_ref_fixtures.py [--big | --spread | --big-extra | --breakdown | --spread256]
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


def run_ref(grid, procs, samples, tol, maxiter, solver, seed, ignore=False, nranks=1, binary=None):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.bin")
        cmd = [binary or REF, "-g", "x".join(map(str, grid)), "-p", "x".join(map(str, procs)),
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


def normwise(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def decompositions(n, ranks):
    """Every p1 x p2 x p3 with p1 p2 p3 = ranks dividing n^3 into sub-domains
    of at least 2 cells per side (what the reference's halo needs)."""
    out = []
    for p1 in range(1, ranks + 1):
        for p2 in range(1, ranks + 1):
            if ranks % (p1 * p2):
                continue
            p3 = ranks // (p1 * p2)
            if all(n % p == 0 and n // p >= 2 for p in (p1, p2, p3)):
                out.append((p1, p2, p3))
    return out


# (label, binary suffix, decomposition, ranks): the -p 1x1x1 SSE2 run first,
# then every decomposition into 2, 4 and 8 ranks with each SIMD width, and
# the 1x1x1 run with the wider SIMD
SIMD = [("sse2", ""), ("avx2", "_avx2"), ("avx512", "_avx512")]
VARIANTS_32 = [("sse2_p111", "", (1, 1, 1), 1)] + \
    [(f"{w}_p{p[0]}{p[1]}{p[2]}", suf, p, r) for r in (2, 4, 8) for p in decompositions(32, r) for w, suf in SIMD] + \
    [("avx2_p111", "_avx2", (1, 1, 1), 1), ("avx512_p111", "_avx512", (1, 1, 1), 1)]
# 13x7x5 has no decomposition into sub-domains of two or more cells per side
# (which the reference's halo needs), so only the SIMD width varies there
VARIANTS_ODD = [("sse2_p111", "", (1, 1, 1), 1), ("avx2_p111", "_avx2", (1, 1, 1), 1),
                ("avx512_p111", "_avx512", (1, 1, 1), 1)]


def spread():
    d = {}
    for tag, grid, samples, tol, maxiter, seed, variants in [
            ("c32", (32, 32, 32), 8, 1e-9, 500, 1234, VARIANTS_32),
            ("odd", (13, 7, 5), 3, 1e-8, 200, 5, VARIANTS_ODD)]:
        d[f"{tag}_variants"] = np.array([v[0] for v in variants])
        for s in (1, 2, 3):
            its, dist = [], []
            base = None
            for label, suffix, procs, nranks in variants:
                r = run_ref(grid, procs, samples, tol, maxiter, s, seed, nranks=nranks, binary=REF + suffix)
                assert np.all(r["codes"] == 0), (tag, s, label)
                base = base or r
                its.append(r["iters"])
                dist.append(normwise(r["D"], base["D"]))
                print(tag, "solver", s, label, r["iters"].tolist(), "D %.2e" % dist[-1], flush=True)
            d[f"{tag}_s{s}_iters"] = np.array(its, dtype=np.int32)
            d[f"{tag}_s{s}_Ddist"] = np.array(dist)
    np.savez_compressed(os.path.join(HERE, "ref_spread.npz"), **d)


VARIANTS_256 = [("sse2_p118", "", (1, 1, 8)), ("sse2_p181", "", (1, 8, 1)), ("sse2_p811", "", (8, 1, 1)),
                ("sse2_p124", "", (1, 2, 4)), ("sse2_p421", "", (4, 2, 1)), ("avx2_p222", "_avx2", (2, 2, 2)),
                ("avx512_p222", "_avx512", (2, 2, 2))]


def spread256():
    """BiCGStab at 256^3 (seed 7, 16 samples, tol 1e-9) under other
    decompositions into 8 ranks and SIMD widths, beside the -p 2x2x2 2-lane
    run of ref_256cube.npz: iteration counts and D at its 4096 strided cells.

    Each variant takes minutes, so the fixture is rewritten after every one
    and a rerun keeps the variants it already holds."""
    path = os.path.join(HERE, "ref_spread_256cube.npz")
    n = 256 ** 3
    idx = np.arange(0, n, n // 4096, dtype=np.int64)[:4096]
    done = {}
    if os.path.exists(path):
        old = np.load(path)
        for j, label in enumerate(old["variants"].tolist()):
            done[label] = (old["s1_iters"][j], old["s1_Dsample"][j])
    for label, suffix, procs in VARIANTS_256:
        if label in done:
            continue
        r = run_ref((256, 256, 256), procs, 16, 1e-9, 2000, 1, 7, ignore=True, nranks=8, binary=REF + suffix)
        done[label] = (r["iters"], r["D"][idx])
        print(label, r["iters"].tolist(), flush=True)
        labels = [v[0] for v in VARIANTS_256 if v[0] in done]
        np.savez_compressed(path, variants=np.array(labels),
                            s1_iters=np.array([done[l][0] for l in labels], dtype=np.int32),
                            s1_Dsample=np.array([done[l][1] for l in labels]), sample_idx=idx)


def big_fixed():
    """BiCGStab at 256^3 in fixed-iteration mode under a second summation
    order (-p 1x1x8 beside the -p 2x2x2 run of big_extra), for 100 iterations,
    and for 20 iterations under both: how far the reference's own trajectory
    moves with the summation order over a fixed number of iterations."""
    path = os.path.join(HERE, "ref_256cube_extra.npz")
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    idx = d["sample_idx"]
    for tag, procs, m in [("fixed100_p118", (1, 1, 8), 100), ("fixed20", (2, 2, 2), 20),
                          ("fixed20_p118", (1, 1, 8), 20)]:
        r = run_ref((256, 256, 256), procs, 16, 1e-9, m, 1, 7, ignore=True, nranks=8)
        D = r["D"]
        d.update({f"s1_{tag}_iters": r["iters"], f"s1_{tag}_codes": r["codes"], f"s1_{tag}_errs": r["errs"],
                  f"s1_{tag}_Dsum": np.array(D.sum()), f"s1_{tag}_Dsq": np.array((D * D).sum()),
                  f"s1_{tag}_Dsample": D[idx]})
        print(tag, r["iters"].tolist(), flush=True)
    np.savez_compressed(path, **d)


def big_extra():
    d = {}
    n = 256 ** 3
    idx = np.arange(0, n, n // 4096, dtype=np.int64)[:4096]
    r = run_ref((256, 256, 256), (2, 2, 2), 16, 1e-9, 100, 1, 7, ignore=True, nranks=8)
    D = r["D"]
    d.update({"s1_fixed100_iters": r["iters"], "s1_fixed100_codes": r["codes"], "s1_fixed100_errs": r["errs"],
              "s1_fixed100_Dsum": np.array(D.sum()), "s1_fixed100_Dsq": np.array((D * D).sum()),
              "s1_fixed100_Dmin": np.array(D.min()), "s1_fixed100_Dmax": np.array(D.max()),
              "s1_fixed100_Dsample": D[idx]})
    print("fixed100", r["iters"].tolist(), flush=True)
    # four single-rank runs at once (one core each)
    jobs = [(1, ""), (1, "_avx512"), (2, ""), (3, "")]

    def one(job):
        s, suffix = job
        return job, run_ref((256, 256, 256), (1, 1, 1), 4, 1e-9, 2000, s, 7, ignore=True, binary=REF + suffix)

    import concurrent.futures
    with concurrent.futures.ThreadPoolExecutor(len(jobs)) as ex:
        for (s, suffix), r in ex.map(one, jobs):
            tag = f"s{s}_p111{suffix}"
            d.update({f"{tag}_iters": r["iters"], f"{tag}_codes": r["codes"], f"{tag}_errs": r["errs"]})
            print(tag, r["iters"].tolist(), flush=True)
    d["sample_idx"] = idx
    np.savez_compressed(os.path.join(HERE, "ref_256cube_extra.npz"), **d)


# (grid, stencil (a, b, c), right-hand side, tol, maxiter): found by a search
# over 1-4 cell grids, stencils and right-hand sides of small integers with
# the oracle, then run on the reference
BREAKDOWN_CASES = [
    ((2, 1, 1), (1.0, 2.0, 0.0), (1.0, -1.0), 1e-12, 6),   # omega = 0 at iteration 1
    ((1, 2, 1), (1.0, 2.0, 0.0), (1.0, -1.0), 1e-12, 6),
    ((1, 1, 2), (1.0, 2.0, 0.0), (1.0, -1.0), 1e-12, 6),
    ((2, 1, 1), (1.0, 0.0, -2.0), (2.0, 2.0), 1e-12, 6),
    ((3, 1, 1), (1.0, 2.0, 1.0), (1.0, 2.0, 2.0), 1e-12, 6),  # rho = 0 at iteration 2
    ((1, 3, 1), (1.0, 2.0, 1.0), (1.0, 2.0, 2.0), 1e-12, 6),
    ((1, 1, 3), (1.0, 2.0, 1.0), (1.0, 2.0, 2.0), 1e-12, 6),
    ((3, 1, 1), (2.0, -2.0, -1.0), (1.0, 2.0, 2.0), 1e-12, 6),
    ((2, 1, 1), (1.0, 2.0, 0.0), (1.0, -1.0), 0.5, 6),     # omega = 0 but err <= tol: converged
]


def breakdown():
    import json
    out = []
    for grid, abc, rhs, tol, maxiter in BREAKDOWN_CASES:
        cmd = [REF + "_cases", "-g", "x".join(map(str, grid)), "-c", ",".join(repr(v) for v in abc),
               "-b", ",".join(repr(v) for v in rhs), "-e", repr(tol), "-m", str(maxiter), "-S", "1", "-f", "1"]
        log = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
        m = re.search(r"result code=(\d+) iter=(-?\d+) err=(\S+)", log)
        x = [float(v) for v in re.search(r"^x (.*)$", log, re.M).group(1).split()]
        hist = [float(v) for v in re.findall(r"BiCgStab: after iter=\d+, err=(\S+)", log)]
        out.append(dict(grid=grid, abc=abc, rhs=rhs, tol=tol, maxiter=maxiter, code=int(m.group(1)),
                        iter=int(m.group(2)), err=float(m.group(3)), hist=hist, x=x))
        print(out[-1], flush=True)
    with open(os.path.join(HERE, "ref_breakdown.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    if "--big" in sys.argv:
        big()
    elif "--spread" in sys.argv:
        spread()
    elif "--big-extra" in sys.argv:
        big_extra()
    elif "--breakdown" in sys.argv:
        breakdown()
    elif "--spread256" in sys.argv:
        spread256()
    elif "--big-fixed" in sys.argv:
        big_fixed()
    else:
        small()
