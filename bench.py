"""Benchmark: probe-solves/s of the estimator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--solver cg|pcg|bicgstab]

Workload (one step): the plain single-probe FP64 drop-in on a 256x256x256 grid
(16,777,216 unknowns), 16 Rademacher probes from mt19937 seed 7, tol 1e-9,
maxiter 2000: one sedi_estimator_run (probe generation + 16 solves +
accumulation, the reference's TIMER_GLOBAL scope).  The probes are generated
on the device, so inputs are resident in HBM when the timed region starts.

Prints one JSON line with the throughput, the roofline of the dominant kernel
(its algorithmic bytes per launch over its average duration, measured with HIP
events on the solver's stream during the timed steps) and the CPU baseline:
the reference's own solver code (oracle/_ref/sedi_ref, built from the
reference's headers) on the host cores under mpirun, on a bounded sample.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GRID = (256, 256, 256)
SAMPLES = 16
SEED = 7
TOL = 1e-9
MAXITER = 2000
NAMES = {"cg": 2, "pcg": 3, "bicgstab": 1}
SOLVER_LABEL = {2: "CG", 3: "PipelinedCG", 1: "BiCGStab"}
# dominant kernel per solver and the vectors it streams per launch (FP64 N cells)
DOMINANT = {2: ("cg_update_kernel", 5), 3: ("pcg_iter_kernel", 16), 1: ("bicgstab_update_kernel", 7)}
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "sedi_ref")
MPIRUN = "/opt/conda/bin/mpirun"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--solver", default="cg", choices=sorted(NAMES))
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-iters", type=int, default=300,
                   help="iterations of the bounded CPU-baseline sample")
    p.add_argument("--stencil", action="store_true",
                   help="time the standalone 7-point stencil kernel (y = A x, FP64) at 256^3 and exit")
    p.add_argument("--pmc-workload", action="store_true",
                   help="run a short fixed workload for rocprofv3 --pmc (1 probe per solver, "
                        "10 iterations each, at 256^3) and exit")
    return p.parse_args()


def pmc_workload():
    """One probe per solver, capped at 10 iterations (-m 10 -i), on the 256^3
    grid: a few dozen dispatches of every kernel, including the estimator's
    streaming accumulate_kernel (reads x, v, t, q; writes t, q), whose known
    traffic calibrates the counters."""
    import sedi_amd
    for kind in (2, 3, 1):
        est = sedi_amd.Estimator(GRID, kind=kind, samples=1, maxiter=10, tol=TOL, seed=SEED,
                                 ignore_errors=True)
        est.run(want_D=False)
        est.close()
    print("pmc workload done", flush=True)


def stencil_only(reps=20):
    """y = A x at 256^3 through sedi_stencil_apply with device-resident x, y
    (allocated by the library's own staging on the first call), kernel time
    from HIP events; algorithmic bytes 2*N*8 per launch (read x, write y)."""
    import numpy as np
    import sedi_amd
    n = GRID[0] * GRID[1] * GRID[2]
    x = np.random.default_rng(0).standard_normal(n)
    st = sedi_amd.Stencil()
    st.apply(GRID, x)  # warm-up
    sedi_amd.profile_reset()
    sedi_amd.profile_enable("stencil_apply_kernel")
    for _ in range(reps):
        st.apply(GRID, x)
    sedi_amd.profile_enable(None)
    launches, ms = sedi_amd.profile_get("stencil_apply_kernel")
    avg = ms / 1e3 / launches
    gbs = 2 * n * 8 / avg / 1e9
    print(json.dumps({"kernel": "stencil_apply_kernel", "grid": list(GRID), "launches": launches,
                      "avg_launch_us": round(avg * 1e6, 2), "achieved_GBs": round(gbs, 1),
                      "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4)}), flush=True)


def traffic_from_profiles(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", "%s_256cube_pmc.txt" % kernel)
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def host_cores():
    n = os.cpu_count() or 1
    return max(1, min(16, n))


def factor3(p, n):
    """A 3-D factorisation p1*p2*p3 = p with each factor dividing n."""
    best = None
    for a in range(1, p + 1):
        if p % a or n % a:
            continue
        for b in range(1, p // a + 1):
            if (p // a) % b or n % b:
                continue
            c = p // a // b
            if n % c:
                continue
            cand = (a, b, c)
            if best is None or max(cand) - min(cand) < max(best) - min(best):
                best = cand
    return best


def cpu_baseline(kind, gpu_iters_mean, iters_cap):
    """The reference's solver on the host cores: 1 probe (seed 7, probe 0) of
    the same 256^3 workload capped at `iters_cap` iterations (-m cap -i),
    scaled to a full probe-solve with the GPU run's mean iteration count."""
    if not os.path.exists(REF_BIN):
        return None
    p = host_cores()
    while p > 1 and factor3(p, GRID[0]) is None:
        p -= 1
    dec = factor3(p, GRID[0])
    cmd = [REF_BIN, "-g", "x".join(map(str, GRID)), "-p", "x".join(map(str, dec)), "-s", "1",
           "-e", repr(TOL), "-m", str(iters_cap), "-i", "-S", str(kind), "-r", str(SEED)]
    if p > 1:
        if not os.path.exists(MPIRUN):
            return None
        cmd = [MPIRUN, "-n", str(p)] + cmd
    try:
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=240).stdout
    except Exception:
        return None
    m = re.search(r"Total solve time: ([0-9.eE+-]+) seconds, iteration count min=(\d+)", out)
    if not m:
        return None
    secs, iters = float(m.group(1)), int(m.group(2))
    per_iter = secs / max(iters, 1)
    probe_secs = per_iter * gpu_iters_mean
    return {"value": 1.0 / probe_secs, "unit": "probe-solves/s", "cores": p, "kind": "reference",
            "sample": ("1 probe of the same workload (256^3, seed 7, probe 0), %s, capped at %d "
                       "iterations (-m %d -i) under mpirun -n %d -p %s: %.2f s = %.2f ms/iteration, "
                       "scaled to the GPU run's mean of %.1f iterations per probe-solve"
                       % (SOLVER_LABEL[kind], iters, iters, p, "x".join(map(str, dec)), secs,
                          per_iter * 1e3, gpu_iters_mean))}


def main():
    args = parse()
    if args.pmc_workload:
        pmc_workload()
        return
    if args.stencil:
        stencil_only()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the whole grid on a single MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import sedi_amd

    kind = NAMES[args.solver]
    kname, nvec = DOMINANT[kind]
    n = GRID[0] * GRID[1] * GRID[2]
    est = sedi_amd.Estimator(GRID, kind=kind, samples=SAMPLES, maxiter=MAXITER, tol=TOL,
                             seed=SEED, ignore_errors=True)
    for _ in range(args.warmup):
        est.run(want_D=False)
    sedi_amd.profile_reset()
    sedi_amd.profile_enable(kname)
    t0 = time.perf_counter()
    runs = [est.run(want_D=False) for _ in range(args.steps)]
    t1 = time.perf_counter()
    sedi_amd.profile_enable(None)
    launches, kms = sedi_amd.profile_get(kname)
    est.close()
    wall = t1 - t0
    loop = sum(r["seconds"] for r in runs)
    iters = np.concatenate([r["iters"] for r in runs])
    solves = SAMPLES * args.steps
    value = solves / wall
    avg_s = kms / 1e3 / max(launches, 1)
    bytes_per_launch = nvec * n * 8
    achieved = bytes_per_launch / avg_s / 1e9
    line = {
        "metric": "probe-solves/s at 256^3 (batched mixed vs plain FP64); % HBM peak; p99 job latency",
        "value": round(value, 4),
        "unit": "probe-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device-generated Rademacher probes, mt19937 seed 7)",
        "config": {"workload": "plain single-probe FP64 %s estimator, 256^3 grid, 16 probes "
                               "(seed 7), tol 1e-9, maxiter 2000" % SOLVER_LABEL[kind],
                   "grid": list(GRID), "samples": SAMPLES, "solver": SOLVER_LABEL[kind],
                   "precision": "fp64", "tol": TOL, "maxiter": MAXITER, "seed": SEED},
        "iterations": {"min": int(iters.min()), "max": int(iters.max()),
                       "mean": round(float(iters.mean()), 2)},
        "estimator_loop_s": round(loop, 4),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(kname),
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": round(avg_s * 1e6, 2), "launches": launches},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(kind, float(iters.mean()), args.cpu_iters)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
