"""Benchmark: probe-solves/s of the stochastic diag(A^-1) estimator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch K] [--mode headline|plain]

Headline workload (one step): the batched mixed-precision CG estimator on a
256x256x256 grid (16,777,216 unknowns): 64 Rademacher probes from mt19937
seed 11 solved as one interleaved block of k = 64 right-hand sides (FP32
vectors and stencil, FP64 dot products and scalars, FP64 residual replacement
every 50 iterations), relative residual 1e-8, maxiter 2000: one
sedi_estimator_run (probe generation + the block solve + accumulation, the
reference's TIMER_GLOBAL scope).  Probes are generated on the device, so the
inputs are resident in HBM when the timed region starts.

Beside it, on the same grid, stream and tolerance: the plain FP64
one-probe-at-a-time CG (the drop-in) on the first 16 probes, the speed-up of
the batched form over it, and the error of the mixed form's diagonal against
the FP64 one on those 16 probes (per cell relative to the exact diag(A^-1),
tests/golden/exact_diag_256cube.npz, and normwise); the p50 / p99 job
latency of the serving stream, on its first 100 jobs at 5 jobs/s (saturated,
--serve-jobs) and its first 40 at 0.5 jobs/s (below saturation,
--serve-sub-jobs/--serve-sub-rate); and the standalone stencil on device
arrays, 256^3 warm and cold and 512^3 cold (--no-stencil-line to skip).

Prints one JSON line with the throughput, the roofline of the dominant kernel
(its algorithmic bytes per launch over its average duration, measured with HIP
events on the solver's stream during the timed steps) and the CPU baseline:
the reference's own solver code (oracle/_ref/sedi_ref, built from the
reference's headers) on the host cores under mpirun, on a bounded sample.

--mode plain measures the plain FP64 drop-in alone (256^3, 16 probes, seed 7,
tol 1e-9, --solver cg|pcg|bicgstab).
--mode ksweep: the batched CG at 256^3 (64 probes, seed 11, rel. residual
1e-8) for k in {8, 16, 32, 64}, FP64 and mixed: probe-solves/s, per-probe
iteration counts and the mixed diagonal's error against the FP64 one.
--mode c512: mixed pipelined CG and BiCGStab at 512^3 (32 probes, seed 23, in
batches of 16, tol 1e-8, maxiter 4000; the replacement policy of each
solver's mixed form) against the plain FP64 GPU solvers on the same probes:
throughput and max / RMS diagonal error.
--mode arith: the arithmetic type of that mixed pipelined CG and BiCGStab:
the default library (FP32 storage, stencil and updates in FP64 registers)
beside its f32compute build (the same kernels in FP32 registers) and the
plain FP64 estimator, on the first --arith-samples probes of the 512^3
configuration in batches of 8.
--mode serve: the serving stream -- Poisson arrivals at --rate jobs/s, grid
side uniform over {64, 96, 128, 192, 256}, samples uniform in [8, 64], tol
uniform over {1e-6, 1e-8, 1e-9}, solver uniform over CG / pipelined CG /
BiCGStab, seed = job id, mixed batched kernels, 200 GB admission budget --
truncated to --jobs jobs: p50 / p99 job latency (submit to result) and
aggregate probe-solves/s.
These print progress lines and a summary line.
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
MAXITER = 2000
NAMES = {"cg": 2, "pcg": 3, "bicgstab": 1}
SOLVER_LABEL = {2: "CG", 3: "PipelinedCG", 1: "BiCGStab"}
# plain FP64 drop-in workload (--mode plain)
PLAIN = dict(samples=16, seed=7, tol=1e-9)
# headline batched mixed workload and its FP64 comparison
BATCH = dict(samples=64, seed=11, tol=1e-8, rep_every=50)
CMP_SAMPLES = 16
# dominant kernel per solver and the FP64 vectors it streams per iteration
# (CG and pipelined CG, whose solution updates come in groups: see
# upd_values, pcg_values)
DOMINANT = {2: ("cg_update_kernel", None), 3: ("pcg_iter_kernel", None), 1: ("bicgstab_update_kernel", 7)}
# batched mixed CG: batch_upd_kernel streams p and r in and r out every
# iteration, and at every fourth iteration also x_lo in and out and the three
# previous directions in (the solution updates of four iterations are applied
# together; the library's default -DSEDI_CG_XGROUP=4): 3 or 8 FP32 values per
# right-hand side and cell.  The plain solver's cg_update_kernel groups them
# the same way: 3 or 8 FP64 values.
BATCH_KERNEL = "batch_upd_kernel"
BATCH_CG_GROUP = 4
PLAIN_CG_GROUP = 4


def upd_values(iters, group):
    """Vectors the CG update pass streams for a right-hand side that ran
    `iters` iterations with solution updates in groups of `group`: 3 per
    iteration, plus x in and out and the group's other directions in at every
    group's last iteration."""
    return 3 * iters + (2 + group - 1) * (iters // group)


def pcg_values(iters, group):
    """Vectors the plain pipelined-CG pass streams for `iters` iterations with
    solution updates in groups of `group`: 7 in and 7 out per iteration (w,
    q, s, u, p, z, r; x waits), plus x in and out and the group's other
    directions in at every group's last iteration."""
    return 14 * iters + (2 + group - 1) * (iters // group)
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "sedi_ref")
MPIRUN = "/opt/conda/bin/mpirun"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--mode", default="headline", choices=["headline", "plain", "ksweep", "c512", "serve", "arith"])
    p.add_argument("--arith-child", type=int, default=0, help=argparse.SUPPRESS)
    p.add_argument("--arith-out", default="", help=argparse.SUPPRESS)
    p.add_argument("--arith-samples", type=int, default=8, help="--mode arith: probes per solver")
    p.add_argument("--jobs", type=int, default=200, help="serve: jobs in the stream")
    p.add_argument("--rate", default="5", help="serve: Poisson arrival rate(s) in jobs/s, comma-separated")
    p.add_argument("--serve-sub-jobs", type=int, default=40,
                   help="headline: jobs of the serving sample below saturation (0 = none)")
    p.add_argument("--serve-sub-rate", type=float, default=0.5, help="headline: its arrival rate, jobs/s")
    p.add_argument("--serve-jobs", type=int, default=100,
                   help="headline: jobs of the serving sample reported beside it (0 = none)")
    p.add_argument("--batch", type=int, default=64, help="right-hand sides per block (headline)")
    p.add_argument("--solver", default="cg", choices=sorted(NAMES), help="--mode plain only")
    p.add_argument("--c512-only", choices=["pcg", "bicgstab"], help="--mode c512: one solver only")
    p.add_argument("--c512-compaction", action="store_true",
                   help="--mode c512: run each mixed estimator with lane compaction off and on (no FP64 run) "
                        "and report both rates and whether D is bitwise the same")
    p.add_argument("--c512-mixed-only", action="store_true", help="--mode c512: the mixed estimator only")
    p.add_argument("--no-compact", action="store_true", help="turn lane compaction of the batched solvers off")
    p.add_argument("--no-graphs", action="store_true",
                   help="launch every kernel directly instead of replaying captured HIP graphs")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-compare", action="store_true",
                   help="headline: skip the plain FP64 comparison run")
    p.add_argument("--cpu-iters", type=int, default=300,
                   help="iterations of the bounded CPU-baseline sample")
    p.add_argument("--stencil-side", type=int, default=0, help="--stencil: grid side (default 256)")
    p.add_argument("--stencil", action="store_true",
                   help="time the standalone 7-point stencil kernel (y = A x, FP64) at 256^3 and exit")
    p.add_argument("--copy", action="store_true",
                   help="time the HIP runtime's device-to-device copy of one vector (the streaming ceiling "
                        "the stencil is compared with; --stencil-side, --cold as for --stencil) and exit")
    p.add_argument("--no-stencil-line", dest="stencil_line", action="store_false",
                   help="headline: skip the standalone stencil measurements reported beside it")
    p.add_argument("--cold", action="store_true",
                   help="--stencil: rotate through (x, y) pairs spanning 1 GiB so x comes from HBM")
    p.add_argument("--pmc-workload", action="store_true",
                   help="run a short fixed workload for rocprofv3 --pmc (1 probe per plain solver "
                        "and one batch of 64 mixed CG probes, 10 iterations each, at 256^3) and exit")
    return p.parse_args()


def pmc_workload():
    """One probe per plain solver and one mixed batch of 64 CG probes, capped
    at 10 iterations (-m 10 -i), and 5 standalone stencil applies, on the 256^3
    grid: a few dozen dispatches of every kernel, including the estimator's
    streaming accumulate_kernel (reads x, v, t, q; writes t, q), whose known
    traffic calibrates the counters."""
    import sedi_amd
    stencil_only(reps=5)
    for kind in (2, 3, 1):
        est = sedi_amd.Estimator(GRID, kind=kind, samples=1, maxiter=10, tol=1e-9, seed=7,
                                 ignore_errors=True)
        est.run(want_D=False)
        est.close()
    est = sedi_amd.Estimator(GRID, kind=2, samples=64, maxiter=10, tol=1e-30, seed=11,
                             ignore_errors=True, precision=sedi_amd.FP32_MIXED, batch_k=64)
    est.run(want_D=False)
    est.close()
    print("pmc workload done", flush=True)


# footprint of a cold measurement: the (x, y) pairs it rotates through span
# at least four times the 256 MB Infinity Cache, and there are at least four
# of them: at 512^3 a launch's time depends on where its pair's buffers lie
# in physical memory (one plan ran 363 us on two of four pairs and 416 us on
# the other two), so the mean is taken over several placements
COLD_BYTES = 1 << 30
COLD_MIN_PAIRS = 4
# y of pair k < 4 is allocated after a spacer of k * COLD_SPACER bytes, so
# the pairs place y at different offsets from x
COLD_SPACER = 34 << 20


def cold_pairs(n, npairs, x0=None):
    """npairs device-resident (x, y) FP64 pairs of n cells for a cold
    measurement, y of pair k < 4 allocated after a k * COLD_SPACER spacer
    (the spacers are returned to be freed with the pairs)."""
    import sedi_amd
    pairs, spacers = [], []
    for k in range(npairs):
        x = sedi_amd.DeviceArray(x0) if x0 is not None else sedi_amd.DeviceArray(n=n)
        if 0 < k < 4:
            spacers.append(sedi_amd.DeviceArray(n=k * COLD_SPACER // 8))
        pairs.append((x, sedi_amd.DeviceArray(n=n)))
    return pairs, spacers


def stencil_only(reps=20, side=None, cold=False):
    """y = A x at 256^3 (or side^3) through sedi_stencil_apply with
    device-resident x, y (sedi_amd.DeviceArray), kernel time from HIP events
    around each launch; algorithmic bytes 2*N*8 per launch (read x, write y).
    Warm: back-to-back launches on one (x, y), so at 256^3 part of x is still
    in the 256 MB Infinity Cache from the previous launch.  Cold: launches
    rotate through four (x, y) pairs (spanning 1 GiB at 256^3, 8 GiB at
    512^3; y at four different offsets from x, see cold_pairs), so each
    launch reads an x untouched since three launches back (evicted) while the
    cache holds the previous launch's dirty y, as in a solver's sweep over
    vectors larger than the cache."""
    import ctypes
    import numpy as np
    import sedi_amd
    grid = (side,) * 3 if side else GRID
    n = grid[0] * grid[1] * grid[2]
    npairs = max(COLD_MIN_PAIRS, -(-COLD_BYTES // (16 * n))) if cold else 1
    x0 = np.random.default_rng(0).standard_normal(n)
    pairs, spacers = cold_pairs(n, npairs, x0)
    g, _ = sedi_amd._grid(grid)
    st = sedi_amd.Stencil()._c()
    L = sedi_amd.lib()

    def apply(k):
        x, y = pairs[k % npairs]
        rc = L.sedi_stencil_apply(ctypes.byref(g), ctypes.byref(st), x.ptr, y.ptr)
        if rc < 0:
            raise RuntimeError(L.sedi_last_error().decode())

    for k in range(npairs):  # warm-up: every pair touched once
        apply(k)
    sedi_amd.profile_reset()
    sedi_amd.profile_enable("stencil_apply_kernel")
    for r in range(reps):
        apply(r)
    sedi_amd.profile_enable(None)
    launches, ms = sedi_amd.profile_get("stencil_apply_kernel")
    avg = ms / 1e3 / launches
    gbs = 2 * n * 8 / avg / 1e9
    for a in [v for p in pairs for v in p] + spacers:
        a.free()
    return {"kernel": "stencil_apply_kernel", "grid": list(grid), "cache": "cold" if cold else "warm",
            "pairs": npairs, "launches": launches, "avg_launch_us": round(avg * 1e6, 2),
            "achieved_GBs": round(gbs, 1), "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4)}


def copy_only(reps=12, side=None, cold=False):
    """The practical streaming ceiling beside the stencil: the HIP runtime's
    own device-to-device copy (hipMemcpyAsync) of one side^3 FP64 vector, 2N*8
    bytes moved, timed with HIP events; warm on one pair, cold rotating
    through pairs spanning 1 GiB as stencil_only does."""
    import ctypes
    import sedi_amd
    H = sedi_amd._hiprt()
    H.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    H.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    H.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    H.hipEventDestroy.argtypes = [ctypes.c_void_p]
    H.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    grid = (side,) * 3 if side else GRID
    n = grid[0] * grid[1] * grid[2]
    npairs = max(COLD_MIN_PAIRS, -(-COLD_BYTES // (16 * n))) if cold else 1
    pairs, spacers = cold_pairs(n, npairs)
    for a, b in pairs:
        a.fill_bytes(1)
        b.fill_bytes(2)
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    H.hipEventCreate(ctypes.byref(e0))
    H.hipEventCreate(ctypes.byref(e1))
    times = []
    for r in range(reps + npairs):
        a, b = pairs[r % npairs]
        H.hipEventRecord(e0, None)
        H.hipMemcpyAsync(b.ptr, a.ptr, 8 * n, 3, None)  # device to device
        H.hipEventRecord(e1, None)
        H.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        H.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        if r >= npairs:
            times.append(ms.value)
    H.hipEventDestroy(e0)
    H.hipEventDestroy(e1)
    for a in [v for p in pairs for v in p] + spacers:
        a.free()
    avg = sum(times) / len(times) * 1e-3
    gbs = 2 * n * 8 / avg / 1e9
    return {"kernel": "hipMemcpyAsync device to device", "grid": list(grid), "cache": "cold" if cold else "warm",
            "pairs": npairs, "launches": len(times), "avg_launch_us": round(avg * 1e6, 2),
            "achieved_GBs": round(gbs, 1), "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4)}


def traffic_ratio_from_profiles(kernel):
    """Measured HBM bytes over algorithmic bytes of a launch, from the committed
    rocprofv3 --pmc summary of the kernel (calibrated FETCH_SIZE + WRITE_SIZE
    of fully active launches), if there is one."""
    path = os.path.join(ROOT, "profiles", "%s_256cube_pmc.txt" % kernel)
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_ratio:\s*([0-9.eE+]+)", open(path).read(), re.M)
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


def cpu_baseline(kind, seed, tol, gpu_iters_mean, iters_cap):
    """The reference's solver on the host cores: 1 probe (probe 0 of `seed`)
    of the same 256^3 workload capped at `iters_cap` iterations (-m cap -i),
    scaled to a full probe-solve with the GPU run's mean iteration count."""
    if not os.path.exists(REF_BIN):
        return None
    p = host_cores()
    while p > 1 and factor3(p, GRID[0]) is None:
        p -= 1
    dec = factor3(p, GRID[0])
    cmd = [REF_BIN, "-g", "x".join(map(str, GRID)), "-p", "x".join(map(str, dec)), "-s", "1",
           "-e", repr(tol), "-m", str(iters_cap), "-i", "-S", str(kind), "-r", str(seed)]
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
    cpu = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            m = re.search(r"^model name\s*:\s*(.+)$", f.read(), re.M)
            cpu = m.group(1).strip() if m else cpu
    except OSError:
        pass
    build = ""
    try:
        with open(os.path.join(ROOT, "oracle", "_ref", "BUILD_LINE.txt")) as f:
            build = f.read().strip()
    except OSError:
        pass
    return {"value": round(1.0 / probe_secs, 5), "unit": "probe-solves/s", "cores": p,
            "kind": "reference", "cpu": cpu,
            "build": build + " (g++ 11 without -march: 2-lane SSE2 dot products, where the reference's "
                             "Makefile builds with icpc -xHost; a bounded, scaled sample, not a tuned CPU run)",
            "sample": ("1 probe of the same workload (256^3, seed %d, probe 0), plain FP64 %s, "
                       "capped at %d iterations (-m %d -i) under mpirun -n %d -p %s: %.2f s = "
                       "%.2f ms/iteration, scaled to the GPU run's mean of %.1f iterations per "
                       "probe-solve" % (seed, SOLVER_LABEL[kind], iters, iters, p,
                                        "x".join(map(str, dec)), secs, per_iter * 1e3,
                                        gpu_iters_mean))}


def base_line(args, value, wall, dtype, config):
    return {
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
        "dtype": dtype,
        "data": "synthetic (device-generated Rademacher probes, mt19937 seed %d)" % config["seed"],
        "config": config,
    }


def roofline(kernel, launches, kms, bytes_total):
    avg_s = kms / 1e3 / max(launches, 1)
    per_launch = bytes_total / max(launches, 1)
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    ratio = traffic_ratio_from_profiles(kernel)
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if ratio is None else int(ratio * per_launch),
            "traffic_over_algorithmic": ratio,
            "algorithmic_bytes_per_launch": int(per_launch), "avg_launch_us": round(avg_s * 1e6, 2),
            "launches": launches}


def run_plain(args):
    import numpy as np
    import sedi_amd
    kind = NAMES[args.solver]
    kname, nvec = DOMINANT[kind]
    n = GRID[0] * GRID[1] * GRID[2]
    est = sedi_amd.Estimator(GRID, kind=kind, samples=PLAIN["samples"], maxiter=MAXITER,
                             tol=PLAIN["tol"], seed=PLAIN["seed"], ignore_errors=True)
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
    iters = np.concatenate([r["iters"] for r in runs])
    config = {"workload": "plain single-probe FP64 %s estimator, 256^3 grid, 16 probes (seed 7), "
                          "tol 1e-9, maxiter 2000" % SOLVER_LABEL[kind],
              "grid": list(GRID), "samples": PLAIN["samples"], "solver": SOLVER_LABEL[kind],
              "precision": "fp64", "tol": PLAIN["tol"], "maxiter": MAXITER, "seed": PLAIN["seed"]}
    line = base_line(args, PLAIN["samples"] * args.steps / wall, wall, "f64", config)
    line["iterations"] = {"min": int(iters.min()), "max": int(iters.max()),
                          "mean": round(float(iters.mean()), 2)}
    line["estimator_loop_s"] = round(sum(r["seconds"] for r in runs), 4)
    # algorithmic bytes of the iterations the probes ran (launches queued after
    # a solve converged return at once and move nothing); plain CG groups its
    # solution updates by four: 3 vectors at iterations 1-3 of a group, 8 at
    # the 4th
    if kind == 2:
        vecs = sum(upd_values(int(m), PLAIN_CG_GROUP) for m in iters)
    elif kind == 3:  # pipelined CG groups them the same way: 14 vectors, 19 at every 4th
        vecs = sum(pcg_values(int(m), PLAIN_CG_GROUP) for m in iters)
    else:
        vecs = int(iters.sum()) * nvec
    line["roofline"] = roofline(kname, launches, kms, vecs * n * 8)
    line["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(kind, PLAIN["seed"], PLAIN["tol"], float(iters.mean()),
                                            args.cpu_iters)
    return line


def run_headline(args):
    import numpy as np
    import sedi_amd
    n = GRID[0] * GRID[1] * GRID[2]
    K = max(4, (args.batch + 3) // 4 * 4)
    stencil = None
    if args.stencil_line:
        # the standalone 7-point stencil on device-resident FP64 vectors: warm
        # (back-to-back on one pair) and cold (x from HBM) at 256^3, cold at
        # 512^3; measured first, on buffers laid out as in a fresh process
        stencil = [stencil_only(side=256), stencil_only(side=256, cold=True), stencil_only(side=512, cold=True)]
    est = sedi_amd.Estimator(GRID, kind=2, samples=BATCH["samples"], maxiter=MAXITER,
                             tol=BATCH["tol"], seed=BATCH["seed"], ignore_errors=True,
                             precision=sedi_amd.FP32_MIXED, batch_k=K,
                             rep_every=BATCH["rep_every"])
    for _ in range(args.warmup):
        est.run(want_D=False)
    sedi_amd.profile_reset()
    sedi_amd.profile_enable(BATCH_KERNEL)
    t0 = time.perf_counter()
    runs = [est.run(want_D=False) for _ in range(args.steps)]
    t1 = time.perf_counter()
    sedi_amd.profile_enable(None)
    launches, kms = sedi_amd.profile_get(BATCH_KERNEL)
    est.close()
    wall = t1 - t0
    iters = np.concatenate([r["iters"] for r in runs])
    codes = np.concatenate([r["codes"] for r in runs])
    # algorithmic bytes: each right-hand side is credited with its own
    # iterations only (a converged one that still rides in an active block of
    # 4 is not counted)
    rhs_values = int(sum(upd_values(int(m), BATCH_CG_GROUP) for m in iters))
    config = {"workload": "batched mixed-precision CG estimator, 256^3 grid, 64 probes (seed 11) "
                          "as one block of k=%d, FP32 vectors / FP64 reductions, residual "
                          "replacement every 50, rel. residual 1e-8, maxiter 2000" % K,
              "grid": list(GRID), "samples": BATCH["samples"], "solver": "CG", "batch_k": K,
              "precision": "fp32 storage + fp64 reductions", "tol": BATCH["tol"],
              "maxiter": MAXITER, "seed": BATCH["seed"], "rep_every": BATCH["rep_every"]}
    value = BATCH["samples"] * args.steps / wall
    line = base_line(args, value, wall, "f32+f64", config)
    line["iterations"] = {"min": int(iters.min()), "max": int(iters.max()),
                          "mean": round(float(iters.mean()), 2),
                          "unconverged": int(np.count_nonzero(codes))}
    line["estimator_loop_s"] = round(sum(r["seconds"] for r in runs), 4)
    line["roofline"] = roofline(BATCH_KERNEL, launches, kms, rhs_values * n * 4)
    if stencil is not None:
        line["stencil"] = stencil
    if not args.no_compare:
        # plain FP64 drop-in and the mixed form on the first 16 probes of the
        # same stream: the speed-up and the diagonal error of the mixed form
        common = dict(kind=2, samples=CMP_SAMPLES, maxiter=MAXITER, tol=BATCH["tol"],
                      seed=BATCH["seed"], ignore_errors=True)
        warm = sedi_amd.Estimator(GRID, **dict(common, samples=1, maxiter=10))
        warm.run(want_D=False)
        warm.close()
        plain = sedi_amd.Estimator(GRID, **common)
        t0 = time.perf_counter()
        rp = plain.run()
        tp = time.perf_counter() - t0
        plain.close()
        mixed = sedi_amd.Estimator(GRID, precision=sedi_amd.FP32_MIXED, batch_k=CMP_SAMPLES,
                                   rep_every=BATCH["rep_every"], **common)
        rm = mixed.run()
        mixed.close()
        pv = CMP_SAMPLES / tp
        line["plain_fp64"] = {"value": round(pv, 4), "unit": "probe-solves/s",
                              "workload": "plain FP64 CG, one probe at a time, probes 0-15 of "
                                          "the same stream, tol 1e-8, maxiter 2000",
                              "iterations_mean": round(float(rp["iters"].mean()), 2)}
        line["speedup_vs_plain_fp64"] = round(value / pv, 3)
        line["mixed_vs_fp64_D"] = dict(d_errors(rm["D"], rp["D"]), probes=CMP_SAMPLES,
                                       iterations_mean_mixed=round(float(rm["iters"].mean()), 2))
    if args.serve_jobs > 0:
        # the metric's third part: job latency of the serving stream, on a
        # bounded sample of it: saturated (5 jobs/s, several times what the
        # GPU completes of this mix) and below saturation (0.5 jobs/s)
        line["serving"] = dict(run_serve(args.serve_jobs, 5.0),
                               workload="the serving stream's first %d jobs at 5 jobs/s (Poisson; sides "
                                        "64-256, 8-64 probes, tol 1e-6/1e-8/1e-9, CG/PipelinedCG/BiCGStab, "
                                        "seed = job id, mixed, 200 GB budget); the GPU is saturated, so "
                                        "latency includes queueing" % args.serve_jobs)
    if args.serve_sub_jobs > 0:
        line["serving_below_saturation"] = dict(
            run_serve(args.serve_sub_jobs, args.serve_sub_rate),
            workload="the same stream's first %d jobs at %g jobs/s (below what the GPU completes of this "
                     "mix): latency is mostly service time" % (args.serve_sub_jobs, args.serve_sub_rate))
    line["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(2, BATCH["seed"], BATCH["tol"], float(iters.mean()),
                                            args.cpu_iters)
    return line


# diag(A^-1) of the (6, 1, 1) operator is smallest at a corner cell,
# 0.185577 on every grid of side >= 6
DIAG_MIN = 0.185577
EXACT_DIAG = os.path.join(ROOT, "tests", "golden", "exact_diag_256cube.npz")


def d_errors(D, ref):
    """Errors of a diagonal estimate against a reference one on the same
    probes.  Per cell relative to the quantity estimated, diag(A^-1)_i:
    max and RMS of |dD_i| / diag(A^-1)_i at the 8192 cells of
    tests/golden/exact_diag_256cube.npz (exact diagonal; 256^3 only), and over
    every cell the bound max|dD| / min diag(A^-1).  Normwise: max|dD| / max|D|
    and ||dD||_2 / ||D||_2.  And the per-cell error relative to the estimate
    itself, |dD_i| / |D_i|, which is ill-posed: with a few dozen probes D
    crosses zero at some cells, so it is large there whatever the precision."""
    import numpy as np
    d = np.abs(D - ref)
    rel = d / np.abs(ref)
    out = {"cell_max_rel_bound_all_cells": float(d.max() / DIAG_MIN)}
    if len(D) == 256 ** 3 and os.path.exists(EXACT_DIAG):
        with np.load(EXACT_DIAG) as z:
            idx = np.concatenate([z["idx_strided"], z["idx_random"]])
            dg = np.concatenate([z["diag_strided"], z["diag_random"]])
        r = d[idx] / dg
        out.update({"cell_max_rel_vs_exact_diag": float(r.max()),
                    "cell_rms_rel_vs_exact_diag": float(np.sqrt(np.mean(r * r))), "exact_diag_cells": len(idx)})
    out.update({"normwise_max": float(d.max() / np.abs(ref).max()),
                "normwise_rms": float(np.sqrt(np.sum(d * d) / np.sum(ref * ref))),
                "cell_max_rel_to_estimate": float(rel.max()),
                "cell_rms_rel_to_estimate": float(np.sqrt(np.mean(rel * rel)))})
    return out


def estimator_run(grid, progress=False, maxiter=MAXITER, **kw):
    import numpy as np
    import sedi_amd
    est = sedi_amd.Estimator(grid, ignore_errors=True, maxiter=maxiter, **kw)
    if progress:  # a line per sample on stderr (long runs)
        est.set_log(lambda line: line.startswith("Solving") and print(line, file=sys.stderr, flush=True))
    t0 = time.perf_counter()
    r = est.run()
    dt = time.perf_counter() - t0
    est.close()
    it = r["iters"]
    r["wall"] = dt
    r["summary"] = {"probe_solves_per_s": round(len(it) / dt, 4), "seconds": round(dt, 3),
                    "iters_min": int(it.min()), "iters_mean": round(float(it.mean()), 2),
                    "iters_max": int(it.max()), "unconverged": int(np.count_nonzero(r["codes"]))}
    return r


def run_ksweep(args):
    import sedi_amd
    samples, seed, tol = BATCH["samples"], BATCH["seed"], BATCH["tol"]
    out = {"mode": "ksweep", "grid": list(GRID), "samples": samples, "seed": seed, "tol": tol, "runs": []}
    ref = None
    for prec in (sedi_amd.FP64, sedi_amd.FP32_MIXED):
        for k in (8, 16, 32, 64):
            r = estimator_run(GRID, kind=2, samples=samples, tol=tol, seed=seed, precision=prec, batch_k=k,
                              rep_every=BATCH["rep_every"])
            line = dict(r["summary"], precision="fp64" if prec == sedi_amd.FP64 else "mixed", k=k,
                        iters=[int(v) for v in r["iters"]])
            if prec == sedi_amd.FP64 and k == 64:
                ref = r["D"]
            if prec == sedi_amd.FP32_MIXED and ref is not None:
                line["D_vs_fp64"] = d_errors(r["D"], ref)
            print(json.dumps(line), flush=True)
            out["runs"].append({kk: v for kk, v in line.items() if kk != "iters"})
    return out


ARITH_LIB = os.path.join(ROOT, "sedi_amd", "variants", "libsedi_amd_f32compute.so")
TUNING_TABLE = os.path.join(ROOT, "sedi_amd", "tuning.txt")


def run_arith(args):
    """The arithmetic type of the mixed pipelined CG and BiCGStab on the 512^3
    configuration (tol 1e-8, seed 23, k = 8, first --arith-samples probes):
    the default library (FP32 storage, stencil and updates computed in FP64
    registers) beside the f32compute build (the same kernels computing in
    FP32 registers; a child process loads it through SEDI_AMD_LIB) and the
    plain FP64 estimator on the same probes.  Reported per form: iterations,
    unconverged probes, probe-solves/s and D against FP64."""
    import numpy as np
    import sedi_amd
    import tempfile
    grid = (512, 512, 512)
    ns = args.arith_samples
    common = dict(samples=ns, tol=1e-8, seed=23)
    out = {"mode": "arith", "grid": list(grid), "samples": ns, "seed": 23, "tol": 1e-8, "batch_k": 8,
           "maxiter": 4000, "runs": []}
    if not os.path.exists(ARITH_LIB):
        raise RuntimeError("%s is missing: build it with sedi_amd.build.build_variant('f32compute')" % ARITH_LIB)
    for kind in (3, 1):
        name = SOLVER_LABEL[kind]
        with tempfile.TemporaryDirectory() as tmp:
            dpath = os.path.join(tmp, "D.npy")
            # the variant lives in its own folder: point it at the package's
            # tuning table, or it would run the untuned default plans
            env = dict(os.environ, SEDI_AMD_LIB=ARITH_LIB, SEDI_TUNING_FILE=TUNING_TABLE)
            res = subprocess.run([sys.executable, os.path.abspath(__file__), "--mode", "arith", "--arith-child",
                                  str(kind), "--arith-out", dpath, "--arith-samples", str(ns)],
                                 env=env, stdout=subprocess.PIPE, check=True)
            f32 = json.loads(res.stdout.decode().strip().splitlines()[-1])
            D32 = np.load(dpath)
        m = estimator_run(grid, progress=True, maxiter=4000, kind=kind, precision=sedi_amd.FP32_MIXED, batch_k=8,
                          rep_every=50, **common)
        print(json.dumps(dict(m["summary"], solver=name, arithmetic="fp64 registers")), flush=True)
        p = estimator_run(grid, progress=True, maxiter=4000, kind=kind, **common)
        print(json.dumps(dict(p["summary"], solver=name, precision="fp64 plain")), flush=True)
        out["runs"].append({"solver": name, "fp64_registers": m["summary"], "fp32_registers": f32,
                            "fp64_plain": p["summary"], "D_vs_fp64": {"fp64_registers": d_errors(m["D"], p["D"]),
                                                                      "fp32_registers": d_errors(D32, p["D"])}})
        print(json.dumps(out["runs"][-1]), flush=True)
        del m, p
    return out


def arith_child(args):
    """run_arith's child: the mixed estimator of one solver with the library
    SEDI_AMD_LIB names; prints its summary, saves D to --arith-out."""
    import numpy as np
    import sedi_amd
    if os.path.realpath(sedi_amd.lib_path) != os.path.realpath(ARITH_LIB):
        raise RuntimeError("the f32compute child loaded %s" % sedi_amd.lib_path)
    kind = args.arith_child
    r = estimator_run((512, 512, 512), progress=True, maxiter=4000, kind=kind, samples=args.arith_samples,
                      tol=1e-8, seed=23, precision=sedi_amd.FP32_MIXED, batch_k=8, rep_every=50)
    np.save(args.arith_out, r["D"])
    print(json.dumps(dict(r["summary"], solver=SOLVER_LABEL[kind], arithmetic="fp32 registers")), flush=True)


def run_c512(args):
    import sedi_amd
    grid = (512, 512, 512)
    out = {"mode": "c512", "grid": list(grid), "samples": 32, "seed": 23, "tol": 1e-8, "batch_k": 16,
           "maxiter": 4000, "runs": []}
    fp64_D = {}
    for kind, name in ((3, "PipelinedCG"), (1, "BiCGStab")):
        if args.c512_only and NAMES[args.c512_only] != kind:
            continue
        if args.c512_compaction:
            ab = {}
            for on in (False, True):
                sedi_amd.batch_compact_enable(on)
                try:
                    ab[on] = estimator_run(grid, progress=True, maxiter=4000, kind=kind, samples=32, tol=1e-8,
                                           seed=23, precision=sedi_amd.FP32_MIXED, batch_k=16, rep_every=50)
                finally:
                    sedi_amd.batch_compact_enable(not args.no_compact)
                print(json.dumps(dict(ab[on]["summary"], solver=name, precision="mixed", compaction=on)),
                      flush=True)
            out["runs"].append({"solver": name, "mixed_no_compaction": ab[False]["summary"],
                                "mixed_compaction": ab[True]["summary"],
                                "compaction_speedup": round(ab[True]["summary"]["probe_solves_per_s"] /
                                                            ab[False]["summary"]["probe_solves_per_s"], 4),
                                "D_bitwise_equal": bool((ab[True]["D"] == ab[False]["D"]).all()),
                                "iters_equal": bool((ab[True]["iters"] == ab[False]["iters"]).all())})
            del ab
            continue
        m = estimator_run(grid, progress=True, maxiter=4000, kind=kind, samples=32, tol=1e-8, seed=23,
                          precision=sedi_amd.FP32_MIXED, batch_k=16, rep_every=50)
        print(json.dumps(dict(m["summary"], solver=name, precision="mixed",
                              iters=[int(v) for v in m["iters"]])), flush=True)
        if args.c512_mixed_only:
            out["runs"].append({"solver": name, "mixed": m["summary"]})
            continue
        p = estimator_run(grid, progress=True, maxiter=4000, kind=kind, samples=32, tol=1e-8, seed=23)
        print(json.dumps(dict(p["summary"], solver=name, precision="fp64 plain")), flush=True)
        out["runs"].append({"solver": name, "mixed": m["summary"], "fp64_plain": p["summary"],
                            "speedup": round(m["summary"]["probe_solves_per_s"] /
                                             p["summary"]["probe_solves_per_s"], 3),
                            "D_vs_fp64": d_errors(m["D"], p["D"])})
        fp64_D[name] = p["D"]
        del m, p
    if len(fp64_D) == 2:
        # how far two FP64 solvers' estimates lie apart at this tolerance: the
        # scale against which a mixed result's distance from FP64 is read
        out["fp64_pcg_vs_fp64_bicgstab_D"] = d_errors(fp64_D["BiCGStab"], fp64_D["PipelinedCG"])
    return out


SERVE_SIDES = (64, 96, 128, 192, 256)
SERVE_TOLS = (1e-6, 1e-8, 1e-9)
SERVE_KINDS = (2, 3, 1)


def serve_stream(jobs, rate, seed=0):
    """The job stream: arrival times and job parameters (seed = job id)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=jobs))
    out = []
    for i in range(jobs):
        out.append(dict(t=float(t[i]), side=int(rng.choice(SERVE_SIDES)), samples=int(rng.integers(8, 65)),
                        tol=float(rng.choice(SERVE_TOLS)), kind=int(rng.choice(SERVE_KINDS)), seed=i))
    return out


def run_serve(jobs, rate, progress=True):
    """Submits the stream in real time from a thread and waits for every job."""
    import threading
    import numpy as np
    import sedi_amd
    stream = serve_stream(jobs, rate)
    srv = sedi_amd.Server(hbm_budget=200e9, precision=sedi_amd.FP32_MIXED)
    ids = [None] * len(stream)
    ready = [threading.Event() for _ in stream]

    def submitter():
        t0 = time.perf_counter()
        for i, j in enumerate(stream):
            dt = j["t"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids[i] = srv.submit((j["side"],) * 3, kind=j["kind"], samples=j["samples"], tol=j["tol"],
                                maxiter=MAXITER, seed=j["seed"])
            ready[i].set()

    th = threading.Thread(target=submitter)
    t0 = time.perf_counter()
    th.start()
    lat, probes, failed, first_submit, last_done = [], 0, 0, None, 0.0
    per_side = {}
    for i, j in enumerate(stream):
        ready[i].wait()
        r = srv.wait(ids[i])
        tm = r["times"]
        first_submit = tm["submit"] if first_submit is None else min(first_submit, tm["submit"])
        last_done = max(last_done, tm["done"])
        if r["status"] != 0:
            failed += 1
            continue
        lat.append(tm["done"] - tm["submit"])
        probes += j["samples"]
        per_side.setdefault(j["side"], []).append(tm["done"] - tm["submit"])
        if progress and (i + 1) % 10 == 0:
            print("serve: %d/%d jobs done, %.1f s" % (i + 1, len(stream), time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    th.join()
    st = srv.stats()
    srv.close()
    lat = np.array(lat)
    span = last_done - first_submit
    return {"jobs": len(stream), "rate_jobs_per_s": rate, "failed": failed,
            "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
            "mean_latency_s": round(float(lat.mean()), 3),
            "probe_solves_per_s": round(probes / span, 3), "span_s": round(span, 2),
            "offered_span_s": round(stream[-1]["t"], 2), "peak_device_GB": round(st["peak_bytes"] / 1e9, 2),
            "p50_latency_by_side_s": {str(k): round(float(np.percentile(v, 50)), 3)
                                      for k, v in sorted(per_side.items())}}


def main():
    args = parse()
    if args.pmc_workload:
        pmc_workload()
        return
    if args.stencil:
        print(json.dumps(stencil_only(side=args.stencil_side or None, cold=args.cold)), flush=True)
        return
    if args.copy:
        print(json.dumps(copy_only(side=args.stencil_side or None, cold=args.cold)), flush=True)
        return
    if args.no_compact:
        import sedi_amd
        sedi_amd.batch_compact_enable(False)
    if args.no_graphs:
        import sedi_amd
        sedi_amd.graphs_enable(False)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the whole grid on a single MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.mode == "arith" and args.arith_child:
        arith_child(args)
        return
    if args.mode == "serve":
        runs = []
        for rate in [float(v) for v in args.rate.split(",")]:
            runs.append(run_serve(args.jobs, rate))
            print(json.dumps(runs[-1]), file=sys.stderr, flush=True)
        line = dict(mode="serve", runs=runs,
                    workload="the serving stream (Poisson arrivals; sides 64/96/128/192/256, 8-64 probes, "
                             "tol 1e-6/1e-8/1e-9, CG/PipelinedCG/BiCGStab uniform, seed = job id; FP32 "
                             "storage / FP64 reductions; 200 GB budget), first %d jobs at each rate" % args.jobs)
    elif args.mode == "ksweep":
        line = run_ksweep(args)
    elif args.mode == "c512":
        line = run_c512(args)
    elif args.mode == "arith":
        line = run_arith(args)
    elif args.mode == "plain":
        line = run_plain(args)
    else:
        line = run_headline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
