"""Offline autotuner: measures, on the GPU it runs on, the z chunk (planes per
tile, i.e. the unroll of the march) of every solver kernel and the batch size
k of the estimator for a set of grids, and writes the table the library reads
at run time (``sedi_amd/tuning.txt``; see csrc/tuning.h).

    python -m sedi_amd.autotune [--sides 64,128,256,512] [--iters 40] [--k-iters 200] [--out sedi_amd/tuning.txt]

1. z chunk and schedule: for each of 8, 16 and 32 planes and each schedule
   (0: the resident workgroups walk XCD-contiguous runs of tiles; 1: up to 8
   workgroups per CU take tiles in launch order; 3 and 4: as 1, with the tiles
   dealt to the XCDs in chunks of 4 or 8 rows of tiles, so a tile's y
   neighbours run on its XCD and its halo rows hit that XCD's L2), a child process runs every
   solver (plain FP64 and batched mixed, CG / pipelined CG / BiCGStab) for a
   fixed number of iterations on each grid, and applies the standalone
   stencil, with that chunk and schedule forced on every kernel
   (``zc:* 0 0 0 Z``, ``sched:* 0 0 0 S``), timing each kernel with HIP
   events; each kernel then gets the pair of least mean launch time on each
   grid.
2. batch size: with those chunks, the batched estimator runs k = 8..64 probes
   (as memory allows) for a fixed number of iterations; each (solver, precision,
   grid) gets the k of the highest probe-iterations per second.  In FP64 the
   plain solver (k = 1) competes too, over at least PLAIN_ITERS iterations: it
   generates each probe sequentially on a side stream while the previous probe
   solves, which real solves hide but short timing runs would not (at 256^3 a
   probe takes ~48 ms, 240 plain iterations).

3. the standalone stencil's plan: the one-cell march with chunk (8, 16, 32) x
   schedule (0, 1, and 2: one workgroup per tile, which only a kernel without
   partial sums can use) x planes of loads in flight (2, 4, 8), and the pair
   march (two cells per lane, 16-byte accesses, one 16-wave workgroup per CU)
   with chunk (8, 16, 32) x schedule (0, 3, 4) x planes in flight (2, 4), each
   forced in a child process and timed cold -- launches rotate through (x, y)
   pairs spanning 1 GiB of device memory, so x comes from HBM as in a sweep
   over vectors larger than the Infinity Cache -- replacing step 1's stencil
   entries.  At 512^3 a launch's time depends on where x and y lie in
   physical memory, so the pairs place y at different offsets from x.

``--k-only`` re-measures the batch sizes alone, with the chunk and schedule
entries of the existing table (--out), which it keeps; ``--stencil-only``
re-measures step 3 alone and keeps everything else.

The table is data: a run's results depend only on it and the grid, so the
same table gives the same results on every run.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

PKG = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(PKG) not in sys.path:
    sys.path.insert(0, os.path.dirname(PKG))

KINDS = {2: "CG", 3: "PipelinedCG", 1: "BiCGStab"}
# kernels whose launches the chunk decides (the replacement and set-up passes
# run rarely and follow the occupancy plan)
PLAIN_KERNELS = {2: ["cg_direction_kernel", "cg_update_kernel"], 3: ["pcg_iter_kernel"],
                 1: ["bicgstab_direction_kernel", "bicgstab_half_kernel", "bicgstab_update_kernel"]}
BATCH_KERNELS = {2: ["batch_dir_kernel", "batch_upd_kernel"], 3: ["batch_pcg_iter_kernel"],
                 1: ["batch_bicg_dir_kernel", "batch_bicg_half_kernel", "batch_bicg_upd_kernel"]}
# device bytes per right-hand side and cell of a batch (FP32 storage + FP64
# solution for the mixed form), for sizing k to the memory: CG b, x_lo, r x2,
# p x4; pipelined CG b, x_lo, r, u, w x2, z, q, s, p x4; BiCGStab b, x_lo, r x2,
# r^, p x2, v x2
BYTES_MIXED = {2: 40, 3: 60, 1: 44}
BYTES_FP64 = {2: 64, 3: 104, 1: 72}
K_CHOICES = (8, 16, 32, 64)
PLAIN_ITERS = 600


def device_bytes():
    try:
        import torch
        free, total = torch.cuda.mem_get_info()
        return float(free)
    except Exception:
        return 250e9


def k_max(side, kind, mixed, free):
    n = side ** 3
    per = (BYTES_MIXED if mixed else BYTES_FP64)[kind] * n
    if not mixed and n * 32 >= 2 ** 32:  # an FP64 4-RHS block must stay below 4 GiB
        return 0
    return int(0.6 * free / per) // 4 * 4


def run_est(side, kind, samples, iters, precision, batch_k):
    import sedi_amd
    est = sedi_amd.Estimator((side,) * 3, kind=kind, samples=samples, maxiter=iters, tol=1e-30, seed=1,
                             ignore_errors=True, precision=precision, batch_k=batch_k)
    try:
        r = est.run(want_D=False)
    finally:
        est.close()
    return r["seconds"], float(r["iters"].sum())


def measure_chunks(sides, iters):
    """Child: per grid, the mean launch time of every tuned kernel."""
    import sedi_amd
    free = device_bytes()
    out = {}
    for side in sides:
        res = {}
        for kind in KINDS:
            for mixed in (False, True):
                names = (BATCH_KERNELS if mixed else PLAIN_KERNELS)[kind]
                k = min(64, k_max(side, kind, True, free)) if mixed else 1
                if mixed and k < 4:
                    continue
                prec = sedi_amd.FP32_MIXED if mixed else sedi_amd.FP64
                run_est(side, kind, k, 2, prec, k)  # first launches load the code objects
                sedi_amd.profile_reset()
                sedi_amd.profile_enable("*")
                run_est(side, kind, k, iters, prec, k)
                sedi_amd.profile_enable(None)
                for name in names:
                    try:
                        launches, ms = sedi_amd.profile_get(name)
                        res[name] = 1e3 * ms / launches
                    except sedi_amd.SediError:
                        pass
        import numpy as np
        x = np.random.default_rng(0).standard_normal(side ** 3)
        st = sedi_amd.Stencil()
        st.apply((side,) * 3, x)
        sedi_amd.profile_reset()
        sedi_amd.profile_enable("stencil_apply_kernel")
        for _ in range(10):
            st.apply((side,) * 3, x)
        sedi_amd.profile_enable(None)
        launches, ms = sedi_amd.profile_get("stencil_apply_kernel")
        res["stencil_apply_kernel"] = 1e3 * ms / launches
        out[side] = res
        print("  side %d: %s" % (side, json.dumps({k: round(v, 1) for k, v in res.items()})), file=sys.stderr,
              flush=True)
    return out


# (pairs, chunk, schedule, planes in flight): the one-cell march, then the pair march
STENCIL_PLANS = ([(0, zc, sched, pf) for zc in (8, 16, 32) for sched in (0, 1, 2) for pf in (2, 4, 8)] +
                 [(1, zc, sched, pf) for zc in (8, 16, 32) for sched in (0, 3, 4) for pf in (2, 4)])
# schedules of the solver kernels (march.h tuned_geometry): 0 XCD-contiguous
# runs, 1 launch order, 3 and 4 launch order with tiles dealt to the XCDs in
# chunks of 4 or 8 rows of tiles
SCHEDS = (0, 1, 3, 4)


def measure_stencil(sides, reps):
    """Child: per grid, the mean launch time of the standalone stencil, cold
    (back-to-back launches rotate through at least four (x, y) pairs spanning
    1 GiB, y at different offsets from x)."""
    import ctypes
    import numpy as np
    import sedi_amd
    out = {}
    L = sedi_amd.lib()
    st = sedi_amd.Stencil()._c()
    # largest grid first: its buffers are then laid out as in a fresh process
    # (after the many small pairs of a small grid were allocated and freed,
    # 512^3 plans measured up to 10% faster here than they run in one)
    for side in sorted(sides, reverse=True):
        n = side ** 3
        g, _ = sedi_amd._grid((side,) * 3)
        # at least four pairs: at 512^3 a launch's time depends on where its
        # pair's buffers lie in physical memory (the offset of y from x), so
        # plans are compared on the mean over several placements
        npairs = max(4, -(-(1 << 30) // (16 * n)))
        x0 = np.random.default_rng(0).standard_normal(n)
        # y of pair k < 4 after a k * 34 MiB spacer: the pairs place y at
        # different offsets from x, and a plan is judged over all of them
        pairs, spacers = [], []
        for k in range(npairs):
            x = sedi_amd.DeviceArray(x0)
            if 0 < k < 4:
                spacers.append(sedi_amd.DeviceArray(n=k * (34 << 20) // 8))
            pairs.append((x, sedi_amd.DeviceArray(n=n)))
        for x, y in pairs:
            if L.sedi_stencil_apply(ctypes.byref(g), ctypes.byref(st), x.ptr, y.ptr) < 0:
                raise RuntimeError(L.sedi_last_error().decode())
        # back-to-back launches, as in a solver's sweep: a launch also pays for
        # writing back what the previous one left dirty in the Infinity Cache
        # (timing launches one at a time, with the device idle in between,
        # ranked the plans differently and 7-12% faster than they sustain)
        sedi_amd.profile_reset()
        sedi_amd.profile_enable("stencil_apply_kernel")
        for r in range(reps):
            x, y = pairs[r % npairs]
            L.sedi_stencil_apply(ctypes.byref(g), ctypes.byref(st), x.ptr, y.ptr)
        sedi_amd.profile_enable(None)
        launches, ms = sedi_amd.profile_get("stencil_apply_kernel")
        out[side] = 1e3 * ms / launches
        for a in [v for p in pairs for v in p] + spacers:
            a.free()
    return out


def stencil_lines(args):
    """Step 3: the stencil's best (march, chunk, schedule, prefetch) per grid."""
    t = {}
    for x2, zc, sched, pf in STENCIL_PLANS:
        print("stencil: %s march, z chunk %d, schedule %d, prefetch %d" % ("pair" if x2 else "one-cell", zc, sched,
                                                                         pf), file=sys.stderr, flush=True)
        t[(x2, zc, sched, pf)] = child(args, ["x2:stencil_apply_kernel 0 0 0 %d" % x2,
                                              "zc:stencil_apply_kernel 0 0 0 %d" % zc,
                                              "sched:stencil_apply_kernel 0 0 0 %d" % sched,
                                              "pf:stencil_apply_kernel 0 0 0 %d" % pf], "stencil")
    lines = []
    for side in args.sides:
        tt = {c: t[c][str(side)] for c in t}
        x2, zc, sched, pf = min(tt, key=tt.get)
        best = sorted(tt, key=tt.get)[:4]
        note = " ".join("%s%d/%d/%d:%.1f" % ("p" if c[0] else "", c[1], c[2], c[3], tt[c]) for c in best)
        one = min((c for c in tt if not c[0]), key=tt.get)
        lines.append("x2:stencil_apply_kernel %d %d %d %d   # cold us per launch (p: pair march; chunk/schedule/"
                     "prefetch), best four: %s; best one-cell plan %d/%d/%d:%.1f" % (
                         side, side, side, x2, note, one[1], one[2], one[3], tt[one]))
        lines.append("zc:stencil_apply_kernel %d %d %d %d" % (side, side, side, zc))
        lines.append("sched:stencil_apply_kernel %d %d %d %d" % (side, side, side, sched))
        lines.append("pf:stencil_apply_kernel %d %d %d %d" % (side, side, side, pf))
    return lines


def measure_batches(sides, iters):
    """Child: per (solver, precision, grid), probe-iterations/s for each k."""
    import sedi_amd
    free = device_bytes()
    out = []
    for side in sides:
        for kind in KINDS:
            for mixed in (True, False):
                prec = sedi_amd.FP32_MIXED if mixed else sedi_amd.FP64
                rates = {}
                cands = ([1] if not mixed else []) + [k for k in K_CHOICES if k <= k_max(side, kind, mixed, free)]
                for k in cands:
                    # untimed first run (code objects, graphs), then the best of two
                    run_est(side, kind, max(k, 4), 2, prec, k)
                    best = 0.0
                    for _ in range(2):
                        secs, its = run_est(side, kind, max(k, 4), max(iters, PLAIN_ITERS) if k == 1 else iters,
                                            prec, k)
                        best = max(best, its / secs)
                    rates[k] = best
                if rates:
                    best = max(rates, key=rates.get)
                    out.append(dict(side=side, kind=kind, precision=prec, k=best,
                                    rates={str(k): round(v, 1) for k, v in rates.items()}))
                    print("  side %d %s %s: k=%d %s" % (side, KINDS[kind], "mixed" if mixed else "fp64", best,
                                                         out[-1]["rates"]), file=sys.stderr, flush=True)
    return out


def child(args, table_lines, mode):
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        f.write("\n".join(table_lines) + "\n")
        path = f.name
    env = dict(os.environ, SEDI_TUNING_FILE=path)
    iters = args.iters if mode == "chunks" else args.k_iters
    cmd = [sys.executable, "-m", "sedi_amd.autotune", "--child", mode, "--sides",
           ",".join(str(s) for s in args.sides), "--iters", str(iters)]
    try:
        res = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, cwd=os.path.dirname(PKG), check=True)
    finally:
        os.unlink(path)
    return json.loads(res.stdout.decode().strip().splitlines()[-1])


def main():
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--sides", default="64,128,256,512", type=lambda s: [int(v) for v in s.split(",")])
    p.add_argument("--iters", type=int, default=40, help="iterations per kernel timing")
    p.add_argument("--k-iters", type=int, default=200,
                   help="iterations per batch-size timing (long enough that probe generation, which the "
                        "plain FP64 path does one probe at a time, is a small part)")
    p.add_argument("--out", default=os.path.join(PKG, "tuning.txt"))
    p.add_argument("--k-only", action="store_true",
                   help="re-measure the batch sizes only, keeping the table's chunk and schedule entries")
    p.add_argument("--stencil-only", action="store_true",
                   help="re-measure the standalone stencil's plan only, keeping every other entry")
    p.add_argument("--child", choices=["chunks", "batches", "stencil"], help=argparse.SUPPRESS)
    args = p.parse_args()
    if args.child == "chunks":
        print(json.dumps(measure_chunks(args.sides, args.iters)))
        return
    if args.child == "stencil":
        print(json.dumps(measure_stencil(args.sides, 16)))
        return
    if args.child == "batches":
        print(json.dumps(measure_batches(args.sides, args.iters)))
        return

    t0 = time.time()
    if args.stencil_only:
        with open(args.out) as f:
            kept = [l.rstrip("\n") for l in f if l.strip() and ":stencil_apply_kernel " not in l]
        head = [l for l in kept if l.startswith("#")]
        lines = [l for l in kept if not l.startswith("#")] + stencil_lines(args)
        head.append("# stencil entries re-measured by --stencil-only (cold, sides %s); %s; %.0f s" % (
            ",".join(str(v) for v in args.sides), time.strftime("%Y-%m-%d"), time.time() - t0))
        with open(args.out, "w") as f:
            f.write("\n".join(head + lines) + "\n")
        print("wrote %s (%d entries)" % (args.out, len(lines)), file=sys.stderr, flush=True)
        return
    if args.k_only:
        with open(args.out) as f:
            kept = [l.rstrip("\n") for l in f if l.strip() and not l.startswith("k:")]
        head = [l for l in kept if l.startswith("#")]
        lines = [l for l in kept if not l.startswith("#")]
        print("batch sizes", file=sys.stderr, flush=True)
        batches = child(args, [l.split("#")[0] for l in lines], "batches")
        for b in batches:
            s = b["side"]
            lines.append("k:%d:%d %d %d %d %d   # probe-iterations/s: %s" % (
                b["kind"], b["precision"], s, s, s, b["k"], " ".join("%s:%s" % kv for kv in b["rates"].items())))
        head.append("# k entries re-measured by --k-only --k-iters %d; %s; %.0f s" % (
            args.k_iters, time.strftime("%Y-%m-%d"), time.time() - t0))
        with open(args.out, "w") as f:
            f.write("\n".join(head + lines) + "\n")
        print("wrote %s (%d entries)" % (args.out, len(lines)), file=sys.stderr, flush=True)
        return
    timings = {}
    for zc in (8, 16, 32):
        for sched in SCHEDS:
            print("z chunk %d, schedule %d" % (zc, sched), file=sys.stderr, flush=True)
            timings[(zc, sched)] = child(args, ["zc:* 0 0 0 %d" % zc, "sched:* 0 0 0 %d" % sched], "chunks")
    lines = []
    for side in args.sides:
        names = sorted(set().union(*(timings[c].get(str(side), {}).keys() for c in timings)))
        for name in names:
            if name == "stencil_apply_kernel":
                continue  # step 3
            t = {c: timings[c][str(side)][name] for c in timings if name in timings[c].get(str(side), {})}
            zc, sched = min(t, key=t.get)
            note = " ".join("%d/%d:%.1f" % (c[0], c[1], t[c]) for c in sorted(t))
            lines.append("zc:%s %d %d %d %d   # us per launch (chunk/schedule): %s" % (name, side, side, side, zc, note))
            lines.append("sched:%s %d %d %d %d" % (name, side, side, side, sched))
    print("batch sizes", file=sys.stderr, flush=True)
    batches = child(args, [l.split("#")[0] for l in lines], "batches")
    for b in batches:
        s = b["side"]
        lines.append("k:%d:%d %d %d %d %d   # probe-iterations/s: %s" % (
            b["kind"], b["precision"], s, s, s, b["k"], " ".join("%s:%s" % kv for kv in b["rates"].items())))
    lines += stencil_lines(args)
    dev = "unknown device"
    try:
        import torch
        dev = torch.cuda.get_device_name(0)
    except Exception:
        pass
    head = ["# sedi_amd tuning table, written by: python -m sedi_amd.autotune --sides %s --iters %d --k-iters %d" % (
                ",".join(str(s) for s in args.sides), args.iters, args.k_iters),
            "# device: %s; %s; %.0f s" % (dev, time.strftime("%Y-%m-%d"), time.time() - t0),
            "# <key> <n1> <n2> <n3> <value>; zc:<kernel> = planes per tile, sched:<kernel> = 1 for tiles in",
            "# launch order (0: XCD-contiguous runs; 2: one workgroup per tile, stencil only; 3/4: launch order",
            "# with tiles dealt to the XCDs in chunks of 4/8 tile rows), pf:<kernel> =",
            "# planes of loads in flight (stencil), x2:<kernel> = 1 for the pair march (stencil),",
            "# wpc:<kernel> = workgroups per CU (0: occupancy), k:<solver>:<precision> = probes per batch",
            "# (solver 1 BiCGStab, 2 CG, 3 PipelinedCG; precision 0 FP64, 1 FP32 storage)"]
    with open(args.out, "w") as f:
        f.write("\n".join(head + lines) + "\n")
    print("wrote %s (%d entries)" % (args.out, len(lines)), file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
