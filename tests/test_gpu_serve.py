"""GPU: the continuous-batching estimation server (sedi_server_*).

Bars: a served job's diagonal estimate, iteration counts and codes equal the
batched estimator's for the same job bitwise, whether the job runs alone or
mixed with jobs of other grids, solvers and tolerances (a job's probes are
accumulated in probe order; its blocks' arithmetic does not depend on the
other blocks in its pool); jobs are admitted within the memory budget; a job
that arrives while a long one runs finishes before it (continuous batching,
not a queue).
"""
import time

import numpy as np
import pytest

sedi = pytest.importorskip("sedi_amd")

pytestmark = pytest.mark.gpu

JOBS = [  # grid, kind, samples, tol, seed
    ((24, 20, 18), sedi.CG, 6, 1e-9, 5),
    ((16, 16, 16), sedi.PCG, 9, 1e-8, 6),
    ((24, 20, 18), sedi.CG, 5, 1e-6, 7),
    ((12, 14, 10), sedi.BICGSTAB, 8, 1e-8, 8),
    ((16, 16, 16), sedi.CG, 4, 1e-9, 9),
    ((24, 20, 18), sedi.PCG, 3, 1e-9, 10),
]


def reference(grid, kind, samples, tol, seed, precision=sedi.FP32_MIXED):
    return sedi.estimate(grid, kind=kind, samples=samples, maxiter=2000, tol=tol, seed=seed,
                         precision=precision, batch_k=8, ignore_errors=True)


def check_same(res, ref):
    assert res["status"] == 0
    assert np.array_equal(res["iters"], ref["iters"]), (res["iters"], ref["iters"])
    assert np.array_equal(res["codes"], ref["codes"])
    assert np.array_equal(res["D"], ref["D"])


def test_single_job_equals_estimator():
    srv = sedi.Server(hbm_budget=8e9)
    try:
        grid, kind, s, tol, seed = JOBS[0]
        jid = srv.submit(grid, kind=kind, samples=s, tol=tol, seed=seed)
        res = srv.wait(jid, timeout=120)
    finally:
        srv.close()
    check_same(res, reference(grid, kind, s, tol, seed))
    t = res["times"]
    assert t["submit"] <= t["admit"] <= t["start"] <= t["done"]


def test_same_alone_and_in_the_mix():
    alone = {}
    for j in (0, 3):
        srv = sedi.Server(hbm_budget=8e9)
        try:
            grid, kind, s, tol, seed = JOBS[j]
            alone[j] = srv.wait(srv.submit(grid, kind=kind, samples=s, tol=tol, seed=seed), timeout=120)
        finally:
            srv.close()
    srv = sedi.Server(hbm_budget=8e9, max_slots=8)
    try:
        ids = [srv.submit(g, kind=k, samples=s, tol=t, seed=sd) for g, k, s, t, sd in JOBS]
        mixed = [srv.wait(i, timeout=300) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    assert st["done"] == len(JOBS) and st["probes"] == sum(j[2] for j in JOBS)
    for j in (0, 3):
        check_same(mixed[j], alone[j])
    for (g, k, s, t, sd), res in zip(JOBS, mixed):
        check_same(res, reference(g, k, s, t, sd))


def test_fp64_server_matches_fp64_estimator():
    srv = sedi.Server(hbm_budget=8e9, precision=sedi.FP64)
    try:
        grid, kind, s, tol, seed = JOBS[1]
        res = srv.wait(srv.submit(grid, kind=kind, samples=s, tol=tol, seed=seed), timeout=120)
    finally:
        srv.close()
    check_same(res, reference(grid, kind, s, tol, seed, precision=sedi.FP64))


def test_admission_within_budget():
    # room for about one pool of this grid and a job's accumulators at a time
    grid = (48, 48, 48)
    n = 48 ** 3
    per_rhs = n * (4 * 5 + 8) + 10000 * 8
    budget = 16 * per_rhs * 1.2 + 3 * 16 * n
    srv = sedi.Server(hbm_budget=budget, max_slots=16)
    try:
        ids = [srv.submit(grid, kind=sedi.CG, samples=8, tol=1e-8, seed=20 + i) for i in range(4)]
        out = [srv.wait(i, timeout=300) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    assert all(r["status"] == 0 for r in out)
    assert st["peak_bytes"] <= budget
    for i, r in enumerate(out):
        check_same(r, reference(grid, sedi.CG, 8, 1e-8, 20 + i))


def test_job_larger_than_the_budget_fails_cleanly():
    srv = sedi.Server(hbm_budget=1e6)
    try:
        res = srv.wait(srv.submit((64, 64, 64), kind=sedi.CG, samples=4, tol=1e-8, seed=1), timeout=60)
        ok = srv.wait(srv.submit((4, 4, 4), kind=sedi.CG, samples=4, tol=1e-8, seed=1), timeout=60)
    finally:
        srv.close()
    assert res["status"] < 0
    assert ok["status"] == 0


def test_continuous_batching_overlaps_jobs():
    """A small job submitted while a large one runs finishes first."""
    srv = sedi.Server(hbm_budget=40e9, max_slots=16)
    try:
        big = srv.submit((128, 128, 128), kind=sedi.CG, samples=32, tol=1e-10, seed=3)
        time.sleep(0.05)
        small = srv.submit((16, 16, 16), kind=sedi.CG, samples=4, tol=1e-8, seed=4)
        rs = srv.wait(small, timeout=300)
        rb = srv.wait(big, timeout=600)
    finally:
        srv.close()
    assert rs["status"] == 0 and rb["status"] == 0
    assert rs["times"]["done"] < rb["times"]["done"]
    check_same(rs, reference((16, 16, 16), sedi.CG, 4, 1e-8, 4))
