"""GPU: the batched multi-RHS CG (FP64 and mixed FP32/FP64) and the
jump-ahead probe batches, through the C ABI, against the oracle and the
compiled reference's fixtures.

Bars:
  * probe batches: bit-exact with the oracle's sequential mt19937 stream;
  * FP64 batch: per right-hand side it runs the Cg.hpp recurrences of the
    plain solver (dot products in a tree order), so fixed-iteration solutions
    and histories agree with the oracle to rounding (1e-12 normwise) and
    converged iteration counts within +-1;
  * mixed batch: converges to the tolerance on the TRUE residual
    ||b - A x|| / ||b|| (FP64 residual replacement), and D stays within 1e-6
    normwise of the FP64 estimate;
  * every right-hand side's result is bitwise the same whichever batch it is
    solved in (alone, at another slot, with other right-hand sides, other K).
"""
import numpy as np
import pytest

import oracle_lib as O
from conftest import d_bar, iteration_bar, normwise

sedi = pytest.importorskip("sedi_amd")

pytestmark = pytest.mark.gpu

S32 = (32, 32, 32)
N32 = 32 ** 3


def probes(seed, shape, k0, count):
    n = int(np.prod(shape))
    return np.stack([O.rademacher(seed, 2 * n * k, n) for k in range(k0, k0 + count)])


def true_rel_residual(shape, b, x):
    return np.linalg.norm(b - O.stencil_apply(shape, x)) / np.linalg.norm(b)


# ------------------------------------------------------------ probe batches --
@pytest.mark.parametrize("seed,shape,k0,count", [(1234, S32, 0, 8), (7, (13, 7, 5), 3, 6),
                                                 (5, (1, 1, 1), 0, 5), (11, (455, 1, 1), 2, 3),
                                                 (23, (40, 30, 20), 17, 4)])
def test_probe_batch_bitwise(seed, shape, k0, count):
    n = int(np.prod(shape))
    v = sedi.probe_batch(seed, n, k0, count)
    assert v.shape == (count, n)
    assert np.array_equal(v, probes(seed, shape, k0, count))


# ------------------------------------------------------------------ FP64 -----
def test_batch_fp64_fixed_iterations_vs_oracle():
    b = probes(1234, S32, 0, 8)
    bs = sedi.Batch(S32, 8, sedi.FP64)
    x, iters, errs, codes = bs.solve(b, maxiter=20, tol=1e-30)
    assert np.all(codes == 1) and np.all(iters == 20)
    for k in range(8):
        code, ox, ofe, ofi, oh = O.solve(O.CG, S32, b[k], maxiter=20, tol=1e-30)
        assert (code, ofi) == (1, 20)
        assert normwise(x[k], ox) < 1e-12
        h = bs.err_history(k)
        assert len(h) == 20
        assert np.max(np.abs(h - oh) / oh) < 1e-11
        assert abs(errs[k] - ofe) <= 1e-11 * ofe


def test_batch_fp64_converged_vs_oracle():
    b = probes(1234, S32, 0, 8)
    x, iters, errs, codes = sedi.Batch(S32, 8, sedi.FP64).solve(b, maxiter=500, tol=1e-9)
    assert np.all(codes == 0)
    for k in range(8):
        code, ox, ofe, ofi, oh = O.solve(O.CG, S32, b[k], maxiter=500, tol=1e-9)
        assert abs(int(iters[k]) - ofi) <= 1
        assert errs[k] <= 1e-9
        assert normwise(x[k], ox) < 1e-7
        assert true_rel_residual(S32, b[k], x[k]) < 2e-9


def test_batch_fp64_odd_grid_and_general_stencil():
    shape = (13, 7, 5)
    st = sedi.Stencil(6.5, 0.75, 1.25)
    b = probes(5, shape, 0, 3)
    x, iters, errs, codes = sedi.Batch(shape, 3, sedi.FP64, stencil=st).solve(b, maxiter=15, tol=1e-30)
    for k in range(3):
        code, ox, ofe, ofi, oh = O.solve(O.CG, shape, b[k], maxiter=15, tol=1e-30,
                                         abc=(6.5, 0.75, 1.25))
        assert (codes[k], iters[k]) == (code, ofi)
        assert normwise(x[k], ox) < 1e-12


@pytest.mark.parametrize("prec", [sedi.FP64, sedi.FP32_MIXED])
def test_batch_zero_rhs_and_maxiter(prec):
    shape = (8, 8, 8)
    b = np.stack([O.rademacher(3, 0, 512), np.zeros(512), O.rademacher(3, 1024, 512)])
    bs = sedi.Batch(shape, 4, prec)
    x, iters, errs, codes = bs.solve(b, maxiter=10, tol=1e-12)
    assert (codes[1], iters[1], errs[1]) == (0, 0, 0.0) and not x[1].any()
    assert len(bs.err_history(1)) == 0
    for m in (0, 1, 3):
        x, iters, errs, codes = bs.solve(b, maxiter=m, tol=1e-12)
        ocode, ox, ofe, ofi, oh = O.solve(O.CG, shape, b[0], maxiter=m, tol=1e-12)
        assert (codes[0], iters[0]) == (ocode, ofi) == (1, m)
        assert (codes[1], iters[1]) == (0, 0)


def test_batch_eigenvector_one_iteration():
    i = np.arange(1, 9)
    s = np.sin(np.pi * 2 * i / 9)
    phi = (s[:, None, None] * s[None, :, None] * s[None, None, :]).reshape(-1)
    x, iters, errs, codes = sedi.Batch((8, 8, 8), 4, sedi.FP64).solve(phi, maxiter=10, tol=1e-12)
    assert codes[0] == 0 and iters[0] == 1
    lam = 6.0 - 6 * np.cos(np.pi * 2 / 9)
    assert np.allclose(x[0], phi / lam, atol=1e-14)


@pytest.mark.parametrize("prec", [sedi.FP64, sedi.FP32_MIXED])
def test_update_pending_when_finishing_at_iteration_one(prec):
    """CG applies solution updates in groups; a right-hand side that finishes
    inside a group has updates pending.  Here the eigenvector finishes at
    iteration 1 while its block-mates go on (the next direction pass applies
    it), and alone in its block (the read-out applies it); the block-mates'
    results equal their solves alone bitwise."""
    shape = (8, 8, 8)
    i = np.arange(1, 9)
    s = np.sin(np.pi * 2 * i / 9)
    phi = (s[:, None, None] * s[None, :, None] * s[None, None, :]).reshape(-1)
    lam = 6.0 - 6 * np.cos(np.pi * 2 / 9)
    others = probes(9, shape, 0, 2)
    b = np.vstack([phi, others])
    bs = sedi.Batch(shape, 4, prec)
    x, iters, errs, codes = bs.solve(b, maxiter=200, tol=1e-10)
    assert np.all(codes == 0) and np.all(iters[1:] > 2)
    if prec == sedi.FP64:
        # the eigenvector is solved exactly by one step (FP32 storage rounds
        # phi, so the mixed form takes a few more, ending where it may)
        assert iters[0] == 1
        assert np.max(np.abs(x[0] - phi / lam)) < 1e-13 * np.max(np.abs(phi / lam))
    else:  # the mixed form solves with b rounded to FP32
        assert np.max(np.abs(x[0] - phi / lam)) < 1e-7 * np.max(np.abs(phi / lam))
    for k in (1, 2):
        x1, it1, e1, c1 = bs.solve(b[k], maxiter=200, tol=1e-10)
        assert np.array_equal(x1[0], x[k]) and (it1[0], c1[0]) == (iters[k], codes[k])
    x0, it0, e0, c0 = bs.solve(phi, maxiter=200, tol=1e-10)
    assert np.array_equal(x0[0], x[0]) and it0[0] == iters[0]


@pytest.mark.parametrize("kind", [O.CG, O.PCG])
@pytest.mark.parametrize("prec", [sedi.FP64, sedi.FP32_MIXED])
@pytest.mark.parametrize("m", [1, 2, 3, 4, 5])
def test_updates_pending_at_every_place_in_a_group(kind, prec, m):
    """CG and pipelined CG apply solution updates in groups of four
    iterations, so a right-hand side that finishes inside a group has one to
    three pending.  A sum of m eigenvectors with distinct eigenvalues is
    solved by m CG iterations in exact arithmetic (pipelined CG stops one
    later): in FP64 it finishes at every place in a group while two random
    probes of its block go on, and alone in its block.  Its solution matches
    the dense solve (FP64: 1e-12 normwise; mixed, which solves with b rounded
    to FP32 and takes more iterations: 1e-8 normwise against the dense solve
    of that rounded b), and every right-hand side's result is bitwise the
    same in the mix and alone."""
    shape = (8, 8, 8)
    n = 512
    i = np.arange(1, 9)
    rhs = np.zeros(n)
    for a in (1, 2, 3, 5, 7)[:m]:
        s = np.sin(np.pi * a * i / 9)
        rhs += (s[:, None, None] * s[None, :, None] * s[None, None, :]).reshape(-1)
    A = np.stack([O.stencil_apply(shape, e) for e in np.eye(n)], axis=1)
    b = np.vstack([rhs, probes(9, shape, 0, 2)])
    bs = sedi.Batch(shape, 4, prec, kind=kind)
    x, iters, errs, codes = bs.solve(b, maxiter=200, tol=1e-10)
    assert np.all(codes == 0) and np.all(iters[1:] > 6)
    if prec == sedi.FP64:
        assert iters[0] == m + (kind == O.PCG)
        assert normwise(x[0], np.linalg.solve(A, rhs)) < 1e-12
    else:
        assert normwise(x[0], np.linalg.solve(A, rhs.astype(np.float32).astype(np.float64))) < 1e-8
    for k in range(3):
        x1, it1, e1, c1 = bs.solve(b[k], maxiter=200, tol=1e-10)
        assert np.array_equal(x1[0], x[k]) and (it1[0], c1[0]) == (iters[k], codes[k])


# ------------------------------------------- pipelined CG and BiCGStab ------
KINDS = [O.CG, O.PCG, O.BICGSTAB]


@pytest.mark.parametrize("kind", [O.PCG, O.BICGSTAB])
def test_batch_fp64_other_solvers_fixed_iterations_vs_oracle(kind):
    b = probes(1234, S32, 0, 6)
    bs = sedi.Batch(S32, 8, sedi.FP64, kind=kind)
    x, iters, errs, codes = bs.solve(b, maxiter=20, tol=1e-30)
    for k in range(6):
        code, ox, ofe, ofi, oh = O.solve(kind, S32, b[k], maxiter=20, tol=1e-30)
        assert (codes[k], iters[k]) == (code, ofi) == (1, 20)
        # BiCGStab's rounding differences grow along its iterations
        bar = 1e-11 if kind == O.PCG else 1e-9
        assert normwise(x[k], ox) < bar
        h = bs.err_history(k)
        assert len(h) == len(oh) and np.max(np.abs(h - oh) / oh) < 10 * bar


@pytest.mark.parametrize("kind", [O.PCG, O.BICGSTAB])
def test_batch_fp64_other_solvers_converged_vs_oracle(ref_spread, kind):
    b = probes(1234, S32, 0, 8)
    x, iters, errs, codes = sedi.Batch(S32, 8, sedi.FP64, kind=kind).solve(b, maxiter=500, tol=1e-9)
    assert np.all(codes == 0)
    for k in range(8):
        code, ox, ofe, ofi, oh = O.solve(kind, S32, b[k], maxiter=500, tol=1e-9)
        if kind == O.PCG:
            assert abs(int(iters[k]) - ofi) <= 1
        else:  # the reference's own spread under summation-order changes
            assert abs(int(iters[k]) - ofi) <= iteration_bar(ref_spread, "c32", kind), (k, iters[k], ofi)
        assert true_rel_residual(S32, b[k], x[k]) < 3e-9


@pytest.mark.parametrize("kind", [O.PCG, O.BICGSTAB])
@pytest.mark.parametrize("prec", [sedi.FP64, sedi.FP32_MIXED])
def test_batch_other_solvers_zero_rhs_and_maxiter(kind, prec):
    shape = (8, 8, 8)
    b = np.stack([O.rademacher(3, 0, 512), np.zeros(512), O.rademacher(3, 1024, 512)])
    bs = sedi.Batch(shape, 4, prec, kind=kind)
    x, iters, errs, codes = bs.solve(b, maxiter=10, tol=1e-12)
    assert (codes[1], iters[1], errs[1]) == (0, 0, 0.0) and not x[1].any()
    for m in (0, 1, 3):
        x, iters, errs, codes = bs.solve(b, maxiter=m, tol=1e-12)
        ocode, ox, ofe, ofi, oh = O.solve(kind, shape, b[0], maxiter=m, tol=1e-12)
        assert (codes[0], iters[0]) == (ocode, ofi), (m, codes[0], iters[0], ocode, ofi)
        assert (codes[1], iters[1]) == (0, 0)


@pytest.mark.parametrize("kind", [O.PCG, O.BICGSTAB])
@pytest.mark.parametrize("shape,tol", [(S32, 1e-8), (S32, 1e-10), ((12, 10, 8), 1e-12)])
def test_batch_mixed_other_solvers_converge_on_true_residual(kind, shape, tol):
    b = probes(77, shape, 0, 8)
    x, iters, errs, codes = sedi.Batch(shape, 8, sedi.FP32_MIXED, kind=kind).solve(b, maxiter=2000, tol=tol)
    xd, itd, _, cd = sedi.Batch(shape, 8, sedi.FP64, kind=kind).solve(b, maxiter=2000, tol=tol)
    assert np.all(codes == 0), (iters, errs, codes)
    for k in range(8):
        t = true_rel_residual(shape, b[k], x[k])
        # the error reported is the replaced (true) residual's; pipelined CG
        # stops on the residual of the previous iteration, so its final x is
        # one update further and its residual smaller still
        assert errs[k] <= tol and t <= 1.05 * tol, (k, errs[k], t)
        # pipelined CG follows FP64's counts closely; BiCGStab's restarts at
        # its (rare) replacements cost some iterations, and its counts move by
        # +-20% under rounding alone
        slack = max(8, itd[k] // 5) if kind == O.PCG else max(8, itd[k] // 2)
        assert iters[k] <= itd[k] + slack, (iters[k], itd[k])
        assert normwise(x[k], xd[k]) < max(1e-6, 1e3 * tol)


@pytest.mark.parametrize("kind", [O.PCG, O.BICGSTAB])
def test_result_same_alone_and_in_the_mix_other_solvers(kind):
    shape = (20, 18, 16)
    b = probes(5, shape, 0, 6)
    x8, it8, e8, c8 = sedi.Batch(shape, 8, sedi.FP32_MIXED, kind=kind).solve(b, maxiter=600, tol=1e-9)
    bs1 = sedi.Batch(shape, 4, sedi.FP32_MIXED, kind=kind)
    for k in (1, 4):
        x1, it1, e1, c1 = bs1.solve(b[k], maxiter=600, tol=1e-9)
        assert np.array_equal(x1[0], x8[k]) and (it1[0], e1[0], c1[0]) == (it8[k], e8[k], c8[k])


@pytest.mark.parametrize("kind", [O.PCG, O.BICGSTAB])
def test_estimator_batched_other_solvers_vs_reference(ref32, ref_spread, kind):
    r = sedi.estimate(S32, kind=kind, samples=8, maxiter=500, tol=1e-9, seed=1234, batch_k=8)
    ref_it = ref32[f"s{kind}_iters"]
    assert r["status"] == 0 and np.all(r["codes"] == 0)
    bar = iteration_bar(ref_spread, "c32", kind)
    print("kind", kind, "iters", r["iters"].tolist(), "ref", ref_it.tolist(), "bar", bar, flush=True)
    assert np.all(np.abs(r["iters"] - ref_it) <= bar), (r["iters"], ref_it)
    assert normwise(r["D"], ref32[f"s{kind}_D"]) < d_bar(ref_spread, "c32", kind,
                                                         np.array_equal(r["iters"], ref_it))
    f = sedi.estimate(S32, kind=kind, samples=8, maxiter=20, tol=1e-9, seed=1234, batch_k=8,
                      ignore_errors=True)
    assert np.array_equal(f["iters"], ref32[f"s{kind}_fixed20_iters"])
    assert normwise(f["D"], ref32[f"s{kind}_fixed20_D"]) < 1e-10


@pytest.mark.parametrize("kind", [O.PCG, O.BICGSTAB])
def test_estimator_mixed_other_solvers_vs_fp64(kind):
    p = sedi.estimate(S32, kind=kind, samples=8, maxiter=500, tol=1e-10, seed=1234)
    m = sedi.estimate(S32, kind=kind, samples=8, maxiter=500, tol=1e-10, seed=1234,
                      precision=sedi.FP32_MIXED, batch_k=8)
    assert m["status"] == 0 and np.all(m["codes"] == 0)
    assert normwise(m["D"], p["D"]) < 1e-6


# ------------------------------------------------------------------ mixed ----
@pytest.mark.parametrize("shape,tol,rep", [(S32, 1e-8, 0), (S32, 1e-9, 0), (S32, 1e-9, 10), (S32, 1e-12, 0),
                                           ((12, 10, 8), 1e-8, 0), ((12, 10, 8), 1e-11, 0),
                                           ((12, 10, 8), 1e-13, 0)])
def test_batch_mixed_converges_on_true_residual(shape, tol, rep):
    # tolerances far below FP32's precision and grids that converge before
    # the first periodic replacement: the tenfold rule keeps the FP32
    # recursion close to the true residual
    b = probes(1234, shape, 0, 8)
    x, iters, errs, codes = sedi.Batch(shape, 8, sedi.FP32_MIXED, rep_every=rep).solve(
        b, maxiter=1000, tol=tol)
    xd, itd, _, _ = sedi.Batch(shape, 8, sedi.FP64).solve(b, maxiter=1000, tol=tol)
    assert np.all(codes == 0), (iters, errs)
    for k in range(8):
        # the reported error is the FP64 replacement's true residual
        t = true_rel_residual(shape, b[k], x[k])
        assert errs[k] <= tol and t <= 1.05 * tol, (k, errs[k], t)
        assert abs(t - errs[k]) <= 0.05 * tol
        # a few more iterations than FP64 at most
        assert iters[k] <= itd[k] + max(5, itd[k] // 10), (iters[k], itd[k])
        assert normwise(x[k], xd[k]) < max(1e-6, 1e3 * tol)


def test_batch_mixed_history_ends_with_true_residual():
    shape = (20, 16, 12)
    b = probes(9, shape, 0, 4)
    bs = sedi.Batch(shape, 4, sedi.FP32_MIXED, rep_every=7)
    x, iters, errs, codes = bs.solve(b, maxiter=500, tol=1e-9)
    for k in range(4):
        h = bs.err_history(k)
        assert len(h) == iters[k] and h[-1] == errs[k]
        # entries at multiples of 7 are replacement (true) residuals
        for it in range(7, iters[k] + 1, 7):
            xk, _, _, _ = sedi.Batch(shape, 4, sedi.FP32_MIXED, rep_every=7).solve(
                b[k], maxiter=it, tol=1e-30)
            t = true_rel_residual(shape, b[k], xk[0])
            assert abs(h[it - 1] - t) <= 1e-6 * t
            break


# ------------------------------------------------- independence of the mix ---
@pytest.mark.parametrize("prec", [sedi.FP64, sedi.FP32_MIXED])
def test_result_same_alone_and_in_the_mix(prec):
    shape = (24, 20, 18)
    b = probes(31, shape, 0, 9)
    x8, it8, e8, c8 = sedi.Batch(shape, 8, prec).solve(b[:8], maxiter=400, tol=1e-9)
    bs1 = sedi.Batch(shape, 1, prec)
    alone_hist = {}
    for k in (0, 3, 6):
        x1, it1, e1, c1 = bs1.solve(b[k], maxiter=400, tol=1e-9)
        assert np.array_equal(x1[0], x8[k]) and (it1[0], e1[0], c1[0]) == (it8[k], e8[k], c8[k])
        alone_hist[k] = bs1.err_history(0)
    # another slot, another K, neighbours that finish earlier and later
    order = [8, 5, 3, 0, 1]
    bs12 = sedi.Batch(shape, 12, prec)
    x12, it12, e12, c12 = bs12.solve(b[order], maxiter=400, tol=1e-9)
    for slot, k in enumerate(order):
        if k < 8:
            assert np.array_equal(x12[slot], x8[k]) and it12[slot] == it8[k] and e12[slot] == e8[k]
        if k in alone_hist:
            assert np.array_equal(bs12.err_history(slot), alone_hist[k])


# -------------------------------------------------------------- estimator ----
def test_estimator_batched_fp64_vs_reference(ref32):
    r = sedi.estimate(S32, kind=sedi.CG, samples=8, maxiter=500, tol=1e-9, seed=1234, batch_k=8)
    ref_it = ref32["s2_iters"]
    assert r["status"] == 0 and np.all(r["codes"] == 0)
    assert np.all(np.abs(r["iters"] - ref_it) <= 1), (r["iters"], ref_it)
    assert normwise(r["D"], ref32["s2_D"]) < (1e-10 if np.array_equal(r["iters"], ref_it) else 1e-8)


def test_estimator_batched_fixed_iterations_vs_reference(ref32):
    r = sedi.estimate(S32, kind=sedi.CG, samples=8, maxiter=20, tol=1e-9, seed=1234,
                      ignore_errors=True, batch_k=4)
    assert np.array_equal(r["iters"], ref32["s2_fixed20_iters"])
    assert np.array_equal(r["codes"], ref32["s2_fixed20_codes"])
    assert normwise(r["D"], ref32["s2_fixed20_D"]) < 1e-10


def test_estimator_mixed_vs_fp64():
    p = sedi.estimate(S32, kind=sedi.CG, samples=8, maxiter=500, tol=1e-9, seed=1234)
    m = sedi.estimate(S32, kind=sedi.CG, samples=8, maxiter=500, tol=1e-9, seed=1234,
                      precision=sedi.FP32_MIXED, batch_k=8)
    assert m["status"] == 0 and np.all(m["codes"] == 0)
    d = m["D"] - p["D"]
    max_rel = float(np.max(np.abs(d) / np.abs(p["D"])))
    rms_rel = float(np.sqrt(np.mean((d / p["D"]) ** 2)))
    print("mixed vs FP64 D: max rel %.3e, rms rel %.3e" % (max_rel, rms_rel))
    assert normwise(m["D"], p["D"]) < 1e-6 and rms_rel < 1e-6


def test_estimator_batched_ragged_samples():
    shape = (16, 12, 10)
    a = sedi.estimate(shape, kind=sedi.CG, samples=7, maxiter=20, tol=1e-9, seed=3,
                      ignore_errors=True)
    for k in (4, 8, 16):
        b = sedi.estimate(shape, kind=sedi.CG, samples=7, maxiter=20, tol=1e-9, seed=3,
                          ignore_errors=True, batch_k=k)
        assert np.array_equal(a["iters"], b["iters"]) and np.array_equal(a["codes"], b["codes"])
        assert normwise(b["D"], a["D"]) < 1e-12


def test_estimator_batched_error_exit():
    kw = dict(kind=sedi.CG, samples=6, maxiter=5, tol=1e-12, seed=1)
    a = sedi.estimate((16, 16, 16), **kw)
    b = sedi.estimate((16, 16, 16), batch_k=4, **kw)
    assert a["status"] == b["status"] == 101
    assert np.array_equal(a["codes"], b["codes"]) and np.array_equal(a["iters"], b["iters"])


def test_estimator_batched_log_lines():
    lines = []
    est = sedi.Estimator((10, 10, 10), kind=sedi.CG, samples=2, maxiter=100, tol=1e-6, seed=4,
                         batch_k=4)
    est.set_log(lines.append)
    r = est.run()
    est.close()
    assert lines[0] == "Solving for sample 1/2..."
    k = lines.index("Solving for sample 2/2...")
    assert k == 1 + r["iters"][0]
    assert lines[1].startswith("Cg: after iter=1, err=")
    assert len(lines) == 2 + int(r["iters"].sum())


def test_estimator_batched_rejects_unknown_solver():
    with pytest.raises(sedi.SediError):
        sedi.Estimator((8, 8, 8), kind=7, samples=4, batch_k=4)


# ------------------------------------------------------------- HIP graphs ---
@pytest.mark.parametrize("kind", KINDS)
def test_graph_replay_matches_direct_launches(kind):
    """Chunks of iterations replay a captured graph from the second chunk on;
    results must equal direct launches bitwise (plain and batched, several
    solves in a row so every replay is checked against eager work)."""
    shape = (24, 20, 18)
    runs = {}
    for on in (False, True):
        sedi.graphs_enable(on)
        try:
            a = sedi.estimate(shape, kind=kind, samples=3, maxiter=400, tol=1e-9, seed=2)
            b = sedi.estimate(shape, kind=kind, samples=6, maxiter=400, tol=1e-9, seed=2,
                              precision=sedi.FP32_MIXED, batch_k=4)
            c = sedi.estimate(shape, kind=kind, samples=3, maxiter=400, tol=1e-9, seed=2)
        finally:
            sedi.graphs_enable(True)
        runs[on] = (a, b, c)
    for x, y in zip(runs[False], runs[True]):
        assert np.array_equal(x["D"], y["D"]) and np.array_equal(x["iters"], y["iters"])
    assert np.array_equal(runs[True][0]["D"], runs[True][2]["D"])


def test_estimator_autotuned_batch_size_changes_nothing_but_speed():
    """batch_k = 0 takes the autotuner's k for the grid; a probe's result does
    not depend on its batch, so D and the iteration counts equal an explicit
    batch's bitwise."""
    k = sedi.tuning_lookup("k:%d:%d" % (O.CG, sedi.FP32_MIXED), S32, 64)
    assert k >= 1
    auto = sedi.estimate(S32, kind=O.CG, samples=12, maxiter=500, tol=1e-9, seed=1234,
                         precision=sedi.FP32_MIXED, batch_k=0)
    fixed = sedi.estimate(S32, kind=O.CG, samples=12, maxiter=500, tol=1e-9, seed=1234,
                          precision=sedi.FP32_MIXED, batch_k=4)
    assert np.array_equal(auto["iters"], fixed["iters"]) and np.array_equal(auto["D"], fixed["D"])
