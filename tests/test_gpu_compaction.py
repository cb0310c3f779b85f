"""GPU: lane compaction of the batched solvers.

A block of 4 right-hand sides runs until its slowest lane converges; between
iterations the live right-hand sides of the emptiest block trade lanes with
finished ones of other blocks and the emptied block stops.  Bar: bitwise the
same solutions, iteration counts, errors, codes and error histories as with
compaction off (every right-hand side goes through the same operations in
any lane), estimator D bitwise the same, and lanes did move.
"""
import numpy as np
import pytest

import oracle_lib as O

sedi = pytest.importorskip("sedi_amd")

pytestmark = pytest.mark.gpu

SHAPE = (24, 20, 18)
KINDS = [O.CG, O.PCG, O.BICGSTAB]


def eigvec(shape, p):
    """Dirichlet eigenvector of the 7-point stencil (i1 fastest): CG,
    pipelined CG and BiCGStab solve it in one or a few iterations."""
    n1, n2, n3 = shape
    s1 = np.sin(np.pi * p[0] * np.arange(1, n1 + 1) / (n1 + 1))
    s2 = np.sin(np.pi * p[1] * np.arange(1, n2 + 1) / (n2 + 1))
    s3 = np.sin(np.pi * p[2] * np.arange(1, n3 + 1) / (n3 + 1))
    return (s3[:, None, None] * s2[None, :, None] * s1[None, None, :]).reshape(-1)


def test_eigvec_is_an_eigenvector():
    v = eigvec(SHAPE, (2, 1, 3))
    av = O.stencil_apply(SHAPE, v)
    lam = float(av @ v / (v @ v))
    assert np.linalg.norm(av - lam * v) < 1e-12 * np.linalg.norm(av)


def mixed_rhs():
    """12 right-hand sides in 3 blocks: block 0 all probes, block 1 three
    eigenvectors and a probe, block 2 two eigenvectors and two probes, so
    blocks 1 and 2 empty early and unevenly."""
    n = int(np.prod(SHAPE))
    pr = [O.rademacher(17, 2 * n * k, n) for k in range(7)]
    ev = [eigvec(SHAPE, p) for p in [(1, 1, 1), (2, 1, 1), (1, 3, 2), (3, 2, 1), (2, 2, 2)]]
    return np.stack([pr[0], pr[1], pr[2], pr[3], ev[0], ev[1], ev[2], pr[4], ev[3], ev[4], pr[5], pr[6]])


def solve_all(kind, prec, b, on):
    sedi.batch_compact_enable(on)
    try:
        bs = sedi.Batch(SHAPE, 12, prec, kind=kind)
        x, iters, errs, codes = bs.solve(b, maxiter=600, tol=1e-9)
        hist = [bs.err_history(k) for k in range(len(b))]
        moves = bs.lane_moves
        bs.close()
    finally:
        sedi.batch_compact_enable(True)
    return x, iters, errs, codes, hist, moves


@pytest.mark.parametrize("prec", [sedi.FP64, sedi.FP32_MIXED])
@pytest.mark.parametrize("kind", KINDS)
def test_compaction_changes_nothing_but_lanes(kind, prec):
    b = mixed_rhs()
    print("case kind %d precision %d" % (kind, prec), flush=True)
    x0, it0, e0, c0, h0, m0 = solve_all(kind, prec, b, False)
    x1, it1, e1, c1, h1, m1 = solve_all(kind, prec, b, True)
    print("iterations", it1.tolist(), "lanes moved", m1, flush=True)
    assert m0 == 0 and m1 > 0
    assert np.all(c1 == 0)
    assert np.array_equal(it0, it1) and np.array_equal(c0, c1) and np.array_equal(e0, e1)
    assert np.array_equal(x0, x1)
    for k in range(len(b)):
        assert np.array_equal(h0[k], h1[k]), k
    # and the answer is the answer (to the right-hand side as stored: FP32
    # in the mixed form)
    bs = b if prec == sedi.FP64 else b.astype(np.float32).astype(np.float64)
    for k in (4, 7, 11):
        r = np.linalg.norm(bs[k] - O.stencil_apply(SHAPE, x1[k])) / np.linalg.norm(bs[k])
        assert r < 1e-8, (k, r)


@pytest.mark.parametrize("kind", KINDS)
def test_compaction_estimator_D_bitwise(kind):
    runs = []
    for on in (False, True):
        sedi.batch_compact_enable(on)
        try:
            runs.append(sedi.estimate(SHAPE, kind=kind, samples=14, maxiter=600, tol=1e-9, seed=5,
                                      precision=sedi.FP32_MIXED, batch_k=16))
        finally:
            sedi.batch_compact_enable(True)
    a, b = runs
    assert a["status"] == b["status"] == 0
    assert np.array_equal(a["iters"], b["iters"]) and np.array_equal(a["codes"], b["codes"])
    assert np.array_equal(a["D"], b["D"])
