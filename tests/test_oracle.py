"""Pins the oracle (oracle/libsedi_oracle.so) before it is trusted as the checker.

Every oracle function is checked against something other than a retyping of
its own formula: the compiled reference's outputs (tests/golden/ref_*.npz,
written by oracle/_ref/sedi_ref from the unmodified reference headers), the
C++ library's own mt19937/uniform_real_distribution (rng_golden.json), the
mt19937 known answer, closed forms (constant field, sine eigenvectors, exact
diag(A^-1)), dense linear algebra on tiny grids, and long-double sums.
"""
import numpy as np
import pytest

import oracle_lib as O
from conftest import normwise, sign_words, wsum

SHAPE32 = (32, 32, 32)


# ---------------------------------------------------------------- RNG -------
def test_mt19937_known_answer(rng_golden):
    m = O.MT19937(5489)
    for _ in range(9999):
        m.next()
    v = m.next()
    assert v == 4123659995 == rng_golden["mt19937_default_10000th"]


@pytest.mark.parametrize("case", range(9))
def test_rademacher_matches_libstdcxx(rng_golden, case):
    c = rng_golden["cases"][case]
    n, seed = c["ncells"], c["seed"]
    if n * c["probes"] <= 64 ** 3 * 8:
        # whole stream: every probe's count and checksum
        for k in range(c["probes"]):
            v = O.rademacher(seed, 2 * n * k, n)
            w = sign_words(v)
            assert int((v > 0).sum()) == c["plus"][k]
            assert wsum(w) == int(c["wsum"][k])
    # head of probe 0 verbatim
    m = min(n, 4096)
    v = O.rademacher(seed, 0, m)
    words = sign_words(v)
    head = [int(h, 16) for h in c["head"]]
    nw = len(words) if m % 32 == 0 else len(words) - 1
    assert [int(x) for x in words[:nw]] == head[:nw]


def test_rademacher_sign_rule_closed_form():
    # +1 iff x1 > 2^31 or (x1 == 2^31 and x0 > 1024): the double rounding of
    # x0 + x1*2^32 at 2^63 (spacing 2048, ties to even).
    m = O.MT19937(1234)
    xs = [m.next() for _ in range(20000)]
    m2 = O.MT19937(1234)
    for k in range(10000):
        x0, x1 = xs[2 * k], xs[2 * k + 1]
        u = m2.uniform01()
        assert (u > 0.5) == (x1 > 0x80000000 or (x1 == 0x80000000 and x0 > 1024))


# ------------------------------------------------------------ stencil -------
def test_stencil_constant_field():
    shape = (7, 5, 4)
    y = O.stencil_apply(shape, np.ones(7 * 5 * 4)).reshape(4, 5, 7)
    i3, i2, i1 = np.meshgrid(np.arange(4), np.arange(5), np.arange(7), indexing="ij")
    nb = ((i1 > 0).astype(int) + (i1 < 6) + (i2 > 0) + (i2 < 4) + (i3 > 0) + (i3 < 3))
    assert np.array_equal(y, 6.0 - nb)


@pytest.mark.parametrize("k", [(1, 1, 1), (2, 3, 1), (5, 1, 4)])
def test_stencil_sine_eigenvectors(k):
    n1, n2, n3 = 9, 7, 6
    i1 = np.arange(1, n1 + 1)
    i2 = np.arange(1, n2 + 1)
    i3 = np.arange(1, n3 + 1)
    phi = (np.sin(np.pi * k[2] * i3 / (n3 + 1))[:, None, None]
           * np.sin(np.pi * k[1] * i2 / (n2 + 1))[None, :, None]
           * np.sin(np.pi * k[0] * i1 / (n1 + 1))[None, None, :])
    lam = 6.0 - 2 * (np.cos(np.pi * k[0] / (n1 + 1)) + np.cos(np.pi * k[1] / (n2 + 1))
                     + np.cos(np.pi * k[2] / (n3 + 1)))
    y = O.stencil_apply((n1, n2, n3), phi.reshape(-1))
    assert np.allclose(y, lam * phi.reshape(-1), atol=1e-13)


def test_stencil_asymmetric_coefficients():
    # b != c: A x picks c from the minus side and b from the plus side
    shape = (5, 4, 3)
    x = np.zeros(60)
    x[1 + 5 * 1 + 20 * 1] = 1.0       # a single spike at (1,1,1)
    y = O.stencil_apply(shape, x, a=2.0, b=0.25, c=0.5).reshape(3, 4, 5)
    assert y[1, 1, 1] == 2.0
    assert y[1, 1, 2] == -0.5 and y[1, 2, 1] == -0.5 and y[2, 1, 1] == -0.5   # (+1 cells see x(-1))
    assert y[1, 1, 0] == -0.25 and y[1, 0, 1] == -0.25 and y[0, 1, 1] == -0.25


def test_dots_vs_long_double():
    rng = np.random.default_rng(0)
    shape = (33, 17, 9)
    xs = [rng.standard_normal(33 * 17 * 9) for _ in range(3)]
    ys = [rng.standard_normal(33 * 17 * 9) for _ in range(3)]
    got = O.dots(shape, xs, ys)
    for d in range(3):
        ref = float(np.sum(xs[d].astype(np.longdouble) * ys[d].astype(np.longdouble)))
        assert abs(got[d] - ref) <= 1e-12 * np.sum(np.abs(xs[d] * ys[d]))


# ------------------------------------------------------------ solvers -------
def dense_matrix(shape, a=6.0, b=1.0, c=1.0):
    n = int(np.prod(shape))
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        M[:, j] = O.stencil_apply(shape, e, a, b, c)
    return M


@pytest.mark.parametrize("kind", [O.CG, O.PCG, O.BICGSTAB])
def test_solvers_vs_dense_solve(kind):
    shape = (6, 5, 4)
    M = dense_matrix(shape)
    rhs = O.rademacher(99, 0, 120)
    code, x, fe, fi, hist = O.solve(kind, shape, rhs, maxiter=300, tol=1e-12)
    assert code == 0
    xe = np.linalg.solve(M, rhs)
    assert np.max(np.abs(x - xe)) / np.max(np.abs(xe)) < 1e-9
    # the reported err is the relative recursive residual: compare with the true one
    true = np.linalg.norm(rhs - M @ x) / np.linalg.norm(rhs)
    assert true < 1e-10


def test_cg_eigenvector_rhs_converges_in_one_iteration():
    # with a scalar preconditioner CG's Krylov space of an eigenvector is 1-D
    n1, n2, n3 = 8, 8, 8
    i = np.arange(1, 9)
    s = np.sin(np.pi * 2 * i / 9)
    phi = (s[:, None, None] * s[None, :, None] * s[None, None, :]).reshape(-1)
    for kind in (O.CG, O.PCG):
        code, x, fe, fi, hist = O.solve(kind, (n1, n2, n3), phi, maxiter=10, tol=1e-12)
        assert code == 0
        lam = 6.0 - 6 * np.cos(np.pi * 2 / 9)
        assert np.allclose(x, phi / lam, atol=1e-14)
        # CG stops after iteration 1; pipelined CG's test lags one iteration
        assert fi == (1 if kind == O.CG else 2)


def test_zero_rhs_returns_immediately():
    for kind in (O.CG, O.PCG, O.BICGSTAB):
        code, x, fe, fi, hist = O.solve(kind, (4, 4, 4), np.zeros(64), maxiter=10, tol=1e-9)
        assert (code, fi, fe) == (0, 0, 0.0) and not x.any()


def test_maxiter_reached_code():
    v = O.rademacher(3, 0, 512)
    for kind in (O.CG, O.PCG, O.BICGSTAB):
        code, x, fe, fi, hist = O.solve(kind, (8, 8, 8), v, maxiter=3, tol=1e-12)
        assert code == 1 and fi == 3 and fe > 1e-12


def test_exact_diag_inv_vs_dense_inverse():
    for shape in [(4, 4, 4), (5, 4, 3), (6, 6, 6)]:
        M = dense_matrix(shape)
        d = np.diag(np.linalg.inv(M))
        cells = np.arange(int(np.prod(shape)))
        e = O.exact_diag_inv(shape, cells)
        assert np.max(np.abs(e - d)) < 1e-13


def test_exact_diag_inv_known_values():
    # corner (0,0,0) approaches 0.185577 for n >= 6; centre of 32^3 = 0.248505
    assert abs(O.exact_diag_inv((8, 8, 8), [0])[0] - 0.185577) < 5e-6
    c = 15 + 32 * 15 + 32 * 32 * 15
    assert abs(O.exact_diag_inv(SHAPE32, [c])[0] - 0.248505) < 5e-6


def test_estimator_is_unbiased():
    # E[D] = diag(A^-1): with 400 probes on 6^3 the estimate is within 4 sigma
    shape = (6, 6, 6)
    st, D, iters, errs, codes = O.estimate(O.CG, shape, 5, 400, 500, 1e-12)
    exact = O.exact_diag_inv(shape, np.arange(216))
    M = dense_matrix(shape)
    Ainv = np.linalg.inv(M)
    # Var of one probe's v_i (A^-1 v)_i is sum_{j != i} (A^-1)_ij^2
    var = (Ainv ** 2).sum(axis=1) - np.diag(Ainv) ** 2
    z = (D - exact) / np.sqrt(var / 400)
    assert np.max(np.abs(z)) < 4.5
    assert abs(np.mean(z)) < 0.5


# ------------------------------------- oracle vs the compiled reference -------
@pytest.mark.parametrize("solver", [1, 2, 3])
def test_oracle_equals_reference_32cube(ref32, solver):
    st, D, iters, errs, codes = O.estimate(solver, SHAPE32, 1234, 8, 500, 1e-9)
    assert st == 0
    assert np.array_equal(iters, ref32[f"s{solver}_iters"])
    assert np.array_equal(codes, ref32[f"s{solver}_codes"])
    assert np.array_equal(errs, ref32[f"s{solver}_errs"])
    assert np.array_equal(D, ref32[f"s{solver}_D"])


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_oracle_err_history_equals_reference_lines(ref32, solver):
    lens = ref32[f"s{solver}_hist_len"]
    flat = ref32[f"s{solver}_hist"]
    off = 0
    for k in range(8):
        v = O.rademacher(1234, 2 * 32768 * k, 32768)
        code, x, fe, fi, hist = O.solve(solver, SHAPE32, v, maxiter=500, tol=1e-9)
        ref = flat[off: off + lens[k]]
        off += lens[k]
        # the reference prints %.10e; compare at that precision
        assert len(hist) >= len(ref) - 1 or solver == 1
        m = min(len(hist), len(ref))
        assert np.all([f"{a:.10e}" == f"{b:.10e}" for a, b in zip(hist[:m], ref[:m])])


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_oracle_equals_reference_fixed_iterations(ref32, solver):
    st, D, iters, errs, codes = O.estimate(solver, SHAPE32, 1234, 8, 20, 1e-9, ignore_errors=True)
    assert np.array_equal(iters, ref32[f"s{solver}_fixed20_iters"])
    assert np.array_equal(codes, ref32[f"s{solver}_fixed20_codes"])
    assert np.array_equal(D, ref32[f"s{solver}_fixed20_D"])


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_oracle_equals_reference_odd_grid(ref_odd, solver):
    st, D, iters, errs, codes = O.estimate(solver, (13, 7, 5), 5, 3, 200, 1e-8)
    assert np.array_equal(iters, ref_odd[f"s{solver}_iters"])
    assert np.array_equal(errs, ref_odd[f"s{solver}_errs"])
    assert np.array_equal(D, ref_odd[f"s{solver}_D"])


def test_decomposed_reference_agrees_to_rounding(ref32):
    # -p 2x2x2 under 8 MPI ranks changes only the summation order of the dots
    from conftest import load_npz
    p8 = load_npz("ref_32cube_p8.npz")
    assert np.array_equal(p8["s2_iters"], ref32["s2_iters"])
    assert normwise(p8["s2_D"], ref32["s2_D"]) < 1e-10


@pytest.mark.parametrize("tag", ["c32", "odd"])
def test_reference_summation_order_spread(ref32, ref_odd, ref_spread, tag):
    """The evidence behind the BiCGStab bars: the reference's own solvers,
    changed only in their summation order (other decompositions, 4- and
    8-lane SIMD dot products), stop CG and pipelined CG at the same iteration
    for every probe with D equal to rounding, while BiCGStab's counts move by
    several iterations and D by ~1e-8 at 32^3.  Variant 0 is the -p 1x1x1
    2-lane run of ref_32cube.npz / ref_odd.npz, which the oracle equals
    bitwise."""
    ref = ref32 if tag == "c32" else ref_odd
    for kind in (1, 2, 3):
        its = ref_spread[f"{tag}_s{kind}_iters"]
        assert np.array_equal(its[0], ref[f"s{kind}_iters"])
        assert ref_spread[f"{tag}_s{kind}_Ddist"][0] == 0.0
        if kind != 1:
            assert np.all(its == its[0])
            assert np.max(ref_spread[f"{tag}_s{kind}_Ddist"]) < 1e-11
    if tag == "c32":
        spread = np.abs(ref_spread["c32_s1_iters"][1:] - ref_spread["c32_s1_iters"][0])
        assert spread.max() == 9 and 1e-8 < np.max(ref_spread["c32_s1_Ddist"]) < 1e-7
        # the decompositions alone, and the SIMD width alone, each move it
        labels = list(ref_spread["c32_variants"])
        simd_only = [labels.index("avx2_p111"), labels.index("avx512_p111")]
        decomp_only = [i for i, l in enumerate(labels) if l.startswith("sse2_") and i > 0]
        assert len(labels) == 60 and len(decomp_only) == 19
        assert spread[[i - 1 for i in simd_only]].max() > 0 and spread[[i - 1 for i in decomp_only]].max() > 0


def _breakdown_cases():
    import json
    import os
    from conftest import GOLDEN
    with open(os.path.join(GOLDEN, "ref_breakdown.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", range(9))
def test_oracle_bicgstab_breakdowns_equal_reference(case):
    """BiCGStab's breakdown outcomes (BiCgStab.hpp:162-163,229-230,237-259) on
    systems where they happen in exact arithmetic: omega = 0 (code 3), rho = 0
    at the second iteration (code 2), each along x, y and z, and omega = 0
    with err <= tol (code 0, the outcome order).  The reference's own solver
    (tests/golden/ref_breakdown.json) and the oracle agree bitwise."""
    c = _breakdown_cases()[case]
    code, x, fe, fi, hist = O.solve(O.BICGSTAB, tuple(c["grid"]), np.array(c["rhs"]), maxiter=c["maxiter"],
                                    tol=c["tol"], abc=tuple(c["abc"]))
    assert (code, fi, fe) == (c["code"], c["iter"], c["err"])
    assert np.array_equal(x, np.array(c["x"]))
    nh = len(c["hist"])
    assert [float("%.10e" % h) for h in hist[:nh]] == c["hist"]
    assert code in (0, 2, 3)


def test_breakdown_cases_cover_every_outcome():
    codes = [c["code"] for c in _breakdown_cases()]
    assert codes.count(3) == 4 and codes.count(2) == 4 and codes.count(0) == 1


def test_exact_diag_fixture_256cube():
    """tests/golden/exact_diag_256cube.npz (diag(A^-1) at 8192 cells of the
    256^3 grid, from oracle_exact_diag_inv): re-computed cells agree, the
    corner is the known 0.185577 and every cell lies between it and Watson's
    limit 0.2527 (the infinite lattice's value, which interior cells approach
    from below)."""
    from conftest import load_npz
    ex = load_npz("exact_diag_256cube.npz")
    assert abs(ex["diag_strided"][0] - 0.185577) < 1e-6
    for name in ("strided", "random"):
        d = ex["diag_" + name]
        assert len(d) == 4096 and d.min() >= 0.185577 - 1e-6 and d.max() < 0.25273
        pick = np.array([0, 1000, 4095])
        again = O.exact_diag_inv((256, 256, 256), ex["idx_" + name][pick])
        assert np.allclose(again, d[pick], rtol=0, atol=1e-15)
    n = 256 ** 3
    assert np.array_equal(ex["idx_strided"], np.arange(0, n, n // 4096)[:4096])


def test_reference_full_size_order_stability():
    """At 256^3 (seed 7) the reference stops CG and pipelined CG at the same
    iteration on one rank (-p 1x1x1) as on eight (-p 2x2x2), while BiCGStab
    moves by tens of iterations between -p 1x1x1, -p 2x2x2 and an 8-lane
    SIMD build (tests/golden/ref_256cube_extra.npz beside ref_256cube.npz)."""
    from conftest import load_npz
    big, extra = load_npz("ref_256cube.npz"), load_npz("ref_256cube_extra.npz")
    for kind in (2, 3):
        assert np.array_equal(extra[f"s{kind}_p111_iters"], big[f"s{kind}_iters"][:4])
    dev = np.concatenate([np.abs(extra["s1_p111_iters"] - big["s1_iters"][:4]),
                          np.abs(extra["s1_p111_avx512_iters"] - extra["s1_p111_iters"])])
    assert dev.max() >= 20
    assert np.all(extra["s1_fixed100_iters"] == 100) and np.all(extra["s1_fixed100_codes"] == 1)
    # in fixed-iteration mode the reference's BiCGStab trajectory depends on
    # the summation order only after many iterations: 20 agree to rounding,
    # 100 do not
    e20 = np.abs(extra["s1_fixed20_p118_errs"] / extra["s1_fixed20_errs"] - 1).max()
    e100 = np.abs(extra["s1_fixed100_p118_errs"] / extra["s1_fixed100_errs"] - 1).max()
    assert e20 < 1e-10 and e100 > 0.1
    assert normwise(extra["s1_fixed20_p118_Dsample"], extra["s1_fixed20_Dsample"]) < 1e-11


def test_reference_bicgstab_spread_256cube():
    """The evidence behind the GPU's BiCGStab bars at 256^3: the reference's
    own BiCGStab (seed 7, 16 probes, tol 1e-9) under other summation orders
    (tests/golden/ref_spread_256cube.npz: -p 1x1x8, 1x8x1, 8x1x1, 1x2x4,
    4x2x1 with 2-lane SIMD and -p 2x2x2 with 4- and 8-lane SIMD) against its
    -p 2x2x2 2-lane run (ref_256cube.npz): every variant converges, their
    counts move by up to a hundred-odd iterations on some probe (but not by
    half the count anywhere), and D moves at the level the tolerance leaves
    (between 1e-9 and 1e-6 normwise on the 4096 strided cells)."""
    from conftest import load_npz
    big, sp = load_npz("ref_256cube.npz"), load_npz("ref_spread_256cube.npz")
    assert np.array_equal(sp["sample_idx"], np.arange(0, 256 ** 3, 256 ** 3 // 4096)[:4096])
    its = sp["s1_iters"]
    assert its.shape == (len(sp["variants"]), 16) and len(sp["variants"]) >= 1
    dev = np.abs(its - big["s1_iters"][None, :])
    assert 50 <= dev.max() < big["s1_iters"].min() // 2
    ddist = [normwise(v, big["s1_Dsample"]) for v in sp["s1_Dsample"]]
    assert all(1e-9 < d < 1e-6 for d in ddist), ddist
