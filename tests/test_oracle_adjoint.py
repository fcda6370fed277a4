"""Pins of the oracle's adjoint part (CPU).

The reference has no adjoint, so nothing here can be checked against its
outputs; instead each piece is tied to something already pinned or to dense
linear algebra on tiny grids:
  * the explicit matrix rows reproduce the pinned forward operator (fapply with
    the closure ghosts) for every boundary type and closure order;
  * the transpose equals the dense transpose of the matrix built column by
    column from the forward operator;
  * the MLMG solve of A^T x = b matches a dense solve;
  * the coefficient gradient equals -lam . (A(c + e) phi - A(c) phi) computed
    with the forward operator itself (A is affine in a and b, so the unit
    difference is the exact derivative), with lam from a dense transpose solve.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

DIR, NEU, REF = 101, 102, 103


def _coefs(n, seed, periodic=(0, 0, 0)):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.5, (n[2], n[1], n[0]))
    bx = rng.uniform(0.5, 2.0, (n[2], n[1], n[0] + 1))
    by = rng.uniform(0.5, 2.0, (n[2], n[1] + 1, n[0]))
    bz = rng.uniform(0.5, 2.0, (n[2] + 1, n[1], n[0]))
    return dict(a=a, bx=bx, by=by, bz=bz)


CASES = [
    # name, n, op, periodic, bc_lo, bc_hi, maxorder, max_grid_size
    ("abec_dir_mo3", (6, 5, 4), "abec", (0, 0, 0), (DIR,) * 3, (DIR,) * 3, 3, 8),
    ("abec_dir_mo2", (6, 5, 4), "abec", (0, 0, 0), (DIR,) * 3, (DIR,) * 3, 2, 8),
    ("abec_mixed_short_boxes", (5, 6, 7), "abec", (0, 0, 0), (DIR, NEU, REF), (REF, DIR, DIR), 3, 2),
    ("abec_periodic_x_mo4", (6, 5, 7), "abec", (1, 0, 0), (0, DIR, DIR), (0, DIR, NEU), 4, 4),
    ("abec_all_periodic", (4, 6, 4), "abec", (1, 1, 1), (0,) * 3, (0,) * 3, 3, 8),
    ("poisson_dir_mo3", (6, 5, 4), "poisson", (0, 0, 0), (DIR,) * 3, (DIR,) * 3, 3, 8),
    ("poisson_periodic_neu", (6, 4, 5), "poisson", (1, 0, 0), (0, NEU, REF), (0, NEU, NEU), 3, 3),
    ("poisson_mixed_dir", (6, 4, 5), "poisson", (1, 0, 0), (0, NEU, DIR), (0, NEU, DIR), 3, 3),
]
IDS = [c[0] for c in CASES]


def _setup(case, seed=3):
    name, n, op, periodic, lo, hi, mo, mgs = case
    prm = po.make_params(n, op=op, periodic=periodic, bc_lo=lo, bc_hi=hi, max_grid_size=mgs,
                         maxorder=mo, alpha=1.3, beta=0.7)
    coef = _coefs(n, seed) if op == "abec" else {}
    return prm, coef, n


def _apply_valid(prm, x, bcv=None, **coef):
    n = tuple(prm.n)
    g = po.grown(n)
    g[1:-1, 1:-1, 1:-1] = x
    out, _ = po.apply(prm, g, bcv=bcv, **coef)
    return out


def _dense(prm, coef, n):
    """Dense matrix of the homogeneous operator, column by column from fapply."""
    nc = n[0] * n[1] * n[2]
    A = np.zeros((nc, nc))
    e = np.zeros(nc)
    for c in range(nc):
        e[:] = 0.0
        e[c] = 1.0
        A[:, c] = _apply_valid(prm, e.reshape(n[2], n[1], n[0]), **coef).ravel()
    return A


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_rows_reproduce_forward_operator(case):
    prm, coef, n = _setup(case)
    x = np.random.default_rng(7).standard_normal((n[2], n[1], n[0]))
    ref = _apply_valid(prm, x, **coef)
    got = po.apply_rows(prm, x, **coef)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_transpose_equals_dense_transpose(case):
    prm, coef, n = _setup(case)
    A = _dense(prm, coef, n)
    y = np.random.default_rng(9).standard_normal((n[2], n[1], n[0]))
    got = po.apply_transpose(prm, y, **coef)
    ref = (A.T @ y.ravel()).reshape(y.shape)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # the cases with a Dirichlet closure beyond first order are not symmetric
    asym = np.abs(A - A.T).max() / np.abs(A).max()
    name = case[0]
    if name in ("abec_dir_mo3", "abec_mixed_short_boxes", "abec_periodic_x_mo4", "poisson_dir_mo3",
                "poisson_mixed_dir"):
        assert asym > 1e-3
    if name in ("poisson_periodic_neu", "abec_dir_mo2"):
        assert asym < 1e-14


@pytest.mark.parametrize("op", ["abec", "poisson"])
def test_transpose_solve_matches_dense(op):
    n = (12, 10, 8)
    prm = po.make_params(n, op=op, bc_lo=(DIR, NEU, DIR), bc_hi=(DIR, DIR, REF), max_grid_size=4,
                         maxorder=3, alpha=1.0, beta=1.0, tol_rel=1e-12, transpose=1)
    coef = _coefs(n, 4) if op == "abec" else {}
    g = np.random.default_rng(5).standard_normal((n[2], n[1], n[0]))
    lam, st = po.solve(prm, g, **coef)
    assert st["status"] == 0
    A = _dense(po.make_params(n, op=op, bc_lo=(DIR, NEU, DIR), bc_hi=(DIR, DIR, REF), max_grid_size=4,
                              maxorder=3), coef, n)
    ref = np.linalg.solve(A.T, g.ravel()).reshape(g.shape)
    assert np.linalg.norm(lam - ref) / np.linalg.norm(ref) <= 1e-10
    # and it is not the forward solution
    fwd = np.linalg.solve(A, g.ravel()).reshape(g.shape)
    assert np.linalg.norm(fwd - ref) / np.linalg.norm(ref) > 1e-4


def test_transpose_solve_mixed_precision():
    n = (32, 32, 32)
    coef = _coefs(n, 6)
    g = np.random.default_rng(2).standard_normal((n[2], n[1], n[0]))
    p64 = po.make_params(n, op="abec", max_grid_size=16, transpose=1, tol_rel=1e-11)
    pmx = po.make_params(n, op="abec", max_grid_size=16, transpose=1, tol_rel=1e-11, precision="mixed")
    l64, s64 = po.solve(p64, g, **coef)
    lmx, smx = po.solve(pmx, g, **coef)
    assert s64["status"] == 0 and smx["status"] == 0
    r = g - po.apply_transpose(p64, l64, **coef)
    assert np.abs(r).max() <= 1e-11 * np.abs(g).max()
    assert np.linalg.norm(lmx - l64) / np.linalg.norm(l64) <= 1e-9


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[3], CASES[4]], ids=[IDS[0], IDS[2], IDS[3], IDS[4]])
@pytest.mark.parametrize("inhomog", [False, True])
def test_coefficient_gradient_matches_operator_difference(case, inhomog):
    prm, coef, n = _setup(case, seed=11)
    rng = np.random.default_rng(12)
    bcv = rng.standard_normal((n[2] + 2, n[1] + 2, n[0] + 2)) if inhomog else None
    A = _dense(prm, coef, n)
    c = _apply_valid(prm, np.zeros((n[2], n[1], n[0])), bcv=bcv, **coef).ravel()  # boundary part
    f = rng.standard_normal(A.shape[0])
    phi = np.linalg.solve(A, f - c)
    target = rng.standard_normal(A.shape[0])
    lam = np.linalg.solve(A.T, phi - target)
    phi3, lam3 = phi.reshape(n[2], n[1], n[0]), lam.reshape(n[2], n[1], n[0])
    ga, gb = po.grad_coeffs(prm, lam3, phi3, bcv=bcv, **coef)

    def op_phi(**cf):
        return _apply_valid(prm, phi3, bcv=bcv, **cf).ravel()

    base = op_phi(**coef)
    # a-gradient at every cell
    ga_ref = np.zeros(A.shape[0])
    for p in range(A.shape[0]):
        cf = dict(coef)
        cf["a"] = coef["a"].copy()
        cf["a"].ravel()[p] += 1.0
        ga_ref[p] = -lam @ (op_phi(**cf) - base)
    np.testing.assert_allclose(ga.ravel(), ga_ref, rtol=0, atol=1e-12 * np.abs(ga_ref).max())
    # b-gradient at every face of every direction
    for d, key in enumerate(("bx", "by", "bz")):
        ref = np.zeros(coef[key].size)
        for q in range(coef[key].size):
            cf = dict(coef)
            cf[key] = coef[key].copy()
            cf[key].ravel()[q] += 1.0
            ref[q] = -lam @ (op_phi(**cf) - base)
        np.testing.assert_allclose(gb[d].ravel(), ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_gradient_matches_central_difference_of_objective():
    # End to end on a tiny grid: J(b) = 0.5 |phi(b) - t|^2 with dense solves;
    # central difference h = 1e-6 against the adjoint gradient.
    case = CASES[0]
    prm, coef, n = _setup(case, seed=21)
    rng = np.random.default_rng(22)
    f = rng.standard_normal(n[0] * n[1] * n[2])
    t = rng.standard_normal(f.size)

    def J(cf):
        A = _dense(prm, cf, n)
        phi = np.linalg.solve(A, f)
        return 0.5 * np.sum((phi - t) ** 2), A, phi

    _, A, phi = J(coef)
    lam = np.linalg.solve(A.T, phi - t)
    _, gb = po.grad_coeffs(prm, lam.reshape(n[2], n[1], n[0]), phi.reshape(n[2], n[1], n[0]), **coef)
    h = 1e-6
    for key, d in (("bx", 0), ("by", 1), ("bz", 2)):
        for q in rng.choice(coef[key].size, 3, replace=False):
            cp, cm = dict(coef), dict(coef)
            cp[key] = coef[key].copy()
            cm[key] = coef[key].copy()
            cp[key].ravel()[q] += h
            cm[key].ravel()[q] -= h
            fd = (J(cp)[0] - J(cm)[0]) / (2 * h)
            assert gb[d].ravel()[q] == pytest.approx(fd, rel=1e-5, abs=1e-8)


# The adjoint's V-cycles on A^T (maxorder <= 3): a colour of the transposed
# sweep is one Gauss-Seidel half-step on the dense A^T with the forward
# sweep's relaxation divisor, read off the forward colour on the same input
# (x' = x + omega / g_m_d (b - A x) on the colour's cells).
TSWEEP = [c for c in CASES if c[6] <= 3]


@pytest.mark.parametrize("rb", [0, 1])
@pytest.mark.parametrize("case", TSWEEP, ids=[c[0] for c in TSWEEP])
def test_transposed_colour_is_a_gauss_seidel_step_on_the_transpose(case, rb):
    prm, coef, n = _setup(case)
    name, _, op, periodic, lo, hi, mo, mgs = case
    A = _dense(prm, coef, n)
    rng = np.random.default_rng(11 + rb)
    x = rng.standard_normal((n[2], n[1], n[0]))
    b = rng.standard_normal(x.shape)
    g = po.grown(n)
    g[1:-1, 1:-1, 1:-1] = x
    kk, jj, ii = np.meshgrid(range(n[2]), range(n[1]), range(n[0]), indexing="ij")
    col = ((ii + jj + kk + rb) % 2) == 0
    fwd = po.gsrb_colour(prm, g, b, rb, **coef)[1:-1, 1:-1, 1:-1]
    res_a = (b.ravel() - A @ x.ravel()).reshape(x.shape)
    gmd = 1.15 * res_a[col] / (fwd - x)[col]
    prm_t = po.make_params(n, op=op, periodic=periodic, bc_lo=lo, bc_hi=hi, max_grid_size=mgs, maxorder=mo,
                           alpha=1.3, beta=0.7, transpose=1)
    got = po.gsrb_colour(prm_t, g, b, rb, **coef)[1:-1, 1:-1, 1:-1]
    res_t = (b.ravel() - A.T @ x.ravel()).reshape(x.shape)
    ref = x.copy()
    ref[col] = x[col] + 1.15 / gmd * res_t[col]
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert np.array_equal(got[~col], x[~col])
    # where A is not symmetric the transposed colour is not the forward one
    if np.abs(A - A.T).max() > 1e-3 * np.abs(A).max():
        assert np.abs(got - fwd).max() > 1e-6 * np.abs(fwd).max()
    else:
        np.testing.assert_allclose(got, fwd, rtol=0, atol=1e-12 * np.abs(fwd).max())


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_adjoint_cycles_take_the_forward_iterations(precision):
    # 40x48x56 ABecLaplacian, homogeneous Dirichlet, maxorder 3: with V-cycles
    # of L the adjoint solve took 16 iterations against the forward 10
    n = (40, 48, 56)
    rng = np.random.default_rng(7)
    coef = dict(a=rng.uniform(0.5, 1.5, (n[2], n[1], n[0])), bx=rng.uniform(0.5, 2.0, (n[2], n[1], n[0] + 1)),
                by=rng.uniform(0.5, 2.0, (n[2], n[1] + 1, n[0])), bz=rng.uniform(0.5, 2.0, (n[2] + 1, n[1], n[0])))
    g = np.random.default_rng(8).standard_normal((n[2], n[1], n[0]))
    kw = dict(op="abec", bc_lo=(DIR,) * 3, bc_hi=(DIR,) * 3, max_grid_size=16, agg_grid_size=32, maxorder=3,
              alpha=1.3, beta=0.7, tol_rel=1e-11, precision=precision)
    _, fst = po.solve(po.make_params(n, **kw), g, **coef)
    lam, ast = po.solve(po.make_params(n, transpose=1, **kw), g, **coef)
    assert fst["status"] == 0 and ast["status"] == 0
    assert ast["niters"] == fst["niters"]
    r = g - po.apply_transpose(po.make_params(n, **kw), lam, **coef)
    assert np.abs(r).max() <= 1e-10 * np.abs(g).max()
