"""Pins of the oracle's face values: MLMG::getGradSolution and MLMG::getFluxes (CPU).

Each check fixes the restatement by something independent of its own formula:
  * exactness: for a linear field with its values on the walls as Dirichlet
    data, the order-3 closure extrapolates exactly, so every face gradient
    (wall faces included) equals the field's slope; a periodic direction with
    zero slope gives zero;
  * conservation: the pinned forward operator satisfies
    L(phi) = alpha*a*phi + beta * sum_d dxinv_d * (F_d(i+1) - F_d(i)) for the
    ABecLaplacian flux F = -b grad phi, and L(phi) = -sum_d dxinv_d * (F(i+1) - F(i))
    for the Poisson flux F = -grad phi, at every cell with every boundary type;
    a dropped b, a wrong sign or an off-by-one face index breaks it;
  * flux = -b * gradient on every face (ABecLaplacian) and -gradient (Poisson).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

DIR, NEU, REF = 101, 102, 103

CASES = [
    # name, n, op, periodic, bc_lo, bc_hi, maxorder, max_grid_size, alpha, beta
    ("abec_dir_mo3", (8, 6, 5), "abec", (0, 0, 0), (DIR,) * 3, (DIR,) * 3, 3, 4, 1.3, 0.7),
    ("abec_mixed_mo2", (6, 7, 5), "abec", (0, 0, 0), (DIR, NEU, REF), (REF, DIR, NEU), 2, 4, 0.9, 2.0),
    ("abec_periodic_x", (8, 5, 6), "abec", (1, 0, 0), (0, DIR, NEU), (0, NEU, DIR), 3, 4, 1.0, 1.0),
    ("poisson_dir", (6, 6, 7), "poisson", (0, 0, 0), (DIR,) * 3, (DIR, NEU, DIR), 3, 4, 0.0, 1.0),
    ("poisson_periodic", (8, 8, 4), "poisson", (1, 1, 1), (0,) * 3, (0,) * 3, 3, 4, 0.0, 1.0),
]
IDS = [c[0] for c in CASES]


def _params(case):
    name, n, op, per, lo, hi, mo, mgs, alpha, beta = case
    return po.make_params(n, op=op, periodic=per, bc_lo=lo, bc_hi=hi, maxorder=mo, max_grid_size=mgs,
                          alpha=alpha, beta=beta)


def _coefs(case, seed=3):
    n, op = case[1], case[2]
    if op != "abec":
        return {}
    rng = np.random.default_rng(seed)
    return dict(a=rng.uniform(0.5, 1.5, (n[2], n[1], n[0])), bx=rng.uniform(0.5, 2.0, (n[2], n[1], n[0] + 1)),
                by=rng.uniform(0.5, 2.0, (n[2], n[1] + 1, n[0])), bz=rng.uniform(0.5, 2.0, (n[2] + 1, n[1], n[0])))


def test_linear_field_gradient_is_exact():
    n = (8, 6, 5)
    h = [1.0 / v for v in n]
    slope = (0.7, -1.3, 2.1)
    p = lambda x, y, z: 0.25 + slope[0] * x + slope[1] * y + slope[2] * z
    cx, cy, cz = [(np.arange(-1, m + 1) + 0.5) * hh for m, hh in zip(n, h)]
    Zg, Yg, Xg = np.meshgrid(cz, cy, cx, indexing="ij")
    bcv = p(np.clip(Xg, 0, 1), np.clip(Yg, 0, 1), np.clip(Zg, 0, 1))
    phi = p(Xg, Yg, Zg)[1:-1, 1:-1, 1:-1]
    for op in ("abec", "poisson"):
        prm = po.make_params(n, op=op, maxorder=3, max_grid_size=4)
        g = po.face_values(prm, phi, "grad", bcv=bcv,
                           **(dict(a=np.ones_like(phi), bx=np.ones((n[2], n[1], n[0] + 1)),
                                   by=np.ones((n[2], n[1] + 1, n[0])), bz=np.ones((n[2] + 1, n[1], n[0])))
                              if op == "abec" else {}))
        for d in range(3):
            np.testing.assert_allclose(g[d], slope[d], rtol=0, atol=1e-11)
    # periodic in x with zero x slope: zero x gradient; y, z exact
    slope = (0.0, -1.3, 2.1)
    phi = p(Xg, Yg, Zg)[1:-1, 1:-1, 1:-1]
    bcv = p(np.clip(Xg, 0, 1), np.clip(Yg, 0, 1), np.clip(Zg, 0, 1))
    prm = po.make_params(n, op="poisson", periodic=(1, 0, 0), bc_lo=(0, DIR, DIR), bc_hi=(0, DIR, DIR), maxorder=3,
                         max_grid_size=4)
    g = po.face_values(prm, phi, "grad", bcv=bcv)
    np.testing.assert_allclose(g[0], 0.0, rtol=0, atol=1e-11)
    np.testing.assert_allclose(g[1], slope[1], rtol=0, atol=1e-11)
    np.testing.assert_allclose(g[2], slope[2], rtol=0, atol=1e-11)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_flux_divergence_reproduces_operator(case):
    name, n, op, per, lo, hi, mo, mgs, alpha, beta = case
    prm = _params(case)
    coef = _coefs(case)
    rng = np.random.default_rng(11)
    phi = rng.standard_normal((n[2], n[1], n[0]))
    bcv = rng.standard_normal((n[2] + 2, n[1] + 2, n[0] + 2))
    x = po.grown(n)
    x[1:-1, 1:-1, 1:-1] = phi
    L, _ = po.apply(prm, x, bcv=bcv, **coef)
    F = po.face_values(prm, phi, "flux", bcv=bcv, **coef)
    dxi = n  # unit cube
    div = (dxi[0] * (F[0][:, :, 1:] - F[0][:, :, :-1]) + dxi[1] * (F[1][:, 1:, :] - F[1][:, :-1, :])
           + dxi[2] * (F[2][1:, :, :] - F[2][:-1, :, :]))
    if op == "abec":
        expect = alpha * coef["a"] * phi + beta * div
    else:
        expect = -div
    scale = np.abs(L).max()
    np.testing.assert_allclose(expect, L, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_flux_is_minus_b_times_gradient(case):
    name, n, op = case[0], case[1], case[2]
    prm = _params(case)
    coef = _coefs(case)
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((n[2], n[1], n[0]))
    bcv = rng.standard_normal((n[2] + 2, n[1] + 2, n[0] + 2))
    G = po.face_values(prm, phi, "grad", bcv=bcv, **coef)
    F = po.face_values(prm, phi, "flux", bcv=bcv, **coef)
    for d in range(3):
        b = coef[("bx", "by", "bz")[d]] if op == "abec" else 1.0
        np.testing.assert_allclose(F[d], -b * G[d], rtol=1e-13, atol=1e-13 * np.abs(G[d]).max())


# ---- two-level hierarchy: coarse periodic level + fine boxes (ratio 2) ----------

# two fine boxes sharing a face and an isolated one, fine index space of a 32^3 coarse level
AMR_BOXES = [(18, 20, 22, 25, 27, 29), (26, 20, 22, 33, 31, 29), (20, 36, 36, 27, 43, 43)]


def _amr_fields(nc, fn):
    def centres(lo, hi, h):
        ax = [(np.arange(lo[d], hi[d] + 1) + 0.5) * h for d in range(3)]
        Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        return X, Y, Z
    return fn(*centres((0, 0, 0), (nc - 1,) * 3, 1.0 / nc)), [fn(*centres(b[:3], b[3:], 0.5 / nc)) for b in AMR_BOXES]


def test_amr_fine_gradient_exact_for_linear_fields():
    # the coarse/fine ghosts (interpolated boundary value + Lagrange closure) are
    # exact for linear data, so every fine face gradient is the slope
    nc = 32
    slope = (1.7, -2.2, 0.9)
    s0, s1 = _amr_fields(nc, lambda X, Y, Z: 0.3 + slope[0] * X + slope[1] * Y + slope[2] * Z)
    _, f = po.amr_faces(po.amr_params(nc, max_grid_size=16), AMR_BOXES, s0, s1, "grad")
    for d in range(3):
        for g in f[d]:
            np.testing.assert_allclose(g, slope[d], rtol=0, atol=1e-10)


def test_amr_faces_close_the_composite_operator():
    # fine: L1(sol1) = rhs1 - res1 (the pinned composite residual) equals
    # -sum_d dxinv_d (F(i+1) - F(i)) over each fine box's own faces, cells at the
    # coarse/fine boundary and at the shared face included; coarse: the faces are
    # the single-level faces of the coarse solution.
    nc = 32
    rng = np.random.default_rng(9)
    s0 = rng.standard_normal((nc,) * 3)
    s1 = [rng.standard_normal((b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)) for b in AMR_BOXES]
    prm = po.amr_params(nc, max_grid_size=16)
    z0, z1 = np.zeros_like(s0), [np.zeros_like(v) for v in s1]
    _, r1 = po.amr_residual(prm, AMR_BOXES, z0, z1, s0, s1)
    c, f = po.amr_faces(prm, AMR_BOXES, s0, s1, "flux")
    dxi = 2 * nc
    for q in range(len(AMR_BOXES)):
        F = [f[d][q] for d in range(3)]
        div = dxi * ((F[0][:, :, 1:] - F[0][:, :, :-1]) + (F[1][:, 1:, :] - F[1][:, :-1, :])
                     + (F[2][1:, :, :] - F[2][:-1, :, :]))
        scale = np.abs(r1[q]).max()
        np.testing.assert_allclose(-div, -r1[q], rtol=0, atol=1e-12 * scale)
    single = po.face_values(po.amr_params(nc, max_grid_size=16), s0, "flux")
    for d in range(3):
        np.testing.assert_array_equal(c[d], single[d])
    g, gf = po.amr_faces(prm, AMR_BOXES, s0, s1, "grad")
    for d in range(3):
        np.testing.assert_array_equal(c[d], -g[d])
        for q in range(len(AMR_BOXES)):
            np.testing.assert_array_equal(f[d][q], -gf[d][q])
