"""Pins of the oracle's two-level MLABecLaplacian composite solve (CPU).

The ABecLaplacian composite path differs from the pinned Poisson one in four
places: the fine level's coefficients in its smoother and residual, the
b-weighted fluxes of the reflux, and the coarse coefficients under the fine
level, replaced by the fine averages (averageDownCoeffsToCoarseAmrLevel).
Each is tied to something fixed independently:
  * with Poisson's scalars (alpha 0, beta -1, b = 1) the solve is the pinned
    Poisson composite solve;
  * the refluxed composite operator with variable b is conservative on a
    periodic domain (the volume-weighted composite residual sums to zero),
    which fixes the b-weighting of both fluxes of every coarse/fine face;
  * coarse coefficients under the fine level do not matter: a solve with
    random ones there equals a solve with the fine averages (computed here);
  * the composite solve reproduces a linear field whose right-hand side is
    the textbook 7-point operator with variable a and b on each level;
  * with no fine boxes it is the single-level ABecLaplacian solve;
  * the fine smoothing leaves the solution of its own equation unchanged (a
    Gauss-Seidel fixed point): with res1 = L_f c for a random c (homogeneous
    coarse/fine boundary, the residual pinned by the linear field), a sweep
    from c returns c.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from amrcoef import averaged as _averaged, coefs as _coefs, lin_apply as _lin_apply

DIR = 101
PER = dict(periodic=(1, 1, 1), bc_lo=(0,) * 3, bc_hi=(0,) * 3)
DIR_ALL = dict(periodic=(0, 0, 0), bc_lo=(DIR,) * 3, bc_hi=(DIR,) * 3)
BOXES = [(20, 16, 16, 35, 31, 31), (36, 18, 20, 49, 41, 43), (6, 40, 8, 17, 51, 19)]
LIN_BOXES = [(2, 2, 2, 17, 13, 21), (24, 30, 20, 41, 45, 61), (44, 8, 40, 55, 19, 51)]


def _params(nc, bc, **kw):
    return po.amr_params(nc, max_grid_size=16, op="abec", **bc, **kw)


def _rhs(nc, boxes, seed):
    rng = np.random.default_rng(seed)
    r0 = rng.standard_normal((nc,) * 3)
    r1 = [rng.standard_normal((b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)) for b in boxes]
    return r0, r1


def test_poisson_scalars_give_the_poisson_composite_solve():
    nc = 32
    r0, r1 = _rhs(nc, BOXES, 3)
    m = r0.mean()
    r0, r1 = r0 - m, [x - m for x in r1]
    ones = (dict(a=np.ones((nc,) * 3), bx=np.ones((nc, nc, nc + 1)), by=np.ones((nc, nc + 1, nc)),
                 bz=np.ones((nc + 1, nc, nc))),
            {"a": [np.ones_like(x) for x in r1],
             "bx": [np.ones((x.shape[0], x.shape[1], x.shape[2] + 1)) for x in r1],
             "by": [np.ones((x.shape[0], x.shape[1] + 1, x.shape[2])) for x in r1],
             "bz": [np.ones((x.shape[0] + 1, x.shape[1], x.shape[2])) for x in r1]})
    p0, p1, ps = po.amr_solve(po.amr_params(nc, max_grid_size=16, tol_rel=1e-11), BOXES, r0, r1)
    a0, a1, st = po.amr_solve(_params(nc, PER, alpha=0.0, beta=-1.0, tol_rel=1e-11), BOXES, r0, r1, coef=ones)
    assert ps["status"] == 0 and st["status"] == 0
    assert abs(st["niters"] - ps["niters"]) <= 1
    sc = np.abs(p0 - p0.mean()).max()
    assert np.abs((a0 - a0.mean()) - (p0 - p0.mean())).max() <= 1e-9 * sc
    for a, p in zip(a1, p1):
        assert np.abs((a - a0.mean()) - (p - p0.mean())).max() <= 1e-9 * sc


@pytest.mark.parametrize("boxes", [BOXES, [(0, 40, 40, 15, 55, 55), (112 // 2, 10, 20, 63, 25, 35)]],
                         ids=["interior", "periodic_faces"])
def test_refluxed_abec_composite_operator_is_conservative(boxes):
    # alpha = 0: L u = -div(beta b grad u) in flux form; on the periodic domain
    # the volume-weighted composite residual of any field with rhs = 0 sums to
    # zero once every coarse/fine face carries the fine fluxes on both sides
    nc = 32
    coef, _, _ = _coefs(nc, boxes, 5, (1, 1, 1))
    u0, u1 = _rhs(nc, boxes, 6)
    z0 = np.zeros_like(u0)
    z1 = [np.zeros_like(x) for x in u1]
    P = _params(nc, PER, alpha=0.0, beta=0.8)
    res0, res1 = po.amr_residual(P, boxes, z0, z1, u0, u1, coef=coef)
    cov = np.zeros((nc,) * 3, bool)
    for b in boxes:
        cov[b[2] // 2:b[5] // 2 + 1, b[1] // 2:b[4] // 2 + 1, b[0] // 2:b[3] // 2 + 1] = True
    total = 8.0 * res0[~cov].sum() + sum(r.sum() for r in res1)
    scale = 8.0 * np.abs(res0[~cov]).sum() + sum(np.abs(r).sum() for r in res1)
    assert abs(total) <= 1e-12 * scale
    # a field with no variation has no residual at all: the test is not trivially met
    assert scale > 1.0


def test_coarse_coefficients_under_the_fine_level_are_the_fine_averages():
    nc = 32
    per = (1, 1, 1)
    boxes = BOXES + [(0, 40, 40, 15, 55, 55)]
    coef, fa, fb = _coefs(nc, boxes, 7, per)
    avg = _averaged(coef, fa, fb, boxes, per)
    r0, r1 = _rhs(nc, boxes, 8)
    P = _params(nc, PER, alpha=1.3, beta=0.7, tol_rel=1e-11)
    s0, s1, st = po.amr_solve(P, boxes, r0, r1, coef=coef)
    t0, t1, tt = po.amr_solve(P, boxes, r0, r1, coef=avg)
    assert st["status"] == 0 and st["niters"] == tt["niters"]
    np.testing.assert_allclose(s0, t0, rtol=0, atol=1e-13 * np.abs(t0).max())
    for a, b in zip(s1, t1):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-13 * np.abs(b).max())
    # and they differ from the coarse values given there: the average matters
    assert np.abs(avg[0]["bx"] - coef[0]["bx"]).max() > 0.1


def test_abec_composite_solve_reproduces_a_linear_field():
    # u = 1 + 0.3 x - 0.2 y + 0.1 z (x in coarse cell widths): every closure and
    # the coarse/fine interpolation are exact for it, and the reflux changes
    # nothing (the coarse b of a coarse/fine face is the mean of its fine
    # faces), so with the textbook operator of u as right-hand side on both
    # levels and u's face values as Dirichlet data the composite solution is u
    nc = 32
    boxes = LIN_BOXES
    per = (0, 0, 0)
    coef, fa, fb = _coefs(nc, boxes, 9, per)
    c0, c1 = _averaged(coef, fa, fb, boxes, per)
    alpha, beta = 1.3, 0.7
    u = lambda x, y, z: 1.0 + 0.3 * x - 0.2 * y + 0.1 * z

    def grown(lo, hi, h):
        zz, yy, xx = np.meshgrid(*[(np.arange(lo[d] - 1, hi[d] + 2) + 0.5) * h for d in (2, 1, 0)], indexing="ij")
        return u(xx, yy, zz)

    ug0 = grown((0, 0, 0), (nc - 1,) * 3, 1.0)
    rhs0 = _lin_apply(ug0, c0["a"], [c0["bx"], c0["by"], c0["bz"]], alpha, beta, float(nc))
    rhs1, u1 = [], []
    for q, bx in enumerate(boxes):
        ug = grown(bx[:3], bx[3:], 0.5)
        u1.append(ug[1:-1, 1:-1, 1:-1])
        rhs1.append(_lin_apply(ug, c1["a"][q], [c1["bx"][q], c1["by"][q], c1["bz"][q]], alpha, beta, 2.0 * nc))
    g = np.arange(-1, nc + 1) + 0.5
    g[0], g[-1] = 0.0, float(nc)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    bcv = u(x, y, z)
    for prec in ("f64", "mixed"):
        P = _params(nc, DIR_ALL, alpha=alpha, beta=beta, tol_rel=1e-12, precision=prec)
        s0, s1, st = po.amr_solve(P, boxes, rhs0, rhs1, bcv=bcv, coef=(c0, c1))
        assert st["status"] == 0 and st["niters"] > 1
        assert np.abs(s0 - ug0[1:-1, 1:-1, 1:-1]).max() < 1e-8
        assert max(np.abs(a - b).max() for a, b in zip(s1, u1)) < 1e-8


def test_no_fine_boxes_is_the_single_level_abec_solve():
    nc = 32
    coef, _, _ = _coefs(nc, [], 11, (1, 1, 1))
    r0, _ = _rhs(nc, [], 12)
    P = _params(nc, PER, alpha=1.3, beta=0.7)
    s0, _, st = po.amr_solve(P, [], r0, [], coef=coef)
    c0 = coef[0]
    ref, rst = po.solve(po.make_params((nc,) * 3, op="abec", max_grid_size=16, agg_grid_size=32, alpha=1.3, beta=0.7,
                                       **PER), r0, c0["a"], c0["bx"], c0["by"], c0["bz"])
    assert st["niters"] == rst["niters"]
    np.testing.assert_allclose(s0, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_mixed_precision_abec_composite_matches_fp64():
    nc = 32
    coef, _, _ = _coefs(nc, BOXES, 13, (1, 1, 1))
    r0, r1 = _rhs(nc, BOXES, 14)
    s0, s1, st = po.amr_solve(_params(nc, PER, alpha=1.3, beta=0.7), BOXES, r0, r1, coef=coef)
    m0, m1, sm = po.amr_solve(_params(nc, PER, alpha=1.3, beta=0.7, precision="mixed"), BOXES, r0, r1, coef=coef)
    assert st["status"] == 0 and sm["status"] == 0 and sm["niters"] <= st["niters"] + 1
    assert np.linalg.norm(m0 - s0) / np.linalg.norm(s0) <= 1e-9
    for a, b in zip(m1, s1):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-9


@pytest.mark.parametrize("boxes", [BOXES, [(0, 40, 40, 15, 55, 55), (56, 10, 20, 63, 25, 35)]],
                         ids=["interior", "periodic_faces"])
def test_fine_smoothing_fixes_the_solution_of_its_equation(boxes):
    nc = 32
    coef, _, _ = _coefs(nc, boxes, 15, (1, 1, 1))
    c0, c1 = _rhs(nc, boxes, 16)
    P = _params(nc, PER, alpha=1.3, beta=0.7)
    z0 = np.zeros_like(c0)
    z1 = [np.zeros_like(x) for x in c1]
    # res1 = L_f c1 with the homogeneous coarse/fine boundary (sol0 = 0)
    _, mr1 = po.amr_residual(P, boxes, z0, z1, z0, c1, coef=coef)
    res1 = [-x for x in mr1]
    got = po.amr_fine_sweeps(P, boxes, res1, c1, 2, coef=coef)
    for g, c in zip(got, c1):
        np.testing.assert_allclose(g, c, rtol=0, atol=1e-11 * np.abs(c).max())
    # from another start the sweeps move (they are not the identity)
    moved = po.amr_fine_sweeps(P, boxes, res1, z1, 1, coef=coef)
    assert max(np.abs(m).max() for m in moved) > 1e-3
