"""Pins of the oracle's patch solve with coarse/fine boundary data
(MLLinOp::setCoarseFineBC): the o3 boundary interpolation
(interpbndrydata_*_o3, AMReX_InterpBndryData_3D_K.H:23-118) and the Dirichlet
closure located 0.5*ratio*dx outside the coarse/fine faces
(AMReX_MLMGBndry.H:138-145, AMReX_MLLinOp_K.H:96-100).

What fixes them:
* o3 interpolation of coarse data sampled from a quadratic returns the
  quadratic at the fine ghost's tangential position and the coarse cell
  centre's normal position, where the full stencil applies; linear data
  where the stencil is one-sided.  A transposed tangential pair or a flipped
  tangential offset fails this.
* The 7-point operators are exact on quadratics, the order-3 closure is exact
  on quadratics, and the coarse cell centre next to a coarse/fine face lies
  0.5*ratio*dx outside it: the patch solve of rhs = L u for a quadratic u with
  coarse data u reproduces u to the solver tolerance.  Placing the boundary
  values at the face (bloc 0) or dropping the normal offset fails it.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def _quad(c):
    def u(x, y, z):
        return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * x + c[5] * y * y + c[6] * z * z
                + c[7] * x * y + c[8] * y * z + c[9] * x * z)
    return u


def _coarse(u, nc):
    hc = [1.0 / v for v in nc]
    K, J, I = np.meshgrid(*[(np.arange(nc[d]) + 0.5) * hc[d] for d in (2, 1, 0)], indexing="ij")
    return u(I, J, K)


def _expected_faces(u, dom_n, region, ratio):
    """Per coarse/fine face: u at the fine ghost's tangential centres and the
    normal centre of its coarse cell."""
    lo, hi = region
    h = [1.0 / v for v in dom_n]
    out = {}
    for f in range(6):
        d, side = f % 3, f // 3
        t1, t2 = (1, 2) if d == 0 else ((0, 2) if d == 1 else (0, 1))
        g = lo[d] - 1 if side == 0 else hi[d] + 1
        xn = (g // ratio + 0.5) * ratio * h[d]
        W, U = np.meshgrid(np.arange(lo[t2], hi[t2] + 1), np.arange(lo[t1], hi[t1] + 1), indexing="ij")
        pos = [None] * 3
        pos[d] = np.full(W.shape, xn)
        pos[t1] = (U + 0.5) * h[t1]
        pos[t2] = (W + 0.5) * h[t2]
        out[f] = u(*pos)
    return out


def test_o3_interpolation_exact_for_quadratics_with_full_stencils():
    dom, region, r = (64, 64, 64), ((16, 20, 12), (47, 43, 51)), 2
    u = _quad([0.3, 1.1, -0.7, 0.4, 0.9, -1.3, 0.6, 0.8, -0.5, 0.35])
    crse = _coarse(u, [v // r for v in dom])
    got = po.cf_boundary_values(dom, (0, 0, 0), region, r, crse)
    want = _expected_faces(u, dom, region, r)
    assert sorted(got) == list(range(6))
    for f in range(6):
        np.testing.assert_allclose(got[f], want[f], rtol=0, atol=1e-13)
    # planted mistakes: a transposed tangential pair, a flipped offset
    for f in (0, 4):
        g = got[f]
        assert np.abs(np.flip(g, axis=1) - want[f]).max() > 1e-4
    wrong = {f: v.copy() for f, v in got.items()}
    assert np.abs(np.flip(wrong[2], axis=0) - want[2]).max() > 1e-4


def test_o3_interpolation_one_sided_at_physical_faces_and_periodic_images():
    # region touching the y-lo physical face and spanning the periodic x
    # direction: y-hi and z faces are coarse/fine; the stencil turns one-sided
    # in y next to y = 0 (exact for data linear in y), full in x (images)
    dom, per, region, r = (32, 64, 64), (1, 0, 0), ((0, 0, 16), (31, 31, 47)), 2
    assert po.cf_faces(dom, per, region) == ["per", "phys", "cf", "per", "cf", "cf"]
    u = _quad([1.0, 0.0, 0.5, 0.3, 0.0, 0.0, 0.7, 0.0, 0.0, 0.0])
    crse = _coarse(u, [v // r for v in dom])
    got = po.cf_boundary_values(dom, per, region, r, crse)
    want = _expected_faces(u, dom, region, r)
    assert sorted(got) == [2, 4, 5]
    for f in got:
        np.testing.assert_allclose(got[f], want[f], rtol=0, atol=1e-13)
    # a y^2 term: next to the physical face (ghost rows y = 0, 1 of the z
    # faces, whose y-2 neighbour lies beyond it) the stencil is one-sided and
    # misses the curvature; further in, and on the y face, it is exact
    u2 = _quad([1.0, 0.0, 0.5, 0.3, 0.0, 0.6, 0.7, 0.0, 0.0, 0.0])
    got2 = po.cf_boundary_values(dom, per, region, r, _coarse(u2, [v // r for v in dom]))
    want2 = _expected_faces(u2, dom, region, r)
    for f in (2, 5):
        err = np.abs(got2[f] - want2[f])
        assert err[:2].min() > 1e-6
        assert err[2:].max() <= 1e-13
    np.testing.assert_allclose(got2[4], want2[4], rtol=0, atol=1e-13)


CASES = [
    # name, dom, periodic, region, bc_lo, bc_hi, op
    ("interior_poisson", (64, 64, 64), (0, 0, 0), ((16, 16, 16), (47, 47, 47)), (101,) * 3, (101,) * 3, "poisson"),
    ("interior_abec", (64, 64, 64), (0, 0, 0), ((16, 16, 16), (47, 47, 47)), (101,) * 3, (101,) * 3, "abec"),
    ("phys_and_periodic", (32, 64, 64), (1, 0, 0), ((0, 0, 16), (31, 31, 47)), (0, 101, 101), (0, 101, 101),
     "poisson"),
]


def _quadratic_patch(case, bloc_scale=1.0):
    name, dom, per, region, bc_lo, bc_hi, op = case
    r = 2
    if per[0]:
        u = _quad([1.0, 0.0, 0.5, 0.3, 0.0, 0.0, 0.7, 0.0, 0.0, 0.0])
        lap = 2 * 0.7
    else:
        u = _quad([0.3, 1.1, -0.7, 0.4, 0.9, -1.3, 0.6, 0.8, -0.5, 0.35])
        lap = 2 * (0.9 - 1.3 + 0.6)
    lo, hi = region
    m = [hi[d] - lo[d] + 1 for d in range(3)]
    h = [1.0 / v for v in dom]
    K, J, I = np.meshgrid(*[(np.arange(lo[d], hi[d] + 1) + 0.5) * h[d] for d in (2, 1, 0)], indexing="ij")
    uf = u(I, J, K)
    alpha, beta, aval, bval = (0.0, 1.0, 0.0, 1.0) if op == "poisson" else (1.3, 0.8, 1.7, 0.6)
    rhs = alpha * aval * uf - beta * bval * lap * np.ones_like(uf)
    if op == "poisson":
        rhs = lap * np.ones_like(uf)
    # physical Dirichlet data: u on the domain faces
    bcd = np.zeros((dom[2] + 2, dom[1] + 2, dom[0] + 2))
    cen = [np.concatenate(([0.0], (np.arange(dom[d]) + 0.5) * h[d], [1.0])) for d in range(3)]
    Kg, Jg, Ig = np.meshgrid(cen[2], cen[1], cen[0], indexing="ij")
    bcd[:] = u(Ig, Jg, Kg)
    crse = _coarse(u, [v // r for v in dom])
    prm = po.cf_patch_params(dom, per, region, r, bc_lo, bc_hi, op=op, max_grid_size=16, agg_grid_size=32,
                             alpha=alpha, beta=beta, tol_rel=1e-12)
    prm.bloc[:] = [v * bloc_scale for v in prm.bloc]
    bcv = po.cf_patch_bcv(dom, per, region, r, crse, bcdata=bcd)
    coef = {}
    if op == "abec":
        coef = dict(a=np.full((m[2], m[1], m[0]), aval), bx=np.full((m[2], m[1], m[0] + 1), bval),
                    by=np.full((m[2], m[1] + 1, m[0]), bval), bz=np.full((m[2] + 1, m[1], m[0]), bval))
    sol, st = po.solve(prm, rhs, bcv=bcv, **coef)
    return sol, uf, st


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_patch_solve_reproduces_quadratics(case):
    sol, uf, st = _quadratic_patch(case)
    assert st["status"] == 0
    err = np.abs(sol - uf).max() / np.abs(uf).max()
    assert err <= 1e-9, err
    # planted mistake: boundary values placed on the face instead of 0.5*ratio*dx out
    bad, uf, _ = _quadratic_patch(case, bloc_scale=0.0)
    assert np.abs(bad - uf).max() / np.abs(uf).max() > 1e-5
