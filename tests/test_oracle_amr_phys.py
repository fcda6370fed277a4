"""Pins of the oracle's two-level composite solve with fine boxes on physical
domain faces (CPU).

A fine box may touch a non-periodic domain face.  Its ghost cells there take
the face's own closure with the fine level's boundary data (setLevelBC(1,
data)): Dirichlet with the value at the face, Neumann.  Positions beyond the
face are neither covered nor coarse/fine: the o3 boundary interpolation
treats them like covered ones (one-sided slopes), the correction
interpolation reads 0 there, and the reflux does not wrap across the face.
Each piece is tied to something fixed independently:
  * a linear field is reproduced on both levels with boxes in the corners of a
    Dirichlet domain (Poisson and ABecLaplacian, FP64 and mixed), its face
    values as the boundary data of both levels; a periodic wrap of the
    coarse/fine interpolation across a physical face, or a coarse/fine
    closure in place of the physical one, breaks it;
  * the refluxed composite operator is conservative with boxes on Neumann
    faces (no flux leaves the domain);
  * one fine GSRB sweep from zero against a unit residual at a red cell on a
    physical face sets that cell to omega / diag, where diag is the closed
    row's diagonal computed here by hand: -6 dh + dh c_1 per physical face
    (c_1 = -2, the Lagrange weight of the first cell for nodes 0, 1/2, 3/2;
    1 for Neumann).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from amrcoef import averaged, coefs, lin_apply

DIR, NEU = 101, 102
# fine boxes (lo xyz, hi xyz) of a 32^3 coarse domain touching its faces: a
# corner box on the three low faces, one on the x-high face, one on the y- and
# z-high faces
CORNER_BOXES = [(0, 0, 0, 15, 11, 21), (48, 30, 20, 63, 45, 41), (20, 52, 44, 35, 63, 63)]


def _u(x, y, z):
    return 1.0 + 0.3 * x - 0.2 * y + 0.1 * z


def _face_data(m, h):
    """Ghost-grown (m+2)^3 array of u at the cell centres (width h), the
    ghosts at the domain faces."""
    g = (np.arange(-1, m + 1) + 0.5) * h
    g[0], g[-1] = 0.0, m * h
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return _u(x, y, z)


def _grown(lo, hi, h):
    zz, yy, xx = np.meshgrid(*[(np.arange(lo[d] - 1, hi[d] + 2) + 0.5) * h for d in (2, 1, 0)], indexing="ij")
    return _u(xx, yy, zz)


DIR_ALL = dict(periodic=(0, 0, 0), bc_lo=(DIR,) * 3, bc_hi=(DIR,) * 3)


@pytest.mark.parametrize("prec", ["f64", "mixed"])
@pytest.mark.parametrize("op", ["poisson", "abec"])
def test_linear_field_with_boxes_on_dirichlet_faces(op, prec):
    nc = 32
    boxes = CORNER_BOXES
    bcv0 = _face_data(nc, 1.0)
    bcv1 = _face_data(2 * nc, 0.5)
    if op == "poisson":
        # u is harmonic: no right-hand side
        rhs0 = np.zeros((nc,) * 3)
        rhs1 = [np.zeros((b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)) for b in boxes]
        coef, P = None, po.amr_params(nc, max_grid_size=16, tol_rel=1e-12, precision=prec, **DIR_ALL)
    else:
        coef, fa, fb = coefs(nc, boxes, 31, (0, 0, 0))
        c0, c1 = averaged(coef, fa, fb, boxes, (0, 0, 0))
        alpha, beta = 1.3, 0.7
        rhs0 = lin_apply(_grown((0, 0, 0), (nc - 1,) * 3, 1.0), c0["a"], [c0["bx"], c0["by"], c0["bz"]], alpha,
                         beta, float(nc))
        rhs1 = [lin_apply(_grown(b[:3], b[3:], 0.5), c1["a"][q], [c1["bx"][q], c1["by"][q], c1["bz"][q]], alpha,
                          beta, 2.0 * nc) for q, b in enumerate(boxes)]
        P = po.amr_params(nc, max_grid_size=16, tol_rel=1e-12, precision=prec, op="abec", alpha=alpha, beta=beta,
                          **DIR_ALL)
    s0, s1, st = po.amr_solve(P, boxes, rhs0, rhs1, bcv=bcv0, coef=coef, fine_bcv=bcv1)
    assert st["status"] == 0 and st["niters"] > 1
    assert np.abs(s0 - bcv0[1:-1, 1:-1, 1:-1]).max() < 1e-8
    for b, f in zip(boxes, s1):
        assert np.abs(f - _grown(b[:3], b[3:], 0.5)[1:-1, 1:-1, 1:-1]).max() < 1e-8
    if op == "poisson" and prec == "f64":
        # without the fine level's data the fine boundary is homogeneous: the
        # solution is no longer u
        t0, _, _ = po.amr_solve(P, boxes, rhs0, rhs1, bcv=bcv0)
        assert np.abs(t0 - bcv0[1:-1, 1:-1, 1:-1]).max() > 1e-3


def test_refluxed_operator_is_conservative_with_boxes_on_neumann_faces():
    # Neumann in x and z, periodic in y; boxes on the x and z faces (one on a
    # z-low / x-high edge) and one across the y wrap.  With rhs = 0 the
    # volume-weighted composite residual of any field sums to zero.
    nc = 32
    boxes = [(0, 8, 10, 13, 23, 25), (50, 40, 0, 63, 55, 17), (24, 0, 40, 39, 9, 63), (20, 56, 20, 31, 63, 31)]
    P = po.amr_params(nc, max_grid_size=16, periodic=(0, 1, 0), bc_lo=(NEU, 0, NEU), bc_hi=(NEU, 0, NEU))
    rng = np.random.default_rng(33)
    u0 = rng.standard_normal((nc,) * 3)
    u1 = [rng.standard_normal((b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)) for b in boxes]
    z0 = np.zeros_like(u0)
    z1 = [np.zeros_like(x) for x in u1]
    res0, res1 = po.amr_residual(P, boxes, z0, z1, u0, u1)
    cov = np.zeros((nc,) * 3, bool)
    for b in boxes:
        cov[b[2] // 2:b[5] // 2 + 1, b[1] // 2:b[4] // 2 + 1, b[0] // 2:b[3] // 2 + 1] = True
    total = 8.0 * res0[~cov].sum() + sum(r.sum() for r in res1)
    scale = 8.0 * np.abs(res0[~cov]).sum() + sum(np.abs(r).sum() for r in res1)
    assert abs(total) <= 1e-12 * scale and scale > 1.0


@pytest.mark.parametrize("bc", [DIR, NEU])
def test_fine_sweep_divisor_on_physical_faces(bc):
    nc = 16
    box = (0, 0, 4, 11, 9, 19)  # on the x-low and y-low faces
    P = po.amr_params(nc, max_grid_size=16, periodic=(0, 0, 0), bc_lo=(bc,) * 3, bc_hi=(bc,) * 3)
    dh = (2.0 * nc) ** 2
    c1 = -2.0 if bc == DIR else 1.0
    omega = 1.15
    shp = (box[5] - box[2] + 1, box[4] - box[1] + 1, box[3] - box[0] + 1)
    # red cells (i + j + k even): on the x face only, and on the x-y edge
    for (i, j, k), nphys in (((0, 4, 8), 1), ((0, 0, 10), 2)):
        res = np.zeros(shp)
        res[k - box[2], j - box[1], i - box[0]] = 1.0
        cor = po.amr_fine_sweeps(P, [box], [res], [np.zeros(shp)], 1)[0]
        diag = -6.0 * dh + nphys * dh * c1
        assert cor[k - box[2], j - box[1], i - box[0]] == pytest.approx(omega / diag, rel=1e-14)
