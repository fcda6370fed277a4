"""Pins of the oracle's two-level AMR composite Poisson solve (CPU).

The reference ships no golden vectors for the composite path, and its build
is not available here, so each piece is tied to what the mathematics fixes:
  * the fine operator through the coarse/fine boundary (o3 tangential
    interpolation of coarse data + Lagrange closure at bcl = dx_f) is exact for
    linear functions, masked one-sided stencils included, and for quadratics
    around isolated boxes;
  * the refluxed composite operator is conservative: on a periodic domain the
    volume-weighted composite residual of any field sums to zero, which fixes
    the flux-register signs, scaling and face set;
  * the AMR correction interpolation reproduces linear functions;
  * with no fine boxes the composite solve is the single-level solve;
  * the composite solution is closer to a uniformly refined solve than the
    coarse-only solution is, on the refined region;
  * fine boxes on the periodic domain faces: the composite problem is
    invariant under periodic translation, so a configuration translated until
    its boxes cross a domain face (split there) converges to the translated
    solution; with such boxes the refluxed operator stays conservative.
"""
import numpy as np
import pytest

from amrex_amd import workload as wl
from oracle import pyoracle as po


def _centres(lo, hi, h):
    x = (np.arange(lo[0], hi[0] + 1) + 0.5) * h
    y = (np.arange(lo[1], hi[1] + 1) + 0.5) * h
    z = (np.arange(lo[2], hi[2] + 1) + 0.5) * h
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def _fields(nc, boxes, fn):
    X, Y, Z = _centres((0, 0, 0), (nc - 1,) * 3, 1.0 / nc)
    c = fn(X, Y, Z)
    f = [fn(*_centres(b[:3], b[3:], 0.5 / nc)) for b in boxes]
    return c, f


# two boxes sharing a face (covered ghosts, one-sided masked stencils) and an isolated one
TOUCH = [(36, 40, 44, 51, 55, 59), (52, 40, 44, 67, 63, 59), (40, 70, 70, 55, 85, 85)]
# boxes on the faces of the 128^3 fine domain: two that meet across the x face, one on the y
# face whose periodic image faces a coarse region, one spanning the whole z extent
PERIODIC_FACES = [(0, 40, 40, 15, 55, 55), (112, 44, 36, 127, 59, 51), (60, 0, 20, 75, 11, 35),
                  (24, 90, 0, 39, 101, 127)]
ISOLATED = [(36, 40, 44, 51, 55, 59), (40, 70, 70, 55, 85, 85)]


def test_fine_operator_exact_for_linear_functions():
    nc = 64
    lin = lambda X, Y, Z: 0.3 + 1.7 * X - 2.2 * Y + 0.9 * Z
    s0, s1 = _fields(nc, TOUCH, lin)
    zero0, zero1 = np.zeros_like(s0), [np.zeros_like(f) for f in s1]
    _, r1 = po.amr_residual(po.amr_params(nc, max_grid_size=32), TOUCH, zero0, zero1, s0, s1)
    scale = (2 * nc) ** 2 * 4.0
    for r in r1:
        assert np.abs(r).max() <= 1e-12 * scale


def test_fine_operator_exact_for_quadratics_around_isolated_boxes():
    nc = 64
    quad = lambda X, Y, Z: 0.5 * X * X - 1.5 * Y * Y + 0.75 * Z * Z + 0.2 * X * Y - 0.4 * Y * Z
    s0, s1 = _fields(nc, ISOLATED, quad)
    lap = 2 * (0.5 - 1.5 + 0.75)
    r0z = np.zeros_like(s0)
    rhs1 = [np.full_like(f, lap) for f in s1]
    _, r1 = po.amr_residual(po.amr_params(nc, max_grid_size=32), ISOLATED, r0z, rhs1, s0, s1)
    scale = (2 * nc) ** 2 * 4.0
    for r in r1:
        assert np.abs(r).max() <= 1e-10 * scale
    # with a touching neighbour the masked stencil is only first order: not exact
    t0, t1 = _fields(nc, TOUCH, quad)
    _, q1 = po.amr_residual(po.amr_params(nc, max_grid_size=32), TOUCH, np.zeros_like(t0),
                            [np.full_like(f, lap) for f in t1], t0, t1)
    assert max(np.abs(r).max() for r in q1) > 1e-6


@pytest.mark.parametrize("boxes", [TOUCH, wl.amr_boxes(64, 5), PERIODIC_FACES],
                         ids=["touching", "seeded", "periodic_faces"])
def test_refluxed_composite_operator_is_conservative(boxes):
    nc = 64
    rng = np.random.default_rng(5)
    s0 = rng.standard_normal((nc, nc, nc))
    s1 = [rng.standard_normal((b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)) for b in boxes]
    z0, z1 = np.zeros_like(s0), [np.zeros_like(f) for f in s1]
    prm = po.amr_params(nc, max_grid_size=32)
    r0, r1 = po.amr_residual(prm, boxes, z0, z1, s0, s1)
    covered = np.zeros((nc, nc, nc), dtype=bool)
    for b in boxes:
        covered[b[2] // 2:b[5] // 2 + 1, b[1] // 2:b[4] // 2 + 1, b[0] // 2:b[3] // 2 + 1] = True
    hc3, hf3 = (1.0 / nc) ** 3, (0.5 / nc) ** 3
    total = r0[~covered].sum() * hc3 + sum(r.sum() for r in r1) * hf3
    mag = np.abs(r0[~covered]).sum() * hc3 + sum(np.abs(r).sum() for r in r1) * hf3
    assert abs(total) <= 1e-13 * mag
    # without the reflux the sum would not vanish: the coarse/fine faces carry flux
    assert mag > 0


def test_amr_interpolation_reproduces_linear_functions():
    nc = 64
    lin = lambda X, Y, Z: -0.7 + 0.4 * X + 1.1 * Y - 2.5 * Z
    c, f = _fields(nc, TOUCH, lin)
    got = po.amr_interp(po.amr_params(nc, max_grid_size=32), TOUCH, c)
    for g, r in zip(got, f):
        np.testing.assert_allclose(g, r, rtol=0, atol=1e-13)


def test_no_fine_boxes_is_the_single_level_solve():
    nc = 32
    rhs, _ = wl.fourier_rhs((nc,) * 3, 1)
    s0, _, st = po.amr_solve(po.amr_params(nc, max_grid_size=16), [], rhs, [])
    ref, rst = po.solve(po.make_params((nc,) * 3, op="poisson", periodic=(1, 1, 1), bc_lo=(0,) * 3,
                                       bc_hi=(0,) * 3, max_grid_size=16, agg_grid_size=32), rhs)
    assert st["niters"] == rst["niters"]
    np.testing.assert_allclose(s0 - s0.mean(), ref - ref.mean(), rtol=0, atol=1e-12 * np.abs(ref).max())


def test_composite_solution_is_closer_to_the_refined_solve():
    nc = 32
    boxes = [(16, 16, 16, 31, 31, 31), (32, 18, 20, 47, 41, 43)]
    # smooth periodic right-hand side with structure inside the refined region
    fn = lambda X, Y, Z: np.exp(-((X - 0.6) ** 2 + (Y - 0.55) ** 2 + (Z - 0.55) ** 2) / (2 * 0.08 ** 2))
    c, f = _fields(nc, boxes, fn)
    mean = c.mean()
    c = c - mean
    f = [x - mean for x in f]
    s0, s1, st = po.amr_solve(po.amr_params(nc, max_grid_size=16, tol_rel=1e-12), boxes, c, f)
    assert st["status"] == 0
    # uniformly refined reference at 2nc
    cf, _ = _fields(2 * nc, [], fn)
    cf = cf - cf.mean()
    uf, ust = po.solve(po.make_params((2 * nc,) * 3, op="poisson", periodic=(1, 1, 1), bc_lo=(0,) * 3,
                                      bc_hi=(0,) * 3, max_grid_size=32, agg_grid_size=32, tol_rel=1e-12), cf)
    cs, _ = po.solve(po.make_params((nc,) * 3, op="poisson", periodic=(1, 1, 1), bc_lo=(0,) * 3, bc_hi=(0,) * 3,
                                    max_grid_size=16, agg_grid_size=32, tol_rel=1e-12), c)
    uf -= uf.mean()
    cs -= cs.mean()
    s_shift = s0.mean()
    err_amr = err_crse = 0.0
    for b, sf in zip(boxes, s1):
        ref = uf[b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1]
        crse = cs[b[2] // 2:b[5] // 2 + 1, b[1] // 2:b[4] // 2 + 1, b[0] // 2:b[3] // 2 + 1]
        crse = np.repeat(np.repeat(np.repeat(crse, 2, 0), 2, 1), 2, 2)
        err_amr = max(err_amr, np.abs(sf - s_shift - ref).max())
        err_crse = max(err_crse, np.abs(crse - ref).max())
    assert err_amr < 0.5 * err_crse, (err_amr, err_crse)


def test_mixed_precision_composite_matches_fp64():
    nc = 64
    boxes = wl.amr_boxes(nc, 5)
    r0, r1 = wl.amr_rhs(nc, boxes)
    s0, s1, st = po.amr_solve(po.amr_params(nc, max_grid_size=32), boxes, r0, r1)
    m0, m1, sm = po.amr_solve(po.amr_params(nc, max_grid_size=32, precision="mixed"), boxes, r0, r1)
    assert st["status"] == 0 and sm["status"] == 0
    assert sm["niters"] <= st["niters"] + 1
    d = (m0 - m0.mean()) - (s0 - s0.mean())
    assert np.linalg.norm(d) / np.linalg.norm(s0 - s0.mean()) <= 1e-9
    shift = m0.mean() - s0.mean()
    for a, b in zip(m1, s1):
        assert np.linalg.norm(a - shift - b) / np.linalg.norm(b) <= 1e-9


def _translate_x(boxes, fields, sc, nc):
    """Boxes and their fine data translated by 2*sc fine cells along x on the
    periodic fine domain; a box that crosses the x face is split there."""
    nf = 2 * nc
    out_b, out_f = [], []
    for b, f in zip(boxes, fields):
        lo, hi = b[0] + 2 * sc, b[3] + 2 * sc
        if lo >= nf:
            lo, hi = lo - nf, hi - nf
        if hi < nf:
            out_b.append((lo, b[1], b[2], hi, b[4], b[5]))
            out_f.append(f)
        else:
            cut = nf - lo
            out_b += [(lo, b[1], b[2], nf - 1, b[4], b[5]), (0, b[1], b[2], hi - nf, b[4], b[5])]
            out_f += [f[:, :, :cut], f[:, :, cut:]]
    return out_b, out_f


def test_composite_solve_is_invariant_under_periodic_translation():
    # The same problem solved as given and translated by 9 coarse cells, which
    # moves one box across the x face (split in two) and another onto it; the
    # converged composite solutions agree after translating back, up to the
    # solve tolerance and the additive constant of the periodic problem.
    nc, sc = 32, 9
    boxes = [(20, 16, 16, 35, 31, 31), (36, 18, 20, 49, 41, 43), (6, 40, 8, 17, 51, 19)]
    fn = lambda X, Y, Z: np.exp(-((X - 0.6) ** 2 + (Y - 0.55) ** 2 + (Z - 0.45) ** 2) / (2 * 0.1 ** 2))
    c, f = _fields(nc, boxes, fn)
    mean = c.mean()
    c, f = c - mean, [x - mean for x in f]
    prm = po.amr_params(nc, max_grid_size=16, tol_rel=1e-12)
    a0, a1, sta = po.amr_solve(prm, boxes, c, f)
    tb, tf = _translate_x(boxes, f, sc, nc)
    assert any(b[0] == 0 for b in tb) and any(b[3] == 2 * nc - 1 for b in tb)
    b0, b1, stb = po.amr_solve(prm, tb, np.roll(c, sc, axis=2), tf)
    assert sta["status"] == 0 and stb["status"] == 0
    ka, kb = a0.mean(), b0.mean()
    scale = np.abs(a0 - ka).max()
    np.testing.assert_allclose(np.roll(a0 - ka, sc, axis=2), b0 - kb, rtol=0, atol=1e-9 * scale)
    _, back = _translate_x(boxes, [x - ka for x in a1], sc, nc)
    for got, want in zip(b1, back):
        np.testing.assert_allclose(got - kb, want, rtol=0, atol=1e-9 * scale)


NEU_ALL = dict(periodic=(0, 0, 0), bc_lo=(102,) * 3, bc_hi=(102,) * 3)
DIR_ALL = dict(periodic=(0, 0, 0), bc_lo=(101,) * 3, bc_hi=(101,) * 3)


def test_refluxed_operator_is_conservative_with_neumann_faces():
    # Homogeneous Neumann faces let no flux leave the domain, so the refluxed
    # composite residual of any field still sums to zero; boxes keep one coarse
    # cell from the faces, one of them at that distance on three faces.
    nc = 32
    boxes = [(2, 2, 2, 17, 13, 21), (24, 30, 20, 41, 45, 61), (44, 8, 40, 55, 19, 51)]
    rng = np.random.default_rng(9)
    s0 = rng.standard_normal((nc, nc, nc))
    s1 = [rng.standard_normal((b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)) for b in boxes]
    z0, z1 = np.zeros_like(s0), [np.zeros_like(f) for f in s1]
    r0, r1 = po.amr_residual(po.amr_params(nc, max_grid_size=16, **NEU_ALL), boxes, z0, z1, s0, s1)
    covered = np.zeros((nc, nc, nc), dtype=bool)
    for b in boxes:
        covered[b[2] // 2:b[5] // 2 + 1, b[1] // 2:b[4] // 2 + 1, b[0] // 2:b[3] // 2 + 1] = True
    hc3, hf3 = (1.0 / nc) ** 3, (0.5 / nc) ** 3
    total = r0[~covered].sum() * hc3 + sum(r.sum() for r in r1) * hf3
    mag = np.abs(r0[~covered]).sum() * hc3 + sum(np.abs(r).sum() for r in r1) * hf3
    assert abs(total) <= 1e-13 * mag


def test_dirichlet_composite_solution_is_closer_to_the_refined_solve():
    nc = 32
    boxes = [(16, 16, 16, 31, 31, 31), (32, 18, 20, 47, 41, 43)]
    fn = lambda X, Y, Z: np.exp(-((X - 0.6) ** 2 + (Y - 0.55) ** 2 + (Z - 0.55) ** 2) / (2 * 0.08 ** 2))
    c, f = _fields(nc, boxes, fn)
    s0, s1, st = po.amr_solve(po.amr_params(nc, max_grid_size=16, tol_rel=1e-12, **DIR_ALL), boxes, c, f)
    assert st["status"] == 0
    cf, _ = _fields(2 * nc, [], fn)
    uf, _ = po.solve(po.make_params((2 * nc,) * 3, op="poisson", max_grid_size=32, agg_grid_size=32,
                                    tol_rel=1e-12, **DIR_ALL), cf)
    cs, _ = po.solve(po.make_params((nc,) * 3, op="poisson", max_grid_size=16, agg_grid_size=32, tol_rel=1e-12,
                                    **DIR_ALL), c)
    err_amr = err_crse = 0.0
    for b, sf in zip(boxes, s1):
        ref = uf[b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1]
        crse = cs[b[2] // 2:b[5] // 2 + 1, b[1] // 2:b[4] // 2 + 1, b[0] // 2:b[3] // 2 + 1]
        crse = np.repeat(np.repeat(np.repeat(crse, 2, 0), 2, 1), 2, 2)
        err_amr = max(err_amr, np.abs(sf - ref).max())
        err_crse = max(err_crse, np.abs(crse - ref).max())
    assert err_amr < 0.5 * err_crse, (err_amr, err_crse)


LIN_BOXES = [(2, 2, 2, 17, 13, 21), (24, 30, 20, 41, 45, 61), (44, 8, 40, 55, 19, 51)]


def _linear(nc, boxes, c):
    """u = c0 + c1 x + c2 y + c3 z (x in coarse cell widths from the low
    corner): coarse cell centres, fine box cell centres, and the ghost-grown
    coarse boundary data holding u at the centres of the domain faces (AMReX
    keeps Dirichlet values in the face ghosts)."""
    u = lambda x, y, z: c[0] + c[1] * x + c[2] * y + c[3] * z
    g = np.arange(nc) + 0.5
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    u0 = u(x, y, z)
    u1 = []
    for b in boxes:
        zz, yy, xx = np.meshgrid(*[(np.arange(b[d], b[3 + d] + 1) + 0.5) / 2 for d in (2, 1, 0)], indexing="ij")
        u1.append(u(xx, yy, zz))
    gg = np.arange(-1, nc + 1) + 0.5
    gg[0], gg[-1] = 0.0, float(nc)
    z, y, x = np.meshgrid(gg, gg, gg, indexing="ij")
    return u0, u1, u(x, y, z)


def _zeros(nc, boxes):
    return np.zeros((nc,) * 3), [np.zeros((b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)) for b in boxes]


def test_inhomogeneous_dirichlet_residual_of_a_linear_field_is_zero():
    # A linear field is harmonic and every closure (maxorder 3 Dirichlet, the
    # coarse/fine boundary values, the fine closure) is exact for it, so its
    # composite residual vanishes only if the coarse level's boundary data
    # enters the coarse operator with the right weight and sign.
    nc = 32
    c = (1.0, 0.3, -0.2, 0.1)
    u0, u1, bcv = _linear(nc, LIN_BOXES, c)
    z0, z1 = _zeros(nc, LIN_BOXES)
    P = po.amr_params(nc, max_grid_size=16, **DIR_ALL)
    r0, r1 = po.amr_residual(P, LIN_BOXES, z0, z1, u0, u1, bcv=bcv)
    assert np.abs(r0).max() < 1e-10 and max(np.abs(r).max() for r in r1) < 1e-10
    h0, _ = po.amr_residual(P, LIN_BOXES, z0, z1, u0, u1)
    assert np.abs(h0).max() > 1.0  # homogeneous closure: the boundary cells see u = 0 outside


def test_inhomogeneous_dirichlet_composite_solve_reproduces_a_linear_field():
    nc = 32
    c = (1.0, 0.3, -0.2, 0.1)
    u0, u1, bcv = _linear(nc, LIN_BOXES, c)
    z0, z1 = _zeros(nc, LIN_BOXES)
    for prec in ("f64", "mixed"):
        s0, s1, st = po.amr_solve(po.amr_params(nc, max_grid_size=16, precision=prec, tol_rel=1e-11, **DIR_ALL),
                                  LIN_BOXES, z0, z1, bcv=bcv)
        assert st["status"] == 0 and st["niters"] > 1
        scale = np.abs(u0).max()
        assert np.abs(s0 - u0).max() < 1e-8 * scale
        assert max(np.abs(a - b).max() for a, b in zip(s1, u1)) < 1e-8 * scale
