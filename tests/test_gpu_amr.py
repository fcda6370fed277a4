"""Two-level AMR composite Poisson on the GPU, through the C ABI.

A periodic coarse level and fine boxes of refinement ratio 2, solved by the
composite MLMG iteration.  Parity is against the oracle's composite solve
(tests/test_oracle_amr.py pins its pieces): the same number of iterations and
relative L2 differences <= 1e-9 on both levels.  The problem is singular
(periodic Poisson), so solutions are compared after removing the coarse
level's mean from both levels.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from amrex_amd import capi, mlmg, workload as wl  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

TOUCH = [(36, 40, 44, 51, 55, 59), (52, 40, 44, 67, 63, 59), (40, 70, 70, 55, 85, 85)]
# on the faces of the 128^3 fine domain: two boxes meeting across the x face, one on the y
# face, one spanning the whole z extent (its z ghosts are its own periodic images)
PERIODIC_FACES = [(0, 40, 40, 15, 55, 55), (112, 44, 36, 127, 59, 51), (60, 0, 20, 75, 11, 35),
                  (24, 90, 0, 39, 101, 127)]


PER = mlmg.LinOpBCType.Periodic


def _compare(nc, mgs, boxes, precision, tol=1e-10, use_graph=True, oracle_prec="f64", bc=None, bcv=None,
             rhs=None):
    print(f"\n[amr] nc={nc} boxes={len(boxes)} precision={precision} graph={use_graph} bc={bc} "
          f"bcv={bcv is not None}", flush=True)
    r0, r1 = wl.amr_rhs(nc, boxes) if rhs is None else rhs
    kw, okw = {}, {}
    if bc is not None:
        lo, hi = bc
        kw = dict(bc_lo=lo, bc_hi=hi)
        per = tuple(1 if lo[d] == PER else 0 for d in range(3))
        okw = dict(periodic=per, bc_lo=tuple(0 if per[d] else lo[d] for d in range(3)),
                   bc_hi=tuple(0 if per[d] else hi[d] for d in range(3)))
    prob = mlmg.AmrProblem(nc, mgs, boxes, precision=precision, use_graph=use_graph, bcv=bcv, **kw)
    g0, g1, _ = prob.solve(r0, r1, tol_rel=tol)
    st = prob.mlmg.stats()
    s0, s1, rst = po.amr_solve(po.amr_params(nc, max_grid_size=mgs, tol_rel=tol, precision=oracle_prec, **okw),
                               boxes, r0, r1, bcv=bcv)
    assert rst["status"] == 0
    m = s0.mean()
    gm = g0.mean()
    e0 = np.linalg.norm((g0 - gm) - (s0 - m)) / np.linalg.norm(s0 - m)
    num = sum(np.linalg.norm((a - gm) - (b - m)) ** 2 for a, b in zip(g1, s1))
    den = sum(np.linalg.norm(b - m) ** 2 for b in s1)
    e1 = np.sqrt(num / den)
    print(f"  iters {st['niters']} vs {rst['niters']}, coarse rel L2 {e0:.3e}, fine rel L2 {e1:.3e}", flush=True)
    return st, rst, e0, e1


@pytest.mark.parametrize("boxes", [wl.amr_boxes(64, 5), TOUCH, PERIODIC_FACES],
                         ids=["seeded", "touching", "periodic_faces"])
def test_composite_fp64_matches_oracle(boxes):
    st, rst, e0, e1 = _compare(64, 32, boxes, "double")
    assert st["niters"] == rst["niters"]
    assert st["final_resnorm"] <= 1e-10 * max(st["rhsnorm"], st["init_resnorm"])
    assert e0 <= 1e-9 and e1 <= 1e-9


DIR, NEU = 101, 102


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", ["dirichlet", "mixed_faces"])
def test_composite_with_physical_faces_matches_oracle(case, precision):
    # A coarse domain with homogeneous physical faces: boxes keep one coarse
    # cell from them (three at that distance); with a periodic y direction two
    # boxes sit on the y faces.
    if case == "dirichlet":
        bc = ((DIR,) * 3, (DIR,) * 3)
        boxes = [(2, 2, 2, 17, 13, 21), (24, 30, 20, 41, 45, 61), (44, 8, 40, 55, 19, 51)]
    else:
        bc = ((DIR, PER, NEU), (NEU, PER, DIR))
        boxes = [(4, 0, 6, 19, 11, 21), (30, 50, 30, 45, 63, 45), (2, 24, 40, 13, 35, 61)]
    st, rst, e0, e1 = _compare(32, 16, boxes, precision, bc=bc)
    if precision == "double":
        assert st["niters"] == rst["niters"]
    else:
        assert st["niters"] <= rst["niters"] + 1
    assert st["final_resnorm"] <= 1e-10 * max(st["rhsnorm"], st["init_resnorm"])
    assert e0 <= 1e-9 and e1 <= 1e-9


def test_composite_mixed_matches_oracle():
    st, rst, e0, e1 = _compare(64, 32, wl.amr_boxes(64, 5), "mixed")
    assert st["niters"] <= rst["niters"] + 1
    assert e0 <= 1e-9 and e1 <= 1e-9
    # against the oracle's own mixed form (FP32 coarse V-cycle and FP32 fine
    # smoothing sweeps): the same iteration count
    st, rst, e0, e1 = _compare(64, 32, wl.amr_boxes(64, 5), "mixed", oracle_prec="mixed")
    assert st["niters"] == rst["niters"]
    assert e0 <= 1e-9 and e1 <= 1e-9


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_composite_solve_from_device_tensors_equals_host_solve(precision):
    # AmrProblem.solve_device takes and returns CUDA tensors; the solve itself
    # is the same, so the results agree bit for bit
    nc, boxes = 64, wl.amr_boxes(64, 5)
    r0, r1 = wl.amr_rhs(nc, boxes)
    prob = mlmg.AmrProblem(nc, 32, boxes, precision=precision)
    h0, h1, _ = prob.solve(r0, r1)
    st = prob.mlmg.stats()
    d0, d1, _ = prob.solve_device(torch.as_tensor(r0, device="cuda"), [torch.as_tensor(b, device="cuda") for b in r1])
    assert prob.mlmg.stats()["niters"] == st["niters"]
    np.testing.assert_array_equal(d0.cpu().numpy(), h0)
    for a, b in zip(d1, h1):
        np.testing.assert_array_equal(a.cpu().numpy(), b)


def test_composite_graph_and_eager_agree():
    boxes = wl.amr_boxes(64, 5)
    r0, r1 = wl.amr_rhs(64, boxes)
    out = []
    for g in (True, False):
        prob = mlmg.AmrProblem(64, 32, boxes, use_graph=g)
        out.append(prob.solve(r0, r1))
        out[-1] = out[-1] + (prob.mlmg.stats()["niters"],)
    assert out[0][3] == out[1][3]
    np.testing.assert_allclose(out[0][0], out[1][0], rtol=0, atol=1e-12 * np.abs(out[1][0]).max())


@pytest.mark.parametrize("boxes", [TOUCH, PERIODIC_FACES], ids=["touching", "periodic_faces"])
@pytest.mark.parametrize("kind", ["grad", "flux"])
def test_composite_faces_match_oracle(kind, boxes):
    # getGradSolution / getFluxes on both levels against the oracle's faces of
    # the same (GPU) solution: coarse faces from the periodic coarse level, fine
    # faces per box with coarse/fine ghosts from the last solve's coarse solution
    r0, r1 = wl.amr_rhs(64, boxes)
    prob = mlmg.AmrProblem(64, 32, boxes)
    g0, g1, _ = prob.solve(r0, r1, tol_rel=1e-10)
    c, f = prob.face_values(kind)
    rc, rf = po.amr_faces(po.amr_params(64, max_grid_size=32), boxes, g0, g1, kind)
    for d in range(3):
        np.testing.assert_allclose(c[d], rc[d], rtol=0, atol=1e-12 * np.abs(rc[d]).max())
        for a, b in zip(f[d], rf[d]):
            assert a.shape == b.shape
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


@pytest.mark.slow
def test_composite_full_size_mixed():
    # coarse 256^3 periodic (max_grid_size 64), 40 seeded fine boxes, blob rhs,
    # mixed precision against the FP64 oracle
    st, rst, e0, e1 = _compare(256, 64, wl.amr_boxes(256, 40), "mixed")
    assert st["niters"] <= rst["niters"] + 1
    assert e0 <= 1e-9 and e1 <= 1e-9


LIN_BOXES = [(2, 2, 2, 17, 13, 21), (24, 30, 20, 41, 45, 61), (44, 8, 40, 55, 19, 51)]


@pytest.mark.parametrize("use_graph", [True, False], ids=["graph", "eager"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_composite_inhomogeneous_dirichlet_matches_oracle(precision, use_graph):
    # setLevelBC(0, data): seeded Dirichlet values on every coarse domain face
    # and a seeded right-hand side on both levels
    nc = 32
    bcv = np.random.default_rng(41).uniform(-2.0, 2.0, (nc + 2,) * 3)
    bc = ((DIR,) * 3, (DIR,) * 3)
    st, rst, e0, e1 = _compare(nc, 16, LIN_BOXES, precision, use_graph=use_graph, bc=bc, bcv=bcv)
    if precision == "double":
        assert st["niters"] == rst["niters"]
    else:
        assert st["niters"] <= rst["niters"] + 1
    assert st["final_resnorm"] <= 1e-10 * max(st["rhsnorm"], st["init_resnorm"])
    assert e0 <= 1e-9 and e1 <= 1e-9


def test_composite_inhomogeneous_dirichlet_reproduces_a_linear_field():
    # u = 1 + 0.3 x - 0.2 y + 0.1 z (x in coarse cell widths) is harmonic and
    # every closure is exact for it: with its face values as boundary data and
    # no right-hand side, the composite solution is u at the cell centres.
    nc = 32
    u = lambda x, y, z: 1.0 + 0.3 * x - 0.2 * y + 0.1 * z
    g = np.arange(-1, nc + 1) + 0.5
    g[0], g[-1] = 0.0, float(nc)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    bcv = u(x, y, z)
    zeros = (np.zeros((nc,) * 3), [np.zeros((b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)) for b in LIN_BOXES])
    for precision in ("double", "mixed"):
        prob = mlmg.AmrProblem(nc, 16, LIN_BOXES, precision=precision, bc_lo=(DIR,) * 3, bc_hi=(DIR,) * 3, bcv=bcv)
        prob.linop.setLevelBC(1, prob.sol1)  # accepted and ignored: no fine box touches a physical face
        g0, g1, _ = prob.solve(*zeros, tol_rel=1e-11)
        assert np.abs(g0 - bcv[1:-1, 1:-1, 1:-1]).max() < 1e-8
        for b, f in zip(LIN_BOXES, g1):
            zz, yy, xx = np.meshgrid(*[(np.arange(b[d], b[3 + d] + 1) + 0.5) / 2 for d in (2, 1, 0)], indexing="ij")
            assert np.abs(f - u(xx, yy, zz)).max() < 1e-8
