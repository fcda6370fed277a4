"""Two-level AMR composite MLABecLaplacian on the GPU, through the C ABI.

Random a and face b on both levels, the coarse ones under the fine level
replaced by the fine averages when the operator is prepared
(averageDownCoeffsToCoarseAmrLevel).  Parity is against the oracle's composite
solve (tests/test_oracle_amr_abec.py pins its ABecLaplacian pieces): FP64 with
the same iteration count and relative L2 differences <= 1e-9 on both levels,
the mixed form within one iteration and the same bound; face fluxes of both
levels at 1e-12 of their largest value.  A manufactured linear field is
reproduced to 1e-8 (the closures and the coarse/fine interpolation are exact
for it).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from amrex_amd import mlmg, workload as wl  # noqa: E402
from oracle import pyoracle as po  # noqa: E402
from amrcoef import averaged, coefs, lin_apply  # noqa: E402

DIR = 101
PER = mlmg.LinOpBCType.Periodic
PERIODIC_FACES = [(0, 40, 40, 15, 55, 55), (112, 44, 36, 127, 59, 51), (60, 0, 20, 75, 11, 35),
                  (24, 90, 0, 39, 101, 127)]
LIN_BOXES = [(2, 2, 2, 17, 13, 21), (24, 30, 20, 41, 45, 61), (44, 8, 40, 55, 19, 51)]
ALPHA, BETA = 1.3, 0.7


def _rel(g0, g1, s0, s1):
    e0 = np.linalg.norm(g0 - s0) / np.linalg.norm(s0)
    num = sum(np.linalg.norm(a - b) ** 2 for a, b in zip(g1, s1))
    den = sum(np.linalg.norm(b) ** 2 for b in s1)
    return e0, np.sqrt(num / den)


def _run(nc, mgs, boxes, precision, seed, periodic=True, bcv=None, use_graph=True, tol=1e-10, oracle_prec="f64"):
    print(f"\n[amr abec] nc={nc} boxes={len(boxes)} precision={precision} graph={use_graph} "
          f"periodic={periodic} seed={seed}", flush=True)
    per = (1, 1, 1) if periodic else (0, 0, 0)
    coef, _, _ = coefs(nc, boxes, seed, per)
    rng = np.random.default_rng(seed + 1)
    r0 = rng.standard_normal((nc,) * 3)
    r1 = [rng.standard_normal((b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)) for b in boxes]
    kw = {} if periodic else dict(bc_lo=(DIR,) * 3, bc_hi=(DIR,) * 3)
    prob = mlmg.AmrProblem(nc, mgs, boxes, precision=precision, use_graph=use_graph, op="abec", alpha=ALPHA,
                           beta=BETA, coef=coef, bcv=bcv, **kw)
    g0, g1, _ = prob.solve(r0, r1, tol_rel=tol)
    st = prob.mlmg.stats()
    okw = dict(periodic=(1, 1, 1), bc_lo=(0,) * 3, bc_hi=(0,) * 3) if periodic else \
        dict(periodic=(0, 0, 0), bc_lo=(DIR,) * 3, bc_hi=(DIR,) * 3)
    P = po.amr_params(nc, max_grid_size=mgs, tol_rel=tol, precision=oracle_prec, op="abec", alpha=ALPHA, beta=BETA,
                      **okw)
    s0, s1, rst = po.amr_solve(P, boxes, r0, r1, bcv=bcv, coef=coef)
    assert rst["status"] == 0
    e0, e1 = _rel(g0, g1, s0, s1)
    print(f"  iters {st['niters']} vs {rst['niters']}, coarse rel L2 {e0:.3e}, fine rel L2 {e1:.3e}", flush=True)
    return prob, coef, (g0, g1, st), (s0, s1, rst), (e0, e1), P


@pytest.mark.parametrize("boxes", [wl.amr_boxes(64, 5), PERIODIC_FACES], ids=["seeded", "periodic_faces"])
def test_abec_composite_fp64_matches_oracle(boxes):
    _, _, (g0, g1, st), (s0, s1, rst), (e0, e1), _ = _run(64, 32, boxes, "double", 21)
    assert st["niters"] == rst["niters"]
    assert e0 <= 1e-9 and e1 <= 1e-9


@pytest.mark.parametrize("use_graph", [True, False], ids=["graph", "eager"])
def test_abec_composite_mixed_matches_oracle(use_graph):
    _, _, (g0, g1, st), (s0, s1, rst), (e0, e1), _ = _run(64, 32, wl.amr_boxes(64, 5), "mixed", 23,
                                                          use_graph=use_graph)
    assert st["niters"] <= rst["niters"] + 1
    assert e0 <= 1e-9 and e1 <= 1e-9


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_abec_composite_inhomogeneous_dirichlet_matches_oracle(precision):
    nc = 32
    bcv = np.random.default_rng(43).uniform(-2.0, 2.0, (nc + 2,) * 3)
    _, _, (g0, g1, st), (s0, s1, rst), (e0, e1), _ = _run(nc, 16, LIN_BOXES, precision, 25, periodic=False, bcv=bcv)
    if precision == "double":
        assert st["niters"] == rst["niters"]
    else:
        assert st["niters"] <= rst["niters"] + 1
    assert e0 <= 1e-9 and e1 <= 1e-9


@pytest.mark.parametrize("kind", ["flux", "grad"])
def test_abec_composite_face_values_match_oracle(kind):
    boxes = wl.amr_boxes(64, 5)
    prob, coef, (g0, g1, _), _, _, P = _run(64, 32, boxes, "double", 27)
    c, f = prob.face_values(kind)
    rc, rf = po.amr_faces(P, boxes, g0, g1, kind, coef=coef)
    for d in range(3):
        np.testing.assert_allclose(c[d], rc[d], rtol=0, atol=1e-12 * np.abs(rc[d]).max())
        for a, b in zip(f[d], rf[d]):
            assert a.shape == b.shape
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_abec_composite_reproduces_a_linear_field():
    # u = 1 + 0.3 x - 0.2 y + 0.1 z (x in coarse cell widths) with the textbook
    # operator of u (each level's own coefficients, the coarse ones averaged
    # under the fine level) as right-hand side and u's face values as
    # Dirichlet data: the composite solution is u on both levels.  The coarse
    # coefficients handed over are the raw random ones: the operator must
    # average the fine ones onto the covered cells and faces itself.
    nc = 32
    boxes = LIN_BOXES
    coef, fa, fb = coefs(nc, boxes, 29, (0, 0, 0))
    c0, c1 = averaged(coef, fa, fb, boxes, (0, 0, 0))
    u = lambda x, y, z: 1.0 + 0.3 * x - 0.2 * y + 0.1 * z

    def grown(lo, hi, h):
        zz, yy, xx = np.meshgrid(*[(np.arange(lo[d] - 1, hi[d] + 2) + 0.5) * h for d in (2, 1, 0)], indexing="ij")
        return u(xx, yy, zz)

    ug0 = grown((0, 0, 0), (nc - 1,) * 3, 1.0)
    rhs0 = lin_apply(ug0, c0["a"], [c0["bx"], c0["by"], c0["bz"]], ALPHA, BETA, float(nc))
    rhs1, u1 = [], []
    for q, bx in enumerate(boxes):
        ug = grown(bx[:3], bx[3:], 0.5)
        u1.append(ug[1:-1, 1:-1, 1:-1])
        rhs1.append(lin_apply(ug, c1["a"][q], [c1["bx"][q], c1["by"][q], c1["bz"][q]], ALPHA, BETA, 2.0 * nc))
    g = np.arange(-1, nc + 1) + 0.5
    g[0], g[-1] = 0.0, float(nc)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    bcv = u(x, y, z)
    for precision in ("double", "mixed"):
        print(f"\n[amr abec linear] precision={precision}", flush=True)
        prob = mlmg.AmrProblem(nc, 16, boxes, precision=precision, bc_lo=(DIR,) * 3, bc_hi=(DIR,) * 3, bcv=bcv,
                               op="abec", alpha=ALPHA, beta=BETA, coef=coef)
        g0, g1, _ = prob.solve(rhs0, rhs1, tol_rel=1e-12)
        assert prob.mlmg.stats()["niters"] > 1
        assert np.abs(g0 - ug0[1:-1, 1:-1, 1:-1]).max() < 1e-8
        assert max(np.abs(a - b).max() for a, b in zip(g1, u1)) < 1e-8
