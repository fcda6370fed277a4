"""F-cycle iterations (MLMG::setMaxFmgIter) of the HIP path, through the C ABI.

Bar: against the oracle's F-cycle (which reads the coarse ghosts of the
interpolation as the reference leaves them), the same number of iterations,
residual histories within 1e-6 relative and solutions within relative L2
1e-9.  Dense and box storage, FP64 and mixed precision, levels with and
without agglomeration, one and two AMR levels.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from amrex_amd import mlmg, workload as wl  # noqa: E402
from oracle import pyoracle as po  # noqa: E402
from refprob import abec_prob2  # noqa: E402

OPREC = {"double": "f64", "mixed": "mixed"}


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _same_history(st, rst, precision="double"):
    # FP64: the same iterations, histories 1e-6 relative down to the rounding
    # floor of the last ones; mixed: the FP32 V-cycles of the two round
    # differently, so 1e-4 relative and at most one iteration apart
    if precision == "double":
        assert st["niters"] == rst["niters"]
    else:
        assert abs(st["niters"] - rst["niters"]) <= 1
    m = min(st["niters"], rst["niters"])
    h = np.array(rst["res_history"][:m])
    rtol = 1e-6 if precision == "double" else 1e-4
    np.testing.assert_allclose(st["res_history"][:m], h, rtol=rtol, atol=1e-12 * h.max())


@pytest.mark.parametrize("dense", [True, False], ids=["dense", "boxes"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("nfmg", [1, 3])
def test_abec_fcycle_matches_oracle(nfmg, precision, dense):
    n, mgs = (40, 48, 56), 16
    print(f"[fmg] abec n={n} nfmg={nfmg} {precision} dense={dense}", flush=True)
    a, bx, by, bz, rhs = wl.abec_coeffs(n, seed=2)
    prob = mlmg.Problem(n, mgs, op="abec", a=a, bx=bx, by=by, bz=bz, precision=precision, dense=dense)
    prob.mlmg.setMaxFmgIter(nfmg)
    sol = prob.solve(rhs).cpu().numpy()
    st = prob.mlmg.stats()
    prm = po.make_params(n, op="abec", max_grid_size=mgs, agg_grid_size=32, precision=OPREC[precision],
                         max_fmg_iters=nfmg)
    ref, rst = po.solve(prm, rhs, a, bx, by, bz)
    _same_history(st, rst, precision)
    assert _rel(sol, ref) <= 1e-9


def test_reference_problem_with_fcycles():
    # ABecLaplacian_C prob_type=2 (inhomogeneous Dirichlet + Neumann, maxorder 2) with two F-cycles
    n = 64
    pr = abec_prob2(n)
    prob = mlmg.Problem((n,) * 3, 32, op="abec", bc_lo=pr["bc_lo"], bc_hi=pr["bc_hi"], maxorder=2,
                        alpha=pr["alpha"], beta=pr["beta"], a=pr["a"], bx=pr["bx"], by=pr["by"],
                        bz=pr["bz"], bcv=pr["bcv"])
    prob.mlmg.setMaxFmgIter(2)
    sol = prob.solve(pr["rhs"]).cpu().numpy()
    st = prob.mlmg.stats()
    prm = po.make_params((n,) * 3, op="abec", bc_lo=pr["bc_lo"], bc_hi=pr["bc_hi"], max_grid_size=32,
                         agg_grid_size=32, maxorder=2, alpha=pr["alpha"], beta=pr["beta"], max_fmg_iters=2)
    ref, rst = po.solve(prm, pr["rhs"], pr["a"], pr["bx"], pr["by"], pr["bz"], bcv=pr["bcv"])
    assert st["niters"] == rst["niters"]
    _same_history(st, rst)
    assert _rel(sol, ref) <= 1e-9
    # the converged solution carries the same discretization error as the V-cycle solve
    assert np.abs(sol - pr["exact"]).max() == pytest.approx(2.369e-3, abs=0.0005e-3)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_periodic_poisson_fcycle_matches_oracle_and_closed_form(precision):
    cfg = wl.CONFIGS["poisson64"]
    rhs, modes = wl.fourier_rhs(cfg.n, cfg.seed)
    prob = mlmg.Problem(cfg.n, cfg.max_grid_size, op="poisson", periodic=(1, 1, 1), precision=precision)
    prob.mlmg.setMaxFmgIter(2)
    sol = prob.solve(rhs).cpu().numpy()
    st = prob.mlmg.stats()
    prm = po.make_params(cfg.n, op="poisson", periodic=(1, 1, 1), bc_lo=(0,) * 3, bc_hi=(0,) * 3,
                         max_grid_size=cfg.max_grid_size, agg_grid_size=32, precision=OPREC[precision],
                         max_fmg_iters=2)
    ref, rst = po.solve(prm, rhs)
    _same_history(st, rst, precision)
    assert _rel(sol - sol.mean(), ref - ref.mean()) <= 1e-9
    exact = wl.fourier_exact(cfg.n, modes)
    assert _rel(sol - sol.mean(), exact) <= 1e-9


@pytest.mark.parametrize("dense", [True, False], ids=["dense", "boxes"])
def test_agglomerated_hierarchy_fcycle_matches_oracle(dense):
    # 64^3 in 16^3 boxes with agg_grid_size 8: MG level 2 is agglomerated, so
    # the interpolation from it reads 0 outside the Dirichlet/Neumann faces
    # while the one from level 1 reads the ghosts level 1's sweeps left
    n = (64, 64, 64)
    DIR, NEU = 101, 102
    a, bx, by, bz, rhs = wl.abec_coeffs(n, seed=4)
    prob = mlmg.Problem(n, 16, op="abec", bc_lo=(DIR, NEU, DIR), bc_hi=(NEU, DIR, DIR), a=a, bx=bx, by=by, bz=bz,
                        agg_grid_size=8, dense=dense)
    prob.mlmg.setMaxFmgIter(2)
    sol = prob.solve(rhs).cpu().numpy()
    st = prob.mlmg.stats()
    prm = po.make_params(n, op="abec", bc_lo=(DIR, NEU, DIR), bc_hi=(NEU, DIR, DIR), max_grid_size=16,
                         agg_grid_size=8, max_fmg_iters=2)
    assert [len(h["boxes"]) for h in po.hierarchy(prm)][:3] == [64, 64, 8]
    ref, rst = po.solve(prm, rhs, a, bx, by, bz)
    assert st["niters"] == rst["niters"]
    _same_history(st, rst)
    assert _rel(sol, ref) <= 1e-9


def test_composite_fcycle_matches_oracle():
    nc, mgs = 64, 32
    boxes = wl.amr_boxes(nc, 5)
    r0, r1 = wl.amr_rhs(nc, boxes)
    prob = mlmg.AmrProblem(nc, mgs, boxes)
    prob.mlmg.setMaxFmgIter(2)
    g0, g1, _ = prob.solve(r0, r1, tol_rel=1e-10)
    st = prob.mlmg.stats()
    s0, s1, rst = po.amr_solve(po.amr_params(nc, max_grid_size=mgs, tol_rel=1e-10, max_fmg_iters=2), boxes, r0, r1)
    assert rst["status"] == 0
    assert st["niters"] == rst["niters"]
    _same_history(st, rst)
    m, gm = s0.mean(), g0.mean()
    assert np.linalg.norm((g0 - gm) - (s0 - m)) <= 1e-9 * np.linalg.norm(s0 - m)
    num = sum(np.linalg.norm((a - gm) - (b - m)) ** 2 for a, b in zip(g1, s1))
    den = sum(np.linalg.norm(b - m) ** 2 for b in s1)
    assert np.sqrt(num / den) <= 1e-9
