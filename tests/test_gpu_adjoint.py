"""Adjoint solve and coefficient gradients on the GPU, through the C ABI.

The reference has no adjoint.  Parity is against the oracle's adjoint part,
which tests/test_oracle_adjoint.py pins to the forward operator and to dense
linear algebra.  Bars: the transposed residual and the gradient kernels agree
with the oracle to 1e-12 relative on arbitrary inputs; adjoint solves agree to
relative L2 1e-9; end-to-end gradients (torch.autograd through the C ABI,
FP32-smoothing / FP64-refinement forward and adjoint) agree with the FP64 CPU
gradient to 1e-6 relative and with central finite differences (h = 1e-6) to
1e-6 relative (components below a tenth of the gradient's max norm measured
on the scale of that tenth).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from amrex_amd import adjoint, mlmg, workload as wl  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

DIR, NEU, REF = 101, 102, 103
DEV = torch.device("cuda")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _t(x):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV)


CASES = [
    # name, n, max_grid_size, op, periodic, bc_lo, bc_hi, maxorder
    ("abec_dir_mo3", (40, 48, 56), 16, "abec", (0, 0, 0), (DIR,) * 3, (DIR,) * 3, 3),
    ("abec_mixed", (64, 48, 40), 16, "abec", (0, 0, 0), (DIR, NEU, REF), (REF, DIR, NEU), 3),
    ("abec_periodic_x", (48, 40, 32), 16, "abec", (1, 0, 0), (0, DIR, DIR), (0, NEU, DIR), 3),
    ("abec_dir_mo4", (32, 40, 48), 16, "abec", (0, 0, 0), (DIR,) * 3, (DIR, NEU, DIR), 4),
    ("poisson_dir_mo3", (32, 32, 48), 16, "poisson", (0, 0, 0), (DIR,) * 3, (DIR, NEU, DIR), 3),
]
IDS = [c[0] for c in CASES]


def _coefs(n, seed):
    rng = np.random.default_rng(seed)
    return dict(a=rng.uniform(0.5, 1.5, (n[2], n[1], n[0])),
                bx=rng.uniform(0.5, 2.0, (n[2], n[1], n[0] + 1)),
                by=rng.uniform(0.5, 2.0, (n[2], n[1] + 1, n[0])),
                bz=rng.uniform(0.5, 2.0, (n[2] + 1, n[1], n[0])))


def _problem(case, coef, bcv=None, precision="double", dense=True):
    name, n, mgs, op, per, lo, hi, mo = case
    kw = coef if op == "abec" else {}
    return mlmg.Problem(n, mgs, op=op, periodic=per, bc_lo=lo, bc_hi=hi, maxorder=mo, alpha=1.3, beta=0.7,
                        bcv=bcv, precision=precision, dense=dense, **kw)


def _params(case, **kw):
    name, n, mgs, op, per, lo, hi, mo = case
    return po.make_params(n, op=op, periodic=per, bc_lo=lo, bc_hi=hi, max_grid_size=mgs, agg_grid_size=32,
                          maxorder=mo, alpha=1.3, beta=0.7, **kw)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_transposed_residual_matches_oracle(case):
    n = case[1]
    print(f"\n[transposed residual] {case[0]} n={n}", flush=True)
    coef = _coefs(n, 3)
    prob = _problem(case, coef)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((n[2], n[1], n[0]))
    b = rng.standard_normal((n[2], n[1], n[0]))
    xg = np.zeros((n[2] + 2, n[1] + 2, n[0] + 2))
    xg[1:-1, 1:-1, 1:-1] = x
    io = _t(xg)
    out = torch.zeros(n[2] * n[1] * n[0] + 2, dtype=torch.float64, device=DEV)
    prob.linop.levelOp(0, 9, in1=_t(b), io=io, out=out, count=1)
    got = out[:-2].cpu().numpy().reshape(b.shape)
    ref = b - po.apply_transpose(_params(case), x, **(coef if case[3] == "abec" else {}))
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert out[-2].item() == pytest.approx(np.abs(ref).max(), rel=1e-12)
    assert out[-1].item() == pytest.approx(ref.sum(), rel=1e-9, abs=1e-9 * np.abs(ref).sum())


# The adjoint cycle's kernels (maxorder <= 3): one colour of the transposed
# sweep, the transposed residual of a correction, the transposed Krylov
# bottom solve; against the oracle's (tests/test_oracle_adjoint.py pins the
# sweep to a Gauss-Seidel step on the dense transpose).
TCASES = [c for c in CASES if c[7] <= 3]


@pytest.mark.parametrize("precision", [0, 1], ids=["f64", "f32"])
@pytest.mark.parametrize("case", TCASES, ids=[c[0] for c in TCASES])
def test_transposed_colour_and_residual_match_oracle(case, precision):
    n = case[1]
    print(f"\n[transposed colour] {case[0]} n={n} precision={precision}", flush=True)
    coef = _coefs(n, 3)
    prob = _problem(case, coef)
    kw = coef if case[3] == "abec" else {}
    prm = _params(case, transpose=1)
    if precision:
        prm.precision = po.PRECISION["f32"]
    tol = 1e-12 if precision == 0 else 2e-6
    rng = np.random.default_rng(12)
    x = rng.standard_normal((n[2], n[1], n[0]))
    b = rng.standard_normal((n[2], n[1], n[0]))
    xg = np.zeros((n[2] + 2, n[1] + 2, n[0] + 2))
    xg[1:-1, 1:-1, 1:-1] = x
    for rb in (0, 1):
        io = _t(xg)
        prob.linop.levelOp(0, 15, precision, in1=_t(b), io=io, count=rb)
        got = io.cpu().numpy()[1:-1, 1:-1, 1:-1]
        ref = po.gsrb_colour(prm, xg, b, rb, **kw)[1:-1, 1:-1, 1:-1]
        np.testing.assert_allclose(got, ref, rtol=0, atol=tol * np.abs(ref).max())
    out = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device=DEV)
    prob.linop.levelOp(0, 16, precision, in1=_t(b), io=_t(xg), out=out)
    ref = b - po.apply_transpose(_params(case), x, **kw)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=0, atol=tol * np.abs(ref).max())


@pytest.mark.parametrize("solver", ["bicgstab", "cgbicg"])
def test_transposed_bottom_solve_matches_oracle(solver):
    # a small Dirichlet maxorder-3 level, where A^T differs from A
    n = (6, 5, 4)
    case = ("abec_dir_mo3_small", n, 8, "abec", (0, 0, 0), (DIR,) * 3, (DIR,) * 3, 3)
    coef = _coefs(n, 21)
    prob = _problem(case, coef)
    code = {"bicgstab": 2, "cgbicg": 5}[solver]
    b = np.random.default_rng(5).standard_normal((n[2], n[1], n[0]))
    x = _t(np.zeros_like(b))
    out = _t(np.zeros(3))
    prob.linop.levelOp(0, 17, in1=_t(b), io=x, out=out, count=code + 16 * 200)
    got = x.cpu().numpy()
    st = out.cpu().numpy().astype(int)
    prm = _params(case, transpose=1)
    prm.bottom_solver = code
    prm.bottom_maxiter = 200
    xo, oo = np.zeros_like(b), np.zeros(3)
    po.level_op(prm, 13, in1=b, io=xo, out=oo, **coef)
    assert (st[0], st[2]) == (int(oo[0]), int(oo[2]))
    assert abs(st[1] - int(oo[1])) <= 1
    assert st[0] == 0
    np.testing.assert_allclose(got, xo, rtol=0, atol=1e-6 * np.abs(xo).max())
    # it solved the transposed system (to the bottom tolerance, 1e-4 relative)
    r = b - po.apply_transpose(_params(case), got, **coef)
    assert np.abs(r).max() <= 1e-3 * np.abs(b).max()


@pytest.mark.parametrize("dense", [True, False], ids=["dense", "boxes"])
@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[2], CASES[3]], ids=IDS[:4])
@pytest.mark.parametrize("inhomog", [False, True])
def test_coefficient_gradient_kernels_match_oracle(case, inhomog, dense):
    # box storage: boxes of max_grid_size 16, each closing its own boundary
    # ghosts and looking its neighbours up in the other boxes
    n = case[1]
    print(f"\n[gradient kernels] {case[0]} inhomog={inhomog} dense={dense}", flush=True)
    coef = _coefs(n, 5)
    rng = np.random.default_rng(6)
    bcv = rng.standard_normal((n[2] + 2, n[1] + 2, n[0] + 2)) if inhomog else None
    prob = _problem(case, coef, bcv=bcv, dense=dense)
    lam = rng.standard_normal((n[2], n[1], n[0]))
    phi = rng.standard_normal((n[2], n[1], n[0]))
    ga = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device=DEV)
    gb = [torch.zeros(c.shape, dtype=torch.float64, device=DEV) for c in (coef["bx"], coef["by"], coef["bz"])]
    prob.coeff_gradient(_t(lam), _t(phi), ga, gb)
    ra, rb = po.grad_coeffs(_params(case), lam, phi, bcv=bcv, **coef)
    np.testing.assert_allclose(ga.cpu().numpy(), ra, rtol=0, atol=1e-13 * np.abs(ra).max())
    for d in range(3):
        np.testing.assert_allclose(gb[d].cpu().numpy(), rb[d], rtol=0, atol=1e-12 * np.abs(rb[d]).max())
    # accumulate adds a second contribution
    prob.coeff_gradient(_t(lam), _t(phi), ga, gb, accumulate=True)
    np.testing.assert_allclose(gb[0].cpu().numpy(), 2 * rb[0], rtol=0, atol=1e-12 * np.abs(rb[0]).max())


# Solve cases avoid reflect-odd faces: the reference's GSRB weight for them
# (+1, AMReX_MLLinOp_K.H:351, while the ghost is -phi) over-relaxes boundary
# cells, and with independent U[0.5, 2] face coefficients the reference
# algorithm itself diverges (the oracle shows it for forward and adjoint alike).
SOLVE_CASES = [
    CASES[0],
    ("abec_mixed_dn", (64, 48, 40), 16, "abec", (0, 0, 0), (DIR, NEU, NEU), (NEU, DIR, DIR), 3),
    CASES[2],
]


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", SOLVE_CASES, ids=[c[0] for c in SOLVE_CASES])
def test_adjoint_solve_matches_oracle(case, precision):
    n = case[1]
    print(f"\n[adjoint solve] {case[0]} {precision}", flush=True)
    coef = _coefs(n, 7)
    prob = _problem(case, coef, precision=precision)
    g = np.random.default_rng(8).standard_normal((n[2], n[1], n[0]))
    lam = prob.solve_adjoint(g, tol_rel=1e-11).cpu().numpy()
    st = prob.mlmg.stats()
    ref, rst = po.solve(_params(case, transpose=1, tol_rel=1e-11), g, **coef)
    assert rst["status"] == 0
    if precision == "double":
        assert st["niters"] == rst["niters"]
    else:
        assert st["niters"] <= rst["niters"] + 1
    assert st["final_resnorm"] <= 1e-11 * max(st["rhsnorm"], st["init_resnorm"])
    assert _rel(lam, ref) <= 1e-9
    # the forward solve on the same object is unaffected by the adjoint graphs,
    # and with the cycles on L^T (maxorder <= 3) the adjoint takes its iterations
    sol = prob.solve(g, tol_rel=1e-11).cpu().numpy()
    fst = prob.mlmg.stats()
    fwd, _ = po.solve(_params(case, tol_rel=1e-11), g, **coef)
    assert _rel(sol, fwd) <= 1e-9
    if case[7] <= 3:
        if precision == "double":
            assert st["niters"] == fst["niters"]
        else:
            assert st["niters"] <= fst["niters"] + 1


def _oracle_gradient(cfg_n, mgs, coef, rhs, target, tol):
    """FP64 CPU gradient of J = sum_i 0.5 |phi_i - t_i|^2 (oracle)."""
    prm = po.make_params(cfg_n, op="abec", max_grid_size=mgs, agg_grid_size=32, tol_rel=tol)
    prt = po.make_params(cfg_n, op="abec", max_grid_size=mgs, agg_grid_size=32, tol_rel=tol, transpose=1)
    gb = None
    lams, phis = [], []
    for i in range(rhs.shape[0]):
        phi, st = po.solve(prm, rhs[i], **coef)
        assert st["status"] == 0
        lam, st = po.solve(prt, phi - target[i], **coef)
        assert st["status"] == 0
        _, g = po.grad_coeffs(prm, lam, phi, **coef)
        gb = g if gb is None else [x + y for x, y in zip(gb, g)]
        lams.append(lam)
        phis.append(phi)
    return np.stack(phis), np.stack(lams), gb


def test_autograd_gradients_match_oracle_64():
    n = (64, 64, 64)
    print("\n[autograd] 64^3 batch 2, mixed precision", flush=True)
    a, bx, by, bz, _ = wl.abec_coeffs(n, seed=2)
    rng = np.random.default_rng(31)
    rhs = rng.standard_normal((2, n[2], n[1], n[0]))
    target = rng.standard_normal((2, n[2], n[1], n[0])) * 1e-3
    prob = mlmg.Problem(n, 32, op="abec", a=a, bx=bx, by=by, bz=bz, precision="mixed")
    tb = [_t(c).requires_grad_(True) for c in (bx, by, bz)]
    tr = _t(rhs).requires_grad_(True)
    phi = adjoint.solve(prob, tr, _t(a), *tb, tol_rel=1e-12)
    J = 0.5 * ((phi - _t(target)) ** 2).sum()
    J.backward()
    coef = dict(a=a, bx=bx, by=by, bz=bz)
    rphi, rlam, rgb = _oracle_gradient(n, 32, coef, rhs, target, 1e-12)
    assert _rel(phi.detach().cpu().numpy(), rphi) <= 1e-9
    assert _rel(tr.grad.cpu().numpy(), rlam) <= 1e-8
    for d in range(3):
        err = _rel(tb[d].grad.cpu().numpy(), rgb[d])
        print(f"  dJ/db_{'xyz'[d]} rel L2 vs oracle = {err:.3e}")
        assert err <= 1e-6


@pytest.mark.slow
def test_adjoint_config_256_batch8_fd_and_oracle():
    # ABecLaplacian 256^3 (homogeneous Dirichlet, max_grid_size 64, the
    # seeded coefficient recipe), batch of 8 seeded RHS fields,
    # J = 0.5 |phi - phi_t|^2 with phi_t N(0,1); mixed-precision forward and
    # adjoint; gradients w.r.t. the RHS and the face coefficients.
    cfg = wl.CONFIGS["abec256"]
    n, mgs = cfg.n, cfg.max_grid_size
    print(f"\n[adjoint config] {n} batch 8", flush=True)
    a, bx, by, bz, _ = wl.make_inputs(cfg)
    rhs, target = wl.adjoint_inputs(n, 8)
    prob = mlmg.Problem(n, mgs, op="abec", a=a, bx=bx, by=by, bz=bz, precision="mixed")
    tb = [_t(c).requires_grad_(True) for c in (bx, by, bz)]
    tr = _t(rhs).requires_grad_(True)
    ta = _t(a)
    tt = _t(target)
    phi = adjoint.solve(prob, tr, ta, *tb, tol_rel=1e-10)
    J = 0.5 * ((phi - tt) ** 2).sum()
    J.backward()
    print("  forward + adjoint done", flush=True)
    gb = [t.grad.detach() for t in tb]
    phi0 = phi.detach()

    # FP64 CPU gradient (oracle): whole arrays
    coef = dict(a=a, bx=bx, by=by, bz=bz)
    rphi, rlam, rgb = _oracle_gradient(n, mgs, coef, rhs, target, 1e-10)
    assert _rel(tr.grad.cpu().numpy(), rlam) <= 1e-6
    for d in range(3):
        err = _rel(gb[d].cpu().numpy(), rgb[d])
        print(f"  dJ/db_{'xyz'[d]} rel L2 vs FP64 CPU gradient = {err:.3e}", flush=True)
        assert err <= 1e-6

    # central differences at 32 seeded face locations: phi(b +- h e) = phi0 + delta
    # with L(b +- h e) delta = rhs - L(b +- h e) phi0, solved in FP64 to 1e-13 of
    # its own norm, so that the two solves' tolerances stay far below the
    # difference they resolve also beside Dirichlet faces, where phi and lambda
    # are small; err = |fd - ad| / |ad| <= 1e-6 at every location.
    h = 1e-6
    fd_prob = mlmg.Problem(n, mgs, op="abec", a=a, bx=bx, by=by, bz=bz, precision="double")
    scale = [float(g.abs().max()) for g in gb]
    worst = 0.0
    rows = []
    for m, (d, loc) in enumerate(wl.fd_locations(n, 32)):
        dJ = 0.0
        deltas = {}
        for sgn in (1.0, -1.0):
            cb = [t.detach().clone() for t in tb]
            cb[d][loc] += sgn * h
            fd_prob.set_coefficients(ta, *cb)
            acc = []
            for i in range(8):
                r = fd_prob.residual(phi0[i], tr.detach()[i])
                acc.append(fd_prob.solve(r, tol_rel=1e-13))
            deltas[sgn] = acc
        for i in range(8):
            dp, dm = deltas[1.0][i], deltas[-1.0][i]
            dJ += float(((dp - dm) * (phi0[i] - tt[i])).sum() + 0.5 * ((dp * dp) - (dm * dm)).sum())
        fd = dJ / (2 * h)
        ad = float(gb[d][loc])
        rel = abs(fd - ad) / max(abs(ad), 1e-300)
        worst = max(worst, rel)
        rows.append((m, d, loc, fd, ad, rel))
        print(f"  {m:2d} d={d} {loc} fd={fd:+.10e} ad={ad:+.10e} cpu={float(rgb[d][loc]):+.10e} err={rel:.2e}",
              flush=True)
    print(f"  FD check: max|dJ/db| = {scale}, worst relative difference over 32 locations = {worst:.3e}",
          flush=True)
    assert worst <= 1e-6, max(rows, key=lambda r: r[5])
