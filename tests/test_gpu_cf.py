"""Single-level solves of a patch with coarse/fine boundary data
(MLLinOp::setCoarseFineBC, AMReX_MLLinOp.H:200-212) on the HIP path, through
the C ABI, against the oracle's patch solve (pinned in test_oracle_cf.py).

The level-0 BoxArray covers one box of the domain; its faces inside the
domain take Dirichlet values located 0.5*ratio*dx outside, interpolated (o3)
from a coarse MultiFab, or homogeneous Neumann.  Bar: the same iteration
count as the oracle and relative L2 difference <= 1e-9; quadratic solutions
reproduced to the solver tolerance (closed form).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from amrex_amd import mlmg  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

DIR, NEU = 101, 102


def _embed(glob_shape, region, arr, nodal=(0, 0, 0)):
    g = np.zeros(glob_shape)
    lo, hi = region
    g[lo[2]:hi[2] + 1 + nodal[2], lo[1]:hi[1] + 1 + nodal[1], lo[0]:hi[0] + 1 + nodal[0]] = arr
    return g


def gpu_patch(dom, per, region, ratio, bc_lo, bc_hi, op, mgs, rhs, crse, bcd=None, coef=None, alpha=1.0, beta=1.0,
              maxorder=3, cf_type=None, tol=1e-10):
    dev = torch.device("cuda")
    lo, hi = region
    m = [hi[d] - lo[d] + 1 for d in range(3)]
    geom = mlmg.Geometry(dom, is_periodic=per)
    ba = mlmg.BoxArray(lo, hi).maxSize(mgs)
    dm = mlmg.DistributionMapping(ba)
    keep = []
    if op == "abec":
        lp = mlmg.MLABecLaplacian(geom, ba, dm)
    else:
        lp = mlmg.MLPoisson(geom, ba, dm)
    blo = [mlmg.LinOpBCType.Periodic if per[d] else bc_lo[d] for d in range(3)]
    bhi = [mlmg.LinOpBCType.Periodic if per[d] else bc_hi[d] for d in range(3)]
    lp.setDomainBC(blo, bhi)
    lp.setMaxOrder(maxorder)
    cmf = None
    if crse is not None:
        nc = [dom[d] // ratio for d in range(3)]
        cba = mlmg.BoxArray((0, 0, 0), [v - 1 for v in nc]).maxSize(16)
        cmf = mlmg.MultiFab(cba, mlmg.DistributionMapping(cba), 1, 0)
        cmf.fromGlobal(torch.as_tensor(crse, device=dev).contiguous(), nc)
        keep.append(cba)
    lp.setCoarseFineBC(cmf, ratio, cf_type)
    bmf = None
    if bcd is not None:
        bmf = mlmg.MultiFab(ba, dm, 1, 1)
        bmf.fromGlobal(torch.as_tensor(bcd, device=dev).contiguous(), dom, gg=1, rg=1)
    lp.setLevelBC(0, bmf)
    if op == "abec":
        lp.setScalars(alpha, beta)
        amf = mlmg.MultiFab(ba, dm, 1, 0)
        amf.fromGlobal(torch.as_tensor(_embed((dom[2], dom[1], dom[0]), region, coef["a"]), device=dev), dom)
        lp.setACoeffs(0, amf)
        bm = []
        for d, key in enumerate(("bx", "by", "bz")):
            nod = [0, 0, 0]
            nod[d] = 1
            shp = [dom[2] + nod[2], dom[1] + nod[1], dom[0] + nod[0]]
            mf = mlmg.MultiFab(ba, dm, 1, 0, nodal=nod)
            mf.fromGlobal(torch.as_tensor(_embed(shp, region, coef[key], nod), device=dev), dom)
            bm.append(mf)
        lp.setBCoeffs(0, bm)
        keep += [amf, bm]
    mg = mlmg.MLMG(lp)
    sol = mlmg.MultiFab(ba, dm, 1, 1)
    rmf = mlmg.MultiFab(ba, dm, 1, 0)
    rmf.fromGlobal(torch.as_tensor(_embed((dom[2], dom[1], dom[0]), region, rhs), device=dev), dom)
    sol.setVal(0.0)
    mg.solve(sol, rmf, tol, 0.0)
    out = torch.zeros((dom[2], dom[1], dom[0]), dtype=torch.float64, device=dev)
    sol.toGlobal(out, dom)
    g = out.cpu().numpy()
    return g[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1], mg.stats(), lp


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [
    # name, dom, periodic, region, bc_lo, bc_hi, op, ratio, maxorder, mgs, cf_type
    ("interior_abec", (64, 64, 64), (0, 0, 0), ((16, 16, 16), (47, 47, 47)), (DIR,) * 3, (DIR,) * 3, "abec", 2, 3,
     16, DIR),
    ("faces_abec_mo2_r4", (64, 64, 64), (0, 0, 0), ((16, 32, 0), (47, 63, 31)), (DIR, NEU, DIR), (DIR, DIR, NEU),
     "abec", 4, 2, 16, DIR),
    ("phys_periodic_poisson", (32, 64, 64), (1, 0, 0), ((0, 0, 16), (31, 31, 47)), (0, DIR, NEU), (0, DIR, DIR),
     "poisson", 2, 3, 16, DIR),
    ("corner_abec_uneven", (48, 40, 56), (0, 0, 0), ((0, 8, 24), (23, 39, 55)), (DIR, NEU, DIR), (NEU, DIR, DIR),
     "abec", 2, 3, 12, DIR),
    ("neumann_cf_abec", (64, 64, 64), (0, 0, 0), ((16, 16, 16), (47, 47, 47)), (DIR,) * 3, (DIR,) * 3, "abec", 2, 3,
     16, NEU),
    ("neumann_cf_poisson_touching_dirichlet", (32, 32, 64), (0, 0, 0), ((0, 8, 16), (15, 23, 47)),
     (DIR, NEU, NEU), (NEU,) * 3, "poisson", 2, 3, 16, NEU),
]


def _inputs(case, seed=5):
    name, dom, per, region, lo_bc, hi_bc, op, ratio, mo, mgs, cft = case
    rng = np.random.default_rng(seed)
    lo, hi = region
    m = [hi[d] - lo[d] + 1 for d in range(3)]
    coef = None
    if op == "abec":
        coef = dict(a=rng.uniform(0.5, 1.5, (m[2], m[1], m[0])), bx=rng.uniform(0.5, 2.0, (m[2], m[1], m[0] + 1)),
                    by=rng.uniform(0.5, 2.0, (m[2], m[1] + 1, m[0])), bz=rng.uniform(0.5, 2.0, (m[2] + 1, m[1], m[0])))
    nc = [dom[d] // ratio for d in range(3)]
    crse = rng.standard_normal((nc[2], nc[1], nc[0]))
    bcd = rng.standard_normal((dom[2] + 2, dom[1] + 2, dom[0] + 2))
    rhs = rng.standard_normal((m[2], m[1], m[0]))
    return coef, crse, bcd, rhs


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_patch_solve_matches_oracle(case):
    name, dom, per, region, lo_bc, hi_bc, op, ratio, mo, mgs, cft = case
    coef, crse, bcd, rhs = _inputs(case)
    alpha, beta = (1.2, 0.9) if op == "abec" else (0.0, 1.0)
    got, st, lp = gpu_patch(dom, per, region, ratio, lo_bc, hi_bc, op, mgs, rhs, crse, bcd=bcd, coef=coef,
                            alpha=alpha, beta=beta, maxorder=mo, cf_type=cft)
    prm = po.cf_patch_params(dom, per, region, ratio, lo_bc, hi_bc, cf_type=cft, op=op, max_grid_size=mgs,
                             agg_grid_size=32, maxorder=mo, alpha=alpha, beta=beta)
    # the oracle restates the patch as its own domain: compare its level extents
    # with the extents of the GPU levels' boxes (their domains are the whole domain's)
    def extent(boxes):
        return tuple(max(b[3 + d] for b in boxes) - min(b[d] for b in boxes) + 1 for d in range(3))
    assert [h["n"] for h in po.hierarchy(prm)] == [extent(lp.levelInfo(l)[1]) for l in range(lp.NMGLevels())]
    bcv = po.cf_patch_bcv(dom, per, region, ratio, crse if cft == DIR else None, bcdata=bcd)
    ref, rst = po.solve(prm, rhs, bcv=bcv, **(coef or {}))
    print(f"{name}: gpu {st['niters']} iterations, oracle {rst['niters']}; rel L2 {_rel(got, ref):.3e}")
    assert st["niters"] == rst["niters"]
    assert _rel(got, ref) <= 1e-9


def test_patch_solve_reproduces_quadratic():
    dom, region, r = (64, 64, 64), ((16, 16, 16), (47, 47, 47)), 2
    c = [0.3, 1.1, -0.7, 0.4, 0.9, -1.3, 0.6, 0.8, -0.5, 0.35]

    def u(x, y, z):
        return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * x + c[5] * y * y + c[6] * z * z
                + c[7] * x * y + c[8] * y * z + c[9] * x * z)
    lap = 2 * (c[4] + c[5] + c[6])
    h = 1.0 / 64
    lo, hi = region
    K, J, I = np.meshgrid(*[(np.arange(lo[d], hi[d] + 1) + 0.5) * h for d in (2, 1, 0)], indexing="ij")
    uf = u(I, J, K)
    hc = 1.0 / 32
    Kc, Jc, Ic = np.meshgrid(*[(np.arange(32) + 0.5) * hc for _ in range(3)], indexing="ij")
    got, st, _ = gpu_patch(dom, (0, 0, 0), region, r, (DIR,) * 3, (DIR,) * 3, "poisson", 16,
                           lap * np.ones_like(uf), u(Ic, Jc, Kc), tol=1e-12)
    err = np.abs(got - uf).max() / np.abs(uf).max()
    print(f"quadratic patch: max rel error {err:.3e}")
    assert err <= 1e-9


def test_coarse_data_read_at_set_level_bc():
    # the reference reads the coarse data when setLevelBC is called; a later
    # change of the coarse MultiFab does not reach the solve until then
    case = CASES[0]
    name, dom, per, region, lo_bc, hi_bc, op, ratio, mo, mgs, cft = case
    coef, crse, bcd, rhs = _inputs(case)
    a, _, _ = gpu_patch(dom, per, region, ratio, lo_bc, hi_bc, op, mgs, rhs, crse, bcd=bcd, coef=coef)
    b, _, _ = gpu_patch(dom, per, region, ratio, lo_bc, hi_bc, op, mgs, rhs, 2.0 * crse, bcd=bcd, coef=coef)
    assert _rel(a, b) > 1e-3
