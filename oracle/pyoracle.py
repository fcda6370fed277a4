"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Arrays are numpy, global layout (z, y, x) with x fastest; ``grown`` arrays
carry one ghost layer on every side, shape (nz+2, ny+2, nx+2).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(_HERE, "_build", "liborc.so")


class Params(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("prob_lo", ctypes.c_double * 3),
        ("prob_hi", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3),
        ("bc_lo", ctypes.c_int * 3),
        ("bc_hi", ctypes.c_int * 3),
        ("max_grid_size", ctypes.c_int),
        ("agg_grid_size", ctypes.c_int),
        ("do_agglomeration", ctypes.c_int),
        ("maxorder", ctypes.c_int),
        ("op", ctypes.c_int),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("nu1", ctypes.c_int),
        ("nu2", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("verbose", ctypes.c_int),
        ("tol_rel", ctypes.c_double),
        ("tol_abs", ctypes.c_double),
        ("bottom_maxiter", ctypes.c_int),
        ("bottom_reltol", ctypes.c_double),
        ("fixed_iters", ctypes.c_int),
        ("max_coarsening_level", ctypes.c_int),
        ("transpose", ctypes.c_int),
        ("max_fmg_iters", ctypes.c_int),
        ("refine_vcycles", ctypes.c_int),
        ("bloc", ctypes.c_double * 6),
        ("bottom_solver", ctypes.c_int),
    ]


class GmresOut(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int),
        ("status", ctypes.c_int),
        ("resnorm", ctypes.c_double),
        ("rnorm0", ctypes.c_double),
        ("res_hist", ctypes.c_double * 512),
        ("nhist", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("niters", ctypes.c_int),
        ("nlevels", ctypes.c_int),
        ("status", ctypes.c_int),
        ("bottom_iters_total", ctypes.c_int),
        ("init_resnorm", ctypes.c_double),
        ("rhsnorm", ctypes.c_double),
        ("final_resnorm", ctypes.c_double),
        ("max_norm", ctypes.c_double),
        ("res_history", ctypes.c_double * 256),
        ("solve_seconds", ctypes.c_double),
        ("setup_seconds", ctypes.c_double),
        ("iter_seconds", ctypes.c_double),
        ("nvcycles", ctypes.c_int),
        ("bottom_solver_final", ctypes.c_int),
    ]


_lib = None


def build():
    """Compile the oracle library (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        lib = ctypes.CDLL(SO)
        assert lib.orc_params_size() == ctypes.sizeof(Params), "orc_params layout mismatch"
        assert lib.orc_stats_size() == ctypes.sizeof(Stats), "orc_stats layout mismatch"
        dp = ctypes.POINTER(ctypes.c_double)
        lib.orc_solve.argtypes = [ctypes.POINTER(Params), dp, dp, dp, dp, dp, dp, dp,
                                  ctypes.POINTER(Stats)]
        lib.orc_level_op.argtypes = [ctypes.POINTER(Params), ctypes.c_int, dp, dp, dp, dp, dp,
                                     dp, dp, dp, ctypes.c_int, ctypes.c_int]
        ip = ctypes.POINTER(ctypes.c_int)
        lib.orc_hierarchy.argtypes = [ctypes.POINTER(Params), ip, ip, ip, ip, ctypes.c_int]
        lib.orc_poly_interp_coeff.argtypes = [ctypes.c_double, dp, ctypes.c_int, dp]
        lib.orc_amr_solve.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ip, dp, dp, dp, dp,
                                      ctypes.POINTER(Stats), dp]
        lib.orc_amr_residual.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ip, dp, dp, dp, dp, dp, dp, dp]
        lib.orc_amr_interp.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ip, dp, dp]
        lib.orc_amr_faces.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ip, dp, dp, ctypes.c_int, dp, dp, dp]
        lib.orc_amr_set_coeffs.argtypes = [dp, dp, dp, dp, dp, dp]
        lib.orc_amr_set_fine_bc.argtypes = [dp]
        lib.orc_amr_fine_sweeps.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ip, dp, dp, ctypes.c_int]
        lib.orc_gmres.argtypes = [ctypes.POINTER(Params), dp, dp, dp, dp, dp, dp, dp, ip, ctypes.POINTER(GmresOut)]
        assert lib.orc_gmres_out_size() == ctypes.sizeof(GmresOut), "orc_gmres_out layout mismatch"
        _lib = lib
    return _lib


OPS = {"abec": 0, "poisson": 1}
PRECISION = {"f64": 0, "f32": 1, "mixed": 2}
# amrex::BottomSolver codes (AMReX_MLLinOp.H:30-32)
BOTTOM = {"default": 0, "smoother": 1, "bicgstab": 2, "cg": 3, "bicgcg": 4, "cgbicg": 5}


def make_params(n, *, op="abec", periodic=(0, 0, 0), bc_lo=(101,) * 3, bc_hi=(101,) * 3,
                max_grid_size=32, agg_grid_size=8, do_agglomeration=1, maxorder=3,
                alpha=1.0, beta=1.0, precision="f64", nu1=2, nu2=2, max_iter=200, verbose=0,
                tol_rel=1e-10, tol_abs=0.0, bottom_maxiter=200, bottom_reltol=-1.0,
                fixed_iters=0, max_coarsening_level=30, prob_lo=(0.0,) * 3, prob_hi=(1.0,) * 3,
                transpose=0, max_fmg_iters=0, refine_vcycles=1, bloc=(0.0,) * 6, bottom_solver="bicgstab"):
    p = Params()
    p.n[:] = list(n)
    p.prob_lo[:] = list(prob_lo)
    p.prob_hi[:] = list(prob_hi)
    p.periodic[:] = list(periodic)
    p.bc_lo[:] = list(bc_lo)
    p.bc_hi[:] = list(bc_hi)
    p.max_grid_size = max_grid_size
    p.agg_grid_size = agg_grid_size
    p.do_agglomeration = do_agglomeration
    p.maxorder = maxorder
    p.op = OPS[op]
    p.alpha, p.beta = alpha, beta
    p.precision = PRECISION[precision]
    p.nu1, p.nu2 = nu1, nu2
    p.max_iter, p.verbose = max_iter, verbose
    p.tol_rel, p.tol_abs = tol_rel, tol_abs
    p.bottom_maxiter, p.bottom_reltol = bottom_maxiter, bottom_reltol
    p.fixed_iters = fixed_iters
    p.max_coarsening_level = max_coarsening_level
    p.transpose = int(transpose)
    p.max_fmg_iters = int(max_fmg_iters)
    p.refine_vcycles = int(refine_vcycles)
    p.bloc[:] = [float(v) for v in bloc]
    p.bottom_solver = BOTTOM[bottom_solver] if isinstance(bottom_solver, str) else int(bottom_solver)
    return p


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def solve(params, rhs, a=None, bx=None, by=None, bz=None, bcv=None, sol0=None):
    """Run MLMG; returns (solution, stats dict)."""
    lib = _load()
    n = tuple(params.n)
    sol = np.zeros((n[2], n[1], n[0])) if sol0 is None else np.array(sol0, dtype=np.float64)
    st = Stats()
    a, bx, by, bz, rhs, bcv = map(_c, (a, bx, by, bz, rhs, bcv))
    lib.orc_solve(ctypes.byref(params), _dp(a), _dp(bx), _dp(by), _dp(bz), _dp(rhs), _dp(bcv),
                  _dp(sol), ctypes.byref(st))
    stats = {
        "niters": st.niters, "nlevels": st.nlevels, "status": st.status,
        "bottom_iters_total": st.bottom_iters_total, "init_resnorm": st.init_resnorm,
        "rhsnorm": st.rhsnorm, "final_resnorm": st.final_resnorm, "max_norm": st.max_norm,
        "res_history": list(st.res_history[: max(st.niters, 0)]),
        "solve_seconds": st.solve_seconds, "setup_seconds": st.setup_seconds,
        "iter_seconds": st.iter_seconds, "nvcycles": st.nvcycles, "bottom_solver": st.bottom_solver_final,
    }
    return sol, stats


def level_op(params, which, *, a=None, bx=None, by=None, bz=None, bcv=None, in1=None, io=None,
             out=None, count=0, level=0):
    lib = _load()
    a, bx, by, bz, bcv, in1 = map(_c, (a, bx, by, bz, bcv, in1))
    r = lib.orc_level_op(ctypes.byref(params), which, _dp(a), _dp(bx), _dp(by), _dp(bz), _dp(bcv),
                         _dp(in1), _dp(io), _dp(out), count, level)
    if r != 0:
        raise RuntimeError(f"orc_level_op({which}) failed: {r}")


def grown(n):
    return np.zeros((n[2] + 2, n[1] + 2, n[0] + 2))


def fmg_interp(params, level, crse_grown, fill=True, **coef):
    """F-cycle interpolation of MG level `level` (grown array, ghosts as given) to
    level-1: (fine valid cells, the coarse data the interpolation read).  fill:
    FillBoundary, or for an agglomerated level a zeroed copy with valid and
    periodic-image values (interpAssign)."""
    dims = [h["n"] for h in hierarchy(params)]
    nf = dims[level - 1]
    io = np.array(crse_grown, dtype=np.float64)
    out = np.zeros((nf[2], nf[1], nf[0]))
    level_op(params, 10, io=io, out=out, count=1 if fill else 0, level=level, **coef)
    return out, io


def smooth(params, phi_grown, rhs, nsweeps=1, **coef):
    """nsweeps x linop.smooth (red then black, homogeneous BC) on level 0."""
    io = np.array(phi_grown, dtype=np.float64)
    level_op(params, 0, in1=rhs, io=io, count=nsweeps, **coef)
    return io


def gsrb_colour(params, phi_grown, rhs, redblack, **coef):
    io = np.array(phi_grown, dtype=np.float64)
    level_op(params, 3, in1=rhs, io=io, count=redblack, **coef)
    return io


def apply(params, x_grown, bcv=None, **coef):
    """out = L(x) with ghosts filled (inhomogeneous Dirichlet if bcv)."""
    n = tuple(params.n)
    io = np.array(x_grown, dtype=np.float64)
    out = np.zeros((n[2], n[1], n[0]))
    level_op(params, 1, io=io, out=out, bcv=bcv, **coef)
    return out, io


def residual(params, x_grown, b, bcv=None, **coef):
    n = tuple(params.n)
    io = np.array(x_grown, dtype=np.float64)
    out = np.zeros((n[2], n[1], n[0]))
    level_op(params, 2, in1=b, io=io, out=out, bcv=bcv, **coef)
    return out


def apply_transpose(params, x, **coef):
    """out = L^T(x), homogeneous boundary data; x holds valid cells (n2, n1, n0)."""
    n = tuple(params.n)
    io = grown(n)
    io[1:-1, 1:-1, 1:-1] = x
    out = np.zeros((n[2], n[1], n[0]))
    level_op(params, 6, io=io, out=out, **coef)
    return out


def apply_rows(params, x, **coef):
    """out = L(x) from the explicit matrix rows, homogeneous boundary data."""
    n = tuple(params.n)
    io = grown(n)
    io[1:-1, 1:-1, 1:-1] = x
    out = np.zeros((n[2], n[1], n[0]))
    level_op(params, 8, io=io, out=out, **coef)
    return out


def grad_coeffs(params, lam, phi, bcv=None, **coef):
    """(dJ/da, [dJ/dbx, dJ/dby, dJ/dbz]) given the adjoint lam and the solution phi."""
    n = tuple(params.n)
    io = grown(n)
    io[1:-1, 1:-1, 1:-1] = phi
    nc = n[0] * n[1] * n[2]
    shapes = [(n[2], n[1], n[0] + 1), (n[2], n[1] + 1, n[0]), (n[2] + 1, n[1], n[0])]
    out = np.zeros(nc + sum(int(np.prod(s)) for s in shapes))
    level_op(params, 7, in1=np.ascontiguousarray(lam), io=io, out=out, bcv=bcv, **coef)
    ga = out[:nc].reshape(n[2], n[1], n[0])
    gb, off = [], nc
    for s in shapes:
        m = int(np.prod(s))
        gb.append(out[off:off + m].reshape(s))
        off += m
    return ga, gb


def face_values(params, phi, mode, bcv=None, **coef):
    """[x, y, z] face arrays of the solution phi (valid cells): mode "grad"
    (MLMG::getGradSolution) or "flux" (MLMG::getFluxes), ghosts filled with the
    level's boundary data (inhomogeneous Dirichlet if bcv)."""
    n = tuple(params.n)
    io = grown(n)
    io[1:-1, 1:-1, 1:-1] = phi
    shapes = [(n[2], n[1], n[0] + 1), (n[2], n[1] + 1, n[0]), (n[2] + 1, n[1], n[0])]
    out = np.zeros(sum(int(np.prod(s)) for s in shapes))
    level_op(params, 9, io=io, out=out, bcv=bcv, count={"grad": 0, "flux": 1}[mode], **coef)
    res, off = [], 0
    for s in shapes:
        m = int(np.prod(s))
        res.append(out[off:off + m].reshape(s))
        off += m
    return res


def hierarchy(params, max_boxes=100000):
    lib = _load()
    nlev = ctypes.c_int()
    dims = (ctypes.c_int * 96)()
    nbox = (ctypes.c_int * 32)()
    boxes = (ctypes.c_int * (6 * max_boxes))()
    tot = lib.orc_hierarchy(ctypes.byref(params), ctypes.byref(nlev), dims, nbox, boxes, max_boxes)
    out = []
    off = 0
    for l in range(nlev.value):
        bl = []
        for _ in range(nbox[l]):
            bl.append(tuple(boxes[6 * off: 6 * off + 6]))
            off += 1
        out.append({"n": tuple(dims[3 * l: 3 * l + 3]), "boxes": bl})
    assert off == tot
    return out


def poly_interp_coeff(x_int, x):
    lib = _load()
    xs = np.ascontiguousarray(x, dtype=np.float64)
    c = np.zeros(len(xs))
    lib.orc_poly_interp_coeff(x_int, _dp(xs), len(xs), _dp(c))
    return c


# ---- two-level AMR composite Poisson (coarse periodic level + fine boxes, ratio 2)

def amr_params(nc, *, max_grid_size=64, agg_grid_size=32, maxorder=3, precision="f64", tol_rel=1e-10,
               tol_abs=0.0, max_iter=200, fixed_iters=0, verbose=0, max_fmg_iters=0, periodic=(1, 1, 1),
               bc_lo=(0,) * 3, bc_hi=(0,) * 3, op="poisson", alpha=1.0, beta=1.0):
    """Two-level parameters: MLPoisson, or op="abec" (MLABecLaplacian with
    scalars alpha, beta; the coefficients go to the amr_* calls as coef)."""
    return make_params((nc,) * 3, op=op, periodic=periodic, bc_lo=bc_lo, bc_hi=bc_hi,
                       max_grid_size=max_grid_size, agg_grid_size=agg_grid_size, maxorder=maxorder,
                       precision=precision, tol_rel=tol_rel, tol_abs=tol_abs, max_iter=max_iter,
                       fixed_iters=fixed_iters, verbose=verbose, max_fmg_iters=max_fmg_iters, alpha=alpha, beta=beta)


class _AmrCoef:
    """Context: hands the two-level ABecLaplacian coefficients and the fine
    level's boundary data to the oracle for the calls inside.  coef = (coarse
    {a, bx, by, bz} global arrays, fine {a: [per box (z, y, x)], bx/by/bz: [per
    box faces (z, y, x)]}); fine_bcv: ghost-grown (2n+2)^3 fine-domain array
    whose face ghosts hold the Dirichlet values of the physical faces."""

    def __init__(self, coef, fine_bcv=None):
        self.coef = coef
        self.fine_bcv = None if fine_bcv is None else np.ascontiguousarray(fine_bcv, dtype=np.float64)
        self.keep = []

    def __enter__(self):
        lib = _load()
        lib.orc_amr_set_fine_bc(_dp(self.fine_bcv) if self.fine_bcv is not None else None)
        if self.coef is None:
            lib.orc_amr_set_coeffs(None, None, None, None, None, None)
            return self
        c0, c1 = self.coef
        arrs = [np.ascontiguousarray(c0[k], dtype=np.float64) for k in ("a", "bx", "by", "bz")]
        arrs.append(_fine_flat(c1["a"]))
        arrs.append(_fine_flat([f for d in ("bx", "by", "bz") for f in c1[d]]))
        self.keep = arrs
        lib.orc_amr_set_coeffs(*[_dp(x) for x in arrs])
        return self

    def __exit__(self, *exc):
        _load().orc_amr_set_coeffs(None, None, None, None, None, None)
        _load().orc_amr_set_fine_bc(None)
        self.keep = []
        return False


def _boxes_arr(boxes):
    flat = np.ascontiguousarray(np.array(boxes, dtype=np.int32).ravel()) if boxes else np.zeros(6, np.int32)
    return flat, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _fine_flat(fine):
    if not fine:
        return np.zeros(1)
    return np.ascontiguousarray(np.concatenate([np.ascontiguousarray(f, dtype=np.float64).ravel() for f in fine]))


def _fine_split(flat, boxes):
    out, off = [], 0
    for b in boxes:
        shp = (b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)
        m = int(np.prod(shp))
        out.append(flat[off:off + m].reshape(shp))
        off += m
    return out


def amr_solve(params, boxes, rhs0, rhs1, sol0=None, sol1=None, bcv=None, coef=None, fine_bcv=None):
    """Composite solve; returns (sol0, [sol1 per box], stats).  bcv: the coarse
    level's boundary data, ghost-grown (n+2)^3 (None: homogeneous); coef: the
    ABecLaplacian coefficients (_AmrCoef)."""
    lib = _load()
    n = tuple(params.n)
    barr, bp = _boxes_arr(boxes)
    r0 = np.ascontiguousarray(rhs0, dtype=np.float64)
    r1 = _fine_flat(rhs1)
    s0 = np.zeros((n[2], n[1], n[0])) if sol0 is None else np.array(sol0, dtype=np.float64)
    s1 = np.zeros_like(r1) if sol1 is None else _fine_flat(sol1).copy()
    st = Stats()
    bcv = _c(bcv)
    with _AmrCoef(coef, fine_bcv):
        r = lib.orc_amr_solve(ctypes.byref(params), len(boxes), bp, _dp(r0), _dp(r1), _dp(s0), _dp(s1),
                              ctypes.byref(st), _dp(bcv))
    if r < 0:
        raise RuntimeError(f"orc_amr_solve failed: {r}")
    stats = {"niters": st.niters, "nlevels": st.nlevels, "status": st.status, "init_resnorm": st.init_resnorm,
             "rhsnorm": st.rhsnorm, "final_resnorm": st.final_resnorm, "max_norm": st.max_norm,
             "res_history": list(st.res_history[: max(st.niters, 0)]), "solve_seconds": st.solve_seconds,
             "setup_seconds": st.setup_seconds, "iter_seconds": st.iter_seconds,
             "bottom_iters_total": st.bottom_iters_total, "bottom_solver": st.bottom_solver_final}
    return s0, _fine_split(s1, boxes), stats


def amr_residual(params, boxes, rhs0, rhs1, sol0, sol1, bcv=None, coef=None, fine_bcv=None):
    """Composite residual (res0 refluxed, res1 with the coarse/fine boundary from sol0)."""
    lib = _load()
    n = tuple(params.n)
    barr, bp = _boxes_arr(boxes)
    r0 = np.ascontiguousarray(rhs0, dtype=np.float64)
    r1 = _fine_flat(rhs1)
    s0 = np.ascontiguousarray(sol0, dtype=np.float64)
    s1 = _fine_flat(sol1)
    o0 = np.zeros((n[2], n[1], n[0]))
    o1 = np.zeros_like(r1)
    bcv = _c(bcv)
    with _AmrCoef(coef, fine_bcv):
        r = lib.orc_amr_residual(ctypes.byref(params), len(boxes), bp, _dp(r0), _dp(r1), _dp(s0), _dp(s1), _dp(o0),
                                 _dp(o1), _dp(bcv))
    if r < 0:
        raise RuntimeError(f"orc_amr_residual failed: {r}")
    return o0, _fine_split(o1, boxes)


def amr_fine_sweeps(params, boxes, res1, cor1, nsweeps=1, coef=None, fine_bcv=None):
    """nsweeps of the composite cycle's fine smoothing (homogeneous coarse/fine
    boundary) of cor1 against res1; returns the new cor1 per box."""
    lib = _load()
    barr, bp = _boxes_arr(boxes)
    r1 = _fine_flat(res1)
    c1 = _fine_flat(cor1).copy()
    with _AmrCoef(coef, fine_bcv):
        r = lib.orc_amr_fine_sweeps(ctypes.byref(params), len(boxes), bp, _dp(r1), _dp(c1), nsweeps)
    if r < 0:
        raise RuntimeError(f"orc_amr_fine_sweeps failed: {r}")
    return _fine_split(c1, boxes)


def amr_interp(params, boxes, crse):
    lib = _load()
    barr, bp = _boxes_arr(boxes)
    c = np.ascontiguousarray(crse, dtype=np.float64)
    nf = sum((b[3] - b[0] + 1) * (b[4] - b[1] + 1) * (b[5] - b[2] + 1) for b in boxes)
    out = np.zeros(nf)
    lib.orc_amr_interp(ctypes.byref(params), len(boxes), bp, _dp(c), _dp(out))
    return _fine_split(out, boxes)


def amr_faces(params, boxes, sol0, sol1, mode, bcv=None, coef=None, fine_bcv=None):
    """Face gradients (mode "grad") or fluxes ("flux") of both levels:
    ([x, y, z] coarse face arrays, [x, y, z] lists of per-box fine face arrays (z, y, x))."""
    lib = _load()
    n = tuple(params.n)
    barr, bp = _boxes_arr(boxes)
    s0 = np.ascontiguousarray(sol0, dtype=np.float64)
    s1 = _fine_flat(sol1)
    shapes0 = [(n[2], n[1], n[0] + 1), (n[2], n[1] + 1, n[0]), (n[2] + 1, n[1], n[0])]
    o0 = np.zeros(sum(int(np.prod(s)) for s in shapes0))
    fshape = [[(b[5] - b[2] + 1 + (d == 2), b[4] - b[1] + 1 + (d == 1), b[3] - b[0] + 1 + (d == 0)) for b in boxes]
              for d in range(3)]
    o1 = np.zeros(max(1, sum(int(np.prod(s)) for sl in fshape for s in sl)))
    bcv = _c(bcv)
    with _AmrCoef(coef, fine_bcv):
        r = lib.orc_amr_faces(ctypes.byref(params), len(boxes), bp, _dp(s0), _dp(s1), {"grad": 0, "flux": 1}[mode],
                              _dp(o0), _dp(o1), _dp(bcv))
    if r < 0:
        raise RuntimeError(f"orc_amr_faces failed: {r}")
    c, off = [], 0
    for s in shapes0:
        m = int(np.prod(s))
        c.append(o0[off:off + m].reshape(s))
        off += m
    f, off = [], 0
    for sl in fshape:
        lev = []
        for s in sl:
            m = int(np.prod(s))
            lev.append(o1[off:off + m].reshape(s))
            off += m
        f.append(lev)
    return c, f


def gmres(params, rhs, a=None, bx=None, by=None, bz=None, bcv=None, sol0=None, use_precond=True,
          precond_niters=1, restart=30, maxiter=2000):
    """GMRESMLMG solve (FP64); returns (sol, {"iters", "status", "resnorm", "rnorm0", "res_history"})."""
    lib = _load()
    n = tuple(params.n)
    r = np.ascontiguousarray(rhs, dtype=np.float64)
    sol = np.zeros((n[2], n[1], n[0])) if sol0 is None else np.array(sol0, dtype=np.float64)
    a, bx, by, bz, bcv = map(_c, (a, bx, by, bz, bcv))
    opts = (ctypes.c_int * 4)(int(bool(use_precond)), precond_niters, restart, maxiter)
    out = GmresOut()
    lib.orc_gmres(ctypes.byref(params), _dp(a), _dp(bx), _dp(by), _dp(bz), _dp(r), _dp(bcv), _dp(sol), opts,
                  ctypes.byref(out))
    return sol, {"iters": out.iters, "status": out.status, "resnorm": out.resnorm, "rnorm0": out.rnorm0,
                 "res_history": list(out.res_hist[:out.nhist])}


# ---- single-level solve of a patch with coarse/fine boundary data
# (MLLinOp::setCoarseFineBC, AMReX_MLLinOp.H:200-212, :1427-1441)

def cf_faces(dom_n, periodic, region):
    """Per face f (dir + 3*side) of the region (lo, hi) of fine cells: "phys"
    where it lies on a non-periodic domain face, "per" where a periodic
    direction is spanned (the periodic image covers the ghosts), else "cf"."""
    lo, hi = region
    out = []
    for f in range(6):
        d, side = f % 3, f // 3
        at = lo[d] == 0 if side == 0 else hi[d] == dom_n[d] - 1
        spans = periodic[d] and lo[d] == 0 and hi[d] == dom_n[d] - 1
        out.append("per" if spans else ("phys" if at and not periodic[d] else "cf"))
    return out


def cf_boundary_values(dom_n, periodic, region, ratio, crse):
    """InterpBndryData::setBndryValues with interpbndrydata_{x,y,z}_o3
    (AMReX_InterpBndryData_3D_K.H:23-118): for every coarse/fine face of the
    region, the ghost values (face cells in tangential order (t2, t1)) from the
    coarse cell data crse (global (z, y, x) on the domain coarsened by ratio).
    The mask of a fine ghost position: beyond a non-periodic face 2 (outside),
    in the region or a periodic image of it 0 (covered), else 1 (not covered)."""
    lo, hi = region
    r = ratio
    nc = [dom_n[d] // r for d in range(3)]

    def cval(c):
        c = list(c)
        for d in range(3):
            if periodic[d]:
                c[d] %= nc[d]
        return crse[c[2], c[1], c[0]]

    def mask(p):
        p = list(p)
        for d in range(3):
            if p[d] < 0 or p[d] >= dom_n[d]:
                if not periodic[d]:
                    return 2
                p[d] %= dom_n[d]
        inside = all(lo[d] <= p[d] <= hi[d] for d in range(3))
        return 0 if inside else 1

    faces = cf_faces(dom_n, periodic, region)
    out = {}
    for f in range(6):
        if faces[f] != "cf":
            continue
        d, side = f % 3, f // 3
        t1, t2 = (1, 2) if d == 0 else ((0, 2) if d == 1 else (0, 1))
        vals = np.zeros((hi[t2] - lo[t2] + 1, hi[t1] - lo[t1] + 1))
        for w in range(lo[t2], hi[t2] + 1):
            for u in range(lo[t1], hi[t1] + 1):
                p = [0, 0, 0]
                p[d] = lo[d] - 1 if side == 0 else hi[d] + 1
                p[t1], p[t2] = u, w
                cc = [p[e] // r for e in range(3)]

                def C(a1, a2):
                    c = list(cc)
                    c[t1] += a1
                    c[t2] += a2
                    return cval(c)

                def M(o1, o2):
                    q = list(p)
                    q[t1] += o1
                    q[t2] += o2
                    return mask(q) == 1

                c0 = C(0, 0)
                dlo, dhi = (-1 if M(-r, 0) else 0), (1 if M(r, 0) else 0)
                fac = 1.0 if dhi == dlo + 1 else 0.5
                d1 = fac * (C(dhi, 0) - C(dlo, 0))
                d11 = 0.5 * (C(1, 0) - 2.0 * c0 + C(-1, 0)) if dhi == dlo + 2 else 0.0
                dlo, dhi = (-1 if M(0, -r) else 0), (1 if M(0, r) else 0)
                fac = 1.0 if dhi == dlo + 1 else 0.5
                d2 = fac * (C(0, dhi) - C(0, dlo))
                d22 = 0.5 * (C(0, 1) - 2.0 * c0 + C(0, -1)) if dhi == dlo + 2 else 0.0
                d12 = (0.25 * (C(1, 1) - C(-1, 1) + C(-1, -1) - C(1, -1))
                       if (M(-r, -r) and M(r, -r) and M(-r, r) and M(r, r)) else 0.0)
                y1 = -0.5 + (p[t1] - cc[t1] * r + 0.5) / r
                y2 = -0.5 + (p[t2] - cc[t2] * r + 0.5) / r
                vals[w - lo[t2], u - lo[t1]] = (c0 + y1 * d1 + y1 * y1 * d11 + y2 * d2 + y2 * y2 * d22
                                                + y1 * y2 * d12)
        out[f] = vals
    return out


def cf_patch_params(dom_n, periodic, region, ratio, bc_lo, bc_hi, cf_type=101, prob_lo=(0.0,) * 3,
                    prob_hi=(1.0,) * 3, **kw):
    """Oracle parameters of the patch solve: the region is the oracle's domain;
    its coarse/fine faces are Dirichlet (or Neumann) faces located
    0.5*ratio level-0 cells outside (bloc), physical faces keep the domain BC,
    periodic only where a periodic direction is spanned."""
    lo, hi = region
    faces = cf_faces(dom_n, periodic, region)
    m = [hi[d] - lo[d] + 1 for d in range(3)]
    dx = [(prob_hi[d] - prob_lo[d]) / dom_n[d] for d in range(3)]
    per = [1 if faces[d] == "per" else 0 for d in range(3)]
    blo, bhi, bloc = [0, 0, 0], [0, 0, 0], [0.0] * 6
    for f in range(6):
        d, side = f % 3, f // 3
        tgt = blo if side == 0 else bhi
        if faces[f] == "phys":
            tgt[d] = (bc_lo if side == 0 else bc_hi)[d]
        elif faces[f] == "cf":
            tgt[d] = cf_type
            bloc[f] = 0.5 * ratio if ratio > 0 else 0.0
    plo = [prob_lo[d] + lo[d] * dx[d] for d in range(3)]
    phi = [plo[d] + m[d] * dx[d] for d in range(3)]
    return make_params(m, periodic=per, bc_lo=blo, bc_hi=bhi, prob_lo=plo, prob_hi=phi, bloc=bloc, **kw)


def cf_patch_bcv(dom_n, periodic, region, ratio, crse, bcdata=None):
    """Grown boundary array of the patch (shape (m2+2, m1+2, m0+2)): physical
    face ghosts from bcdata (grown global array on the domain, or zero), coarse/
    fine face ghosts from cf_boundary_values (zero without coarse data)."""
    lo, hi = region
    m = [hi[d] - lo[d] + 1 for d in range(3)]
    out = np.zeros((m[2] + 2, m[1] + 2, m[0] + 2))
    if bcdata is not None:
        out[:] = bcdata[lo[2]:hi[2] + 3, lo[1]:hi[1] + 3, lo[0]:hi[0] + 3]
    faces = cf_faces(dom_n, periodic, region)
    vals = cf_boundary_values(dom_n, periodic, region, ratio, crse) if crse is not None else {}
    for f in range(6):
        if faces[f] != "cf":
            continue
        d, side = f % 3, f // 3
        idx = [slice(1, m[2] + 1), slice(1, m[1] + 1), slice(1, m[0] + 1)]
        idx[2 - d] = 0 if side == 0 else m[d] + 1
        v = vals.get(f)
        out[tuple(idx)] = 0.0 if v is None else v
    return out
