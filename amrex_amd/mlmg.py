"""Python mirror of the reference's MLMG call surface over the C ABI.

Class and method names follow AMReX (``BoxArray.maxSize``, ``MLABecLaplacian
.setScalars``, ``MLMG.solve`` ...; see Src/LinearSolvers/MLMG/AMReX_MLMG.H and
the Fortran-interface shims in Src/F_Interfaces/LinearSolvers).  MultiFab data
live in GPU memory; ``fromGlobal`` / ``toGlobal`` move data between a MultiFab
and a torch CUDA tensor holding the whole domain (x fastest, shape (z, y, x)).
"""
from __future__ import annotations

import ctypes
from enum import IntEnum

import numpy as np

from . import capi
from .capi import c_vp, ivec, dvec


class LinOpBCType(IntEnum):
    """AMReX_LO_BCTYPES.H:21-35."""
    interior = 0
    Dirichlet = 101
    Neumann = 102
    reflect_odd = 103
    Periodic = 200


class BottomSolver(IntEnum):
    """Codes of amrex_fi_multigrid_set_bottom_solver (AMReX_multigrid_fi.cpp:82-97)."""
    smoother = 0
    bicgstab = 1
    cg = 2
    hypre = 3
    petsc = 4


def _lib():
    return capi.load()


class BoxArray:
    def __init__(self, lo=None, hi=None, boxes=None):
        h = c_vp()
        if boxes is not None:
            flat = []
            for b in boxes:
                flat += [b[0], b[3], b[1], b[4], b[2], b[5]]
            arr = (ctypes.c_int * len(flat))(*flat)
            _lib().amrex_fi_new_boxarray_from_bxfarr(ctypes.byref(h), arr, 2, 3, len(boxes))
        else:
            _lib().amrex_fi_new_boxarray(ctypes.byref(h), ivec(lo), ivec(hi))
        capi.check("BoxArray: ")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.amrex_fi_delete_boxarray(self.h)
            self.h = None

    def maxSize(self, n):
        sz = n if hasattr(n, "__len__") else (n, n, n)
        _lib().amrex_fi_boxarray_maxsize(self.h, ivec(sz))
        capi.check("maxSize: ")
        return self

    def size(self):
        return int(_lib().amrex_fi_boxarray_nboxes(self.h))

    def numPts(self):
        return int(_lib().amrex_fi_boxarray_numpts(self.h))

    def box(self, i):
        lo, hi = ivec((0, 0, 0)), ivec((0, 0, 0))
        _lib().amrex_fi_boxarray_get_box(self.h, i, lo, hi)
        return tuple(lo) + tuple(hi)

    def boxes(self):
        return [self.box(i) for i in range(self.size())]


class Geometry:
    def __init__(self, n, prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 1.0), is_periodic=(0, 0, 0)):
        h = c_vp()
        self.n = tuple(int(v) for v in n)
        _lib().mgx_new_geometry(ctypes.byref(h), ivec((0, 0, 0)), ivec([v - 1 for v in self.n]),
                                dvec(prob_lo), dvec(prob_hi), ivec(is_periodic))
        capi.check("Geometry: ")
        self.h = h
        self.is_periodic = tuple(is_periodic)

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.amrex_fi_delete_geometry(self.h)
            self.h = None


class DistributionMapping:
    def __init__(self, ba: BoxArray):
        h = c_vp()
        _lib().amrex_fi_new_distromap(ctypes.byref(h), ba.h)
        capi.check("DistributionMapping: ")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.amrex_fi_delete_distromap(self.h)
            self.h = None


class MultiFab:
    def __init__(self, ba: BoxArray, dm: DistributionMapping, ncomp=1, ngrow=0, nodal=(0, 0, 0)):
        h = c_vp()
        bah, dmh = c_vp(ba.h.value), c_vp(dm.h.value)
        _lib().amrex_fi_new_multifab(ctypes.byref(h), ctypes.byref(bah), ctypes.byref(dmh), ncomp,
                                     ivec((ngrow,) * 3), ivec(nodal))
        capi.check("MultiFab: ")
        self.h = h
        self.ncomp = ncomp
        self.ngrow = ngrow
        self.nodal = tuple(nodal)
        self._ba = ba  # keep the BoxArray alive
        self._own = None

    @classmethod
    def _from_handle(cls, h):
        """Wrap a MultiFab the library created (it is deleted with this object)."""
        self = cls.__new__(cls)
        self.h = h
        nb, nc, ng = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        nod = (ctypes.c_int * 3)()
        _lib().mgx_multifab_info(h, ctypes.byref(nb), ctypes.byref(nc), ctypes.byref(ng), nod)
        capi.check("MultiFab: ")
        self.ncomp, self.ngrow, self.nodal = nc.value, ng.value, tuple(nod)
        self._ba = None
        own = []
        for i in range(nb.value):
            lo, hi = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
            _lib().mgx_multifab_box(h, i, lo, hi)
            own.append(tuple(lo) + tuple(hi))
        self._own = own
        return self

    @classmethod
    def read(cls, name):
        """A MultiFab defined and filled from the VisMF file `name` (VisMF::Read)."""
        h = c_vp()
        _lib().mgx_read_multifab_new(ctypes.byref(h), str(name).encode())
        capi.check("VisMF::Read: ")
        return cls._from_handle(h)

    def write(self, name, set_ghost=False):
        """VisMF::Write (version-1 header, one data file).  set_ghost: every ghost cell
        is first set to (min + max)/2 of its box's valid cells, in this MultiFab too."""
        if set_ghost:
            _lib().mgx_write_multifab_set_ghost(self.h, str(name).encode(), 1)
        else:
            _lib().amrex_fi_write_multifab(self.h, str(name).encode())
        capi.check("VisMF::Write: ")

    def readInto(self, name):
        """VisMF::Read into this (defined) MultiFab."""
        _lib().amrex_fi_read_multifab(self.h, str(name).encode())
        capi.check("VisMF::Read: ")

    def copyFrom(self, src, srccomp, dstcomp, numcomp, nghost=0):
        """MultiFab::Copy(self, src, ...)."""
        _lib().mgx_multifab_copy(self.h, src.h, srccomp, dstcomp, numcomp, nghost)
        capi.check("MultiFab::Copy: ")
        return self

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.amrex_fi_delete_multifab(self.h)
            self.h = None

    def setVal(self, v, comp=0, ncomp=None):
        nc = self.ncomp - comp if ncomp is None else ncomp
        _lib().amrex_fi_multifab_setval(self.h, float(v), comp, nc, ivec((self.ngrow,) * 3))
        capi.check("setVal: ")

    def sum(self, comp=0):
        r = _lib().amrex_fi_multifab_sum(self.h, comp)
        capi.check("sum: ")
        return r

    def norminf(self, comp=0):
        r = _lib().amrex_fi_multifab_norm0(self.h, comp)
        capi.check("norm0: ")
        return r

    def _global(self, t, dom_n, gg, rg, to_mf):
        import torch
        assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
        n = [dom_n[d] + self.nodal[d] for d in range(3)]
        expect = (n[2] + 2 * gg, n[1] + 2 * gg, n[0] + 2 * gg)
        if tuple(t.shape) != expect:
            raise ValueError(f"global array shape {tuple(t.shape)} != {expect}")
        torch.cuda.synchronize()
        st = _lib().mgx_multifab_copy_global(self.h, c_vp(t.data_ptr()), ivec(n), gg, rg, to_mf)
        capi.check("copy_global: ")
        return st

    def boxes(self):
        """Every box in the MultiFab's own index type, as (lo..., hi...)."""
        if self._own is not None:
            return list(self._own)
        out = []
        for b in self._ba.boxes():
            lo, hi = list(b[:3]), list(b[3:])
            for d in range(3):
                hi[d] += self.nodal[d]
            out.append(tuple(lo) + tuple(hi))
        return out

    def _layout(self):
        """(box, offset, ghost-grown shape (ncomp, z, y, x)) of every box of the packed storage."""
        out, off = [], 0
        for b in self.boxes():
            lo, hi = b[:3], b[3:]
            shp = (self.ncomp,) + tuple(hi[d] - lo[d] + 1 + 2 * self.ngrow for d in (2, 1, 0))
            out.append((b, off, shp))
            off += int(np.prod(shp))
        return out, off

    def getBoxes(self, ghosts=False):
        """Valid cells (or the whole FABs) of every box as numpy arrays (z, y, x), or
        (ncomp, z, y, x) for more than one component."""
        lay, n = self._layout()
        host = np.zeros(n)
        _lib().mgx_multifab_host(self.h, host.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, 0)
        capi.check("getBoxes: ")
        g = 0 if ghosts else self.ngrow
        sl = slice(g, -g) if g else slice(None)
        out = []
        for _, off, shp in lay:
            a = host[off:off + int(np.prod(shp))].reshape(shp)[:, sl, sl, sl].copy()
            out.append(a[0] if self.ncomp == 1 else a)
        return out

    def setBoxes(self, arrays, ghosts=False):
        """Set the valid cells (or the whole FABs) of every box from numpy arrays (z, y, x) or
        (ncomp, z, y, x); other cells become 0."""
        lay, n = self._layout()
        host = np.zeros(n)
        g = 0 if ghosts else self.ngrow
        sl = slice(g, -g) if g else slice(None)
        for (_, off, shp), a in zip(lay, arrays):
            a = np.asarray(a)
            host[off:off + int(np.prod(shp))].reshape(shp)[:, sl, sl, sl] = a if a.ndim == 4 else a[None]
        _lib().mgx_multifab_host(self.h, host.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, 1)
        capi.check("setBoxes: ")
        return self

    def fromGlobal(self, t, dom_n, gg=0, rg=0):
        """Copy a global (z, y, x) CUDA tensor into the MultiFab (boxes grown by rg)."""
        self._global(t, dom_n, gg, rg, 1)
        return self

    def toGlobal(self, t, dom_n, gg=0, rg=0):
        self._global(t, dom_n, gg, rg, 0)
        return t

    def deviceView(self):
        """The MultiFab's device storage as a flat torch float64 tensor (no copy):
        every box's ghost-grown FAB, x fastest, components one after another."""
        import torch
        n = ctypes.c_longlong()
        ptr = _lib().mgx_multifab_data(self.h, ctypes.byref(n))
        return torch.as_tensor(_CudaArray(ptr, n.value), device="cuda")

    def boxViews(self, ghosts=False):
        """Per box, a torch view (z, y, x) -- (ncomp, z, y, x) for several
        components -- of the valid cells (or the whole FAB) in device storage."""
        flat = self.deviceView()
        lay, _ = self._layout()
        g = 0 if ghosts else self.ngrow
        sl = slice(g, -g) if g else slice(None)
        out = []
        for _, off, shp in lay:
            v = flat[off:off + int(np.prod(shp))].view(shp)[:, sl, sl, sl]
            out.append(v[0] if self.ncomp == 1 else v)
        return out


class _CudaArray:
    """A device pointer and a length as a CUDA array interface, for torch.as_tensor."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (int(n),), "typestr": "<f8", "data": (int(ptr or 0), False),
                                         "version": 2}


class _LinOp:
    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.amrex_fi_delete_linop(self.h)
            self.h = None

    def setDomainBC(self, lo, hi):
        _lib().amrex_fi_linop_set_domain_bc(self.h, ivec([int(v) for v in lo]), ivec([int(v) for v in hi]))
        capi.check("setDomainBC: ")

    def setCoarseFineBC(self, crse, ratio, bc_type=None):
        """MLLinOp::setCoarseFineBC: coarse/fine boundary data of level-0 boxes that
        leave part of the domain uncovered (call before setLevelBC)."""
        if bc_type is None:
            _lib().amrex_fi_linop_set_coarse_fine_bc(self.h, crse.h if crse is not None else None, int(ratio))
        else:
            _lib().mgx_linop_set_coarse_fine_bc_type(self.h, crse.h if crse is not None else None, int(ratio),
                                                     int(bc_type))
        capi.check("setCoarseFineBC: ")
        self._cfmf = crse

    def setLevelBC(self, amrlev, mf):
        _lib().amrex_fi_linop_set_level_bc(self.h, amrlev, mf.h if mf is not None else None)
        capi.check("setLevelBC: ")
        self._bcmf = mf

    def setMaxOrder(self, o):
        _lib().amrex_fi_linop_set_maxorder(self.h, o)
        capi.check("setMaxOrder: ")

    def NMGLevels(self):
        return _lib().mgx_linop_num_mg_levels(self.h)

    def levelInfo(self, mglev):
        n = ivec((0, 0, 0))
        nb = ctypes.c_int()
        _lib().mgx_linop_level_info(self.h, mglev, n, ctypes.byref(nb), None, 0)
        boxes = (ctypes.c_int * (6 * nb.value))()
        _lib().mgx_linop_level_info(self.h, mglev, n, ctypes.byref(nb), boxes, nb.value)
        capi.check("levelInfo: ")
        return tuple(n), [tuple(boxes[6 * i: 6 * i + 6]) for i in range(nb.value)]

    def levelOp(self, mglev, which, precision=0, in1=None, io=None, out=None, count=0):
        """Single kernel operation on MG level mglev; torch CUDA float64 global arrays."""
        import torch
        ptr = lambda t: c_vp(t.data_ptr()) if t is not None else None
        torch.cuda.synchronize()
        _lib().mgx_linop_level_op(self.h, mglev, which, precision, ptr(in1), ptr(io), ptr(out), count)
        capi.check("levelOp: ")


def time_smoother(linop, mglev=0, precision="double", nsweeps=10):
    """(ms per sweep, ms per GSRB colour launch, ms per ghost-fill launch) on level mglev."""
    out = (ctypes.c_double * 3)()
    _lib().mgx_linop_time_smoother(linop.h, mglev, 0 if precision == "double" else 1, nsweeps, out)
    capi.check("time_smoother: ")
    return tuple(out)


def time_sweep_pairs(linop, mglev=0, precision="double", npairs=10):
    """ms per launch of the two-sweep GSRB march on level mglev (two sweeps per launch)."""
    out = ctypes.c_double()
    _lib().mgx_linop_time_sweep_pairs(linop.h, mglev, 0 if precision == "double" else 1, npairs, ctypes.byref(out))
    capi.check("time_sweep_pairs: ")
    return out.value


def set_pair_sweeps(on):
    """Whether V-cycles run two consecutive sweeps as one launch where a level allows
    (the default); returns the previous setting.  The results are the same either way."""
    return bool(_lib().mgx_set_pair_sweeps(int(bool(on))))


class MLABecLaplacian(_LinOp):
    def __init__(self, geom: Geometry, ba: BoxArray, dm: DistributionMapping = None, agg_grid_size=-1,
                 agglomeration=1, max_coarsening_level=30, dense=True):
        h = c_vp()
        _lib().mgx_set_dense_storage(1 if dense else 0)
        _lib().mgx_new_abeclaplacian(ctypes.byref(h), geom.h, ba.h, agg_grid_size, agglomeration,
                                     max_coarsening_level)
        capi.check("MLABecLaplacian: ")
        self.h = h

    def setScalars(self, a, b):
        _lib().amrex_fi_abeclap_set_scalars(self.h, a, b)
        capi.check("setScalars: ")

    def setACoeffs(self, amrlev, mf):
        _lib().amrex_fi_abeclap_set_acoeffs(self.h, amrlev, mf.h)
        capi.check("setACoeffs: ")

    def setBCoeffs(self, amrlev, mfs):
        arr = (c_vp * 3)(*[m.h for m in mfs])
        _lib().amrex_fi_abeclap_set_bcoeffs(self.h, amrlev, arr)
        capi.check("setBCoeffs: ")

    def compCoeffGradient(self, lam: MultiFab, phi: MultiFab, grad_a=None, grad_b=None, accumulate=False):
        """dJ/da into grad_a and dJ/db_d into grad_b[d] given the adjoint lam (extension)."""
        gb = None
        if grad_b is not None:
            gb = (c_vp * 3)(*[m.h if m is not None else None for m in grad_b])
        _lib().mgx_abeclap_coeff_gradient(self.h, lam.h, phi.h, grad_a.h if grad_a is not None else None, gb,
                                          1 if accumulate else 0)
        capi.check("compCoeffGradient: ")


class MLPoisson(_LinOp):
    def __init__(self, geom: Geometry, ba: BoxArray, dm: DistributionMapping = None, agg_grid_size=-1,
                 agglomeration=1, max_coarsening_level=30, dense=True):
        h = c_vp()
        _lib().mgx_set_dense_storage(1 if dense else 0)
        _lib().mgx_new_poisson(ctypes.byref(h), geom.h, ba.h, agg_grid_size, agglomeration, max_coarsening_level)
        capi.check("MLPoisson: ")
        self.h = h


class MLMG:
    def __init__(self, linop):
        h = c_vp()
        _lib().amrex_fi_new_multigrid(ctypes.byref(h), linop.h)
        capi.check("MLMG: ")
        self.h = h
        self._linop = linop
        self.setVerbose(0)

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.amrex_fi_delete_multigrid(self.h)
            self.h = None

    def solve(self, sol: MultiFab, rhs: MultiFab, tol_rel, tol_abs=0.0):
        s = (c_vp * 1)(sol.h)
        r = (c_vp * 1)(rhs.h)
        res = _lib().amrex_fi_multigrid_solve(self.h, s, r, tol_rel, tol_abs)
        capi.check("MLMG.solve: ")
        return res

    def solveAdjoint(self, sol: MultiFab, rhs: MultiFab, tol_rel, tol_abs=0.0):
        """Solve L^T sol = rhs (homogeneous boundary data); extension."""
        res = _lib().mgx_multigrid_solve_adjoint(self.h, (c_vp * 1)(sol.h), (c_vp * 1)(rhs.h), tol_rel, tol_abs)
        capi.check("MLMG.solveAdjoint: ")
        return res

    @staticmethod
    def _faces(faces):
        # [mfx, mfy, mfz] for one level, or one such triple per AMR level
        levels = [faces] if isinstance(faces[0], MultiFab) else list(faces)
        flat = [mf.h for lev in levels for mf in lev]
        if any(len(lev) != 3 for lev in levels):
            raise ValueError("three face MultiFabs (x, y, z) per AMR level expected")
        return (c_vp * len(flat))(*flat)

    def getGradSolution(self, grad):
        """Face gradients of the last solution into grad ([x, y, z] face MultiFabs per level)."""
        _lib().amrex_fi_multigrid_get_grad_solution(self.h, self._faces(grad))
        capi.check("getGradSolution: ")

    def getFluxes(self, flux, sol=None):
        """Face fluxes of the last solution, or of sol (one MultiFab or one per level).  flux:
        [x, y, z] face MultiFabs (per level); or cell MultiFabs of >= 3 components (one, or one
        per level), which receive the fluxes averaged to cell centres."""
        cells = [flux] if isinstance(flux, MultiFab) else list(flux)
        if all(isinstance(m, MultiFab) and m.nodal == (0, 0, 0) for m in cells):
            sols = None if sol is None else ([sol] if isinstance(sol, MultiFab) else list(sol))
            _lib().mgx_multigrid_get_cell_fluxes(self.h, (c_vp * len(cells))(*[m.h for m in cells]),
                                                 None if sols is None else (c_vp * len(sols))(*[m.h for m in sols]))
            capi.check("getFluxes: ")
            return
        if sol is None:
            _lib().amrex_fi_multigrid_get_fluxes(self.h, self._faces(flux))
            capi.check("getFluxes: ")
        else:
            sols = [sol] if isinstance(sol, MultiFab) else list(sol)
            _lib().mgx_multigrid_get_fluxes_sol(self.h, self._faces(flux), (c_vp * len(sols))(*[m.h for m in sols]))
            capi.check("getFluxes: ")

    def checkPoint(self, sol: MultiFab, rhs: MultiFab, tol_rel, tol_abs, name):
        """MLMG::checkPoint; reports (after writing) that the operator's part is not implemented."""
        _lib().mgx_multigrid_checkpoint(self.h, (c_vp * 1)(sol.h), (c_vp * 1)(rhs.h), tol_rel, tol_abs,
                                        str(name).encode())
        capi.check("checkPoint: ")

    def compResidual(self, res: MultiFab, sol: MultiFab, rhs: MultiFab):
        _lib().amrex_fi_multigrid_comp_residual(self.h, (c_vp * 1)(res.h), (c_vp * 1)(sol.h), (c_vp * 1)(rhs.h))
        capi.check("compResidual: ")

    def setVerbose(self, v):
        _lib().amrex_fi_multigrid_set_verbose(self.h, v)

    def setMaxIter(self, n):
        _lib().amrex_fi_multigrid_set_max_iter(self.h, n)

    def setFixedIter(self, n):
        _lib().amrex_fi_multigrid_set_fixed_iter(self.h, n)

    def setMaxFmgIter(self, n):
        """Outer iterations that use the F-cycle instead of the V-cycle."""
        _lib().amrex_fi_multigrid_set_max_fmg_iter(self.h, n)

    def setBottomSolver(self, s):
        _lib().amrex_fi_multigrid_set_bottom_solver(self.h, int(s))
        capi.check("setBottomSolver: ")

    def setAlwaysUseBNorm(self, f):
        _lib().amrex_fi_multigrid_set_always_use_bnorm(self.h, int(f))

    def setPrecision(self, p):
        """'double' (reference algorithm in FP64) or 'mixed' (FP32 V-cycle, FP64 refinement)."""
        _lib().mgx_multigrid_set_precision(self.h, 1 if p == "mixed" else 0)

    def setRefinementVcycles(self, n):
        """Mixed precision: FP32 V-cycles per FP64 refinement step (1, 2, or 0: chosen per step)."""
        _lib().mgx_multigrid_set_refinement_vcycles(self.h, int(n))
        capi.check("setRefinementVcycles: ")

    def setFinalFillBC(self, flag):
        """After solve, also fill the face ghost cells of the solution (MLMG::setFinalFillBC)."""
        _lib().amrex_fi_multigrid_set_final_fill_bc(self.h, int(flag))

    def setUseGraph(self, flag):
        _lib().mgx_multigrid_set_use_graph(self.h, 1 if flag else 0)

    BOTTOM_CODES = {"default": 0, "smoother": 1, "bicgstab": 2, "cg": 3, "bicgcg": 4, "cgbicg": 5}

    def setBottomSolverCode(self, code):
        """MLMG::setBottomSolver with an amrex::BottomSolver code or name (0 default, 1 smoother,
        2 bicgstab, 3 cg, 4 bicgcg, 5 cgbicg)."""
        c = self.BOTTOM_CODES[code] if isinstance(code, str) else int(code)
        _lib().mgx_multigrid_set_bottom_solver_code(self.h, c)
        capi.check("setBottomSolver: ")

    def getBottomSolverCode(self):
        """The bottom solver in force (after a solve: Default resolved, bicgcg/cgbicg switches taken over)."""
        return _lib().mgx_multigrid_get_bottom_solver_code(self.h)

    def setBottomTolerance(self, rel, abs_=-1.0):
        _lib().mgx_multigrid_set_bottom_tolerance(self.h, rel, abs_)

    def setBottomMaxIter(self, n):
        _lib().mgx_multigrid_set_bottom_max_iter(self.h, n)

    def stats(self):
        st = capi.SolveStats()
        _lib().mgx_multigrid_get_stats(self.h, ctypes.byref(st))
        return {
            "niters": st.niters, "nlevels": st.nlevels, "bottom_status": st.last_bottom_status,
            "bottom_iters": st.last_bottom_iters, "init_resnorm": st.init_resnorm,
            "rhsnorm": st.rhsnorm, "final_resnorm": st.final_resnorm,
            "solve_seconds": st.solve_seconds, "iter_seconds": st.iter_seconds,
            "res_history": list(st.res_history[: st.niters]), "nvcycles": st.nvcycles,
        }


def WriteMultiLevelPlotfile(name, mfs, varnames, geoms, time=0.0, level_steps=None, ref_ratio=None):
    """amrex_fi_write_plotfile: one MultiFab per level holding len(varnames) components."""
    n = len(mfs)
    steps = (ctypes.c_int * n)(*(level_steps or [0] * n))
    rr = (ctypes.c_int * max(n - 1, 1))(*((ref_ratio or [2] * (n - 1)) or [0]))
    vn = (ctypes.c_char_p * len(varnames))(*[v.encode() for v in varnames])
    _lib().amrex_fi_write_plotfile(str(name).encode(), n, (c_vp * n)(*[m.h for m in mfs]), vn,
                                   (c_vp * n)(*[g.h for g in geoms]), float(time), steps, rr)
    capi.check("write_plotfile: ")


def WriteSingleLevelPlotfile(name, mf, varnames, geom, time=0.0, level_step=0):
    WriteMultiLevelPlotfile(name, [mf], varnames, [geom], time, [level_step])


class PlotFileData:
    """Reader of a plotfile (amrex::PlotFileData)."""

    def __init__(self, name):
        h = c_vp()
        _lib().mgx_new_plotfile_data(ctypes.byref(h), str(name).encode())
        capi.check("PlotFileData: ")
        self.h = h
        nc, fl, t, cs = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        _lib().mgx_plotfile_data_info(h, ctypes.byref(nc), ctypes.byref(fl), ctypes.byref(t), ctypes.byref(cs))
        self.ncomp, self.finest_level, self.time, self.coord = nc.value, fl.value, t.value, cs.value
        self.varnames = []
        for i in range(self.ncomp):
            buf = ctypes.create_string_buffer(512)
            _lib().mgx_plotfile_data_varname(h, i, buf, 512)
            self.varnames.append(buf.value.decode())

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.mgx_delete_plotfile_data(self.h)
            self.h = None

    def level(self, lev):
        """{"domain": (lo, hi), "prob_lo", "prob_hi", "dx", "ref_ratio", "level_step"} of a level."""
        lo, hi = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        plo, phi, dx = (ctypes.c_double * 3)(), (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        rr, ls = ctypes.c_int(), ctypes.c_int()
        _lib().mgx_plotfile_data_level(self.h, lev, lo, hi, plo, phi, dx, ctypes.byref(rr), ctypes.byref(ls))
        capi.check("PlotFileData.level: ")
        return {"domain": (tuple(lo), tuple(hi)), "prob_lo": tuple(plo), "prob_hi": tuple(phi), "dx": tuple(dx),
                "ref_ratio": rr.value, "level_step": ls.value}

    def get(self, lev, varname=None):
        """A MultiFab with every component of a level, or the one named."""
        h = c_vp()
        _lib().mgx_plotfile_data_get(self.h, lev, None if varname is None else varname.encode(), ctypes.byref(h))
        capi.check("PlotFileData.get: ")
        return MultiFab._from_handle(h)


class GMRESMLMG:
    """GMRES with the MLMG V-cycle as preconditioner (amrex::GMRESMLMG), one AMR level."""

    def __init__(self, mlmg: MLMG):
        h = c_vp()
        _lib().mgx_new_gmres_mlmg(ctypes.byref(h), mlmg.h)
        capi.check("GMRESMLMG: ")
        self.h = h
        self._mlmg = mlmg

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.mgx_delete_gmres_mlmg(self.h)
            self.h = None

    def solve(self, sol: MultiFab, rhs: MultiFab, tol_rel, tol_abs=0.0):
        r = _lib().mgx_gmres_mlmg_solve(self.h, sol.h, rhs.h, tol_rel, tol_abs)
        capi.check("GMRESMLMG.solve: ")
        return r

    def setVerbose(self, v):
        _lib().mgx_gmres_mlmg_set_verbose(self.h, int(v))

    def setMaxIters(self, n):
        _lib().mgx_gmres_mlmg_set_max_iters(self.h, int(n))

    def setRestartLength(self, n):
        _lib().mgx_gmres_mlmg_set_restart_length(self.h, int(n))

    def setPrecondNumIters(self, n):
        _lib().mgx_gmres_mlmg_set_precond_num_iters(self.h, int(n))

    def usePrecond(self, flag):
        return bool(_lib().mgx_gmres_mlmg_use_precond(self.h, int(bool(flag))))

    def setPropertyOfZero(self, flag):
        _lib().mgx_gmres_mlmg_set_property_of_zero(self.h, int(bool(flag)))

    def stats(self):
        it, st, rn = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        hist = (ctypes.c_double * 4096)()
        nh = _lib().mgx_gmres_mlmg_stats(self.h, ctypes.byref(it), ctypes.byref(st), ctypes.byref(rn), hist, 4096)
        return {"niters": it.value, "status": st.value, "resnorm": rn.value,
                "res_history": list(hist[:min(nh, 4096)])}


def synchronize():
    _lib().mgx_synchronize()
    capi.check("synchronize: ")


class Problem:
    """One single-level solve set up from global torch CUDA tensors.

    Holds the Geometry, BoxArray, operator, solver and MultiFabs so tests and
    the benchmark can drive the C ABI the way an application would.
    """

    def __init__(self, n, max_grid_size, op="abec", periodic=(0, 0, 0), bc_lo=(101,) * 3,
                 bc_hi=(101,) * 3, maxorder=3, alpha=1.0, beta=1.0, a=None, bx=None, by=None, bz=None,
                 bcv=None, agg_grid_size=-1, precision="double", use_graph=True, dense=True):
        import torch
        self.n = tuple(int(v) for v in n)
        self.geom = Geometry(self.n, is_periodic=periodic)
        self.ba = BoxArray((0, 0, 0), [v - 1 for v in self.n]).maxSize(max_grid_size)
        self.dm = DistributionMapping(self.ba)
        if op == "abec":
            self.linop = MLABecLaplacian(self.geom, self.ba, self.dm, agg_grid_size=agg_grid_size, dense=dense)
        else:
            self.linop = MLPoisson(self.geom, self.ba, self.dm, agg_grid_size=agg_grid_size, dense=dense)
        lo = [LinOpBCType.Periodic if periodic[d] else bc_lo[d] for d in range(3)]
        hi = [LinOpBCType.Periodic if periodic[d] else bc_hi[d] for d in range(3)]
        self.linop.setDomainBC(lo, hi)
        self.linop.setMaxOrder(maxorder)
        dev = torch.device("cuda")
        if bcv is not None:
            self.bcmf = MultiFab(self.ba, self.dm, 1, 1)
            self.bcmf.fromGlobal(torch.as_tensor(bcv, device=dev).contiguous(), self.n, gg=1, rg=1)
            self.linop.setLevelBC(0, self.bcmf)
        else:
            self.linop.setLevelBC(0, None)
        self.op = op
        if op == "abec":
            self.linop.setScalars(alpha, beta)
            self.amf = MultiFab(self.ba, self.dm, 1, 0)
            self.bmf = []
            for d in range(3):
                nodal = [0, 0, 0]
                nodal[d] = 1
                self.bmf.append(MultiFab(self.ba, self.dm, 1, 0, nodal=nodal))
            self.set_coefficients(a, bx, by, bz)
        self.mlmg = MLMG(self.linop)
        self.mlmg.setPrecision(precision)
        self.mlmg.setUseGraph(use_graph)
        self.sol = MultiFab(self.ba, self.dm, 1, 1)
        self.rhs = MultiFab(self.ba, self.dm, 1, 0)

    def set_coefficients(self, a=None, bx=None, by=None, bz=None):
        """(Re)set the ABecLaplacian coefficients from global arrays (None: 1)."""
        import torch
        dev = torch.device("cuda")
        if a is not None:
            self.amf.fromGlobal(torch.as_tensor(a, device=dev, dtype=torch.float64).contiguous(), self.n)
        else:
            self.amf.setVal(1.0)
        self.linop.setACoeffs(0, self.amf)
        for m, bd in zip(self.bmf, (bx, by, bz)):
            if bd is not None:
                m.fromGlobal(torch.as_tensor(bd, device=dev, dtype=torch.float64).contiguous(), self.n)
            else:
                m.setVal(1.0)
        self.linop.setBCoeffs(0, self.bmf)
        self.coef_version = getattr(self, "coef_version", 0) + 1

    def _global_cells(self):
        import torch
        return torch.zeros((self.n[2], self.n[1], self.n[0]), dtype=torch.float64, device="cuda")

    def solve_adjoint(self, g, tol_rel=1e-10, tol_abs=0.0):
        """Solve L^T lam = g (global tensor); returns lam as a global tensor."""
        import torch
        self.rhs.fromGlobal(torch.as_tensor(g, device="cuda", dtype=torch.float64).contiguous(), self.n)
        self.sol.setVal(0.0)
        self.mlmg.solveAdjoint(self.sol, self.rhs, tol_rel, tol_abs)
        return self.sol.toGlobal(self._global_cells(), self.n)

    def coeff_gradient(self, lam, phi, grad_a=None, grad_b=None, accumulate=False):
        """dJ/da, dJ/db (global tensors, updated in place) from lam and phi (global tensors)."""
        import torch
        if not hasattr(self, "_glam"):
            self._glam = MultiFab(self.ba, self.dm, 1, 0)
            self._gphi = MultiFab(self.ba, self.dm, 1, 0)
            self._ga = MultiFab(self.ba, self.dm, 1, 0)
            self._gb = []
            for d in range(3):
                nodal = [0, 0, 0]
                nodal[d] = 1
                self._gb.append(MultiFab(self.ba, self.dm, 1, 0, nodal=nodal))
        self._glam.fromGlobal(torch.as_tensor(lam, device="cuda", dtype=torch.float64).contiguous(), self.n)
        self._gphi.fromGlobal(torch.as_tensor(phi, device="cuda", dtype=torch.float64).contiguous(), self.n)
        if accumulate:
            if grad_a is not None:
                self._ga.fromGlobal(grad_a, self.n)
            if grad_b is not None:
                for m, t in zip(self._gb, grad_b):
                    if t is not None:
                        m.fromGlobal(t, self.n)
        self.linop.compCoeffGradient(self._glam, self._gphi, self._ga if grad_a is not None else None,
                                     None if grad_b is None else
                                     [m if t is not None else None for m, t in zip(self._gb, grad_b)],
                                     accumulate)
        if grad_a is not None:
            self._ga.toGlobal(grad_a, self.n)
        if grad_b is not None:
            for m, t in zip(self._gb, grad_b):
                if t is not None:
                    m.toGlobal(t, self.n)

    def residual(self, sol, rhs):
        """rhs - L(sol) for global tensors (level BC of the operator)."""
        import torch
        self.sol.fromGlobal(torch.as_tensor(sol, device="cuda", dtype=torch.float64).contiguous(), self.n)
        self.rhs.fromGlobal(torch.as_tensor(rhs, device="cuda", dtype=torch.float64).contiguous(), self.n)
        if not hasattr(self, "_res"):
            self._res = MultiFab(self.ba, self.dm, 1, 0)
        self.mlmg.compResidual(self._res, self.sol, self.rhs)
        return self._res.toGlobal(self._global_cells(), self.n)

    def face_values(self, kind="flux", sol=None):
        """Face gradients (kind "grad") or fluxes ("flux") of the last solution, or
        of sol (a global tensor; fluxes only), as global face tensors [x, y, z] of
        shapes (n2, n1, n0+1), (n2, n1+1, n0), (n2+1, n1, n0)."""
        import torch
        if not hasattr(self, "_fmf"):
            self._fmf = []
            for d in range(3):
                nodal = [0, 0, 0]
                nodal[d] = 1
                self._fmf.append(MultiFab(self.ba, self.dm, 1, 0, nodal=nodal))
        if kind == "grad":
            if sol is not None:
                raise ValueError("getGradSolution takes the last solution only")
            self.mlmg.getGradSolution(self._fmf)
        elif sol is None:
            self.mlmg.getFluxes(self._fmf)
        else:
            if not hasattr(self, "_fsol"):
                self._fsol = MultiFab(self.ba, self.dm, 1, 0)
            self._fsol.fromGlobal(torch.as_tensor(sol, device="cuda", dtype=torch.float64).contiguous(), self.n)
            self.mlmg.getFluxes(self._fmf, self._fsol)
        out = []
        for d in range(3):
            shp = [self.n[2], self.n[1], self.n[0]]
            shp[2 - d] += 1
            t = torch.zeros(shp, dtype=torch.float64, device="cuda")
            self._fmf[d].toGlobal(t, self.n)
            out.append(t)
        return out

    def solve_gmres(self, rhs, tol_rel=1e-10, tol_abs=0.0, sol0=None, use_precond=True, precond_niters=1,
                    restart=30, max_iters=2000):
        """GMRESMLMG solve with a global rhs tensor; returns (global solution tensor, stats)."""
        import torch
        if not hasattr(self, "gmres"):
            self.gmres = GMRESMLMG(self.mlmg)
        g = self.gmres
        g.usePrecond(use_precond)
        g.setPrecondNumIters(precond_niters)
        g.setRestartLength(restart)
        g.setMaxIters(max_iters)
        self.rhs.fromGlobal(torch.as_tensor(rhs, device="cuda", dtype=torch.float64).contiguous(), self.n)
        if sol0 is None:
            self.sol.setVal(0.0)
        else:
            self.sol.fromGlobal(torch.as_tensor(sol0, device="cuda", dtype=torch.float64).contiguous(), self.n)
        g.solve(self.sol, self.rhs, tol_rel, tol_abs)
        return self.sol.toGlobal(self._global_cells(), self.n), g.stats()

    def solve(self, rhs, tol_rel=1e-10, tol_abs=0.0, sol0=None):
        """Solve with a global rhs tensor; returns the global solution tensor."""
        import torch
        rhs = torch.as_tensor(rhs, device="cuda").contiguous()
        self.rhs.fromGlobal(rhs, self.n)
        if sol0 is None:
            self.sol.setVal(0.0)
        else:
            self.sol.fromGlobal(torch.as_tensor(sol0, device="cuda").contiguous(), self.n)
        self.mlmg.solve(self.sol, self.rhs, tol_rel, tol_abs)
        out = torch.zeros((self.n[2], self.n[1], self.n[0]), dtype=torch.float64, device="cuda")
        self.sol.toGlobal(out, self.n)
        return out


class AmrProblem:
    """Two-level composite problem, MLPoisson (op "poisson") or MLABecLaplacian
    (op "abec" with scalars alpha, beta and coefficients coef): a coarse level
    (n^3 cells, boxes of max_grid_size; periodic, or with physical boundaries
    bc_lo/bc_hi whose Dirichlet values, if any, come from bcv: a ghost-grown
    (n+2)^3 coarse array, the values in the face ghosts; None is homogeneous)
    and fine boxes of refinement ratio 2 given in the fine index space.
    coef = (coarse {"a", "bx", "by", "bz"}: global (z, y, x) arrays, the faces
    nodal along their direction; fine {"a": [per box], "bx"/"by"/"bz": [per box
    faces]}), a missing entry 1.  Fine boxes may touch the domain faces; the fine
    level's Dirichlet values there come from fine_bcv, a ghost-grown (2n+2)^3
    array over the fine domain with the values in its face ghosts (None:
    homogeneous).  Solutions come back as (coarse global array, [fine box
    arrays])."""

    def __init__(self, ncoarse, max_grid_size, boxes, precision="double", use_graph=True, maxorder=3,
                 bc_lo=(LinOpBCType.Periodic,) * 3, bc_hi=(LinOpBCType.Periodic,) * 3, bcv=None, op="poisson",
                 alpha=1.0, beta=1.0, coef=None, fine_bcv=None):
        import torch
        n = int(ncoarse)
        self.n = (n, n, n)
        self.boxes = [tuple(int(v) for v in b) for b in boxes]
        per = tuple(1 if bc_lo[d] == LinOpBCType.Periodic else 0 for d in range(3))
        self.geom0 = Geometry(self.n, is_periodic=per)
        self.geom1 = Geometry((2 * n,) * 3, is_periodic=per)
        self.ba0 = BoxArray((0, 0, 0), (n - 1,) * 3).maxSize(max_grid_size)
        self.ba1 = BoxArray(boxes=self.boxes)
        self.dm0 = DistributionMapping(self.ba0)
        self.dm1 = DistributionMapping(self.ba1)
        h = c_vp()
        _lib().mgx_set_dense_storage(1)
        g = (c_vp * 2)(self.geom0.h, self.geom1.h)
        b = (c_vp * 2)(self.ba0.h, self.ba1.h)
        d = (c_vp * 2)(self.dm0.h, self.dm1.h)
        if op == "abec":
            _lib().amrex_fi_new_abeclaplacian(ctypes.byref(h), 2, g, b, d, -1, -1, -1, 30)
            capi.check("MLABecLaplacian (2 levels): ")
        else:
            _lib().amrex_fi_new_poisson(ctypes.byref(h), 2, g, b, d, -1, -1, -1, 30)
            capi.check("MLPoisson (2 levels): ")
        self.op = op
        self.linop = _LinOp()
        self.linop.h = h
        if op == "abec":
            _lib().amrex_fi_abeclap_set_scalars(h, alpha, beta)
            capi.check("setScalars: ")
            self.set_coefficients(coef)
        self.linop.setDomainBC(list(bc_lo), list(bc_hi))
        self.linop.setMaxOrder(maxorder)
        if bcv is not None:
            self.bcmf = MultiFab(self.ba0, self.dm0, 1, 1)
            self.bcmf.fromGlobal(torch.as_tensor(np.ascontiguousarray(bcv), device="cuda"), self.n, gg=1, rg=1)
            self.linop.setLevelBC(0, self.bcmf)
        else:
            self.linop.setLevelBC(0, None)
        if fine_bcv is not None:
            g = np.asarray(fine_bcv, dtype=np.float64)
            self.fbcmf = MultiFab(self.ba1, self.dm1, 1, 1)
            self.fbcmf.setBoxes([g[b[2]:b[5] + 3, b[1]:b[4] + 3, b[0]:b[3] + 3] for b in self.boxes], ghosts=True)
            self.linop.setLevelBC(1, self.fbcmf)
        else:
            self.linop.setLevelBC(1, None)
        self.mlmg = MLMG(self.linop)
        self.mlmg.setPrecision(precision)
        self.mlmg.setUseGraph(use_graph)
        self.sol0 = MultiFab(self.ba0, self.dm0, 1, 1)
        self.rhs0 = MultiFab(self.ba0, self.dm0, 1, 0)
        self.sol1 = MultiFab(self.ba1, self.dm1, 1, 1)
        self.rhs1 = MultiFab(self.ba1, self.dm1, 1, 0)

    def set_coefficients(self, coef=None):
        """(Re)set the ABecLaplacian coefficients of both levels (see the class)."""
        import torch
        c0, c1 = coef if coef is not None else ({}, {})
        self._amf = [MultiFab(self.ba0, self.dm0, 1, 0), MultiFab(self.ba1, self.dm1, 1, 0)]
        self._bmf = [[], []]
        for d in range(3):
            nodal = [0, 0, 0]
            nodal[d] = 1
            self._bmf[0].append(MultiFab(self.ba0, self.dm0, 1, 0, nodal=nodal))
            self._bmf[1].append(MultiFab(self.ba1, self.dm1, 1, 0, nodal=nodal))
        for lev, src in ((0, c0), (1, c1)):
            for m, key in zip([self._amf[lev]] + self._bmf[lev], ("a", "bx", "by", "bz")):
                v = src.get(key)
                if v is None:
                    m.setVal(1.0)
                elif lev == 0:
                    m.fromGlobal(torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device="cuda"), self.n)
                else:
                    m.setBoxes([np.asarray(x, dtype=np.float64) for x in v])
            _lib().amrex_fi_abeclap_set_acoeffs(self.linop.h, lev, self._amf[lev].h)
            capi.check("setACoeffs: ")
            arr = (c_vp * 3)(*[m.h for m in self._bmf[lev]])
            _lib().amrex_fi_abeclap_set_bcoeffs(self.linop.h, lev, arr)
            capi.check("setBCoeffs: ")

    def solve(self, rhs0, rhs1, tol_rel=1e-10, tol_abs=0.0):
        import torch
        self.rhs0.fromGlobal(torch.as_tensor(np.ascontiguousarray(rhs0), device="cuda"), self.n)
        self.rhs1.setBoxes(rhs1)
        self.sol0.setVal(0.0)
        self.sol1.setVal(0.0)
        s = (c_vp * 2)(self.sol0.h, self.sol1.h)
        r = (c_vp * 2)(self.rhs0.h, self.rhs1.h)
        res = _lib().amrex_fi_multigrid_solve(self.mlmg.h, s, r, tol_rel, tol_abs)
        capi.check("MLMG.solve (2 levels): ")
        out = torch.zeros((self.n[2], self.n[1], self.n[0]), dtype=torch.float64, device="cuda")
        self.sol0.toGlobal(out, self.n)
        return out.cpu().numpy(), self.sol1.getBoxes(), res

    def solve_device(self, rhs0, rhs1, tol_rel=1e-10, tol_abs=0.0):
        """solve() with inputs and outputs in device memory: rhs0 a CUDA tensor
        (z, y, x) of the coarse domain, rhs1 one CUDA tensor (z, y, x) per fine
        box.  Returns (coarse solution tensor, [fine box solution views]); the
        views alias this problem's solution MultiFab until the next solve."""
        import torch
        self.rhs0.fromGlobal(rhs0.contiguous(), self.n)
        for v, r in zip(self.rhs1.boxViews(), rhs1):
            v.copy_(r)
        torch.cuda.synchronize()
        self.sol0.setVal(0.0)
        self.sol1.setVal(0.0)
        s = (c_vp * 2)(self.sol0.h, self.sol1.h)
        r = (c_vp * 2)(self.rhs0.h, self.rhs1.h)
        res = _lib().amrex_fi_multigrid_solve(self.mlmg.h, s, r, tol_rel, tol_abs)
        capi.check("MLMG.solve (2 levels): ")
        out = torch.empty((self.n[2], self.n[1], self.n[0]), dtype=torch.float64, device="cuda")
        self.sol0.toGlobal(out, self.n)
        return out, self.sol1.boxViews(), res

    def face_values(self, kind="flux"):
        """Face gradients (kind "grad") or fluxes ("flux") of the last solve on both
        levels: ([x, y, z] coarse global face arrays, [x, y, z] lists of per-box
        fine face arrays (z, y, x))."""
        import torch
        if not hasattr(self, "_f0"):
            self._f0, self._f1 = [], []
            for d in range(3):
                nodal = [0, 0, 0]
                nodal[d] = 1
                self._f0.append(MultiFab(self.ba0, self.dm0, 1, 0, nodal=nodal))
                self._f1.append(MultiFab(self.ba1, self.dm1, 1, 0, nodal=nodal))
        faces = [self._f0, self._f1]
        if kind == "grad":
            self.mlmg.getGradSolution(faces)
        else:
            self.mlmg.getFluxes(faces)
        crse = []
        for d in range(3):
            shp = [self.n[2], self.n[1], self.n[0]]
            shp[2 - d] += 1
            t = torch.zeros(shp, dtype=torch.float64, device="cuda")
            self._f0[d].toGlobal(t, self.n)
            crse.append(t.cpu().numpy())
        return crse, [self._f1[d].getBoxes() for d in range(3)]
