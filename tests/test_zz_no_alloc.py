"""mgx_multifab_sum_norm0_noalloc: the MultiFab sum and max-norm that allocate
nothing (partials in a fixed device array per GPU, results into the caller's
device buffer).

Bar: bit-identical to amrex_fi_multifab_sum / amrex_fi_multifab_norm0 (the
same kernels over the same grid, so the same summation order), on two calls in
a row, the second over a smaller MultiFab, so a stale partial left by the
first call would show; sizes beyond the fixed maximum are refused.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from amrex_amd import capi, mlmg  # noqa: E402


def _filled_mf(n, mgs, ncomp, ngrow, seed):
    ba = mlmg.BoxArray((0, 0, 0), [v - 1 for v in n]).maxSize(mgs)
    dm = mlmg.DistributionMapping(ba)
    mf = mlmg.MultiFab(ba, dm, ncomp, ngrow)
    rng = np.random.default_rng(seed)
    lay, _ = mf._layout()
    fabs, valid = [], []
    g = ngrow
    for _, _, shp in lay:
        # ghost cells far larger than any valid value: a reduction that reads them fails
        f = np.full(shp, 1.0e6)
        v = rng.standard_normal((shp[0],) + tuple(s - 2 * g for s in shp[1:]))
        sl = slice(g, -g) if g else slice(None)
        f[:, sl, sl, sl] = v
        fabs.append(f)
        valid.append(v)
    mf.setBoxes(fabs, ghosts=True)
    return mf, valid, (ba, dm)


def _noalloc(mf, comp, ncomp, out):
    lib = capi.load()
    torch.cuda.synchronize()
    st = lib.mgx_multifab_sum_norm0_noalloc(mf.h, comp, ncomp, ctypes.c_void_p(out.data_ptr()))
    if st != 0:
        capi.check("noalloc: ")
    assert st == 0
    assert lib.mgx_synchronize() == 0
    return out.cpu().numpy()


def _check(mf, valid, comp, ncomp, out):
    got = _noalloc(mf, comp, ncomp, out)
    for c in range(ncomp):
        s_ref, m_ref = mf.sum(comp + c), mf.norminf(comp + c)
        assert got[2 * c] == s_ref, (c, got[2 * c], s_ref)
        assert got[2 * c + 1] == m_ref, (c, got[2 * c + 1], m_ref)
        # and the values themselves, from the host copy of the valid cells
        vals = np.concatenate([v[comp + c].ravel() for v in valid])
        assert abs(got[2 * c] - vals.sum()) <= 1e-12 * np.abs(vals).sum()
        assert got[2 * c + 1] == np.abs(vals).max()


def test_noalloc_reduction_matches_allocating_entry_points_two_calls():
    out = torch.full((32,), float("nan"), dtype=torch.float64, device="cuda")
    big, vbig, keep_big = _filled_mf((96, 80, 72), 32, 3, 1, seed=5)
    _check(big, vbig, 0, 3, out)
    small, vsmall, keep_small = _filled_mf((10, 12, 8), 8, 1, 2, seed=6)
    _check(small, vsmall, 0, 1, out)
    # and a component subset of the first MultiFab again
    _check(big, vbig, 1, 2, out)


def test_noalloc_reduction_refuses_beyond_maximum():
    lib = capi.load()
    mf, _, keep = _filled_mf((8, 8, 8), 8, 1, 0, seed=7)
    out = torch.zeros(64, dtype=torch.float64, device="cuda")
    st = lib.mgx_multifab_sum_norm0_noalloc(mf.h, 0, 17, ctypes.c_void_p(out.data_ptr()))
    assert st != 0
    buf = ctypes.create_string_buffer(512)
    lib.mgx_last_error(buf, 512)
    assert b"components" in buf.value
    lib.mgx_clear_error()
    st = lib.mgx_multifab_sum_norm0_noalloc(mf.h, 0, 2, ctypes.c_void_p(out.data_ptr()))
    assert st != 0  # the MultiFab has one component
    lib.mgx_clear_error()
