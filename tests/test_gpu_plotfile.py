"""VisMF, plotfiles, MultiFab::Copy and MLMG::checkPoint of the library, through the C ABI.

Bar: files the library writes from HBM are byte-identical to the oracle's
restatement of AMReX's writers for the same data (headers, FAB headers,
offsets, extrema, data); the library reads the oracle's files back exactly,
into new MultiFabs or defined ones with other ghost widths; PlotFileData
returns the variables, levels and geometry written.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from amrex_amd import capi, mlmg  # noqa: E402
from oracle import plotfile as pf  # noqa: E402


def _fab_shapes(ba, ncomp, ngrow, nodal=(0, 0, 0)):
    out = []
    for b in ba.boxes():
        lo, hi = b[:3], [b[3 + d] + nodal[d] for d in range(3)]
        out.append((ncomp,) + tuple(hi[d] - lo[d] + 1 + 2 * ngrow for d in (2, 1, 0)))
    return out


def _mf(n, mgs, ncomp, ngrow, seed, nodal=(0, 0, 0)):
    ba = mlmg.BoxArray((0, 0, 0), [v - 1 for v in n]).maxSize(mgs)
    dm = mlmg.DistributionMapping(ba)
    m = mlmg.MultiFab(ba, dm, ncomp, ngrow, nodal=nodal)
    rng = np.random.default_rng(seed)
    fabs = [rng.standard_normal(s) for s in _fab_shapes(ba, ncomp, ngrow, nodal)]
    m.setBoxes(fabs, ghosts=True)
    return ba, dm, m, fabs


def _same_files(a, b):
    assert open(a, "rb").read() == open(b, "rb").read(), (a, b)


@pytest.mark.parametrize("ncomp,ngrow,nodal", [(1, 0, (0, 0, 0)), (2, 1, (0, 0, 0)), (1, 2, (0, 1, 0))])
def test_vismf_write_is_byte_identical_to_oracle(tmp_path, ncomp, ngrow, nodal):
    ba, dm, m, fabs = _mf((12, 10, 8), 6, ncomp, ngrow, seed=ncomp + ngrow, nodal=nodal)
    m.write(str(tmp_path / "lib"))
    pf.vismf_write(str(tmp_path / "orc"), m.boxes(), fabs, ngrow, itype=nodal)
    # the data file is named after each MultiFab; compare the headers with that name swapped
    assert open(tmp_path / "lib_H").read().replace("lib_D_", "orc_D_") == open(tmp_path / "orc_H").read()
    _same_files(tmp_path / "lib_D_00000", tmp_path / "orc_D_00000")


def test_vismf_set_ghost_fills_the_multifab_and_the_file(tmp_path):
    ba, dm, m, fabs = _mf((12, 12, 6), 6, 2, 2, seed=4)
    m.write(str(tmp_path / "lib"), set_ghost=True)
    pf.vismf_write(str(tmp_path / "orc"), m.boxes(), fabs, 2, set_ghost=True)
    assert open(tmp_path / "lib_H").read().replace("lib_D_", "orc_D_") == open(tmp_path / "orc_H").read()
    _same_files(tmp_path / "lib_D_00000", tmp_path / "orc_D_00000")
    for a, b in zip(m.getBoxes(ghosts=True), pf.set_ghost_cells(fabs, 2)):
        np.testing.assert_array_equal(a, b)


def test_vismf_read_back_and_other_ghost_widths(tmp_path):
    ba, dm, m, fabs = _mf((16, 12, 10), 8, 2, 1, seed=11)
    name = str(tmp_path / "orc")
    pf.vismf_write(name, m.boxes(), fabs, 1)
    got = mlmg.MultiFab.read(name)
    assert (got.ncomp, got.ngrow) == (2, 1)
    for a, b in zip(got.getBoxes(ghosts=True), fabs):
        np.testing.assert_array_equal(a, b)
    for ng in (0, 2):
        d = mlmg.MultiFab(ba, dm, 2, ng)
        d.setVal(-7.0)
        d.readInto(name)
        g = min(ng, 1)
        for a, b in zip(d.getBoxes(ghosts=True), fabs):
            sa = a[:, ng - g:a.shape[1] - (ng - g), ng - g:a.shape[2] - (ng - g), ng - g:a.shape[3] - (ng - g)]
            sb = b[:, 1 - g:b.shape[1] - (1 - g), 1 - g:b.shape[2] - (1 - g), 1 - g:b.shape[3] - (1 - g)]
            np.testing.assert_array_equal(sa, sb)
            if ng == 2:  # the outer ghost layer the file does not have is left as it was
                assert np.all(a[:, 0] == -7.0)


def test_plotfile_byte_identical_and_read_back(tmp_path):
    n = (16, 8, 8)
    ba, dm, m, fabs = _mf(n, 8, 2, 1, seed=5)
    geom = mlmg.Geometry(n, prob_lo=(-1.0, 0.0, 0.0), prob_hi=(1.0, 0.5, 0.5))
    lib = str(tmp_path / "plt_lib")
    mlmg.WriteSingleLevelPlotfile(lib, m, ["phi", "rhs"], geom, time=2.5, level_step=12)
    valid = [f[:, 1:-1, 1:-1, 1:-1] for f in fabs]
    g = dict(domain=((0, 0, 0), (15, 7, 7)), prob_lo=(-1.0, 0.0, 0.0), prob_hi=(1.0, 0.5, 0.5),
             dx=(2.0 / 16, 0.5 / 8, 0.5 / 8), coord=0)
    orc = str(tmp_path / "plt_orc")
    pf.plotfile_write(orc, [(m.boxes(), valid)], ["phi", "rhs"], [g], time=2.5, level_steps=[12])
    for rel in ("Header", "Level_0/Cell_H", "Level_0/Cell_D_00000"):
        _same_files(os.path.join(lib, rel), os.path.join(orc, rel))
    p = mlmg.PlotFileData(orc)
    assert p.varnames == ["phi", "rhs"] and p.time == 2.5 and p.finest_level == 0 and p.ncomp == 2
    lv = p.level(0)
    assert lv["domain"] == ((0, 0, 0), (15, 7, 7)) and lv["level_step"] == 12
    assert lv["dx"] == (0.125, 0.0625, 0.0625) and lv["prob_lo"] == (-1.0, 0.0, 0.0)
    rhs = p.get(0, "rhs")
    assert rhs.ncomp == 1 and rhs.ngrow == 0
    for a, b in zip(rhs.getBoxes(), valid):
        np.testing.assert_array_equal(a, b[1])
    both = p.get(0)
    for a, b in zip(both.getBoxes(), valid):
        np.testing.assert_array_equal(a, b)


def test_two_level_plotfile_matches_oracle(tmp_path):
    nc = (8, 8, 8)
    ba0, dm0, m0, f0 = _mf(nc, 4, 1, 0, seed=1)
    ba1 = mlmg.BoxArray(boxes=[(4, 4, 4, 9, 7, 11)])
    dm1 = mlmg.DistributionMapping(ba1)
    m1 = mlmg.MultiFab(ba1, dm1, 1, 0)
    f1 = [np.random.default_rng(2).standard_normal((1, 8, 4, 6))]
    m1.setBoxes(f1, ghosts=True)
    g0 = mlmg.Geometry(nc)
    g1 = mlmg.Geometry((16, 16, 16))
    lib = str(tmp_path / "plt2_lib")
    mlmg.WriteMultiLevelPlotfile(lib, [m0, m1], ["u"], [g0, g1], time=0.125, level_steps=[4, 8], ref_ratio=[2])
    geoms = [dict(domain=((0, 0, 0), (7, 7, 7)), prob_lo=(0.0,) * 3, prob_hi=(1.0,) * 3, dx=(0.125,) * 3, coord=0),
             dict(domain=((0, 0, 0), (15, 15, 15)), prob_lo=(0.0,) * 3, prob_hi=(1.0,) * 3, dx=(0.0625,) * 3,
                  coord=0)]
    orc = str(tmp_path / "plt2_orc")
    pf.plotfile_write(orc, [(m0.boxes(), f0), (m1.boxes(), f1)], ["u"], geoms, time=0.125, level_steps=[4, 8])
    for rel in ("Header", "Level_0/Cell_H", "Level_0/Cell_D_00000", "Level_1/Cell_H", "Level_1/Cell_D_00000"):
        _same_files(os.path.join(lib, rel), os.path.join(orc, rel))
    p = mlmg.PlotFileData(lib)
    assert p.finest_level == 1 and p.level(0)["ref_ratio"] == 2
    np.testing.assert_array_equal(p.get(1, "u").getBoxes()[0], f1[0][0])
    # writing again over an existing plotfile keeps the old one under a new name
    mlmg.WriteMultiLevelPlotfile(lib, [m0, m1], ["u"], [g0, g1], time=0.125, level_steps=[4, 8], ref_ratio=[2])
    assert any(d.startswith("plt2_lib.old.") for d in os.listdir(tmp_path))


def test_multifab_copy_components_and_ghosts():
    ba, dm, src, fabs = _mf((12, 12, 12), 6, 3, 2, seed=9)
    dst = mlmg.MultiFab(ba, dm, 2, 1)
    dst.setVal(0.0)
    dst.copyFrom(src, 1, 0, 2, nghost=1)
    for a, b in zip(dst.getBoxes(ghosts=True), fabs):
        np.testing.assert_array_equal(a, b[1:3, 1:-1, 1:-1, 1:-1])
    with pytest.raises(capi.MgxError):
        dst.copyFrom(src, 2, 0, 2)  # components past the end
    with pytest.raises(capi.MgxError):
        dst.copyFrom(src, 0, 0, 1, nghost=2)  # wider than dst's ghost layer


def test_checkpoint_layout_and_error(tmp_path):
    n = (16, 16, 16)
    rng = np.random.default_rng(3)
    prob = mlmg.Problem(n, 8, op="abec", a=rng.uniform(0.5, 1.5, (16, 16, 16)))
    rhs = rng.standard_normal((16, 16, 16))
    sol = prob.solve(rhs).cpu().numpy()
    name = str(tmp_path / "chk")
    with pytest.raises(capi.MgxError, match="checkPoint: not implemented"):
        prob.mlmg.checkPoint(prob.sol, prob.rhs, 1e-10, 0.0, name)
    head = open(os.path.join(name, "Header")).read().split("\n")
    assert head[:4] == ["MLABecLaplacian", "a_tol_rel = 1e-10", "a_tol_abs = 0", "verbose = 0"]
    assert "max_fmg_iters = 0" in head and "namrlevs = 1" in head and "finest_amr_lev = 0" in head
    assert "solve_called = 1" in head and "linop_prepared = 1" in head
    got = mlmg.MultiFab.read(os.path.join(name, "Level_0", "sol"))
    back = np.zeros_like(sol)
    for b, a in zip(got.boxes(), got.getBoxes()):
        back[b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1] = a
    np.testing.assert_array_equal(back, sol)
    assert os.path.exists(os.path.join(name, "Level_0", "rhs_H"))
