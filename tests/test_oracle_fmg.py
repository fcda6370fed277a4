"""Pins of the oracle's F-cycle (MLMG::mgFcycle with setMaxFmgIter) on the CPU.

The reference holds no golden vectors for the F-cycle, so its pieces are tied to
what the mathematics and the reference's data movement fix:
  * the interpolation is the tensor product of the 1-D (3/4, 1/4) cell-centred
    linear rule (a separable numpy restatement, boundary and ghost cells
    included);
  * the coarse data it reads: between levels of one BoxArray lineage,
    FillBoundary copies every ghost cell whose periodic image is a valid cell
    and leaves the others; from an agglomerated level, a zeroed copy holds
    valid and periodic-image values only;
  * one F-cycle solves a smooth periodic Poisson problem to its
    discretization error (the textbook property of full multigrid), while one
    V-cycle from zero does not: a dropped held correction, a wrong offset
    or weight of the interpolation, or a skipped residual update break it;
  * converged F-cycle solves equal converged V-cycle solves.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

DIR, NEU = 101, 102


def _interp_1d(nf, nc):
    # rows: fine cells, columns: grown coarse cells (ghost at column 0)
    W = np.zeros((nf, nc + 2))
    for i in range(nf):
        ic, o = i // 2, (1 if i % 2 else -1)
        W[i, ic + 1] += 0.75
        W[i, ic + o + 1] += 0.25
    return W


def test_interpolation_is_the_tensor_product_linear_rule():
    n = (16, 12, 20)
    prm = po.make_params(n, op="poisson", max_grid_size=8, agg_grid_size=8)
    dims = [h["n"] for h in po.hierarchy(prm)]
    nc = dims[1]
    rng = np.random.default_rng(1)
    crse = rng.standard_normal((nc[2] + 2, nc[1] + 2, nc[0] + 2))
    fine, _ = po.fmg_interp(prm, 1, crse, fill=False)
    Wx, Wy, Wz = (_interp_1d(dims[0][d], nc[d]) for d in range(3))
    ref = np.einsum("ka,jb,ic,abc->kji", Wz, Wy, Wx, crse)
    np.testing.assert_allclose(fine, ref, rtol=0, atol=1e-14)


def test_agglomerated_level_reads_zero_outside_physical_faces():
    # 64^3 in 16^3 boxes, agg_grid_size 8: MG level 1 is the coarsened BoxArray
    # (64 boxes), level 2 an agglomerated one (8 boxes)
    n = (64, 64, 64)
    prm = po.make_params(n, op="poisson", periodic=(0, 1, 0), bc_lo=(DIR, 0, NEU), bc_hi=(NEU, 0, DIR),
                         max_grid_size=16, agg_grid_size=8)
    hier = po.hierarchy(prm)
    assert len(hier[1]["boxes"]) == 64 and len(hier[2]["boxes"]) == 8
    rng = np.random.default_rng(3)
    for level, zeroed in ((1, False), (2, True)):
        # interpolation from MG level `level` to level-1
        nc = hier[level]["n"]
        crse = rng.standard_normal((nc[2] + 2, nc[1] + 2, nc[0] + 2))
        _, seen = po.fmg_interp(prm, level, crse, fill=True)
        inner = (slice(1, -1),) * 3
        np.testing.assert_array_equal(seen[inner], crse[inner])
        phys = np.zeros(seen.shape, dtype=bool)
        phys[:, :, 0] = phys[:, :, -1] = True  # x faces (non-periodic)
        phys[0, :, :] = phys[-1, :, :] = True  # z faces (non-periodic)
        if zeroed:
            assert np.all(seen[phys] == 0.0)
        else:
            assert np.array_equal(seen[phys], crse[phys])
        # y is periodic: ghosts inside the x-z range take the periodic image
        np.testing.assert_array_equal(seen[1:-1, 0, 1:-1], crse[1:-1, -2, 1:-1])
        np.testing.assert_array_equal(seen[1:-1, -1, 1:-1], crse[1:-1, 1, 1:-1])


def test_fillboundary_copies_periodic_images_only():
    n = (16, 16, 16)
    prm = po.make_params(n, op="poisson", periodic=(1, 0, 1), bc_lo=(0, DIR, 0), bc_hi=(0, NEU, 0),
                         max_grid_size=8, agg_grid_size=4)
    hier = po.hierarchy(prm)
    assert len(hier[1]["boxes"]) == 8  # the coarsened BoxArray, not an agglomerated one
    nc = hier[1]["n"]
    rng = np.random.default_rng(2)
    crse = rng.standard_normal((nc[2] + 2, nc[1] + 2, nc[0] + 2))
    _, filled = po.fmg_interp(prm, 1, crse, fill=True)
    per = (1, 0, 1)
    for k in range(nc[2] + 2):
        for j in range(nc[1] + 2):
            for i in range(nc[0] + 2):
                p = [i - 1, j - 1, k - 1]
                out = [d for d in range(3) if not 0 <= p[d] < nc[d]]
                if out and all(per[d] for d in out):
                    q = [p[d] % nc[d] + 1 for d in range(3)]
                    assert filled[k, j, i] == crse[q[2], q[1], q[0]]
                else:
                    assert filled[k, j, i] == crse[k, j, i]


def _smooth_periodic(n):
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    s1 = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y) * np.sin(4 * np.pi * Z)
    s2 = np.cos(2 * np.pi * X) * np.cos(6 * np.pi * Y)
    phi = s1 + 0.5 * s2
    lap = -24 * np.pi ** 2 * s1 - 0.5 * 40 * np.pi ** 2 * s2
    return phi, lap


def test_one_fcycle_reaches_the_discretization_error():
    n = 64
    phi, rhs = _smooth_periodic(n)

    def err(fmg, fixed):
        prm = po.make_params((n,) * 3, op="poisson", periodic=(1, 1, 1), bc_lo=(0,) * 3, bc_hi=(0,) * 3,
                             max_grid_size=32, agg_grid_size=32, max_fmg_iters=fmg, fixed_iters=fixed)
        s, _ = po.solve(prm, rhs)
        return np.abs((s - s.mean()) - (phi - phi.mean())).max()

    e_disc = err(0, 0)
    e_f1 = err(1, 1)
    e_v1 = err(0, 1)
    assert e_f1 <= 1.5 * e_disc
    assert e_v1 >= 5 * e_disc


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_fcycle_solve_converges_to_the_vcycle_solution(precision):
    n = (32, 40, 24)
    rng = np.random.default_rng(7)
    a = rng.uniform(0.5, 1.5, (n[2], n[1], n[0]))
    bx = rng.uniform(0.5, 2.0, (n[2], n[1], n[0] + 1))
    by = rng.uniform(0.5, 2.0, (n[2], n[1] + 1, n[0]))
    bz = rng.uniform(0.5, 2.0, (n[2] + 1, n[1], n[0]))
    bcv = rng.standard_normal((n[2] + 2, n[1] + 2, n[0] + 2))
    rhs = rng.standard_normal((n[2], n[1], n[0]))
    out = {}
    for fmg in (0, 1, 3):
        prm = po.make_params(n, op="abec", bc_lo=(DIR, NEU, DIR), bc_hi=(DIR, DIR, NEU), max_grid_size=16,
                             agg_grid_size=16, precision=precision, max_fmg_iters=fmg, tol_rel=1e-11)
        s, st = po.solve(prm, rhs, a, bx, by, bz, bcv=bcv)
        assert st["status"] == 0
        out[fmg] = (s, st)
    ref = out[0][0]
    for fmg, (s, st) in out.items():
        assert np.linalg.norm(s - ref) <= 1e-9 * np.linalg.norm(ref)
    # the F-cycle's first iteration reduces the residual more than a V-cycle
    assert out[3][1]["res_history"][0] < out[0][1]["res_history"][0]
