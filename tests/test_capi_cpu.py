"""CPU checks of the native boundary: the C-ABI library loads, exports every
symbol include/amrex_mlmg.h declares, and its host-only calls (BoxArray
chopping, error reporting) behave like the reference.  No GPU is touched."""
import ctypes

import numpy as np

from amrex_amd import capi, mlmg
from oracle import pyoracle as po


def test_library_exports_every_header_symbol():
    lib = capi.load()
    syms = capi.header_symbols()
    assert len(syms) >= 60
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) == set(capi._SIGS), set(syms) ^ set(capi._SIGS)
    assert lib.mgx_abi_version() == 1


def test_reference_fi_names_are_exported():
    # Entry points of Src/F_Interfaces/LinearSolvers/*_fi.cpp that bind this path.
    names = ["amrex_fi_new_abeclaplacian", "amrex_fi_abeclap_set_scalars", "amrex_fi_abeclap_set_acoeffs",
             "amrex_fi_abeclap_set_bcoeffs", "amrex_fi_new_poisson", "amrex_fi_delete_linop",
             "amrex_fi_linop_set_maxorder", "amrex_fi_linop_set_domain_bc", "amrex_fi_linop_set_level_bc",
             "amrex_fi_new_multigrid", "amrex_fi_multigrid_solve", "amrex_fi_multigrid_comp_residual",
             "amrex_fi_multigrid_set_max_iter", "amrex_fi_multigrid_set_bottom_solver"]
    lib = capi.load()
    for nm in names:
        assert hasattr(lib, nm), nm


def test_boxarray_maxsize_matches_oracle_chopping():
    for n, mgs in [((100, 64, 48), 32), ((512, 512, 512), 128), ((40, 48, 56), 16), ((20, 12, 28), 8)]:
        ba = mlmg.BoxArray((0, 0, 0), [v - 1 for v in n]).maxSize(mgs)
        prm = po.make_params(n, max_grid_size=mgs)
        ref = po.hierarchy(prm)[0]["boxes"]
        assert sorted(ba.boxes()) == sorted(ref)
        assert ba.numPts() == int(np.prod(n))


def test_boxarray_from_boxes_roundtrip():
    boxes = [(0, 0, 0, 7, 7, 7), (8, 0, 0, 15, 7, 7)]
    ba = mlmg.BoxArray(boxes=boxes)
    assert ba.boxes() == boxes


def test_error_is_recorded_not_aborting():
    lib = capi.load()
    lib.mgx_clear_error()
    lib.amrex_fi_linop_set_coarse_fine_bc(None, None, 2)
    buf = ctypes.create_string_buffer(256)
    assert lib.mgx_last_error(buf, 256) != 0
    assert b"null operator" in buf.value
    lib.mgx_clear_error()
    assert lib.mgx_last_error(buf, 256) == 0
