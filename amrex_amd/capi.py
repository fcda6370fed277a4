"""ctypes binding of libmgx.so (the C ABI declared in include/amrex_mlmg.h).

Loading fails loudly when the HIP library is missing: there is no CPU
fallback for the solver.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmgx.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "amrex_mlmg.h")

c_int, c_double, c_ll, c_vp = ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_void_p
P = ctypes.POINTER
IA3 = P(c_int)
DP = P(c_double)


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("niters", c_int), ("nlevels", c_int), ("last_bottom_status", c_int),
        ("last_bottom_iters", c_int), ("init_resnorm", c_double), ("rhsnorm", c_double),
        ("final_resnorm", c_double), ("solve_seconds", c_double), ("iter_seconds", c_double),
        ("res_history", c_double * 256), ("nvcycles", c_int),
    ]


_SIGS = {
    "amrex_fi_new_boxarray": (None, [P(c_vp), IA3, IA3]),
    "amrex_fi_new_boxarray_from_bxfarr": (None, [P(c_vp), IA3, c_int, c_int, c_int]),
    "amrex_fi_delete_boxarray": (None, [c_vp]),
    "amrex_fi_clone_boxarray": (None, [P(c_vp), c_vp]),
    "amrex_fi_boxarray_maxsize": (None, [c_vp, IA3]),
    "amrex_fi_boxarray_nboxes": (c_ll, [c_vp]),
    "amrex_fi_boxarray_get_box": (None, [c_vp, c_int, IA3, IA3]),
    "amrex_fi_boxarray_numpts": (c_ll, [c_vp]),
    "amrex_fi_new_geometry": (None, [P(c_vp), IA3, IA3]),
    "amrex_fi_delete_geometry": (None, [c_vp]),
    "amrex_fi_geometry_get_intdomain": (None, [c_vp, IA3, IA3]),
    "amrex_fi_new_distromap": (None, [P(c_vp), c_vp]),
    "amrex_fi_delete_distromap": (None, [c_vp]),
    "amrex_fi_new_multifab": (None, [P(c_vp), P(c_vp), P(c_vp), c_int, IA3, IA3]),
    "amrex_fi_delete_multifab": (None, [c_vp]),
    "amrex_fi_multifab_ncomp": (c_int, [c_vp]),
    "amrex_fi_multifab_ngrow": (None, [c_vp, IA3]),
    "amrex_fi_multifab_dataptr_int": (None, [c_vp, c_int, P(DP), IA3, IA3]),
    "amrex_fi_multifab_sum": (c_double, [c_vp, c_int]),
    "amrex_fi_multifab_norm0": (c_double, [c_vp, c_int]),
    "mgx_multifab_sum_norm0_noalloc": (c_int, [c_vp, c_int, c_int, c_vp]),
    "mgx_linop_time_sweep_pairs": (c_int, [c_vp, c_int, c_int, c_int, P(c_double)]),
    "mgx_set_pair_sweeps": (c_int, [c_int]),
    "mgx_hbm_probe": (c_int, [c_int, c_vp, c_vp, c_ll, c_int, P(c_double)]),
    "mgx_linop_set_coarse_fine_bc_type": (c_int, [c_vp, c_vp, c_int, c_int]),
    "amrex_fi_multifab_setval": (None, [c_vp, c_double, c_int, c_int, IA3]),
    "amrex_fi_new_abeclaplacian": (None, [P(c_vp), c_int, P(c_vp), P(c_vp), P(c_vp), c_int, c_int, c_int, c_int]),
    "amrex_fi_abeclap_set_scalars": (None, [c_vp, c_double, c_double]),
    "amrex_fi_abeclap_set_acoeffs": (None, [c_vp, c_int, c_vp]),
    "amrex_fi_abeclap_set_bcoeffs": (None, [c_vp, c_int, P(c_vp)]),
    "amrex_fi_new_poisson": (None, [P(c_vp), c_int, P(c_vp), P(c_vp), P(c_vp), c_int, c_int, c_int, c_int]),
    "amrex_fi_delete_linop": (None, [c_vp]),
    "amrex_fi_linop_set_maxorder": (None, [c_vp, c_int]),
    "amrex_fi_linop_set_domain_bc": (None, [c_vp, IA3, IA3]),
    "amrex_fi_linop_set_coarse_fine_bc": (None, [c_vp, c_vp, c_int]),
    "amrex_fi_linop_set_level_bc": (None, [c_vp, c_int, c_vp]),
    "amrex_fi_new_multigrid": (None, [P(c_vp), c_vp]),
    "amrex_fi_delete_multigrid": (None, [c_vp]),
    "amrex_fi_multigrid_solve": (c_double, [c_vp, P(c_vp), P(c_vp), c_double, c_double]),
    "amrex_fi_multigrid_get_grad_solution": (None, [c_vp, P(c_vp)]),
    "amrex_fi_multigrid_get_fluxes": (None, [c_vp, P(c_vp)]),
    "amrex_fi_multigrid_comp_residual": (None, [c_vp, P(c_vp), P(c_vp), P(c_vp)]),
    "amrex_fi_multigrid_set_verbose": (None, [c_vp, c_int]),
    "amrex_fi_multigrid_set_max_iter": (None, [c_vp, c_int]),
    "amrex_fi_multigrid_set_max_fmg_iter": (None, [c_vp, c_int]),
    "amrex_fi_multigrid_set_fixed_iter": (None, [c_vp, c_int]),
    "amrex_fi_multigrid_set_bottom_solver": (None, [c_vp, c_int]),
    "amrex_fi_multigrid_set_bottom_verbose": (None, [c_vp, c_int]),
    "amrex_fi_multigrid_set_always_use_bnorm": (None, [c_vp, c_int]),
    "amrex_fi_multigrid_set_final_fill_bc": (None, [c_vp, c_int]),
    "mgx_last_error": (c_int, [ctypes.c_char_p, c_int]),
    "mgx_clear_error": (None, []),
    "mgx_set_abort_on_error": (None, [c_int]),
    "mgx_set_stream": (c_int, [c_vp]),
    "mgx_get_stream": (c_vp, []),
    "mgx_synchronize": (c_int, []),
    "mgx_abi_version": (c_int, []),
    "mgx_set_dense_storage": (c_int, [c_int]),
    "mgx_new_geometry": (None, [P(c_vp), IA3, IA3, DP, DP, IA3]),
    "mgx_new_abeclaplacian": (None, [P(c_vp), c_vp, c_vp, c_int, c_int, c_int]),
    "mgx_new_poisson": (None, [P(c_vp), c_vp, c_vp, c_int, c_int, c_int]),
    "mgx_multifab_data": (c_vp, [c_vp, P(c_ll)]),
    "mgx_multifab_copy_global": (c_int, [c_vp, c_vp, IA3, c_int, c_int, c_int]),
    "mgx_multigrid_set_precision": (c_int, [c_vp, c_int]),
    "mgx_multigrid_set_refinement_vcycles": (c_int, [c_vp, c_int]),
    "mgx_multigrid_set_use_graph": (c_int, [c_vp, c_int]),
    "mgx_multigrid_set_bottom_solver_code": (c_int, [c_vp, c_int]),
    "mgx_multigrid_get_bottom_solver_code": (c_int, [c_vp]),
    "mgx_multigrid_set_bottom_tolerance": (c_int, [c_vp, c_double, c_double]),
    "mgx_multigrid_set_bottom_max_iter": (c_int, [c_vp, c_int]),
    "mgx_multigrid_set_pre_post_smooth": (c_int, [c_vp, c_int, c_int]),
    "mgx_multigrid_get_stats": (c_int, [c_vp, P(SolveStats)]),
    "mgx_linop_num_mg_levels": (c_int, [c_vp]),
    "mgx_linop_level_info": (c_int, [c_vp, c_int, IA3, P(c_int), IA3, c_int]),
    "mgx_linop_level_op": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_int]),
    "mgx_linop_time_smoother": (c_int, [c_vp, c_int, c_int, c_int, DP]),
    "mgx_multigrid_solve_adjoint": (c_double, [c_vp, P(c_vp), P(c_vp), c_double, c_double]),
    "mgx_multigrid_get_fluxes_sol": (c_int, [c_vp, P(c_vp), P(c_vp)]),
    "mgx_multigrid_get_cell_fluxes": (c_int, [c_vp, P(c_vp), P(c_vp)]),
    "amrex_fi_write_multifab": (None, [c_vp, ctypes.c_char_p]),
    "amrex_fi_read_multifab": (None, [c_vp, ctypes.c_char_p]),
    "amrex_fi_write_plotfile": (None, [ctypes.c_char_p, c_int, P(c_vp), P(ctypes.c_char_p), P(c_vp), c_double,
                                       P(c_int), P(c_int)]),
    "mgx_read_multifab_new": (c_int, [P(c_vp), ctypes.c_char_p]),
    "mgx_write_multifab_set_ghost": (c_int, [c_vp, ctypes.c_char_p, c_int]),
    "mgx_multifab_info": (c_int, [c_vp, P(c_int), P(c_int), P(c_int), IA3]),
    "mgx_multifab_box": (c_int, [c_vp, c_int, IA3, IA3]),
    "mgx_multifab_copy": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int]),
    "mgx_new_plotfile_data": (c_int, [P(c_vp), ctypes.c_char_p]),
    "mgx_delete_plotfile_data": (None, [c_vp]),
    "mgx_plotfile_data_info": (c_int, [c_vp, P(c_int), P(c_int), P(c_double), P(c_int)]),
    "mgx_plotfile_data_varname": (c_int, [c_vp, c_int, ctypes.c_char_p, c_int]),
    "mgx_plotfile_data_level": (c_int, [c_vp, c_int, IA3, IA3, P(c_double), P(c_double), P(c_double), P(c_int),
                                        P(c_int)]),
    "mgx_plotfile_data_get": (c_int, [c_vp, c_int, ctypes.c_char_p, P(c_vp)]),
    "mgx_multigrid_checkpoint": (c_int, [c_vp, P(c_vp), P(c_vp), c_double, c_double, ctypes.c_char_p]),
    "mgx_new_gmres_mlmg": (c_int, [P(c_vp), c_vp]),
    "mgx_delete_gmres_mlmg": (None, [c_vp]),
    "mgx_gmres_mlmg_solve": (c_double, [c_vp, c_vp, c_vp, c_double, c_double]),
    "mgx_gmres_mlmg_set_verbose": (None, [c_vp, c_int]),
    "mgx_gmres_mlmg_set_max_iters": (None, [c_vp, c_int]),
    "mgx_gmres_mlmg_set_restart_length": (None, [c_vp, c_int]),
    "mgx_gmres_mlmg_set_precond_num_iters": (None, [c_vp, c_int]),
    "mgx_gmres_mlmg_use_precond": (c_int, [c_vp, c_int]),
    "mgx_gmres_mlmg_set_property_of_zero": (None, [c_vp, c_int]),
    "mgx_gmres_mlmg_stats": (c_int, [c_vp, P(c_int), P(c_int), P(c_double), P(c_double), c_int]),
    "mgx_multifab_host": (c_int, [c_vp, DP, ctypes.c_longlong, c_int]),
    "mgx_abeclap_coeff_gradient": (c_int, [c_vp, c_vp, c_vp, c_vp, P(c_vp), c_int]),
}

_lib = None


def header_symbols():
    """Function names declared in include/amrex_mlmg.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:[a-z_ ]+\**)\s*\**\b((?:amrex_fi|mgx)_\w+)\s*\(", text, re.M)))


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build(); there is no CPU fallback")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class MgxError(RuntimeError):
    pass


def check(prefix=""):
    """Raise if the last library call recorded an error."""
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    st = lib.mgx_last_error(buf, 1024)
    if st != 0:
        lib.mgx_clear_error()
        raise MgxError(f"{prefix}{buf.value.decode()} (status {st})")


def ivec(v):
    return (c_int * 3)(*[int(x) for x in v])


def dvec(v):
    return (c_double * 3)(*[float(x) for x in v])
