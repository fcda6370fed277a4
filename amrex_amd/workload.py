"""Seeded synthetic inputs for the MLMG configurations.

Thin ctypes wrapper over ``csrc/workload.cpp`` (SplitMix64 counter-based
draws).  Arrays use the global layouts documented there and are returned as
numpy arrays shaped (nz, ny, nx) (x fastest in memory); face arrays of
direction d carry one extra point along d.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
WORKLOAD_SO = os.path.join(LIB_DIR, "libmgx_workload.so")

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(WORKLOAD_SO):
            raise RuntimeError(f"{WORKLOAD_SO} missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(WORKLOAD_SO)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        u64 = ctypes.c_uint64
        lib.mgx_wl_splitmix64.argtypes = [u64]
        lib.mgx_wl_splitmix64.restype = u64
        lib.mgx_wl_uniform.argtypes = [u64, ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                                       ctypes.c_double, dp]
        lib.mgx_wl_normal.argtypes = [u64, ctypes.c_int, ctypes.c_int64, dp]
        lib.mgx_wl_bface.argtypes = [u64, ip, ctypes.c_int, dp]
        lib.mgx_wl_abec.argtypes = [u64, ip, dp, dp, dp, dp, dp]
        lib.mgx_wl_fourier_rhs.argtypes = [u64, ip, ctypes.c_int, dp, dp]
        lib.mgx_wl_amr_boxes.argtypes = [u64, ctypes.c_int, ctypes.c_int, ip]
        lib.mgx_wl_amr_boxes.restype = ctypes.c_int
        lib.mgx_wl_blobs.argtypes = [u64, ctypes.c_int, dp]
        lib.mgx_wl_blob_field.argtypes = [ctypes.c_int, dp, ctypes.c_double, ip, ip, ctypes.c_double,
                                          ctypes.c_double, dp]
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _ip(t):
    arr = (ctypes.c_int * len(t))(*t)
    return arr


def face_shape(n, d):
    """numpy shape (z, y, x) of the face array of direction d on domain n=(nx,ny,nz)."""
    m = list(n)
    m[d] += 1
    return (m[2], m[1], m[0])


def uniform(seed, array_id, count, lo=0.0, hi=1.0):
    out = np.empty(count, dtype=np.float64)
    _load().mgx_wl_uniform(seed, array_id, count, lo, hi, _dp(out))
    return out


def normal(seed, array_id, count):
    out = np.empty(count, dtype=np.float64)
    _load().mgx_wl_normal(seed, array_id, count, _dp(out))
    return out


def abec_coeffs(n, seed):
    """Variable-coefficient ABecLaplacian recipe (a, bx, by, bz, rhs)."""
    n = tuple(int(v) for v in n)
    a = np.empty((n[2], n[1], n[0]))
    rhs = np.empty_like(a)
    b = [np.empty(face_shape(n, d)) for d in range(3)]
    _load().mgx_wl_abec(seed, _ip(n), _dp(a), _dp(b[0]), _dp(b[1]), _dp(b[2]), _dp(rhs))
    return a, b[0], b[1], b[2], rhs


def fourier_rhs(n, seed, nmodes=8):
    """Periodic Poisson right-hand side and its modes (kx, ky, kz, is_sin, A)."""
    n = tuple(int(v) for v in n)
    rhs = np.empty((n[2], n[1], n[0]))
    modes = np.empty((nmodes, 5))
    _load().mgx_wl_fourier_rhs(seed, _ip(n), nmodes, _dp(rhs), _dp(modes))
    return rhs, modes


def fourier_exact(n, modes):
    """Exact discrete solution of L phi = rhs for the periodic 7-point Laplacian.

    Each Fourier mode of the cell-centred grid is an eigenvector of the
    discrete Laplacian with eigenvalue -sum_d (4/h_d^2) sin^2(pi k_d h_d); the
    solution is therefore sum_m A_m trig_m / lambda_m (zero mean).
    """
    nx, ny, nz = n
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    z = (np.arange(nz) + 0.5) / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    phi = np.zeros((nz, ny, nx))
    h = (1.0 / nx, 1.0 / ny, 1.0 / nz)
    for kx, ky, kz, is_sin, amp in modes:
        lam = -sum(4.0 / hd ** 2 * np.sin(np.pi * kd * hd) ** 2 for kd, hd in zip((kx, ky, kz), h))
        arg = 2 * np.pi * (kx * X + ky * Y + kz * Z)
        phi += amp * (np.sin(arg) if is_sin else np.cos(arg)) / lam
    return phi


@dataclass
class Config:
    """One benchmark configuration."""
    name: str
    n: tuple
    max_grid_size: int
    op: str                 # "abec" or "poisson"
    periodic: tuple
    bc_lo: tuple
    bc_hi: tuple
    seed: int
    alpha: float = 1.0
    beta: float = 1.0
    maxorder: int = 3
    tol_rel: float = 1e-10


DIRICHLET, NEUMANN, REFLECT_ODD, INTERIOR = 101, 102, 103, 0

CONFIGS = {
    # Plain FP64 periodic Poisson, 64^3, 8 boxes.
    "poisson64": Config("poisson64", (64, 64, 64), 32, "poisson", (1, 1, 1),
                        (INTERIOR,) * 3, (INTERIOR,) * 3, seed=1),
    # Variable-coefficient ABecLaplacian 256^3, homogeneous Dirichlet, 64 boxes.
    "abec256": Config("abec256", (256, 256, 256), 64, "abec", (0, 0, 0),
                      (DIRICHLET,) * 3, (DIRICHLET,) * 3, seed=2),
    # Same recipe at 512^3, 64 boxes of 128^3.
    "abec512": Config("abec512", (512, 512, 512), 128, "abec", (0, 0, 0),
                      (DIRICHLET,) * 3, (DIRICHLET,) * 3, seed=3),
}


# Adjoint-gradient workload: the abec256 operator, a batch of 8 right-hand
# sides N(0,1) (seeds 501..508), phi_target N(0,1) (seed 599), and the face
# locations of the finite-difference check (seed 777).
ADJOINT_RHS_SEEDS = tuple(range(501, 509))
ADJOINT_TARGET_SEED = 599
ADJOINT_FD_SEED = 777
_ID_ADJ_RHS, _ID_ADJ_TARGET, _ID_FD_DIR, _ID_FD_LOC = 30, 31, 32, 33


def adjoint_inputs(n, batch=8):
    """(rhs[batch], phi_target[batch]) of the adjoint-gradient workload, shape (batch, nz, ny, nx)."""
    n = tuple(int(v) for v in n)
    nc = n[0] * n[1] * n[2]
    rhs = np.stack([normal(s, _ID_ADJ_RHS, nc).reshape(n[2], n[1], n[0]) for s in ADJOINT_RHS_SEEDS[:batch]])
    tgt = normal(ADJOINT_TARGET_SEED, _ID_ADJ_TARGET, batch * nc).reshape(batch, n[2], n[1], n[0])
    return rhs, tgt


def fd_locations(n, count=32):
    """Face locations (direction d, (k, j, i) index in the face array of d) of the
    finite-difference check, drawn uniformly."""
    u = uniform(ADJOINT_FD_SEED, _ID_FD_DIR, count)
    v = uniform(ADJOINT_FD_SEED, _ID_FD_LOC, 3 * count)
    out = []
    for m in range(count):
        d = min(int(u[m] * 3), 2)
        shape = face_shape(n, d)
        loc = tuple(min(int(v[3 * m + e] * shape[e]), shape[e] - 1) for e in range(3))
        out.append((d, loc))
    return out


def make_inputs(cfg: Config, n=None):
    """(a, bx, by, bz, rhs) for a configuration, optionally at another size."""
    n = tuple(n or cfg.n)
    if cfg.op == "abec":
        return abec_coeffs(n, cfg.seed)
    rhs, _ = fourier_rhs(n, cfg.seed)
    return None, None, None, None, rhs


# Two-level AMR workload: coarse 256^3 periodic Poisson, 40 fine boxes (ratio 2)
# placed with seed 40 inside the central half, RHS = 3 Gaussian blobs
# (sigma 0.05, seed 4), the coarse-level mean subtracted on both levels.
AMR_BOX_SEED = 40
AMR_BLOB_SEED = 4


def amr_boxes(ncoarse=256, nbox=40, seed=AMR_BOX_SEED):
    """Fine boxes (lo xyz, hi xyz) in the fine index space of an ncoarse^3 domain."""
    out = np.zeros(6 * nbox, dtype=np.int32)
    placed = _load().mgx_wl_amr_boxes(seed, ncoarse, nbox, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if placed != nbox:
        raise RuntimeError(f"placed only {placed} of {nbox} boxes")
    return [tuple(int(v) for v in out[6 * i: 6 * i + 6]) for i in range(nbox)]


def amr_rhs(ncoarse, boxes, seed=AMR_BLOB_SEED, nblob=3, sigma=0.05):
    """(coarse rhs (n^3, z-y-x), [fine rhs per box]) of the blob right-hand side."""
    lib = _load()
    blobs = np.zeros(4 * nblob)
    lib.mgx_wl_blobs(seed, nblob, _dp(blobs))
    n = ncoarse
    crse = np.zeros((n, n, n))
    lib.mgx_wl_blob_field(nblob, _dp(blobs), sigma, _ip((0, 0, 0)), _ip((n - 1,) * 3), 1.0 / n, 0.0, _dp(crse))
    shift = float(crse.mean())
    crse -= shift
    fine = []
    for b in boxes:
        shp = (b[5] - b[2] + 1, b[4] - b[1] + 1, b[3] - b[0] + 1)
        f = np.zeros(shp)
        lib.mgx_wl_blob_field(nblob, _dp(blobs), sigma, _ip(b[:3]), _ip(b[3:]), 0.5 / n, shift, _dp(f))
        fine.append(f)
    return crse, fine
