"""Manufactured-solution problems of the reference's own linear-solver test.

Restates ``Tests/LinearSolvers/ABecLaplacian_C/initProb_K.H`` (prob_type 2,
``actual_init_abeclap`` lines 66-147) and the set-up in ``MyTest.cpp``
(domain [0,1]^3, A=1e-3, B=1 from ``MyTest.H:86-87``; BCs Dirichlet/Neumann
from ``MyTest.cpp:163-168``; face b by ``average_cellcenter_to_face``).
The reference test prints the max-norm error of the converged solution; those
printed values are the known answers these problems are checked against.
"""
import numpy as np

DIRICHLET, NEUMANN = 101, 102


def abec_prob2(n):
    """Returns dict with a, bx, by, bz, rhs, bcv (ghost-grown), exact, scalars, BCs."""
    w, sigma = 0.05, 10.0
    theta = 0.5 * np.log(3.0) / (w + 1.0e-50)
    pi = 3.1415926535897932
    tpi, fpi = 2.0 * pi, 4.0 * pi
    fac = 3 * 4 * pi * pi
    a_s, b_s = 1.0e-3, 1.0
    dx = 1.0 / n
    c = (np.arange(-1, n + 1) + 0.5) * dx            # grown cell centres
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    xc = yc = zc = 0.5
    r = np.sqrt((X - xc) ** 2 + (Y - yc) ** 2 + (Z - zc) ** 2)
    beta = (sigma - 1.0) / 2.0 * np.tanh(theta * (r - 0.25)) + (sigma + 1.0) / 2.0
    tmp = np.cosh(theta * (r - 0.25))
    dbdrfac = (sigma - 1.0) / 2.0 / (tmp * tmp) * theta / r * b_s
    cx, cy, cz = np.cos(tpi * X), np.cos(tpi * Y), np.cos(tpi * Z)
    fx, fy, fz = np.cos(fpi * X), np.cos(fpi * Y), np.cos(fpi * Z)
    exact = cx * cy * cz + 0.25 * fx * fy * fz
    rhs = (beta * b_s * fac * (cx * cy * cz + fx * fy * fz)
           + dbdrfac * ((X - xc) * (tpi * np.sin(tpi * X) * cy * cz + pi * np.sin(fpi * X) * fy * fz)
                        + (Y - yc) * (tpi * cx * np.sin(tpi * Y) * cz + pi * fx * np.sin(fpi * Y) * fz)
                        + (Z - zc) * (tpi * cx * cy * np.sin(tpi * Z) + pi * fx * fy * np.sin(fpi * Z)))
           + a_s * (cx * cy * cz + 0.25 * fx * fy * fz))
    # Dirichlet values: exact solution at the clamped (face) position.
    xb, yb, zb = (np.clip(V, 0.0, 1.0) for V in (X, Y, Z))
    bcv = (np.cos(tpi * xb) * np.cos(tpi * yb) * np.cos(tpi * zb)
           + 0.25 * np.cos(fpi * xb) * np.cos(fpi * yb) * np.cos(fpi * zb))
    bcv[1:-1, 1:-1, 1:-1] = 0.0
    s = slice(1, -1)
    bx = 0.5 * (beta[s, s, :-1] + beta[s, s, 1:])      # (n, n, n+1)
    by = 0.5 * (beta[s, :-1, s] + beta[s, 1:, s])      # (n, n+1, n)
    bz = 0.5 * (beta[:-1, s, s] + beta[1:, s, s])      # (n+1, n, n)
    return {
        "a": np.ones((n, n, n)), "bx": np.ascontiguousarray(bx), "by": np.ascontiguousarray(by),
        "bz": np.ascontiguousarray(bz), "rhs": np.ascontiguousarray(rhs[s, s, s]),
        "bcv": np.ascontiguousarray(bcv), "exact": np.ascontiguousarray(exact[s, s, s]),
        "alpha": a_s, "beta": b_s,
        "bc_lo": (DIRICHLET, NEUMANN, NEUMANN), "bc_hi": (NEUMANN, DIRICHLET, NEUMANN),
    }
