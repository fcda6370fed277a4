"""Coefficients of two-level ABecLaplacian test problems (shared by the oracle
and GPU tests): random fields cut into the coarse level and the fine boxes,
the coarse ones with the fine averages under the fine level, and the textbook
7-point operator for manufactured right-hand sides."""
import numpy as np


def coefs(nc, boxes, seed, periodic):
    """(coarse {a, bx, by, bz}, fine {a, bx, by, bz: per-box arrays}) cut from
    global random fields; on a periodic direction the two domain faces agree."""
    rng = np.random.default_rng(seed)

    def faces(n, d):
        f = rng.uniform(0.5, 2.0, (n + (d == 2), n + (d == 1), n + (d == 0)))
        if periodic[d]:
            idx = [slice(None)] * 3
            first = [slice(None)] * 3
            idx[2 - d], first[2 - d] = -1, 0
            f[tuple(idx)] = f[tuple(first)]
        return f

    nf = 2 * nc
    fa = rng.uniform(0.5, 1.5, (nf, nf, nf))
    fb = [faces(nf, d) for d in range(3)]
    ca = rng.uniform(0.5, 1.5, (nc, nc, nc))
    cb = [faces(nc, d) for d in range(3)]
    c1 = {"a": [fa[b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1] for b in boxes]}
    for d, k in enumerate(("bx", "by", "bz")):
        c1[k] = [fb[d][b[2]:b[5] + 1 + (d == 2), b[1]:b[4] + 1 + (d == 1), b[0]:b[3] + 1 + (d == 0)] for b in boxes]
    return (dict(a=ca, bx=cb[0], by=cb[1], bz=cb[2]), c1), fa, fb


def averaged(coef, fa, fb, boxes, periodic):
    """The coarse coefficients with the fine averages under the fine boxes: a
    as the mean of the 8 children, b on every face of a coarsened box (its
    boundary faces included) as the mean of the 4 fine faces, and on a
    periodic direction at the image of a domain face too."""
    c0, c1 = coef
    a = c0["a"].copy()
    b = [c0["bx"].copy(), c0["by"].copy(), c0["bz"].copy()]
    nc = a.shape[0]
    for bx in boxes:
        lo = [bx[d] // 2 for d in range(3)]
        hi = [bx[3 + d] // 2 for d in range(3)]
        m = [hi[d] - lo[d] + 1 for d in range(3)]
        f = fa[2 * lo[2]:2 * hi[2] + 2, 2 * lo[1]:2 * hi[1] + 2, 2 * lo[0]:2 * hi[0] + 2]
        a[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = f.reshape(m[2], 2, m[1], 2, m[0], 2).mean(axis=(1, 3, 5))
        for d in range(3):
            ax = 2 - d  # array axis of direction d (arrays are (z, y, x))
            fs = [slice(2 * lo[2], 2 * hi[2] + 2), slice(2 * lo[1], 2 * hi[1] + 2), slice(2 * lo[0], 2 * hi[0] + 2)]
            fs[ax] = slice(2 * lo[d], 2 * hi[d] + 3, 2)
            f = fb[d][tuple(fs)]
            shp = []
            for i in range(3):
                shp += [f.shape[i], 1] if i == ax else [f.shape[i] // 2, 2]
            r = f.reshape(shp).mean(axis=(1, 3, 5))
            cs = [slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1)]
            cs[ax] = slice(lo[d], hi[d] + 2)
            b[d][tuple(cs)] = r
            if periodic[d]:
                for src, dst in ((0, nc), (nc, 0)):
                    if lo[d] <= src <= hi[d] + 1:
                        k = src - lo[d]
                        ss = [slice(None)] * 3
                        ss[ax] = k
                        ds = list(cs)
                        ds[ax] = dst
                        b[d][tuple(ds)] = r[tuple(ss)]
    return (dict(a=a, bx=b[0], by=b[1], bz=b[2]), c1)


def lin_apply(u, a, b, alpha, beta, dxinv):
    """7-point ABecLaplacian alpha a u - beta div(b grad u) on the interior of a
    ghost-grown array u (textbook form, mlabeclap_adotx)."""
    c = u[1:-1, 1:-1, 1:-1]
    dh = beta * dxinv * dxinv
    y = alpha * a * c
    y = y - dh * (b[0][:, :, 1:] * (u[1:-1, 1:-1, 2:] - c) - b[0][:, :, :-1] * (c - u[1:-1, 1:-1, :-2]))
    y = y - dh * (b[1][:, 1:, :] * (u[1:-1, 2:, 1:-1] - c) - b[1][:, :-1, :] * (c - u[1:-1, :-2, 1:-1]))
    y = y - dh * (b[2][1:, :, :] * (u[2:, 1:-1, 1:-1] - c) - b[2][:-1, :, :] * (c - u[:-2, 1:-1, 1:-1]))
    return y
