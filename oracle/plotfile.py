"""Oracle (test infrastructure only): AMReX's on-disk formats restated in numpy.

Only tests use this module; it checks the library's VisMF and plotfile
writers and readers, it is not one of them.

* VisMF, header version 1 (VisMF::Write / VisMF::Header operator<<,
  AMReX_VisMF.cpp:126-330, 934-1126): <name>_H and one data file
  <name>_D_00000 holding, per box, the FAB header line of FABio_binary
  (AMReX_FArrayBox.cpp:95-105, 905-912) and the FAB's components (x fastest,
  components one after another) as IEEE doubles in the machine's byte order
  (FPC::NativeRealDescriptor, AMReX_FPC.cpp:61-110; RealDescriptor output,
  AMReX_FabConv.cpp:864-897).  The header holds the BoxArray
  (BoxArray::writeOn, AMReX_BoxArray.cpp:475-505), FabOnDisk entries with file
  offsets and per-box minima and maxima of the valid cells in scientific
  notation with 17 digits.
* Plotfiles (WriteGenericPlotfileHeader / WriteMultiLevelPlotfile,
  AMReX_PlotFileUtil.cpp:73-248): <plt>/Header, <plt>/Level_<l>/Cell_H and
  Cell_D_00000, reals at precision 17 in the default float format.
* Readers mirror VisMF::Read and PlotFileDataImpl (AMReX_PlotFileDataImpl.cpp:16-103).

Boxes are (lo0, lo1, lo2, hi0, hi1, hi2) in the data's own index type with a
type vector; FAB arrays are numpy arrays (ncomp, z, y, x) over the box grown by
ngrow.
"""
import os
import re
import sys

import numpy as np

_RD = ("((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" if sys.byteorder == "little"
       else "((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))")


def box_str(b, t=(0, 0, 0)):
    return "((%d,%d,%d) (%d,%d,%d) (%d,%d,%d))" % (b[0], b[1], b[2], b[3], b[4], b[5], t[0], t[1], t[2])


def grow(b, g):
    return (b[0] - g, b[1] - g, b[2] - g, b[3] + g, b[4] + g, b[5] + g)


def fab_header(b, t, ncomp):
    return "FAB " + _RD + box_str(b, t) + " " + str(ncomp) + "\n"


def _sci(v):
    return "%.17e" % v


def set_ghost_cells(fabs, ngrow):
    """Copies of fabs with every ghost cell = (min + max)/2 of the box's valid
    cells, per component (VisMF::Write with set_ghost, AMReX_VisMF.cpp:945-970)."""
    out = []
    for a in fabs:
        a = np.array(a, dtype=np.float64)
        g = ngrow
        if g:
            for c in range(a.shape[0]):
                v = a[c, g:-g, g:-g, g:-g]
                val = (v.min() + v.max()) / 2.0
                keep = v.copy()
                a[c] = val
                a[c, g:-g, g:-g, g:-g] = keep
        out.append(a)
    return out


def vismf_write(name, boxes, fabs, ngrow, itype=(0, 0, 0), how=1, set_ghost=False):
    """Write fabs[i] (ncomp, z, y, x over boxes[i] grown by ngrow) as VisMF
    version 1; returns the byte count the reference reports."""
    if set_ghost:
        fabs = set_ghost_cells(fabs, ngrow)
    ncomp = fabs[0].shape[0] if fabs else 0
    dfile = name + "_D_00000"
    offsets, mins, maxs = [], [], []
    nbytes = 0
    with open(dfile, "wb") as f:
        for b, a in zip(boxes, fabs):
            a = np.ascontiguousarray(a, dtype=np.float64)
            fb = grow(b, ngrow)
            assert a.shape == (ncomp, fb[5] - fb[2] + 1, fb[4] - fb[1] + 1, fb[3] - fb[0] + 1)
            offsets.append(f.tell())
            h = fab_header(fb, itype, ncomp).encode()
            f.write(h)
            f.write(a.tobytes())
            nbytes += len(h) + a.nbytes
            g = ngrow
            v = a[:, g:a.shape[1] - g, g:a.shape[2] - g, g:a.shape[3] - g] if g else a
            mins.append([v[c].min() for c in range(ncomp)])
            maxs.append([v[c].max() for c in range(ncomp)])
    lines = ["1", str(how), str(ncomp), str(ngrow), "(%d 0" % len(boxes)]
    lines += [box_str(b, itype) for b in boxes]
    lines += [")", str(len(boxes))]
    lines += ["FabOnDisk: %s %d" % (os.path.basename(dfile), o) for o in offsets]
    lines += [""]
    for t in (mins, maxs):
        lines += ["%d,%d" % (len(boxes), ncomp)]
        lines += ["".join(_sci(x) + "," for x in row) for row in t]
        lines += [""]
    text = "\n".join(lines) + "\n"
    with open(name + "_H", "w") as f:
        f.write(text)
    return nbytes + len(text)


def _parse_box(tok):
    # "((a,b,c) (d,e,f) (t,u,v))"
    nums = tok.replace("(", " ").replace(")", " ").replace(",", " ").split()
    v = [int(x) for x in nums]
    return tuple(v[:6]), tuple(v[6:9])


def vismf_read_header(name):
    with open(name + "_H") as f:
        text = f.read()
    lines = text.split("\n")
    h = {"version": int(lines[0]), "how": int(lines[1]), "ncomp": int(lines[2]), "ngrow": int(lines[3])}
    nb = int(lines[4].strip("(").split()[0])
    boxes, types = [], []
    for i in range(nb):
        b, t = _parse_box(lines[5 + i])
        boxes.append(b)
        types.append(t)
    assert lines[5 + nb] == ")"
    assert int(lines[6 + nb]) == nb
    fod = []
    for i in range(nb):
        tag, fname, off = lines[7 + nb + i].split()
        assert tag == "FabOnDisk:"
        fod.append((fname, int(off)))
    k = 7 + 2 * nb
    assert lines[k] == ""
    mm = []
    k += 1
    for _ in range(2):
        n, m = (int(x) for x in lines[k].split(","))
        rows = [[float(x) for x in lines[k + 1 + i].rstrip(",").split(",")] if m else [] for i in range(n)]
        mm.append(rows)
        k += n + 2
    h.update(boxes=boxes, types=types, fod=fod, min=mm[0], max=mm[1], text=text)
    return h


def vismf_read(name):
    """(header, [fab arrays (ncomp, z, y, x) over the grown boxes])."""
    h = vismf_read_header(name)
    d = os.path.dirname(name)
    fabs = []
    for b, t, (fname, off) in zip(h["boxes"], h["types"], h["fod"]):
        with open(os.path.join(d, fname), "rb") as f:
            f.seek(off)
            line = f.readline().decode()
            fb = grow(b, h["ngrow"])
            assert line == fab_header(fb, t, h["ncomp"]), line
            shp = (h["ncomp"], fb[5] - fb[2] + 1, fb[4] - fb[1] + 1, fb[3] - fb[0] + 1)
            fabs.append(np.frombuffer(f.read(8 * int(np.prod(shp))), dtype=np.float64).reshape(shp).copy())
    return h, fabs


def _g(v):
    return "%.17g" % v


def plotfile_write(name, levels, varnames, geoms, time=0.0, level_steps=None, ref_ratio=None,
                   version="HyperCLaw-V1.1"):
    """levels[l] = (boxes, fabs (ncomp, z, y, x) of the valid cells); geoms[l] = dict(domain=(lo3, hi3),
    prob_lo, prob_hi, dx, coord).  Writes <name>/Header and <name>/Level_<l>/Cell."""
    nl = len(levels)
    level_steps = level_steps or [0] * nl
    ref_ratio = ref_ratio or [2] * (nl - 1)
    os.makedirs(name, exist_ok=True)
    out = [version, str(len(varnames))] + list(varnames)
    out += ["3", _g(time), str(nl - 1)]
    g0 = geoms[0]
    out += ["".join(_g(x) + " " for x in g0["prob_lo"])]
    out += ["".join(_g(x) + " " for x in g0["prob_hi"])]
    out += ["".join(str(r) + " " for r in ref_ratio[:nl - 1])]
    out += ["".join(box_str(tuple(g["domain"][0]) + tuple(g["domain"][1])) + " " for g in geoms[:nl])]
    out += ["".join(str(s) + " " for s in level_steps[:nl])]
    for g in geoms[:nl]:
        out += ["".join(_g(x) + " " for x in g["dx"])]
    out += [str(g0.get("coord", 0)), "0"]
    for l, (boxes, fabs) in enumerate(levels):
        g = geoms[l]
        dlo = g["domain"][0]
        out += ["%d %d %s" % (l, len(boxes), _g(time)), str(level_steps[l])]
        for b in boxes:
            for d in range(3):
                lo = g["prob_lo"][d] + g["dx"][d] * float(b[d] - dlo[d])
                hi = g["prob_lo"][d] + g["dx"][d] * float(b[3 + d] - dlo[d] + 1)
                out += [_g(lo) + " " + _g(hi)]
        out += ["Level_%d/Cell" % l]
        os.makedirs(os.path.join(name, "Level_%d" % l), exist_ok=True)
        vismf_write(os.path.join(name, "Level_%d" % l, "Cell"), boxes, fabs, 0)
    with open(os.path.join(name, "Header"), "w") as f:
        f.write("\n".join(out) + "\n")


def plotfile_read(name):
    """{"version", "varnames", "time", "finest_level", "prob_lo", "prob_hi", "ref_ratio", "domains",
    "level_steps", "dx", "coord", "levels": [(header, fabs)]}."""
    with open(os.path.join(name, "Header")) as f:
        lines = f.read().split("\n")
    r = {"version": lines[0]}
    nc = int(lines[1])
    r["varnames"] = [lines[2 + i].strip() for i in range(nc)]
    k = 2 + nc
    assert int(lines[k]) == 3
    r["time"] = float(lines[k + 1])
    fl = r["finest_level"] = int(lines[k + 2])
    r["prob_lo"] = [float(x) for x in lines[k + 3].split()]
    r["prob_hi"] = [float(x) for x in lines[k + 4].split()]
    r["ref_ratio"] = [int(x) for x in lines[k + 5].split()]
    doms = re.findall(r"\(\([^()]*\) \([^()]*\) \([^()]*\)\)", lines[k + 6])
    r["domains"] = [_parse_box(x)[0] for x in doms]
    r["level_steps"] = [int(x) for x in lines[k + 7].split()]
    r["dx"] = [[float(x) for x in lines[k + 8 + l].split()] for l in range(fl + 1)]
    k = k + 9 + fl
    r["coord"] = int(lines[k])
    k += 2
    r["levels"] = []
    for l in range(fl + 1):
        lev, ngrids, _ = lines[k].split()
        k += 2 + 3 * int(ngrids)
        rel = lines[k]
        k += 1
        r["levels"].append(vismf_read(os.path.join(name, rel)))
    return r
