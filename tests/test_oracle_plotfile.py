"""Pins of the oracle's VisMF and plotfile restatement (CPU).

The reference ships no plotfile or VisMF files, so the layouts are pinned by
known answers spelled out here from the reference's writers (not from the
oracle): the text of a VisMF header and of a plotfile Header for tiny cases
(scientific notation with 17 digits for the per-box extrema, precision-17
default notation in the plotfile Header, blank lines after the FabOnDisk and
extrema blocks, trailing spaces on the number rows); the binary FAB layout
decoded independently (the FAB header line, then the components as IEEE
doubles, x fastest); and round trips.
"""
import os

import numpy as np

from oracle import plotfile as pf

RD = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"


def _two_boxes(ngrow, ncomp=1, seed=0):
    boxes = [(0, 0, 0, 1, 1, 0), (2, 0, 0, 3, 1, 0)]
    rng = np.random.default_rng(seed)
    fabs = []
    for b in boxes:
        shp = (ncomp, b[5] - b[2] + 1 + 2 * ngrow, b[4] - b[1] + 1 + 2 * ngrow, b[3] - b[0] + 1 + 2 * ngrow)
        fabs.append(rng.standard_normal(shp))
    return boxes, fabs


def test_vismf_header_known_answer(tmp_path):
    boxes, fabs = _two_boxes(1)
    # valid cells of box 0: 1.5 .. 2.25, box 1: -3 .. 0.5 (ghosts are not counted)
    fabs[0][0, 1, 1:3, 1:3] = [[1.5, 2.0], [2.25, 1.75]]
    fabs[1][0, 1, 1:3, 1:3] = [[-3.0, 0.5], [0.0, 0.25]]
    fabs[0][0, 0] = 99.0  # a ghost plane
    name = str(tmp_path / "t")
    nbytes = pf.vismf_write(name, boxes, fabs, 1)
    hdr0 = "FAB " + RD + "((-1,-1,-1) (2,2,1) (0,0,0)) 1\n"
    off1 = len(hdr0) + 48 * 8
    expect = ("1\n1\n1\n1\n(2 0\n((0,0,0) (1,1,0) (0,0,0))\n((2,0,0) (3,1,0) (0,0,0))\n)\n2\n"
              "FabOnDisk: t_D_00000 0\nFabOnDisk: t_D_00000 %d\n\n"
              "2,1\n1.50000000000000000e+00,\n-3.00000000000000000e+00,\n\n"
              "2,1\n2.25000000000000000e+00,\n5.00000000000000000e-01,\n\n" % off1)
    text = open(name + "_H").read()
    assert text == expect
    hdr1 = "FAB " + RD + "((1,-1,-1) (4,2,1) (0,0,0)) 1\n"
    assert nbytes == len(hdr0) + len(hdr1) + 2 * 48 * 8 + len(expect)
    # the data file: FAB header line, then the doubles x fastest
    raw = open(name + "_D_00000", "rb").read()
    assert raw[:len(hdr0)].decode() == hdr0
    a0 = np.frombuffer(raw[len(hdr0):len(hdr0) + 48 * 8], dtype="<f8").reshape(3, 4, 4)
    np.testing.assert_array_equal(a0, fabs[0][0])
    assert raw[off1:off1 + len(hdr1)].decode() == hdr1
    a1 = np.frombuffer(raw[off1 + len(hdr1):], dtype="<f8").reshape(3, 4, 4)
    np.testing.assert_array_equal(a1, fabs[1][0])


def test_vismf_round_trip_components_and_index_types(tmp_path):
    for ngrow, ncomp, itype in ((0, 1, (0, 0, 0)), (2, 3, (0, 0, 0)), (1, 2, (0, 1, 0))):
        boxes, fabs = _two_boxes(ngrow, ncomp, seed=ngrow + ncomp)
        if itype[1]:  # nodal in y: one more point along y
            boxes = [(b[0], b[1], b[2], b[3], b[4] + 1, b[5]) for b in boxes]
            fabs = [np.random.default_rng(7).standard_normal(
                (ncomp, b[5] - b[2] + 1 + 2 * ngrow, b[4] - b[1] + 1 + 2 * ngrow, b[3] - b[0] + 1 + 2 * ngrow))
                for b in boxes]
        name = str(tmp_path / ("m%d%d" % (ngrow, ncomp)))
        pf.vismf_write(name, boxes, fabs, ngrow, itype=itype)
        h, back = pf.vismf_read(name)
        assert (h["version"], h["how"], h["ncomp"], h["ngrow"]) == (1, 1, ncomp, ngrow)
        assert h["boxes"] == boxes and all(t == itype for t in h["types"])
        for a, b in zip(back, fabs):
            np.testing.assert_array_equal(a, b)
        g = ngrow
        for i, a in enumerate(fabs):
            v = a[:, g:a.shape[1] - g, g:a.shape[2] - g, g:a.shape[3] - g] if g else a
            assert h["min"][i] == [v[c].min() for c in range(ncomp)]
            assert h["max"][i] == [v[c].max() for c in range(ncomp)]


def test_plotfile_header_known_answer(tmp_path):
    boxes, fabs = _two_boxes(0, 2)
    geom = dict(domain=((0, 0, 0), (3, 1, 0)), prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 1.0),
                dx=(0.25, 0.5, 1.0), coord=0)
    name = str(tmp_path / "plt00007")
    pf.plotfile_write(name, [(boxes, fabs)], ["phi", "rhs"], [geom], time=0.5, level_steps=[7])
    expect = ("HyperCLaw-V1.1\n2\nphi\nrhs\n3\n0.5\n0\n0 0 0 \n1 1 1 \n\n((0,0,0) (3,1,0) (0,0,0)) \n7 \n"
              "0.25 0.5 1 \n0\n0\n0 2 0.5\n7\n0 0.5\n0 1\n0 1\n0.5 1\n0 1\n0 1\nLevel_0/Cell\n")
    assert open(os.path.join(name, "Header")).read() == expect
    r = pf.plotfile_read(name)
    assert r["varnames"] == ["phi", "rhs"] and r["time"] == 0.5 and r["level_steps"] == [7]
    assert r["domains"] == [(0, 0, 0, 3, 1, 0)] and r["dx"] == [[0.25, 0.5, 1.0]]
    h, back = r["levels"][0]
    for a, b in zip(back, fabs):
        np.testing.assert_array_equal(a, b)


def test_two_level_plotfile_round_trip(tmp_path):
    c_boxes, c_fabs = _two_boxes(0, 1, seed=3)
    f_boxes = [(2, 0, 0, 5, 3, 1)]
    f_fabs = [np.random.default_rng(4).standard_normal((1, 2, 4, 4))]
    geoms = [dict(domain=((0, 0, 0), (3, 1, 0)), prob_lo=(-1.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 0.5),
                  dx=(0.5, 0.5, 0.5), coord=0),
             dict(domain=((0, 0, 0), (7, 3, 1)), prob_lo=(-1.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 0.5),
                  dx=(0.25, 0.25, 0.25), coord=0)]
    name = str(tmp_path / "plt2")
    pf.plotfile_write(name, [(c_boxes, c_fabs), (f_boxes, f_fabs)], ["u"], geoms, time=1.25, level_steps=[3, 6])
    lines = open(os.path.join(name, "Header")).read().split("\n")
    assert lines[5] == "1" and lines[6] == "-1 0 0 " and lines[8] == "2 "
    assert lines[9] == "((0,0,0) (3,1,0) (0,0,0)) ((0,0,0) (7,3,1) (0,0,0)) "
    r = pf.plotfile_read(name)
    assert r["finest_level"] == 1 and r["ref_ratio"] == [2]
    np.testing.assert_array_equal(r["levels"][1][1][0], f_fabs[0])
    # the fine grid's physical extent: x from -1 + 2*0.25 to -1 + 6*0.25
    i = lines.index("1 1 1.25")
    assert lines[i + 2] == "-0.5 0.5" and lines[i + 3] == "0 1" and lines[i + 4] == "0 0.5"


def test_set_ghost_known_answer(tmp_path):
    # one box 2x1x1 grown by 1: valid cells 3 and -1 in component 0, 10 and 10 in component 1;
    # every ghost cell becomes (3 + -1)/2 = 1 and 10, the valid cells stay, the header min/max
    # are those of the valid cells
    a = np.full((2, 3, 3, 4), 99.0)
    a[0, 1, 1, 1:3] = (3.0, -1.0)
    a[1, 1, 1, 1:3] = (10.0, 10.0)
    (g,) = pf.set_ghost_cells([a], 1)
    assert a[0, 0, 0, 0] == 99.0  # the input is not changed
    ghost = np.ones(g.shape[1:], bool)
    ghost[1, 1, 1:3] = False
    assert np.all(g[0][ghost] == 1.0) and np.all(g[1][ghost] == 10.0)
    assert list(g[0, 1, 1, 1:3]) == [3.0, -1.0] and list(g[1, 1, 1, 1:3]) == [10.0, 10.0]
    name = str(tmp_path / "sg")
    pf.vismf_write(name, [(0, 0, 0, 1, 0, 0)], [a], 1, set_ghost=True)
    h, fabs = pf.vismf_read(name)
    np.testing.assert_array_equal(fabs[0], g)
    assert h["min"] == [[-1.0, 10.0]] and h["max"] == [[3.0, 10.0]]
