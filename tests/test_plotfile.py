"""The AMReX plotfile writer (alamo_write_plotfile(_host); Integrator::WritePlotFile's nodal output,
Integrator.cpp:808-1054, through amrex::WriteMultiLevelPlotfile) read back by an independent parser
of the HyperCLaw-V1.1 layout: the top Header (version, names, dimension, time, finest level,
problem extent, refinement ratios, level domains, steps, cell sizes, coordinates, per-level box
extents and multifab paths), each level's VisMF header (version, layout, components, ghost width,
nodal box array, data file offsets, per-box minima and maxima), the FAB headers (little-endian
IEEE double, nodal boxes) and the data bit for bit, and the Checkpoint box arrays.  AMReX and yt
are absent here, so agreement with their readers is unpinned beyond this layout; the host variant
needs no GPU."""
import os
import re

import numpy as np
import pytest

from alamo_amd.amr import write_plotfile
from alamo_amd._lib import AlamoError

BOX = re.compile(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \((\d+),(\d+),(\d+)\)\)")


def parse_boxes(text):
    return [tuple(int(x) for x in m.groups()) for m in BOX.finditer(text)]


def read_header(path):
    L = open(os.path.join(path, "Header")).read().split("\n")
    h = {"version": L[0], "ncomp": int(L[1])}
    i = 2
    h["names"] = L[i:i + h["ncomp"]]
    i += h["ncomp"]
    h["dim"] = int(L[i]); h["time"] = float(L[i + 1]); h["finest"] = int(L[i + 2])
    h["prob_lo"] = [float(x) for x in L[i + 3].split()]
    h["prob_hi"] = [float(x) for x in L[i + 4].split()]
    h["ref"] = [int(x) for x in L[i + 5].split()]
    h["domains"] = parse_boxes(L[i + 6])
    h["steps"] = [int(x) for x in L[i + 7].split()]
    i += 8
    nl = h["finest"] + 1
    h["dx"] = [[float(x) for x in L[i + l].split()] for l in range(nl)]
    i += nl
    h["coord"] = int(L[i]); h["bwidth"] = int(L[i + 1])
    i += 2
    h["levels"] = []
    for l in range(nl):
        lev, nb, t = L[i].split()
        step = int(L[i + 1])
        i += 2
        ext = []
        for b in range(int(nb)):
            ext.append([tuple(float(x) for x in L[i + d].split()) for d in range(3)])
            i += 3
        h["levels"].append({"level": int(lev), "nbox": int(nb), "time": float(t), "step": step, "extents": ext,
                            "path": L[i]})
        i += 1
    return h


def read_vismf(path, level):
    d = os.path.join(path, "Level_%d" % level)
    L = open(os.path.join(d, "Cell_H")).read().split("\n")
    v = {"vers": int(L[0]), "how": int(L[1]), "ncomp": int(L[2]), "ngrow": int(L[3])}
    nb, zero = L[4].lstrip("(").split()
    nb = int(nb)
    v["boxes"] = parse_boxes("\n".join(L[5:5 + nb]))
    assert L[5 + nb] == ")" and int(zero) == 0
    i = 6 + nb
    assert int(L[i]) == nb
    v["fod"] = []
    for b in range(nb):
        tag, name, off = L[i + 1 + b].split()
        assert tag == "FabOnDisk:"
        v["fod"].append((name, int(off)))
    i += 1 + nb
    assert L[i] == ""
    mins, maxs = [], []
    for dest in (mins, maxs):
        i += 1
        n, nc = (int(x) for x in L[i].split(","))
        assert n == nb and nc == v["ncomp"]
        for b in range(nb):
            vals = L[i + 1 + b].split(",")
            assert vals[-1] == ""
            dest.append([float(x) for x in vals[:-1]])
        i += nb
        if dest is mins:
            i += 1
            assert L[i] == ""
    v["min"], v["max"] = mins, maxs
    # FABs
    v["fabs"] = []
    for (name, off), box in zip(v["fod"], v["boxes"]):
        with open(os.path.join(d, name), "rb") as f:
            f.seek(off)
            line = f.readline().decode()
            m = re.match(r"FAB \(\(8, \(64 11 52 0 1 12 0 1023\)\),\(8, \(1 2 3 4 5 6 7 8\)\)\)(.*) (\d+)\n$", line)
            assert m, line
            fbox = parse_boxes(m.group(1))[0]
            assert fbox == box
            nc = int(m.group(2))
            shape = [fbox[3 + q] - fbox[q] + 1 for q in range(3)]
            data = np.frombuffer(f.read(8 * nc * int(np.prod(shape))), dtype="<f8")
            v["fabs"].append((fbox, data.reshape(nc, shape[2], shape[1], shape[0])))
    return v


def _levels(n0, nl, ncomp, seed):
    r = np.random.default_rng(seed)
    return [r.standard_normal(ncomp * ((n0 << l) + 1) ** 3) for l in range(nl)]


def test_two_level_plotfile_round_trip(tmp_path):
    n0 = 8
    boxes = [[((4, 4, 4), (9, 11, 7)), ((10, 0, 8), (15, 5, 15))]]
    names = ["disp001", "disp002", "disp003", "eta"]
    f = _levels(n0, 2, len(names), 3)
    p = tmp_path / "plt00010node"
    write_plotfile(p, (n0,) * 3, (0.0, 0.0, 0.0), (1.0, 2.0, 1.0), boxes, names, f, time=0.25, steps=[10, 20])
    h = read_header(p)
    assert h["version"] == "HyperCLaw-V1.1" and h["names"] == names and h["dim"] == 3
    assert h["time"] == 0.25 and h["finest"] == 1 and h["ref"] == [2]
    assert h["prob_lo"] == [0.0, 0.0, 0.0] and h["prob_hi"] == [1.0, 2.0, 1.0]
    assert h["domains"] == [(0, 0, 0, 7, 7, 7, 0, 0, 0), (0, 0, 0, 15, 15, 15, 0, 0, 0)]
    assert h["steps"] == [10, 20] and h["coord"] == 0 and h["bwidth"] == 0
    assert h["dx"] == [[1 / 8, 2 / 8, 1 / 8], [1 / 16, 2 / 16, 1 / 16]]
    assert [lv["path"] for lv in h["levels"]] == ["Level_0/Cell", "Level_1/Cell"]
    # level 1 box extents: lo dx .. (hi + 1) dx
    assert h["levels"][1]["extents"][0] == [(4 / 16, 10 / 16), (8 / 16, 24 / 16), (4 / 16, 8 / 16)]
    for l, bl in enumerate([[((0, 0, 0), (7, 7, 7))]] + boxes):
        v = read_vismf(p, l)
        assert v["vers"] == 1 and v["ncomp"] == 4 and v["ngrow"] == 0
        assert v["boxes"] == [tuple(b[0]) + tuple(x + 1 for x in b[1]) + (1, 1, 1) for b in bl]
        N = (n0 << l) + 1
        full = f[l].reshape(4, N, N, N)
        for (box, data), mn, mx in zip(v["fabs"], v["min"], v["max"]):
            ref = full[:, box[2]:box[5] + 1, box[1]:box[4] + 1, box[0]:box[3] + 1]
            np.testing.assert_array_equal(data, ref)
            assert mn == list(ref.reshape(4, -1).min(axis=1)) and mx == list(ref.reshape(4, -1).max(axis=1))
    chk = open(p / "Checkpoint").read()
    assert parse_boxes(chk) == [(0, 0, 0, 7, 7, 7, 0, 0, 0), (4, 4, 4, 9, 11, 7, 0, 0, 0), (10, 0, 8, 15, 5, 15, 0, 0, 0)]
    assert chk.startswith("(1 0\n") and "(2 0\n" in chk


def test_plotfile_rejects_bad_boxes(tmp_path):
    f = _levels(4, 2, 1, 1)
    with pytest.raises(AlamoError, match="leaves the level's domain"):
        write_plotfile(tmp_path / "p", (4, 4, 4), (0, 0, 0), (1, 1, 1), [[((0, 0, 0), (8, 1, 1))]], ["u"], f)
    with pytest.raises(ValueError, match="fewer than"):
        write_plotfile(tmp_path / "p", (4, 4, 4), (0, 0, 0), (1, 1, 1), [[((0, 0, 0), (1, 1, 1))]], ["u", "v"], f)


@pytest.mark.gpu
def test_device_fields_write_the_same_files(cuda, tmp_path):
    """alamo_write_plotfile (device fields) writes byte-identical files to the host variant"""
    import torch
    n0 = 8
    boxes = [[((2, 2, 2), (13, 13, 5))]]
    f = _levels(n0, 2, 3, 5)
    write_plotfile(tmp_path / "h", (n0,) * 3, (0, 0, 0), (1, 1, 1), boxes, ["ux", "uy", "uz"], f, time=1.5)
    d = [torch.from_numpy(x).to(cuda) for x in f]
    write_plotfile(tmp_path / "d", (n0,) * 3, (0, 0, 0), (1, 1, 1), boxes, ["ux", "uy", "uz"], d, time=1.5)
    for rel in ("Header", "Checkpoint", "Level_0/Cell_H", "Level_0/Cell_D_00000", "Level_1/Cell_H",
                "Level_1/Cell_D_00000"):
        assert (tmp_path / "h" / rel).read_bytes() == (tmp_path / "d" / rel).read_bytes(), rel
