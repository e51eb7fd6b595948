"""CPU: the VTK plotters' file layouts (the reference's
vtk|vtu::Cartesian|Legendre|Lobatto::vertices|cells::ascii|binary family,
ExaHyPE/exahype/plotters/VTK/ADERDG2{Cartesian,Legendre,Lobatto}VTK.cpp).
Pinned by exact polynomial reproduction: a solution that is a polynomial of
degree <= order in each direction is written at every vertex (or sub-cell
centre) of every patch as that polynomial's value there, and by the patch
geometry (order^dim axis-aligned sub-cells per cell, vertices at the node
family's positions, Lobatto nodes against their closed forms).  The files are
read back by independent readers of the legacy ASCII / big-endian binary and
the XML (ascii / base64 with UInt32 headers) formats.
"""
import base64
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import plotters, problems as P


def read_vtk(path):
    lines = open(path).read().split("\n")
    out, i = {}, 0
    while i < len(lines):
        tok = lines[i].split()
        if tok and tok[0] == "POINTS":
            m = int(tok[1])
            out["points"] = np.array([list(map(float, l.split())) for l in lines[i + 1:i + 1 + m]])
            i += m
        elif tok and tok[0] == "CELLS":
            m = int(tok[1])
            out["cells"] = np.array([list(map(int, l.split())) for l in lines[i + 1:i + 1 + m]])
            i += m
        elif tok and tok[0] == "CELL_TYPES":
            m = int(tok[1])
            out["types"] = np.array([int(l) for l in lines[i + 1:i + 1 + m]])
            i += m
        elif tok and tok[0] in ("Q", "time") and len(tok) == 4:
            ncomp, m = int(tok[1]), int(tok[2])
            out[tok[0]] = np.array([list(map(float, l.split())) for l in lines[i + 1:i + 1 + m]]).reshape(m, ncomp)
            i += m
        i += 1
    return out


@pytest.mark.parametrize("D,order,n", [(3, 3, 2), (3, 2, 3), (2, 4, 3), (2, 1, 4)])
def test_cartesian_vtk_reproduces_polynomials(D, order, n, tmp_path):
    rng = np.random.default_rng(10 * D + order)
    B, nd, h = order + 1, 5, 1.0 / n
    coef = rng.normal(size=(nd,) + (B,) * D)

    def poly(x):  # x[..., D] -> [..., nd]
        val = np.zeros(x.shape[:-1] + (nd,))
        for idx in np.ndindex(*(B,) * D):
            mono = np.ones(x.shape[:-1])
            for d in range(D):
                mono = mono * x[..., d] ** idx[d]
            val += mono[..., None] * coef[(slice(None),) + idx]
        return val

    luh = poly(P.node_points(n, order, D))  # [cells][nodes][nd]
    pl = plotters.CartesianVTK(str(tmp_path / "sol"), 4, order, D, E.tables(order)["x"])
    path = pl.write(luh.ravel(), n, h, nd, 0.25)
    assert path.endswith("sol-0.vtk")
    f = read_vtk(path)
    npts = n ** D * B ** D
    assert f["points"].shape == (npts, 3) and f["Q"].shape == (npts, 4)
    np.testing.assert_array_equal(f["time"], 0.25)
    exact = poly(f["points"][:, :D])[:, :4]
    assert np.abs(f["Q"] - exact).max() <= 1e-11 * max(1.0, np.abs(exact).max())
    # geometry: order^D sub-cells per patch, each an axis-aligned box of edge h / order
    ncorner = 8 if D == 3 else 4
    assert f["cells"].shape == (n ** D * order ** D, ncorner + 1)
    assert np.all(f["cells"][:, 0] == ncorner)
    np.testing.assert_array_equal(f["types"], 12 if D == 3 else 9)
    for c in f["cells"][:: max(1, len(f["cells"]) // 7)]:
        q = f["points"][c[1:]]
        ext = q.max(0) - q.min(0)
        np.testing.assert_allclose(ext[:D], h / order, rtol=1e-12)
        np.testing.assert_allclose(q[0] + np.array([h / order, 0, 0]), q[1], rtol=0, atol=1e-14)


@pytest.mark.parametrize("D,order,n", [(3, 3, 2), (2, 4, 3)])
def test_legendre_vtk_writes_the_nodes(D, order, n, tmp_path):
    """Vertices at the Gauss-Legendre nodes (ADERDG2LegendreVTK.cpp:296), values
    the nodal values, in the solution's node order (x fastest)."""
    rng = np.random.default_rng(D + order)
    B, nd, h = order + 1, 5, 1.0 / n
    luh = rng.normal(size=(n ** D, B ** D, nd))
    pl = plotters.LegendreVTK(str(tmp_path / "leg"), nd, order, D, E.tables(order)["x"])
    f = read_vtk(pl.write(luh.ravel(), n, h, nd, 1.5))
    np.testing.assert_allclose(f["points"][:, :D], P.node_points(n, order, D).reshape(-1, D), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(f["Q"], luh.reshape(-1, nd))
    assert f["cells"].shape[0] == n ** D * order ** D


# ---- the whole vtk|vtu x Cartesian|Legendre|Lobatto x vertices|cells x ascii|binary family ----

def read_legacy_binary(path):
    raw = open(path, "rb").read()
    out, pos = {}, 0

    def line():
        nonlocal pos
        e = raw.index(b"\n", pos)
        s = raw[pos:e].decode("ascii")
        pos = e + 1
        return s

    def block(count, dt):
        nonlocal pos
        nb = count * np.dtype(dt).itemsize
        a = np.frombuffer(raw[pos:pos + nb], dtype=dt)
        pos += nb + 1  # trailing newline
        return a

    assert line().startswith("# vtk DataFile")
    line()
    assert line() == "BINARY"
    while pos < len(raw):
        tok = line().split()
        if not tok:
            continue
        if tok[0] == "POINTS":
            out["points"] = block(3 * int(tok[1]), ">f8").reshape(-1, 3)
        elif tok[0] == "CELLS":
            out["cells"] = block(int(tok[2]), ">i4").reshape(int(tok[1]), -1)
        elif tok[0] == "CELL_TYPES":
            out["types"] = block(int(tok[1]), ">i4")
        elif tok[0] in ("POINT_DATA", "CELL_DATA"):
            out["where"] = tok[0]
        elif tok[0] in ("Q", "time") and len(tok) == 4:
            out[tok[0]] = block(int(tok[1]) * int(tok[2]), ">f8").reshape(int(tok[2]), int(tok[1]))
    return out


def read_vtu(path):
    root = ET.parse(path).getroot()
    assert root.get("type") == "UnstructuredGrid" and root.get("header_type") == "UInt32"
    piece = root.find("UnstructuredGrid/Piece")

    def arr(el):
        dt = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}[el.get("type")]
        nc = int(el.get("NumberOfComponents", "1"))
        if el.get("format") == "binary":
            b = base64.b64decode(el.text.strip())
            nbytes = int(np.frombuffer(b[:4], dtype="<u4")[0])
            a = np.frombuffer(b[4:4 + nbytes], dtype=dt)
            assert len(b) == 4 + nbytes
        else:
            a = np.array(el.text.split(), dtype=np.float64).astype(dt)
        return a.reshape(-1, nc) if nc > 1 else a

    out = {"points": arr(piece.find("Points/DataArray"))}
    cells = {el.get("Name"): arr(el) for el in piece.find("Cells")}
    k = cells["offsets"][0]
    out["cells"] = np.hstack([np.full((len(cells["offsets"]), 1), k), cells["connectivity"].reshape(-1, k)])
    out["types"] = cells["types"]
    for where in ("PointData", "CellData"):
        node = piece.find(where)
        if node is not None:
            out["where"] = "POINT_DATA" if where == "PointData" else "CELL_DATA"
            for el in node:
                a = arr(el)
                out[el.get("Name")] = a.reshape(len(a), -1)
    assert int(piece.get("NumberOfPoints")) == len(out["points"])
    assert int(piece.get("NumberOfCells")) == len(out["cells"])
    return out


def read_any(path, plotter):
    if plotter.container == "vtu":
        return read_vtu(path)
    if plotter.encoding == "binary":
        return read_legacy_binary(path)
    f = read_vtk(path)
    txt = open(path).read()
    f["where"] = "CELL_DATA" if "CELL_DATA" in txt else "POINT_DATA"
    return f


def poly_state(D, order, n, nd, seed):
    rng = np.random.default_rng(seed)
    B = order + 1
    coef = rng.normal(size=(nd,) + (B,) * D)

    def poly(x):
        val = np.zeros(x.shape[:-1] + (nd,))
        for idx in np.ndindex(*(B,) * D):
            mono = np.ones(x.shape[:-1])
            for d in range(D):
                mono = mono * x[..., d] ** idx[d]
            val += mono[..., None] * coef[(slice(None),) + idx]
        return val

    return poly(P.node_points(n, order, D)), poly


IDENTS = ["%s::%s::%s::%s" % (c, k, w, e) for c in ("vtk", "vtu") for k in ("Cartesian", "Legendre", "Lobatto")
          for w in ("vertices", "cells") for e in ("ascii", "binary")]


@pytest.mark.parametrize("ident", IDENTS)
@pytest.mark.parametrize("D,order,n", [(3, 2, 2), (2, 3, 3)])
def test_every_vtk_plotter_reproduces_polynomials(ident, D, order, n, tmp_path):
    """Vertex mode: the cell polynomial at the patch vertices; cell mode: at the
    sub-cell centres (ADERDG2CartesianVTK.cpp:344-390, ADERDG2LegendreVTK.cpp:427,
    ADERDG2LobattoVTK.cpp:424-426).  A polynomial of degree <= order per direction
    is reproduced exactly, so the values equal the polynomial at the written
    positions (vertices) or at the centroid of each sub-cell's corners."""
    nd, h = 5, 1.0 / n
    luh, poly = poly_state(D, order, n, nd, seed=len(ident) + 7 * D)
    pl = plotters.PatchVTK.from_identifier(ident, str(tmp_path / "sol"), 3, order, D, E.tables(order)["x"])
    assert pl.identifier == ident
    path = pl.write(luh.ravel(), n, h, nd, 0.75)
    assert path.endswith("sol-0." + pl.container)
    f = read_any(path, pl)
    B = order + 1
    npts, ncells = n ** D * B ** D, n ** D * order ** D
    ncorner = 8 if D == 3 else 4
    assert f["points"].shape == (npts, 3) and f["cells"].shape == (ncells, ncorner + 1)
    np.testing.assert_array_equal(f["types"], 12 if D == 3 else 9)
    # vertex positions within each patch follow the node family
    fr = {"Cartesian": np.arange(B) / order, "Legendre": np.array(E.tables(order)["x"]),
          "Lobatto": plotters.lobatto_nodes(order)}[pl.kind]
    np.testing.assert_allclose(np.unique(np.round(f["points"][:, 0] / h - np.floor(f["points"][:, 0] / h + 1e-9), 12)),
                               np.unique(np.round(np.where(fr >= 1 - 1e-12, 0.0, fr), 12)), atol=1e-12)
    if pl.cells:
        assert f["where"] == "CELL_DATA" and f["Q"].shape == (ncells, 3)
        centre = f["points"][f["cells"][:, 1:]].mean(1)[:, :D]
        exact = poly(centre)[:, :3]
    else:
        assert f["where"] == "POINT_DATA" and f["Q"].shape == (npts, 3)
        exact = poly(f["points"][:, :D])[:, :3]
    assert np.abs(f["Q"] - exact).max() <= 1e-11 * max(1.0, np.abs(exact).max())
    np.testing.assert_array_equal(f["time"].ravel(), 0.75)


def test_vtu_time_series_and_mapping(tmp_path):
    D, order, n, nd = 2, 2, 3, 4
    luh, poly = poly_state(D, order, n, nd, seed=3)
    pl = plotters.LobattoVTK(str(tmp_path / "ts"), 2, order, D, E.tables(order)["x"], container="vtu",
                             encoding="binary", mapping=lambda x, t, Q: np.stack([Q[:, 0] * t, x[:, 0]], -1))
    for k, t in enumerate((0.0, 0.5, 1.25)):
        pl.write(luh.ravel(), n, 1.0 / n, nd, t)
    root = ET.parse(str(tmp_path / "ts.pvd")).getroot()
    rows = root.findall("Collection/DataSet")
    assert [r.get("file") for r in rows] == ["ts-0.vtu", "ts-1.vtu", "ts-2.vtu"]
    assert [float(r.get("timestep")) for r in rows] == [0.0, 0.5, 1.25]
    f = read_vtu(str(tmp_path / "ts-2.vtu"))
    exact = poly(f["points"][:, :D])
    np.testing.assert_allclose(f["Q"][:, 0], exact[:, 0] * 1.25, rtol=0, atol=1e-11 * np.abs(exact).max())
    np.testing.assert_array_equal(f["Q"][:, 1], f["points"][:, 0])


def test_lobatto_nodes_closed_forms():
    np.testing.assert_array_equal(plotters.lobatto_nodes(1), [0.0, 1.0])
    np.testing.assert_allclose(plotters.lobatto_nodes(2), [0.0, 0.5, 1.0], atol=1e-15)
    r5 = 1 / np.sqrt(5.0)  # order 3: interior nodes (1 -+ 1/sqrt(5)) / 2
    np.testing.assert_allclose(plotters.lobatto_nodes(3), [0.0, (1 - r5) / 2, (1 + r5) / 2, 1.0], atol=1e-15)
    r = np.sqrt(3.0 / 7.0)  # order 4: 0, (1 -+ sqrt(3/7)) / 2, 1/2
    np.testing.assert_allclose(plotters.lobatto_nodes(4), [0.0, (1 - r) / 2, 0.5, (1 + r) / 2, 1.0], atol=1e-15)


def test_cartesian_resolution_subpatches(tmp_path):
    """resolution r: 3^r sub-patches per direction (ADERDG2CartesianVTK.cpp:402-431),
    each an equidistant patch of the cell's polynomial."""
    D, order, n, nd = 2, 2, 2, 3
    luh, poly = poly_state(D, order, n, nd, seed=5)
    pl = plotters.CartesianVTK(str(tmp_path / "res"), nd, order, D, E.tables(order)["x"], resolution=1)
    f = read_vtk(pl.write(luh.ravel(), n, 0.5, nd, 0.0))
    B = order + 1
    assert f["points"].shape[0] == n ** D * 9 * B ** D
    assert f["cells"].shape[0] == n ** D * 9 * order ** D
    exact = poly(f["points"][:, :D])
    assert np.abs(f["Q"] - exact).max() <= 1e-11 * np.abs(exact).max()
    q = f["points"][f["cells"][0, 1:]]
    np.testing.assert_allclose((q.max(0) - q.min(0))[:D], 0.5 / 3 / order, rtol=1e-12)
    with pytest.raises(ValueError):
        plotters.LegendreVTK("x", 1, order, D, E.tables(order)["x"], resolution=1)
    with pytest.raises(ValueError):
        plotters.PatchVTK.from_identifier("vtk::Carpet::cells::ascii", "x", 1, order, D, E.tables(order)["x"])
