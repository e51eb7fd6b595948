"""Host-side plotters over the device mesh.

ProbeAscii writes the reference's ``probe::ascii`` format
(ExaHyPE/exahype/plotters/ADERDG2ProbeAscii.cpp): a file ``<filename>.probe``
whose first line is ``# plot-time, real-time,Q0,Q1,...`` and which gets one
line per plot, ``<plot time>, <cell time stamp>, <value 0>, <value 1>, ...``,
numbers in the C++ stream default format (6 significant digits, as ``%g``).
The values are the solution interpolated at the probe point
(exa_mesh_probe), optionally passed through ``mapping(x, t, Q) -> values``
(the reference's mapQuantities post-processing; identity by default).

The VTK plotters write the reference's ``vtk|vtu::Cartesian|Legendre|Lobatto::
vertices|cells::ascii|binary`` family (ExaHyPE/exahype/plotters/VTK/
ADERDG2{Cartesian,Legendre,Lobatto}VTK.cpp) from the solution array: every cell
is a patch of (order + 1)^dim vertices and order^dim sub-cells; vertex mode
writes the cell polynomial at the vertices, cell mode at the sub-cell centres.
Containers: legacy VTK (``<filename>-<k>.vtk``, ASCII or big-endian binary)
and XML VTU (``<filename>-<k>.vtu``, ASCII or base64 binary, plus the time
series ``<filename>.pvd`` listing every snapshot with its time).
"""
import base64
import os
import re

import numpy as np


def _g(v):
    return "%g" % v


class ProbeAscii:
    def __init__(self, filename, x, written_unknowns, mapping=None):
        self.filename = filename + ".probe"
        self.x = np.asarray(x, dtype=np.float64)
        self.written = int(written_unknowns)
        self.mapping = mapping
        self._out = None

    def _open(self):
        if self._out is None:
            self._out = open(self.filename, "w")
            self._out.write("# plot-time, real-time" + "".join(",Q%d" % i for i in range(self.written)) + "\n")

    def plot(self, mesh, plot_time=None):
        """One line for the mesh's current state; plot_time defaults to the mesh time."""
        t, _ = mesh.time()
        Q = mesh.probe(self.x[None, :])[0]
        vals = self.mapping(self.x, t, Q) if self.mapping else Q[: self.written]
        self._open()
        self._out.write(_g(t if plot_time is None else plot_time) + ", " + _g(t)
                        + "".join(", " + _g(v) for v in vals) + "\n")
        self._out.flush()

    def close(self):
        if self._out is not None:
            self._out.close()
            self._out = None


def lobatto_nodes(order):
    """Gauss-Lobatto nodes on [0, 1], ascending: the end points and the roots of
    P'_order(2x - 1) (the reference stores them descending,
    kernels/generated/GaussLobattoBasis.csnippet, and reads nodes[order][order - i])."""
    order = int(order)
    if order < 1:
        raise ValueError("Lobatto nodes need order >= 1")
    inner = np.polynomial.legendre.Legendre.basis(order).deriv().roots() if order > 1 else np.array([])
    return np.concatenate([[0.0], np.sort((np.real(inner) + 1.0) / 2.0), [1.0]])


def _lagrange(nodes, pts):
    """M[i][j] = phi_j(pts[i]), the Lagrange basis of `nodes`."""
    nodes = np.asarray(nodes, dtype=np.float64)
    pts = np.asarray(pts, dtype=np.float64)
    M = np.ones((len(pts), len(nodes)))
    for j in range(len(nodes)):
        for m in range(len(nodes)):
            if m != j:
                M[:, j] *= (pts - nodes[m]) / (nodes[j] - nodes[m])
    return M


_IDENT = re.compile(r"^(vtk|vtu)::(Cartesian|Legendre|Lobatto)::(vertices|cells)::(ascii|binary)$")


class PatchVTK:
    """One of the reference's ADER-DG VTK plotters
    (ExaHyPE/exahype/plotters/VTK/ADERDG2CartesianVTK.cpp, ADERDG2LegendreVTK.cpp,
    ADERDG2LobattoVTK.cpp), selected by kind, cells, container and encoding:

    * kind ``Cartesian``: vertices o + i h / order (ADERDG2CartesianVTK.cpp:293-335),
      cell centres o + (i + 1/2) h / order (:344-390); ``resolution`` r splits
      each cell into 3^r sub-patches per direction (:402-431: the reference
      prolongates the solution to each sub-patch, which evaluates the same
      polynomial there);
    * kind ``Legendre``: vertices at the Gauss-Legendre nodes o + x_i h with the
      nodal values (ADERDG2LegendreVTK.cpp:288-300), cell centres halfway
      between consecutive nodes (:427);
    * kind ``Lobatto``: vertices at the Gauss-Lobatto nodes (ADERDG2LobattoVTK.cpp:
      339-345), cell centres halfway between consecutive nodes (:424-426);
    * ``cells`` False writes point data, True cell data: ``Q`` (the first
      written_unknowns quantities, or ``mapping(x, t, Q)`` -- the reference's
      mapQuantities, vectorised over points) and ``time``;
    * container ``vtk`` (legacy; encoding ``ascii`` or ``binary``, big-endian)
      or ``vtu`` (XML; ``ascii`` or inline base64 ``binary`` with UInt32
      headers) plus ``<filename>.pvd`` (ADERDG2LobattoVTK.cpp:205-216).

    Each patch has its own vertices (values on shared faces come from each
    cell's own polynomial); sub-cells are hexahedra (quads in 2D) with the
    vertex index x fastest.
    """

    def __init__(self, filename, written_unknowns, order, dim, gauss_nodes, kind="Cartesian", cells=False,
                 container="vtk", encoding="ascii", resolution=0, mapping=None):
        if kind not in ("Cartesian", "Legendre", "Lobatto") or container not in ("vtk", "vtu") \
                or encoding not in ("ascii", "binary"):
            raise ValueError("unknown VTK plotter variant")
        if resolution and kind != "Cartesian":
            raise ValueError("resolution applies to the Cartesian plotter")
        self.filename = filename
        self.written = int(written_unknowns)
        self.order, self.dim = int(order), int(dim)
        if self.order < 1:
            raise ValueError("VTK patches need order >= 1")
        self.kind, self.cells, self.container, self.encoding = kind, bool(cells), container, encoding
        self.sub = 3 ** int(resolution)
        self.mapping = mapping
        self.gauss = np.asarray(gauss_nodes, dtype=np.float64)
        B = self.order + 1
        if kind == "Cartesian":
            self.frac = np.arange(B) / self.order
            centres = (np.arange(self.order) + 0.5) / self.order
        else:
            self.frac = self.gauss.copy() if kind == "Legendre" else lobatto_nodes(self.order)
            centres = 0.5 * (self.frac[:-1] + self.frac[1:])
        self.centres = centres
        self.count = 0
        self.series = []

    @classmethod
    def from_identifier(cls, identifier, filename, written_unknowns, order, dim, gauss_nodes, **kw):
        """The reference's plotter identifier, e.g. ``vtu::Lobatto::cells::binary``."""
        m = _IDENT.match(identifier)
        if not m:
            raise ValueError("unknown plotter identifier %r" % identifier)
        container, kind, what, enc = m.groups()
        return cls(filename, written_unknowns, order, dim, gauss_nodes, kind=kind, cells=(what == "cells"),
                   container=container, encoding=enc, **kw)

    @property
    def identifier(self):
        return "%s::%s::%s::%s" % (self.container, self.kind, "cells" if self.cells else "vertices", self.encoding)

    # ---- geometry and values --------------------------------------------------
    def _axis(self, local):
        """[S][len(local)] positions within the cell (units of h) of each sub-patch."""
        S = self.sub
        return (np.arange(S)[:, None] + np.asarray(local)[None, :]) / S

    def _values(self, luh, n, nd, local):
        """[cells][S^D patches][len(local)^D][nd]: the cell polynomial at the local positions of every
        sub-patch (patch index and point index x fastest)."""
        B, D, S = self.order + 1, self.dim, self.sub
        ax = self._axis(local)  # [S][m]
        m = ax.shape[1]
        L = _lagrange(self.gauss, ax.ravel()).reshape(S, m, B)  # [S][m][B]
        if D == 3:
            u = np.asarray(luh, dtype=np.float64).reshape(n ** 3, B, B, B, nd)
            v = np.einsum("rcz,qby,pax,nzyxu->nrqpcbau", L, L, L, u, optimize=True)
            return v.reshape(n ** 3, S ** 3, m ** 3, nd)
        u = np.asarray(luh, dtype=np.float64).reshape(n ** 2, B, B, nd)
        v = np.einsum("qby,pax,nyxu->nqpbau", L, L, u, optimize=True)
        return v.reshape(n ** 2, S ** 2, m ** 2, nd)

    def _positions(self, n, h, local, origin):
        """[cells][S^D][len(local)^D][3] positions, same ordering as _values."""
        D, S = self.dim, self.sub
        ax = self._axis(local) * h  # [S][m]
        m = ax.shape[1]
        c1 = np.arange(n) * h
        o = np.zeros(3) if origin is None else np.concatenate([np.asarray(origin, dtype=np.float64), np.zeros(3)])[:3]
        if D == 3:
            P = np.zeros((n, n, n, S, S, S, m, m, m, 3))
            P[..., 0] = (c1[None, None, :, None, None, None, None, None, None]
                         + ax[None, None, None, None, None, :, None, None, :]) + o[0]
            P[..., 1] = (c1[None, :, None, None, None, None, None, None, None]
                         + ax[None, None, None, None, :, None, None, :, None]) + o[1]
            P[..., 2] = (c1[:, None, None, None, None, None, None, None, None]
                         + ax[None, None, None, :, None, None, :, None, None]) + o[2]
            return P.reshape(n ** 3, S ** 3, m ** 3, 3)
        P = np.zeros((n, n, S, S, m, m, 3))
        P[..., 0] = (c1[None, :, None, None, None, None] + ax[None, None, None, :, None, :]) + o[0]
        P[..., 1] = (c1[:, None, None, None, None, None] + ax[None, None, :, None, :, None]) + o[1]
        P[..., 2] = o[2]
        return P.reshape(n ** 2, S ** 2, m ** 2, 3)

    def _connectivity(self, npatch):
        B, D, o = self.order + 1, self.dim, self.order
        if D == 3:
            corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
            sub = [(a, b, c) for c in range(o) for b in range(o) for a in range(o)]
            local = np.array([[(s[2] + dz) * B * B + (s[1] + dy) * B + (s[0] + dx) for (dx, dy, dz) in corners]
                              for s in sub])
            ctype = 12  # VTK_HEXAHEDRON
        else:
            corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
            sub = [(a, b) for b in range(o) for a in range(o)]
            local = np.array([[(s[1] + dy) * B + (s[0] + dx) for (dx, dy) in corners] for s in sub])
            ctype = 9  # VTK_QUAD
        nv = B ** D
        conn = (np.arange(npatch)[:, None, None] * nv + local[None, :, :]).reshape(-1, len(corners))
        return conn, ctype

    def vertex_values(self, luh, n, nd):
        """[cells][vertex][nd] values at the patch vertices (without sub-patches)."""
        v = self._values(luh, n, nd, self.frac)
        return v.reshape(v.shape[0], -1, nd)

    # ---- output -----------------------------------------------------------------
    def write(self, luh, n, h, nd, time, path=None, origin=None):
        """Write one snapshot of the periodic n^dim mesh of edge h (cell (0,..) at origin);
        the default path is ``<filename>-<counter>.<vtk|vtu>``."""
        D = self.dim
        if path is None:
            path = "%s-%d.%s" % (self.filename, self.count, self.container)
        self.count += 1
        verts = self._positions(n, h, self.frac, origin).reshape(-1, 3)
        npatch = n ** D * self.sub ** D
        conn, ctype = self._connectivity(npatch)
        local = self.centres if self.cells else self.frac
        vals = self._values(luh, n, nd, local).reshape(-1, nd)
        if self.mapping is not None:
            x = self._positions(n, h, local, origin).reshape(-1, 3)[:, :D]
            data = np.asarray(self.mapping(x, float(time), vals), dtype=np.float64).reshape(len(vals), -1)
        else:
            data = vals[:, :self.written]
        tarr = np.full(len(data), float(time))
        if self.container == "vtk":
            self._write_legacy(path, verts, conn, ctype, data, tarr, time)
        else:
            self._write_vtu(path, verts, conn, ctype, data, tarr)
            self.series.append((float(time), os.path.basename(path)))
            self._write_pvd()
        return path

    def _write_legacy(self, path, verts, conn, ctype, data, tarr, time):
        ncorner = conn.shape[1]
        binary = self.encoding == "binary"
        where = "CELL_DATA" if self.cells else "POINT_DATA"
        cells = np.hstack([np.full((len(conn), 1), ncorner), conn])
        with open(path, "wb") as f:
            def text(s):
                f.write(s.encode("ascii"))

            def block(a, fmt, be):
                if binary:
                    f.write(np.ascontiguousarray(a, dtype=be).tobytes())
                    text("\n")
                else:
                    np.savetxt(f, a, fmt=fmt)

            text("# vtk DataFile Version 2.0\nExaHyPE %s %s, time %.17g\n%s\n"
                 % (self.kind, "cells" if self.cells else "vertices", time, "BINARY" if binary else "ASCII"))
            text("DATASET UNSTRUCTURED_GRID\nPOINTS %d double\n" % len(verts))
            block(verts, "%.17g", ">f8")
            text("CELLS %d %d\n" % (len(conn), cells.size))
            block(cells, "%d", ">i4")
            text("CELL_TYPES %d\n" % len(conn))
            block(np.full(len(conn), ctype), "%d", ">i4")
            text("%s %d\nFIELD FieldData 2\n" % (where, len(data)))
            text("Q %d %d double\n" % (data.shape[1], len(data)))
            block(data, "%.17g", ">f8")
            text("time 1 %d double\n" % len(tarr))
            block(tarr, "%.17g", ">f8")

    def _array(self, name, a, vtype, ncomp):
        a = np.asarray(a)
        attrs = 'type="%s"%s NumberOfComponents="%d"' % (vtype, (' Name="%s"' % name) if name else "", ncomp)
        if self.encoding == "binary":
            le = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}[vtype]
            raw = np.ascontiguousarray(a, dtype=le).tobytes()
            payload = base64.b64encode(np.uint32(len(raw)).tobytes() + raw).decode("ascii")
            return '<DataArray %s format="binary">\n%s\n</DataArray>\n' % (attrs, payload)
        fmt = "%.17g" if vtype == "Float64" else "%d"
        flat = a.reshape(-1)
        lines = "\n".join(" ".join(fmt % v for v in flat[k:k + 12]) for k in range(0, len(flat), 12))
        return '<DataArray %s format="ascii">\n%s\n</DataArray>\n' % (attrs, lines)

    def _write_vtu(self, path, verts, conn, ctype, data, tarr):
        where = "CellData" if self.cells else "PointData"
        offs = np.arange(1, len(conn) + 1) * conn.shape[1]
        parts = ['<?xml version="1.0"?>\n',
                 '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" header_type="UInt32">\n',
                 '<UnstructuredGrid>\n<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (len(verts), len(conn)),
                 '<%s Scalars="time">\n' % where,
                 self._array("Q", data, "Float64", data.shape[1]),
                 self._array("time", tarr, "Float64", 1),
                 '</%s>\n<Points>\n' % where,
                 self._array(None, verts, "Float64", 3),
                 '</Points>\n<Cells>\n',
                 self._array("connectivity", conn, "Int64", 1),
                 self._array("offsets", offs, "Int64", 1),
                 self._array("types", np.full(len(conn), ctype), "UInt8", 1),
                 '</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n']
        with open(path, "w") as f:
            f.write("".join(parts))

    def _write_pvd(self):
        rows = "".join('<DataSet timestep="%.17g" group="" part="0" file="%s"/>\n' % (t, fn) for t, fn in self.series)
        with open(self.filename + ".pvd", "w") as f:
            f.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n'
                    "<Collection>\n%s</Collection>\n</VTKFile>\n" % rows)


class CartesianVTK(PatchVTK):
    """``vtk::Cartesian::vertices::ascii`` by default (ADERDG2CartesianVTK.cpp:293-335, 435-456):
    every cell of edge h at offset o as its own grid of (order + 1)^dim equidistant
    vertices o + i h / order with the cell polynomial's values there."""

    def __init__(self, filename, written_unknowns, order, dim, gauss_nodes, **kw):
        super().__init__(filename, written_unknowns, order, dim, gauss_nodes, kind="Cartesian", **kw)


class LegendreVTK(PatchVTK):
    """``vtk::Legendre::vertices::ascii`` by default (ADERDG2LegendreVTK.cpp:288-300, 333-383): the
    patch vertices are the cell's Gauss-Legendre nodes o + x_i h and carry the nodal values; the
    order^dim sub-cells join neighbouring nodes (they do not reach the cell faces)."""

    def __init__(self, filename, written_unknowns, order, dim, gauss_nodes, **kw):
        super().__init__(filename, written_unknowns, order, dim, gauss_nodes, kind="Legendre", **kw)


class LobattoVTK(PatchVTK):
    """``vtk::Lobatto::vertices::ascii`` by default (ADERDG2LobattoVTK.cpp:326-410): vertices at the
    Gauss-Lobatto nodes (cell corners included) with the cell polynomial's values there."""

    def __init__(self, filename, written_unknowns, order, dim, gauss_nodes, **kw):
        super().__init__(filename, written_unknowns, order, dim, gauss_nodes, kind="Lobatto", **kw)
