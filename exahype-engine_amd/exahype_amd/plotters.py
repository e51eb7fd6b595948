"""Host-side plotters over the device mesh.

ProbeAscii writes the reference's ``probe::ascii`` format
(ExaHyPE/exahype/plotters/ADERDG2ProbeAscii.cpp): a file ``<filename>.probe``
whose first line is ``# plot-time, real-time,Q0,Q1,...`` and which gets one
line per plot, ``<plot time>, <cell time stamp>, <value 0>, <value 1>, ...``,
numbers in the C++ stream default format (6 significant digits, as ``%g``).
The values are the solution interpolated at the probe point
(exa_mesh_probe), optionally passed through ``mapping(x, t, Q) -> values``
(the reference's mapQuantities post-processing; identity by default).

CartesianVTK writes the reference's ``vtk::Cartesian::vertices::ascii``
layout (ExaHyPE/exahype/plotters/VTK/ADERDG2CartesianVTK.cpp) from the
solution array: every cell as an equidistant vertex patch.
"""
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


class CartesianVTK:
    kind = "Cartesian"

    """The reference's ``vtk::Cartesian::vertices::ascii`` plotter
    (ExaHyPE/exahype/plotters/VTK/ADERDG2CartesianVTK.cpp:293-335, 435-456):
    every cell (patch) of edge h at offset o is written as its own grid of
    (order + 1)^dim equidistant vertices o + i h / order, i in [0, order]^dim,
    and order^dim sub-cells; the vertex values are the cell polynomial there,
    sum over Gauss-Legendre nodes of prod_d equidistantGridProjector[order][ii_d][i_d]
    * u[ii] (the projector is the Lagrange basis of the Gauss nodes at i / order,
    so values on shared faces come from each patch's own polynomial).  Point
    data: ``Q`` (the first written_unknowns quantities) and ``time``.  Legacy
    VTK ASCII, unstructured grid of hexahedra (quads in 2D); the data arrays
    are a FIELD (``Q`` may have more than four components).
    """

    def __init__(self, filename, written_unknowns, order, dim, gauss_nodes):
        self.filename = filename
        self.written = int(written_unknowns)
        self.order, self.dim = int(order), int(dim)
        x = np.asarray(gauss_nodes, dtype=np.float64)
        B = self.order + 1
        e = np.arange(B) / self.order
        # proj[ii][i] = phi_ii(i / order), the Lagrange basis of the Gauss nodes
        self.proj = np.ones((B, B))
        for ii in range(B):
            for m in range(B):
                if m != ii:
                    self.proj[ii] *= (e - x[m]) / (x[ii] - x[m])
        self.frac = e  # vertex positions within the cell, in units of h
        self.count = 0

    def vertex_values(self, luh, n, nd):
        """[cells][vertex][nd] values (vertex index x fastest) of luh[cells][nodes][nd]."""
        B, D = self.order + 1, self.dim
        P = self.proj
        if D == 3:
            u = np.asarray(luh).reshape(n ** 3, B, B, B, nd)
            v = np.einsum("zc,yb,xa,nzyxu->ncbau", P, P, P, u, optimize=True)
            return v.reshape(n ** 3, B ** 3, nd)
        u = np.asarray(luh).reshape(n ** 2, B, B, nd)
        v = np.einsum("yb,xa,nyxu->nbau", P, P, u, optimize=True)
        return v.reshape(n ** 2, B ** 2, nd)

    def write(self, luh, n, h, nd, time, path=None):
        """Write one snapshot; the default path is ``<filename>-<counter>.vtk``."""
        B, D, o = self.order + 1, self.dim, self.order
        if path is None:
            path = "%s-%d.vtk" % (self.filename, self.count)
        self.count += 1
        vals = self.vertex_values(luh, n, nd)[:, :, :self.written]
        ncells = n ** D
        k = self.frac * h
        c1 = np.arange(n) * h
        if D == 3:
            cz, cy, cx = np.meshgrid(c1, c1, c1, indexing="ij")
            kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
            pts = np.stack([cx.reshape(-1, 1) + kx.reshape(1, -1), cy.reshape(-1, 1) + ky.reshape(1, -1),
                            cz.reshape(-1, 1) + kz.reshape(1, -1)], -1)
            corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
            ctype = 12  # VTK_HEXAHEDRON
            sub = [(a, b, c) for c in range(o) for b in range(o) for a in range(o)]
            local = np.array([[(s[2] + dz) * B * B + (s[1] + dy) * B + (s[0] + dx) for (dx, dy, dz) in corners]
                              for s in sub])
        else:
            cy, cx = np.meshgrid(c1, c1, indexing="ij")
            ky, kx = np.meshgrid(k, k, indexing="ij")
            pts = np.stack([cx.reshape(-1, 1) + kx.reshape(1, -1), cy.reshape(-1, 1) + ky.reshape(1, -1),
                            np.zeros((ncells, B * B))], -1)
            corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
            ctype = 9  # VTK_QUAD
            sub = [(a, b) for b in range(o) for a in range(o)]
            local = np.array([[(s[1] + dy) * B + (s[0] + dx) for (dx, dy) in corners] for s in sub])
        nv = B ** D
        conn = (np.arange(ncells)[:, None, None] * nv + local[None, :, :]).reshape(-1, len(corners))
        npts = ncells * nv
        with open(path, "w") as f:
            f.write("# vtk DataFile Version 2.0\nExaHyPE %s vertices, time %.17g\nASCII\n" % (self.kind, time))
            f.write("DATASET UNSTRUCTURED_GRID\nPOINTS %d double\n" % npts)
            np.savetxt(f, pts.reshape(-1, 3), fmt="%.17g")
            f.write("CELLS %d %d\n" % (len(conn), len(conn) * (len(corners) + 1)))
            np.savetxt(f, np.hstack([np.full((len(conn), 1), len(corners)), conn]), fmt="%d")
            f.write("CELL_TYPES %d\n" % len(conn))
            np.savetxt(f, np.full(len(conn), ctype), fmt="%d")
            f.write("POINT_DATA %d\nFIELD FieldData 2\n" % npts)
            f.write("Q %d %d double\n" % (self.written, npts))
            np.savetxt(f, vals.reshape(npts, self.written), fmt="%.17g")
            f.write("time 1 %d double\n" % npts)
            np.savetxt(f, np.full(npts, float(time)), fmt="%.17g")
        return path


class LegendreVTK(CartesianVTK):
    """The reference's ``vtk::Legendre::vertices::ascii`` plotter
    (ExaHyPE/exahype/plotters/VTK/ADERDG2LegendreVTK.cpp:288-300, 333-383): the
    patch vertices are the cell's Gauss-Legendre nodes o + x_i h and carry the
    nodal values themselves; the order^dim sub-cells join neighbouring nodes
    (they do not reach the cell faces)."""

    kind = "Legendre"

    def __init__(self, filename, written_unknowns, order, dim, gauss_nodes):
        super().__init__(filename, written_unknowns, order, dim, gauss_nodes)
        self.proj = np.eye(self.order + 1)
        self.frac = np.asarray(gauss_nodes, dtype=np.float64)
