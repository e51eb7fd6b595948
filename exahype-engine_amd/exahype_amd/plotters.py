"""Host-side plotters over the device mesh.

ProbeAscii writes the reference's ``probe::ascii`` format
(ExaHyPE/exahype/plotters/ADERDG2ProbeAscii.cpp): a file ``<filename>.probe``
whose first line is ``# plot-time, real-time,Q0,Q1,...`` and which gets one
line per plot, ``<plot time>, <cell time stamp>, <value 0>, <value 1>, ...``,
numbers in the C++ stream default format (6 significant digits, as ``%g``).
The values are the solution interpolated at the probe point
(exa_mesh_probe), optionally passed through ``mapping(x, t, Q) -> values``
(the reference's mapQuantities post-processing; identity by default).
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
