"""Reader for Gmsh MSH 2.2 ASCII meshes into a device StaticMesh.

Follows solvcon's reader (cpp/solvcon/inout/gmsh.hpp:68-295,
cpp/solvcon/inout/gmsh.cpp:10-118; Python name ``solvcon.Gmsh``,
wrap_Gmsh.cpp:22-41):

- The text is read section by section in the order $MeshFormat,
  optional $PhysicalNames, $Nodes, $Elements.  A line holding '$' that is not
  the next expected section header ends the parse.
- A file without the format, node and element sections, or of another
  version than 2.2, raises ValueError with the reference's messages.
- Node and element numbers are converted from 1-based to 0-based.  The
  corner nodes of every element are placed in solvcon's cell order.
- to_block() keeps only the elements of the highest dimension as cells and
  builds the interior (with metrics and edges), boundary and ghosts.

Parsing is host work on the text; the mesh itself is built by the device
builder of StaticMesh.  The node table keeps the first ndim coordinates.
"""

import numpy as np

from .mesh import StaticMesh

__all__ = ['Gmsh', 'ELEMENTS']

# Gmsh element type -> (dimension, node count, solvcon cell type, solvcon
# position of each corner node) (GmshElementDef::by_id, gmsh.hpp:297-336).
# The positions are the format's node-order facts, so they match the
# reference's table.
ELEMENTS = {
    1: (1, 2, 2, (0, 1)),
    2: (2, 3, 4, (0, 1, 2)),
    3: (2, 4, 3, (0, 1, 2, 3)),
    4: (3, 4, 6, (0, 1, 2, 3)),
    5: (3, 8, 5, (0, 1, 2, 3, 4, 5, 6, 7)),
    6: (3, 6, 7, (0, 2, 1, 3, 5, 4)),
    7: (3, 5, 8, (0, 1, 2, 3, 4)),
    8: (1, 3, 2, (0, 1)),
    9: (2, 6, 4, (0, 1, 2)),
    10: (2, 9, 3, (0, 1, 2, 3)),
    11: (3, 10, 6, (0, 1, 2, 3)),
    12: (3, 27, 5, (0, 1, 2, 3, 4, 5, 6, 7)),
    13: (3, 18, 7, (0, 2, 1, 3, 5, 4)),
    14: (3, 14, 8, (0, 1, 2, 3, 4)),
    15: (0, 1, 1, (0,)),
    16: (2, 8, 3, (0, 1, 2, 3)),
    17: (3, 20, 5, (0, 1, 2, 3, 4, 5, 6, 7)),
    18: (3, 15, 7, (0, 2, 1, 3, 5, 4)),
    19: (3, 13, 8, (0, 1, 2, 3, 4)),
    20: (2, 9, 4, (0, 1, 2)),
    21: (2, 10, 4, (0, 1, 2)),
    22: (2, 12, 4, (0, 1, 2)),
    23: (2, 15, 4, (0, 1, 2)),
    24: (2, 15, 4, (0, 1, 2)),
    25: (2, 21, 4, (0, 1, 2)),
    26: (1, 4, 2, (0, 1)),
    27: (1, 5, 2, (0, 1)),
    28: (1, 6, 2, (0, 1)),
    29: (3, 20, 6, (0, 1, 2, 3)),
    30: (3, 35, 6, (0, 1, 2, 3)),
    31: (3, 56, 6, (0, 1, 2, 3)),
    92: (3, 64, 5, (0, 1, 2, 3, 4, 5, 6, 7)),
    93: (3, 125, 5, (0, 1, 2, 3, 4, 5, 6, 7)),
}
_UNKNOWN = (0, 0, 0, ())

# Section headers allowed after each parse state.
_NEXT = {
    'begin': ('$MeshFormat',),
    'meta': ('$PhysicalNames', '$Nodes'),
    'physical': ('$PhysicalNames', '$Nodes'),
    'nodes': ('$Elements',),
    'elements': (),
}


def _fmt_version(v):
    # The shortest round-trip form, without a trailing ".0".
    s = repr(float(v))
    return s[:-2] if s.endswith('.0') else s


class Gmsh(object):
    """A parsed Gmsh MSH 2.2 text (``data``: bytes or str)."""

    def __init__(self, data):
        if isinstance(data, (bytes, bytearray)):
            data = bytes(data).decode('latin-1')
        self._lines = data.replace('\r', '').split('\n')
        self._pos = 0
        self.state = 'begin'
        self.msh_ver = 0.0
        self.msh_file_type = 0
        self.msh_data_size = 0
        self.nds = np.zeros((0, 3), dtype=np.float64)
        self.cltpn = np.zeros(0, dtype=np.int32)
        self.elgrp = np.zeros(0, dtype=np.int64)
        self.elgeo = np.zeros(0, dtype=np.int64)
        self.eldim = np.zeros(0, dtype=np.int64)
        self.elems = []
        entered = set()
        handlers = {'$MeshFormat': self._load_meta, '$Nodes': self._load_nodes,
                    '$Elements': self._load_elements,
                    '$PhysicalNames': self._load_physical}
        while True:
            line = self._getline()
            if line is None:
                break
            if '$' not in line:
                continue
            if line not in handlers or line not in _NEXT[self.state]:
                break
            handlers[line]()
            entered.add(line)
        if not {'$MeshFormat', '$Nodes', '$Elements'} <= entered:
            raise ValueError('Incorrect msh file format.')

    def _getline(self):
        if self._pos >= len(self._lines):
            return None
        line = self._lines[self._pos]
        self._pos += 1
        return line

    def _body(self):
        """Lines up to (excluding) the next line holding '$', and that line
        (None at the end of the text)."""
        out = []
        while True:
            line = self._getline()
            if line is None or '$' in line:
                return out, line
            out.append(line)

    def _load_meta(self):
        body, end = self._body()
        for line in body:
            tok = line.split()
            self.msh_ver = float(tok[0])
            if self.msh_ver != 2.2:
                raise ValueError('solvcon does not support msh file ver %s.'
                                 % _fmt_version(self.msh_ver))
            self.msh_file_type = int(tok[1])
            self.msh_data_size = int(tok[2])
        if end == '$EndMeshFormat':
            self.state = 'meta'

    def _load_physical(self):
        body, end = self._body()
        if end == '$EndPhysicalNames':
            self.state = 'physical'

    def _load_nodes(self):
        nnode = int(self._getline())
        self.nds = np.zeros((nnode, 3), dtype=np.float64)
        body, end = self._body()
        for line in body:
            tok = line.split()
            self.nds[int(tok[0]) - 1] = [float(t) for t in tok[1:4]]
        if end == '$EndNodes':
            self.state = 'nodes'

    def _load_elements(self):
        nelem = int(self._getline())
        self.cltpn = np.zeros(nelem, dtype=np.int32)
        self.elgrp = np.zeros(nelem, dtype=np.int64)
        self.elgeo = np.zeros(nelem, dtype=np.int64)
        self.eldim = np.zeros(nelem, dtype=np.int64)
        self.elems = []
        end = None
        idx = 0
        while True:
            line = self._getline()
            if line is None or '$' in line or idx >= nelem:
                end = line
                break
            tok = [int(t) for t in line.split()]
            ndim, _nnds, tpn, order = ELEMENTS.get(tok[1], _UNKNOWN)
            ntag = tok[2]
            tags = tok[3:3 + ntag]
            nds = tok[3 + ntag:]
            self.cltpn[idx] = tpn
            self.elgrp[idx] = tags[0]
            self.elgeo[idx] = tags[1]
            self.eldim[idx] = ndim
            row = [0] * (len(nds) + 1)
            for i, pos in enumerate(order):
                row[pos + 1] = nds[i] - 1
            row[0] = len(order)
            self.elems.append(row[:len(order) + 1])
            idx += 1
        if end == '$EndElements':
            self.state = 'elements'

    def definition(self):
        """(ndim, ndcrd, cltpn, clnds) of the cells: the elements of the
        highest dimension, nodes kept as numbered in the file."""
        topdim = int(self.eldim.max()) if self.eldim.size else 0
        cells = [i for i in range(self.eldim.size) if self.eldim[i] == topdim]
        width = max([len(self.elems[i]) for i in cells] + [1])
        clnds = np.zeros((len(cells), width), dtype=np.int32)
        for c, i in enumerate(cells):
            row = self.elems[i]
            clnds[c, :len(row)] = row
        cltpn = self.cltpn[cells].astype(np.int32)
        ndcrd = np.ascontiguousarray(self.nds[:, :topdim])
        return topdim, ndcrd, cltpn, clnds

    def to_block(self):
        """The device StaticMesh: interior with metrics and edges, boundary
        and ghosts built."""
        ndim, ndcrd, cltpn, clnds = self.definition()
        mh = StaticMesh(ndim=ndim, nnode=ndcrd.shape[0], nface=0,
                        ncell=cltpn.shape[0])
        mh.ndcrd.ndarray[...] = ndcrd
        mh.cltpn.ndarray[...] = cltpn
        mh.clnds.ndarray[:, :clnds.shape[1]] = clnds
        mh.build_interior(do_metric=True, build_edge=True)
        mh.build_boundary()
        mh.build_ghost()
        return mh
