"""Unstructured mixed-element mesh built on the GPU.

Drop-in for solvcon's ``StaticMesh`` binding
(cpp/solvcon/mesh/pymod/wrap_StaticMesh.cpp:61-187): the same constructor,
counts, builders, boundary-condition groups, cell-type constants and the 15
tables, which live in HBM and are built by the kernels of mesh.hip.
"""

import ctypes

import numpy as np

from . import _lib as L
from .array import GhostArray

__all__ = ['StaticMesh', 'StaticMeshBc']

L.register({
    'scmesh_create': (L.c_int, [L.c_int, L.c_i32, L.c_i32, L.c_i32, L.P_vp]),
    'scmesh_destroy': (None, [L.c_vp]),
    'scmesh_counts': (L.c_int, [L.c_vp, L.P_i64]),
    'scmesh_table_info': (L.c_int, [L.c_vp, L.c_int, L.P_i64]),
    'scmesh_table_ptr': (L.c_int, [L.c_vp, L.c_int, L.P_vp]),
    'scmesh_download': (L.c_int, [L.c_vp, L.c_int, L.c_vp, L.c_i64]),
    'scmesh_upload': (L.c_int, [L.c_vp, L.c_int, L.c_vp, L.c_i64]),
    'scmesh_build_interior': (L.c_int, [L.c_vp, L.c_int, L.c_int]),
    'scmesh_build_edge': (L.c_int, [L.c_vp]),
    'scmesh_build_boundary': (L.c_int, [L.c_vp]),
    'scmesh_build_ghost': (L.c_int, [L.c_vp]),
    'scmesh_add_bc': (L.c_int, [L.c_vp, ctypes.c_char_p, L.P_i32, L.c_i64]),
    'scmesh_bc': (L.c_int, [L.c_vp, L.c_int, ctypes.c_char_p, L.c_int,
                            L.P_i32, L.P_i32, L.P_i64]),
})

TABLES = ('ndcrd', 'fccnd', 'fcnml', 'fcara', 'clcnd', 'clvol', 'fctpn',
          'cltpn', 'clgrp', 'fcnds', 'fccls', 'clnds', 'clfcs', 'ednds',
          'bndfcs')


class _Table(np.ndarray):
    """A host table indexed like a numpy array that also offers the
    ``.ndarray`` view of the reference's SimpleArray."""

    @property
    def ndarray(self):
        return self.view(np.ndarray)


class StaticMeshBc(object):
    """One boundary-condition group: a name and its faces.  facn columns
    are (face index, row in bndfcs, related-block face = -1)
    (cpp/solvcon/mesh/StaticMesh.hpp:159-251)."""

    NONAME = '<NONAME>'

    def __init__(self, name, facn0, facn1):
        self.name = name
        facn = np.full((len(facn0), 3), -1, dtype='int32')
        facn[:, 0] = facn0
        facn[:, 1] = facn1
        self.facn = facn.view(_Table)

    @property
    def nbound(self):
        return self.facn.shape[0]


class StaticMesh(object):
    """Static unstructured mesh of mixed element types on the GPU."""

    FCMND = 4
    CLMND = 8
    CLMFC = 6
    FCREL = 4
    BFREL = 3

    NONCELLTYPE = 0
    POINT = 1
    LINE = 2
    QUADRILATERAL = 3
    TRIANGLE = 4
    HEXAHEDRON = 5
    TETRAHEDRON = 6
    PRISM = 7
    PYRAMID = 8

    def __init__(self, ndim, nnode, nface=0, ncell=0):
        L.require_gpu()
        h = ctypes.c_void_p()
        L.check(L.lib().scmesh_create(int(ndim), int(nnode), int(nface),
                                      int(ncell), ctypes.byref(h)))
        self._h = h
        self._mirrors = {}
        self._refresh(structural=True)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and L._lib is not None:
            L._lib.scmesh_destroy(h)
            self._h = None

    # -- host mirrors -------------------------------------------------------

    def _info(self, tid):
        out = (ctypes.c_int64 * 4)()
        L.check(L.lib().scmesh_table_info(self._h, tid, out))
        rows, cols, ngh, is_int = list(out)
        shape = (rows, cols) if cols > 0 else (rows,)
        return shape, int(ngh), np.int32 if is_int else np.float64

    def _mk(self, tid, shape, ngh, dtype):
        def up(a):
            L.check(L.lib().scmesh_upload(self._h, tid, a.ctypes.data,
                                          a.nbytes))

        def down(a):
            L.check(L.lib().scmesh_download(self._h, tid, a.ctypes.data,
                                            a.nbytes))
        return L.Mirror(shape, dtype, up, down, nghost=ngh)

    def _refresh(self, structural):
        for tid, name in enumerate(TABLES):
            shape, ngh, dtype = self._info(tid)
            m = self._mirrors.get(name)
            if m is None or structural and (m.shape != tuple(shape)
                                            or m.nghost != ngh):
                self._mirrors[name] = self._mk(tid, shape, ngh, dtype)
            else:
                m.after()

    def _mirror(self, name):
        return self._mirrors[name]

    def _before(self):
        for m in self._mirrors.values():
            m.before()

    def _call(self, fn, *args, structural=True):
        self._before()
        try:
            L.check(fn(self._h, *args))
        finally:
            self._refresh(structural)

    # -- counts -------------------------------------------------------------

    def _counts(self):
        out = (ctypes.c_int64 * 10)()
        L.lib().scmesh_counts(self._h, out)
        return list(out)

    ndim = property(lambda self: self._counts()[0])
    nnode = property(lambda self: self._counts()[1])
    nface = property(lambda self: self._counts()[2])
    ncell = property(lambda self: self._counts()[3])
    nbound = property(lambda self: self._counts()[4])
    ngstnode = property(lambda self: self._counts()[5])
    ngstface = property(lambda self: self._counts()[6])
    ngstcell = property(lambda self: self._counts()[7])
    nedge = property(lambda self: self._counts()[8])
    nbcs = property(lambda self: self._counts()[9])

    # -- builders -----------------------------------------------------------

    def build_interior(self, do_metric=True, build_edge=True):
        self._call(L.lib().scmesh_build_interior, int(bool(do_metric)),
                   int(bool(build_edge)))

    def build_edge(self):
        self._call(L.lib().scmesh_build_edge)

    def build_boundary(self):
        self._call(L.lib().scmesh_build_boundary)

    def build_ghost(self):
        self._call(L.lib().scmesh_build_ghost)

    # -- boundary-condition groups -------------------------------------------

    def add_bc(self, name, faces):
        f = np.ascontiguousarray(np.asarray(faces, dtype=np.int64)
                                 .astype(np.int32))
        self._before()
        L.check(L.lib().scmesh_add_bc(self._h, str(name).encode(),
                                      L.iptr(f), f.size))
        return self.bc(self.nbcs - 1)

    def bc(self, key):
        if isinstance(key, str):
            for b in self.bcs:
                if b.name == key:
                    return b
            raise KeyError(key)
        ibc = int(key)
        n = ctypes.c_int64()
        L.check(L.lib().scmesh_bc(self._h, ibc, None, 0, None, None,
                                  ctypes.byref(n)))
        name = ctypes.create_string_buffer(512)
        f0 = np.zeros(n.value, dtype=np.int32)
        f1 = np.zeros(n.value, dtype=np.int32)
        L.check(L.lib().scmesh_bc(self._h, ibc, name, 512, L.iptr(f0),
                                  L.iptr(f1), ctypes.byref(n)))
        return StaticMeshBc(name.value.decode(), f0, f1)

    @property
    def bcs(self):
        return [self.bc(i) for i in range(self.nbcs)]


def _table_property(name):
    return property(lambda self: GhostArray(self, name))


for _name in TABLES:
    setattr(StaticMesh, _name, _table_property(_name))
del _name
