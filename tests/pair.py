"""Build the same mesh / solver on the GPU library and on the oracle, and
compare their tables."""

import numpy as np

MESH_TABLES = ('ndcrd', 'fccnd', 'fcnml', 'fcara', 'clcnd', 'clvol', 'fctpn',
               'cltpn', 'clgrp', 'fcnds', 'fccls', 'clnds', 'clfcs', 'ednds',
               'bndfcs')
EULER_TABLES = ('cevol', 'cecnd', 'sfcnd', 'sfnml', 'so0c', 'so0n', 'so0t',
                'so1c', 'so1n', 'cflo', 'cflc', 'gamma')


def as_arrays(defn):
    ndim, ndcrd, cltpn, clnds = defn
    ndcrd = np.asarray(ndcrd, dtype=np.float64)
    cltpn = np.asarray(cltpn, dtype=np.int32)
    if not isinstance(clnds, np.ndarray):
        w = max(len(r) for r in clnds)
        a = np.zeros((len(clnds), w), dtype=np.int32)
        for i, r in enumerate(clnds):
            a[i, :len(r)] = r
        clnds = a
    return ndim, ndcrd, cltpn, clnds


def gpu_mesh(defn, bcs=(), edge=True):
    import solvcon_amd as sc
    ndim, ndcrd, cltpn, clnds = as_arrays(defn)
    mh = sc.StaticMesh(ndim=ndim, nnode=ndcrd.shape[0], nface=0,
                       ncell=cltpn.shape[0])
    mh.ndcrd[:, :] = ndcrd
    mh.cltpn[:] = cltpn
    mh.clnds[:, :clnds.shape[1]] = clnds
    mh.build_interior(do_metric=True, build_edge=edge)
    for name, faces in bcs:
        mh.add_bc(name, faces)
    mh.build_boundary()
    mh.build_ghost()
    return mh


def orc_mesh(orc, defn, bcs=(), edge=True):
    ndim, ndcrd, cltpn, clnds = as_arrays(defn)
    m = orc.Mesh(ndim, ndcrd.shape[0], 0, cltpn.shape[0])
    m.ndcrd[:, :] = ndcrd
    m.cltpn[:] = cltpn
    m.clnds[:, :clnds.shape[1]] = clnds
    m.build_interior(do_metric=True, build_edge=edge)
    for name, faces in bcs:
        m.add_bc(name, faces)
    m.build_boundary()
    m.build_ghost()
    return m


def same_mesh(g, o):
    """Assert every table and count of the two meshes is identical."""
    for k in ('ndim', 'nnode', 'nface', 'ncell', 'nbound', 'ngstnode',
              'ngstface', 'ngstcell', 'nedge', 'nbcs'):
        assert getattr(g, k) == getattr(o, k), k
    for name in MESH_TABLES:
        a = getattr(g, name).ndarray
        b = getattr(o, name)
        if name == 'clnds' or name == 'clfcs':
            pass
        assert a.shape == b.shape, (name, a.shape, b.shape)
        assert np.array_equal(a, b), name
    for i in range(o.nbcs):
        name, f0, f1 = o.bc(i)
        bc = g.bc(i)
        assert bc.name == name
        assert bc.facn[:, 0].tolist() == f0.tolist()
        assert bc.facn[:, 1].tolist() == f1.tolist()


def same_euler(g, o, names=EULER_TABLES):
    for name in names:
        a = getattr(g, name).ndarray
        b = getattr(o, name)
        assert a.shape == b.shape, (name, a.shape, b.shape)
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            i = tuple(bad[0])
            raise AssertionError('%s differs at %d places, first %s: %r vs %r'
                                 % (name, len(bad), i, a[i], b[i]))
