"""Pin the oracle's StaticMesh restatement.

Golden values come from the reference tests (tests/test_mesh.py:63-182,
tests/test_mesh_2d.py); the rest is pinned by closed forms and by an
independent set-based face extraction written here.
"""

import numpy as np
import pytest
from numpy.testing import assert_almost_equal

import meshdefs as MD


def _mk(orc, d, **kw):
    return orc.make_mesh(*d, **kw)


def test_triangle_fan_goldens(orc):
    # tests/test_mesh.py:75-150
    m = _mk(orc, MD.TRIANGLES, build=False)
    m.build_interior(do_metric=False, build_edge=False)
    assert (m.nface, m.nedge) == (6, 0)
    assert (m.fccnd == 0).all() and (m.clvol == 0).all()
    m = _mk(orc, MD.TRIANGLES, build=False)
    m.build_interior()
    assert (m.nface, m.nedge) == (6, 6)
    assert_almost_equal(m.fccnd, [[-0.5, -0.5], [0.0, -1.0], [0.5, -0.5],
                                  [0.5, 0.0], [0.0, 0.5], [-0.5, 0.0]])
    assert_almost_equal(m.fcnml, [[-0.7071068, 0.7071068], [0.0, -1.0],
                                  [0.7071068, 0.7071068],
                                  [0.8944272, 0.4472136], [-1.0, -0.0],
                                  [-0.8944272, 0.4472136]])
    assert_almost_equal(m.fcara, [1.4142136, 2.0, 1.4142136, 2.236068, 1.0,
                                  2.236068])
    assert_almost_equal(m.clcnd, [[0.0, -0.6666667], [0.3333333, 0.0],
                                  [-0.3333333, 0.0]])
    assert_almost_equal(m.clvol, [1.0, 0.5, 0.5])
    assert m.nbcs == 0
    m.build_boundary()
    assert (m.nbcs, m.nbound) == (1, 3)
    assert m.bndfcs.tolist() == [[1, 0, -1], [3, 0, -1], [5, 0, -1]]
    m.build_ghost()
    assert (m.ngstnode, m.ngstface, m.ngstcell) == (3, 6, 3)
    assert m.fcnds.shape == (12, 5) and m.clnds.shape == (6, 9)
    assert m.ndcrd.shape == (7, 2) and m.clfcs.shape == (6, 7)


def test_single_tet_goldens(orc):
    # tests/test_mesh.py:121-182
    m = _mk(orc, MD.TETRA_MESHTEST, build=False)
    m.build_interior()
    assert (m.nface, m.nedge) == (4, 6)
    assert_almost_equal(m.fccnd, [[-0.3333333, 0.6666667, 0.],
                                  [0., 0.6666667, 0.3333333],
                                  [-0.3333333, 0.6666667, 0.3333333],
                                  [-0.3333333, 1., 0.3333333]])
    assert_almost_equal(m.fcnml, [[0., 0., -1.], [1., 0., 0.],
                                  [-0.5773503, -0.5773503, 0.5773503],
                                  [0., 1., 0.]])
    assert_almost_equal(m.fcara, [0.5, 0.5, 0.8660254, 0.5])
    assert_almost_equal(m.clcnd, [[-0.25, 0.75, 0.25]])
    assert_almost_equal(m.clvol, [0.1666667])
    m.build_boundary()
    assert m.bndfcs.tolist() == [[0, 0, -1], [1, 0, -1], [2, 0, -1],
                                 [3, 0, -1]]
    m.build_ghost()
    assert (m.ngstnode, m.ngstface, m.ngstcell) == (4, 12, 4)


def test_single_line(orc):
    # tests/test_mesh.py:184-206
    m = orc.Mesh(1, 2, 0, 1)
    m.ndcrd[:] = [[0], [1]]
    m.cltpn[:] = MD.LINE
    m.clnds[:, :3] = [(2, 0, 1)]
    m.build_interior(do_metric=False, build_edge=False)
    assert (m.nface, m.ncell) == (2, 1)


def test_unit_quad(orc):
    # tests/test_mesh_2d.py: unit-square normals, areas, volume, centroid.
    m = _mk(orc, MD.QUADS)
    ng = m.ngstface
    assert_almost_equal(m.fcara[ng:], [1.0] * 4)
    assert_almost_equal(m.fcnml[ng:], [[0, -1], [1, 0], [0, 1], [-1, 0]])
    assert_almost_equal(m.clvol[m.ngstcell:], [1.0])
    assert_almost_equal(m.clcnd[m.ngstcell:], [[0.5, 0.5]])
    # Ghost cells mirror the interior one across each face.
    assert_almost_equal(m.clcnd[:m.ngstcell],
                        [[-0.5, 0.5], [0.5, 1.5], [1.5, 0.5], [0.5, -0.5]])
    assert_almost_equal(m.clvol[:m.ngstcell], [1.0] * 4)


@pytest.mark.parametrize('name', ['hex', 'prism', 'pyramid', 'tet'])
def test_single_3d_volume(orc, name):
    # Closed-form volumes: unit hex 1, right prism 1/2, pyramid 1/3, tet 1/6.
    expect = {'hex': 1.0, 'prism': 0.5, 'pyramid': 1 / 3, 'tet': 1 / 6}[name]
    m = _mk(orc, MD.SINGLES[name])
    assert_almost_equal(m.clvol[m.ngstcell:], [expect], decimal=12)
    # Ghost cells are mirror images: same volume.
    assert_almost_equal(m.clvol[:m.ngstcell], [expect] * m.ngstcell,
                        decimal=12)
    # Closed surface: sum of outward area vectors vanishes.
    ng = m.ngstface
    s = (m.fcnml[ng:] * m.fcara[ng:, None]).sum(axis=0)
    assert_almost_equal(s, np.zeros(3), decimal=12)


def _brute_faces(cltpn, clnds):
    """Independent face extraction: dict keyed by (type, node set); first
    occurrence in (cell, local face) order numbers the face."""
    tab = {
        MD.TRI: [(2, (0, 1)), (2, (1, 2)), (2, (2, 0))],
        MD.QUAD: [(2, (0, 1)), (2, (1, 2)), (2, (2, 3)), (2, (3, 0))],
        MD.TET: [(4, (0, 2, 1)), (4, (0, 1, 3)), (4, (0, 3, 2)),
                 (4, (1, 2, 3))],
        MD.HEX: [(3, (0, 3, 2, 1)), (3, (1, 2, 6, 5)), (3, (4, 5, 6, 7)),
                 (3, (0, 4, 7, 3)), (3, (0, 1, 5, 4)), (3, (2, 3, 7, 6))],
        MD.PRISM: [(4, (0, 1, 2)), (4, (3, 5, 4)), (3, (0, 3, 4, 1)),
                   (3, (0, 2, 5, 3)), (3, (1, 4, 5, 2))],
        MD.PYRAMID: [(4, (0, 4, 3)), (4, (1, 4, 0)), (4, (2, 4, 1)),
                     (4, (3, 4, 2)), (3, (0, 3, 2, 1))],
    }
    key2id = {}
    fcnds = []
    fccls = []
    clfcs = []
    for icl, (tp, row) in enumerate(zip(cltpn, clnds)):
        ids = []
        for ftp, loc in tab[tp]:
            nodes = tuple(row[1 + k] for k in loc)
            key = (ftp, frozenset(nodes))
            if key not in key2id:
                key2id[key] = len(fcnds)
                fcnds.append(nodes)
                fccls.append([icl, -1])
            else:
                f = key2id[key]
                if fccls[f][1] == -1:
                    fccls[f][1] = icl
            ids.append(key2id[key])
        clfcs.append(ids)
    return fcnds, fccls, clfcs


def _random_tri_mesh(n, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    nodes = [(i * h, j * h) for j in range(n + 1) for i in range(n + 1)]
    cells = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            b, c, d = a + 1, a + n + 2, a + n + 1
            if rng.integers(2):
                cells += [(3, a, b, c), (3, a, c, d)]
            else:
                cells += [(3, a, b, d), (3, b, c, d)]
    perm = rng.permutation(len(cells))
    cells = [cells[p] for p in perm]
    return (2, nodes, [MD.TRI] * len(cells), cells)


@pytest.mark.parametrize('seed', [1, 2, 3])
def test_face_extraction_matches_bruteforce(orc, seed):
    d = _random_tri_mesh(7, seed)
    m = orc.make_mesh(*d, build=False)
    m.build_interior(do_metric=False, build_edge=False)
    fcnds, fccls, clfcs = _brute_faces(d[2], d[3])
    assert m.nface == len(fcnds)
    for f, nodes in enumerate(fcnds):
        assert m.fcnds[f, 0] == len(nodes)
        assert tuple(m.fcnds[f, 1:1 + len(nodes)]) == nodes
        assert list(m.fccls[f, :2]) == fccls[f]
    for c, ids in enumerate(clfcs):
        assert m.clfcs[c, 0] == len(ids)
        assert list(m.clfcs[c, 1:1 + len(ids)]) == ids


def test_face_extraction_hex_block(orc):
    d = MD.hex_block(3, 2, 2)
    m = orc.make_mesh(*d, build=False)
    m.build_interior(do_metric=False, build_edge=False)
    fcnds, fccls, clfcs = _brute_faces(d[2], d[3])
    assert m.nface == len(fcnds)
    for f in range(m.nface):
        assert list(m.fccls[f, :2]) == fccls[f]


def test_normals_point_out_of_first_cell(orc):
    d = _random_tri_mesh(6, 9)
    m = orc.make_mesh(*d)
    ngf, ngc = m.ngstface, m.ngstcell
    fccnd, fcnml, clcnd = m.fccnd, m.fcnml, m.clcnd
    for f in range(m.nface):
        c0 = m.fccls[ngf + f, 0]
        v = (fccnd[ngf + f] - clcnd[ngc + c0]) @ fcnml[ngf + f]
        assert v > 0
    # Body area sums to the unit square; ghost cells mirror their neighbours.
    assert_almost_equal(m.clvol[ngc:].sum(), 1.0, decimal=12)
    for g in range(m.ngstcell):
        f = m.bndfcs[g, 0]
        assert m.fccls[ngf + f, 1] == -(g + 1)
        ci = m.fccls[ngf + f, 0]
        assert_almost_equal(m.clvol[ngc - 1 - g], m.clvol[ngc + ci],
                            decimal=12)


def test_user_bc_ordering(orc):
    # build_boundary: user groups first in add order, then a catch-all group
    # of the remaining boundary faces in ascending order
    # (StaticMesh_boundary.cpp:103-182).
    m = orc.make_mesh(*MD.quad_channel(3, 2, 3.0, 2.0), build=False)
    m.build_interior()
    bnd = [f for f in range(m.nface) if m.fccls[f, 1] < 0]
    m.add_bc('left', [bnd[5], bnd[1]])
    m.add_bc('right', [bnd[3]])
    with pytest.raises(ValueError):
        m.add_bc('left', [bnd[0]])
    with pytest.raises(ValueError):
        m.add_bc('dup', [bnd[3]])
    with pytest.raises(ValueError):
        m.add_bc('interior', [f for f in range(m.nface)
                              if m.fccls[f, 1] >= 0][:1])
    m.build_boundary()
    rest = [f for f in bnd if f not in (bnd[5], bnd[1], bnd[3])]
    assert m.bndfcs[:, 0].tolist() == [bnd[5], bnd[1], bnd[3]] + rest
    assert m.bndfcs[:, 1].tolist() == [0, 0, 1] + [2] * len(rest)
    name, f0, f1 = m.bc(2)
    assert name == '<NONAME>' and f1.tolist() == list(range(3, len(bnd)))
    with pytest.raises(RuntimeError):
        m.build_boundary()
