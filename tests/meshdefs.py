"""Tiny mesh definitions shared by the oracle and GPU tests.

Each entry is (ndim, ndcrd, cltpn, clnds) with clnds rows in the reference
layout (nnode, n0, n1, ...).  They are the fixtures of the reference tests:
tests/test_multidim.py:152-281 and :810-827, tests/test_mesh.py:63-182.
"""

QUAD, TRI, HEX, TET, PRISM, PYRAMID, LINE = 3, 4, 5, 6, 7, 8, 2

TRIANGLES = (2, [(0, 0), (-1, -1), (1, -1), (0, 1)], [TRI] * 3,
             [(3, 0, 1, 2), (3, 0, 2, 3), (3, 0, 3, 1)])
QUADS = (2, [(0, 0), (1, 0), (1, 1), (0, 1)], [QUAD], [(4, 0, 1, 2, 3)])
MIXED = (2, [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)],
         [QUAD, TRI, TRI], [(4, 0, 1, 4, 3), (3, 1, 2, 4, 0), (3, 2, 5, 4, 0)])
TETRA = (3, [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)], [TET],
         [(4, 0, 1, 2, 3)])
TETRA_MESHTEST = (3, [(0, 0, 0), (0, 1, 0), (-1, 1, 0), (0, 1, 1)], [TET],
                  [(4, 0, 1, 2, 3)])
HEXA = (3, [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
            (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], [HEX],
        [(8, 0, 1, 2, 3, 4, 5, 6, 7)])
PRISMS = (3, [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1),
              (0, 1, 1)], [PRISM], [(6, 0, 1, 2, 3, 4, 5)])
PYRAMIDS = (3, [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0.5, 0.5, 1)],
            [PYRAMID], [(5, 0, 1, 2, 3, 4)])

SINGLES = {'triangles': TRIANGLES, 'quad': QUADS, 'mixed': MIXED,
           'tet': TETRA, 'hex': HEXA, 'prism': PRISMS, 'pyramid': PYRAMIDS}


def quad_channel(nx, ny, lx, ly):
    """Structured nx-by-ny quad grid (tests/test_multidim.py:810-827)."""
    ndcrd = [(i * lx / nx, j * ly / ny)
             for j in range(ny + 1) for i in range(nx + 1)]

    def nid(i, j):
        return j * (nx + 1) + i

    clnds = [(4, nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1))
             for j in range(ny) for i in range(nx)]
    return (2, ndcrd, [QUAD] * (nx * ny), clnds)


def hex_block(nx, ny, nz, jitter=0.0, seed=0):
    """A small structured hexahedral block on the unit cube, optionally with
    interior nodes jittered; used for 3D march parity on several cells."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ndcrd = []
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                x = np.array([i / nx, j / ny, k / nz])
                if jitter and 0 < i < nx and 0 < j < ny and 0 < k < nz:
                    x = x + rng.uniform(-jitter, jitter, 3) / max(nx, ny, nz)
                ndcrd.append(tuple(x))

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    clnds = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                clnds.append((8, nid(i, j, k), nid(i + 1, j, k),
                              nid(i + 1, j + 1, k), nid(i, j + 1, k),
                              nid(i, j, k + 1), nid(i + 1, j, k + 1),
                              nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)))
    return (3, ndcrd, [HEX] * len(clnds), clnds)
