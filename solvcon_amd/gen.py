"""Seeded synthetic meshes and initial states for the benchmark
configurations (numpy ``Generator(PCG64(seed))``).

2D: the unit square as an m-by-m quad grid, interior nodes jittered
uniformly by up to ``jitter * h`` per axis, each quad split into two
counter-clockwise triangles along a seeded diagonal; cells are numbered in
(row j, column i, triangle k) order.
"""

import numpy as np

__all__ = ['tri_mesh_def', 'riemann_box', 'build_static_mesh', 'cfl_dt',
           'sod_tube', 'mixed_mesh_def', 'blast_state', 'rect_mesh_def',
           'edge_faces']


def tri_mesh_def(m, seed=2, jitter=0.2):
    """(ndim, ndcrd[nnode, 2], cltpn[ncell], clnds[ncell, 4]) of the
    jittered triangle mesh with 2*m*m cells."""
    rng = np.random.Generator(np.random.PCG64(seed))
    h = 1.0 / m
    j, i = np.meshgrid(np.arange(m + 1), np.arange(m + 1), indexing='ij')
    x = i * h
    y = j * h
    jx = rng.uniform(-jitter * h, jitter * h, size=x.shape)
    jy = rng.uniform(-jitter * h, jitter * h, size=y.shape)
    inner = (i > 0) & (i < m) & (j > 0) & (j < m)
    x = np.where(inner, x + jx, x)
    y = np.where(inner, y + jy, y)
    ndcrd = np.stack([x.ravel(), y.ravel()], axis=1)
    diag = rng.integers(0, 2, size=(m, m))
    jj, ii = np.meshgrid(np.arange(m), np.arange(m), indexing='ij')
    a = jj * (m + 1) + ii
    b = a + 1
    c = a + m + 2
    d = a + m + 1
    t0 = np.where(diag == 0, np.stack([a, b, c]), np.stack([a, b, d]))
    t1 = np.where(diag == 0, np.stack([a, c, d]), np.stack([b, c, d]))
    tri = np.stack([t0, t1], axis=-1)  # [3, m, m, 2]
    tri = tri.transpose(1, 2, 3, 0).reshape(-1, 3)
    clnds = np.empty((tri.shape[0], 4), dtype=np.int32)
    clnds[:, 0] = 3
    clnds[:, 1:] = tri
    cltpn = np.full(tri.shape[0], 4, dtype=np.int32)
    return 2, ndcrd, cltpn, clnds


def riemann_box(seed=2, lo=0.1, hi=1.5):
    """Four quadrant states (rho, p) drawn from U(lo, hi), u = v = 0,
    quadrants ordered (x<0.5,y<0.5), (x>=0.5,y<0.5), (x<0.5,y>=0.5),
    (x>=0.5,y>=0.5)."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    rho = rng.uniform(lo, hi, size=4)
    p = rng.uniform(lo, hi, size=4)
    return rho, p


def quadrant_state(centroids, rho, p, gamma=1.4, split=0.5):
    """Conserved rows [ncell, 4] of the Riemann box at cell centroids."""
    q = (centroids[:, 0] >= split).astype(int) \
        + 2 * (centroids[:, 1] >= split).astype(int)
    out = np.zeros((centroids.shape[0], 4))
    out[:, 0] = rho[q]
    out[:, 3] = p[q] / (gamma - 1.0)
    return out


def build_static_mesh(defn, cls=None, bcs=()):
    """Create and build a StaticMesh (the GPU class by default) from a mesh
    definition."""
    if cls is None:
        from .mesh import StaticMesh as cls
    ndim, ndcrd, cltpn, clnds = defn
    mh = cls(ndim=ndim, nnode=ndcrd.shape[0], nface=0, ncell=cltpn.shape[0])
    mh.ndcrd[:, :] = ndcrd
    mh.cltpn[:] = cltpn
    mh.clnds[:, :clnds.shape[1]] = clnds
    mh.build_interior(do_metric=True, build_edge=False)
    for name, faces in bcs:
        mh.add_bc(name, faces)
    mh.build_boundary()
    mh.build_ghost()
    return mh


def cfl_dt(cecnd, ncell, ngst, ndim, nfc, state, gamma, cfl=0.8):
    """Time step that puts the largest calc_cfl number at ``cfl``:
    dt = cfl / max(wspd / (2 dist)) with wspd = sqrt(gamma p / rho) + |u|
    and dist the shortest CCE-to-BCE centroid distance of the cell."""
    ce = np.asarray(cecnd)[ngst:ngst + ncell]
    cen = ce[:, :ndim]
    dist = np.full(ncell, np.inf)
    for f in range(1, nfc + 1):
        dd = np.sqrt(((ce[:, f * ndim:(f + 1) * ndim] - cen) ** 2).sum(1))
        dist = np.minimum(dist, dd)
    rho = state[:, 0]
    mom = state[:, 1:1 + ndim]
    en = state[:, ndim + 1]
    ke = (mom ** 2).sum(1) / (2 * rho)
    p = np.maximum((gamma - 1) * (en - ke), 0)
    wspd = np.sqrt(gamma * p / rho) + np.sqrt((mom ** 2).sum(1)) / rho
    return cfl / np.max(0.5 * wspd / dist)


def _orient(ndcrd, cells, kind):
    """Reorder the nodes of cells with negative reference orientation."""
    x = ndcrd[cells]
    if kind == 'tet':
        v = np.einsum('ij,ij->i', np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]),
                      x[:, 3] - x[:, 0])
        neg = v < 0
        cells[neg] = cells[neg][:, [0, 2, 1, 3]]
    elif kind == 'prism':
        v = np.einsum('ij,ij->i', np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]),
                      x[:, 3] - x[:, 0])
        neg = v < 0
        cells[neg] = cells[neg][:, [0, 2, 1, 3, 5, 4]]
    elif kind == 'pyramid':
        v = np.einsum('ij,ij->i', np.cross(x[:, 1] - x[:, 0], x[:, 3] - x[:, 0]),
                      x[:, 4] - x[:, 0])
        neg = v < 0
        cells[neg] = cells[neg][:, [0, 3, 2, 1, 4]]
    return cells


def mixed_mesh_def(n, seed=3, jitter=0.1, tet_frac=0.3, prism_frac=0.2):
    """A conforming hex/tet/prism/pyramid mesh of the unit cube from an n^3
    hexahedral block.

    Splitting a hexahedron into tetrahedra triangulates all six faces, so a
    tetrahedral region can share faces only with triangulated faces.  The
    block is therefore split in z-slabs: the bottom layers are hexes split
    into 5 tetrahedra (alternating parity, so face diagonals match); one
    layer of transition hexes (5 pyramids and 2 tetrahedra about a centre
    node) joins them to the hexahedral slab; the top layers are hexes split
    into 2 prisms extruded along x (their triangles lie on x = const faces,
    which meet only other prisms or the domain boundary).  round(tet_frac*n)
    layers are tetrahedral including the transition layer and
    round(prism_frac*n) layers are prismatic.  Interior nodes are jittered by
    up to jitter*h per axis from the seed.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    h = 1.0 / n
    m = n + 1
    k_, j_, i_ = np.meshgrid(np.arange(m), np.arange(m), np.arange(m),
                             indexing='ij')
    crd = np.stack([i_ * h, j_ * h, k_ * h], axis=-1).reshape(-1, 3)
    inner = ((i_ > 0) & (i_ < n) & (j_ > 0) & (j_ < n) & (k_ > 0)
             & (k_ < n)).ravel()
    jit = rng.uniform(-jitter * h, jitter * h, size=crd.shape)
    crd = crd + jit * inner[:, None]
    kp = int(round(prism_frac * n))
    kt = max(int(round(tet_frac * n)) - 1, 0)
    ktr = kt if kt > 0 else -1
    kh0 = kt + 1 if kt > 0 else 0

    def nid(i, j, k):
        return (k * m + j) * m + i

    def corners(k):
        j, i = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
        j = j.ravel()
        i = i.ravel()
        v = {}
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    v[(a, b, c)] = nid(i + a, j + b, k + c)
        return i, j, v

    cells = []  # (type, node array [nc, w]) in layer order
    extra = []
    nextra = crd.shape[0]
    for k in range(n):
        i, j, v = corners(k)
        V = lambda a, b, c: v[(a, b, c)]  # noqa: E731
        if k < kt:
            par = (i + j + k) % 2
            ta = [(V(0, 0, 0), V(1, 0, 0), V(0, 1, 0), V(0, 0, 1)),
                  (V(1, 1, 0), V(0, 1, 0), V(1, 0, 0), V(1, 1, 1)),
                  (V(1, 0, 1), V(0, 0, 1), V(1, 1, 1), V(1, 0, 0)),
                  (V(0, 1, 1), V(1, 1, 1), V(0, 0, 1), V(0, 1, 0)),
                  (V(1, 0, 0), V(0, 1, 0), V(0, 0, 1), V(1, 1, 1))]
            tb = [(V(1, 0, 0), V(0, 0, 0), V(1, 1, 0), V(1, 0, 1)),
                  (V(0, 1, 0), V(1, 1, 0), V(0, 0, 0), V(0, 1, 1)),
                  (V(0, 0, 1), V(1, 0, 1), V(0, 1, 1), V(0, 0, 0)),
                  (V(1, 1, 1), V(0, 1, 1), V(1, 0, 1), V(1, 1, 0)),
                  (V(0, 0, 0), V(1, 1, 0), V(1, 0, 1), V(0, 1, 1))]
            tets = np.stack([np.where(par == 0, np.stack(a), np.stack(b))
                             for a, b in zip(ta, tb)], axis=1)  # [3? ...]
            tets = tets.transpose(2, 1, 0).reshape(-1, 4)
            cells.append(('tet', tets))
        elif k == ktr:
            c = nextra + j * n + i
            cen = np.zeros((n * n, 3))
            for key in v:
                cen += crd[v[key]]
            extra.append(cen / 8.0)
            q = (i + j + k - 1) % 2
            pyr = np.stack([
                np.stack([V(0, 0, 1), V(1, 0, 1), V(1, 1, 1), V(0, 1, 1), c]),
                np.stack([V(0, 0, 0), V(0, 1, 0), V(0, 1, 1), V(0, 0, 1), c]),
                np.stack([V(1, 0, 0), V(1, 1, 0), V(1, 1, 1), V(1, 0, 1), c]),
                np.stack([V(0, 0, 0), V(1, 0, 0), V(1, 0, 1), V(0, 0, 1), c]),
                np.stack([V(0, 1, 0), V(1, 1, 0), V(1, 1, 1), V(0, 1, 1), c]),
            ], axis=0)  # [5, 5, nc]
            t1 = np.where(q == 0,
                          np.stack([V(0, 0, 0), V(1, 0, 0), V(1, 1, 0), c]),
                          np.stack([V(1, 0, 0), V(1, 1, 0), V(0, 1, 0), c]))
            t2 = np.where(q == 0,
                          np.stack([V(0, 0, 0), V(1, 1, 0), V(0, 1, 0), c]),
                          np.stack([V(1, 0, 0), V(0, 1, 0), V(0, 0, 0), c]))
            cells.append(('transition', (pyr.transpose(2, 0, 1),
                                         np.stack([t1, t2], axis=0)
                                         .transpose(2, 0, 1))))
            nextra += n * n
        elif k < n - kp:
            hexa = np.stack([V(0, 0, 0), V(1, 0, 0), V(1, 1, 0), V(0, 1, 0),
                             V(0, 0, 1), V(1, 0, 1), V(1, 1, 1), V(0, 1, 1)],
                            axis=1)
            cells.append(('hex', hexa))
        else:
            dg = rng.integers(0, 2, size=n)[j]
            pa = np.where(dg == 0,
                          np.stack([V(0, 0, 0), V(0, 1, 0), V(0, 1, 1),
                                    V(1, 0, 0), V(1, 1, 0), V(1, 1, 1)]),
                          np.stack([V(0, 0, 0), V(0, 1, 0), V(0, 0, 1),
                                    V(1, 0, 0), V(1, 1, 0), V(1, 0, 1)]))
            pb = np.where(dg == 0,
                          np.stack([V(0, 0, 0), V(0, 1, 1), V(0, 0, 1),
                                    V(1, 0, 0), V(1, 1, 1), V(1, 0, 1)]),
                          np.stack([V(0, 1, 0), V(0, 1, 1), V(0, 0, 1),
                                    V(1, 1, 0), V(1, 1, 1), V(1, 0, 1)]))
            pr = np.stack([pa, pb], axis=0).transpose(2, 0, 1).reshape(-1, 6)
            cells.append(('prism', pr))
    if extra:
        crd = np.concatenate([crd] + extra, axis=0)
    tpn = {'tet': 6, 'hex': 5, 'prism': 7, 'pyramid': 8}
    out_t = []
    out_n = []
    for kind, arr in cells:
        if kind == 'transition':
            pyr, tt = arr
            nc = pyr.shape[0]
            pyr = _orient(crd, pyr.reshape(-1, 5).copy(), 'pyramid')
            tt = _orient(crd, tt.reshape(-1, 4).copy(), 'tet')
            blk = np.full((nc, 7, 9), -1, dtype=np.int64)
            typ = np.zeros((nc, 7), dtype=np.int32)
            blk[:, :5, 0] = 5
            blk[:, :5, 1:6] = pyr.reshape(nc, 5, 5)
            typ[:, :5] = 8
            blk[:, 5:, 0] = 4
            blk[:, 5:, 1:5] = tt.reshape(nc, 2, 4)
            typ[:, 5:] = 6
            out_t.append(typ.reshape(-1))
            out_n.append(blk.reshape(-1, 9))
            continue
        arr = _orient(crd, arr.copy(), kind)
        w = arr.shape[1]
        blk = np.full((arr.shape[0], 9), -1, dtype=np.int64)
        blk[:, 0] = w
        blk[:, 1:1 + w] = arr
        out_t.append(np.full(arr.shape[0], tpn[kind], dtype=np.int32))
        out_n.append(blk)
    cltpn = np.concatenate(out_t)
    clnds = np.concatenate(out_n).astype(np.int32)
    clnds[clnds < 0] = 0
    return 3, crd, cltpn, clnds


def blast_state(centroids, gamma=1.4, r=0.1, center=(0.5, 0.5, 0.5),
                p_in=10.0, p_out=0.1, rho=1.0):
    """Spherical blast: rho = 1, p = p_in inside r, p_out outside, u = 0."""
    d = np.sqrt(((centroids - np.asarray(center)) ** 2).sum(1))
    out = np.zeros((centroids.shape[0], 5))
    out[:, 0] = rho
    out[:, 4] = np.where(d < r, p_in, p_out) / (gamma - 1.0)
    return out


SOD_LEFT = (1.0, 0.0, 1.0)
SOD_RIGHT = (0.125, 0.0, 0.1)


def sod_tube(ncoord=2001, time_increment=0.4 / 600, xmin=-1.0, xmax=1.0,
             left=SOD_LEFT, right=SOD_RIGHT, gamma=1.4, split=0.0):
    """A GPU Euler1DCore holding a shock tube ready to march.

    ``ncoord`` uniform points on [xmin, xmax]; the primitive states
    ``left`` / ``right`` (rho, u, p) meet at ``split`` (a point at the split
    takes the right state); gradients start at zero; the initial CFL
    numbers are set by ``setup_march``."""
    from .onedim import Euler1DCore
    core = Euler1DCore(ncoord=ncoord, time_increment=time_increment)
    x = np.linspace(xmin, xmax, num=ncoord)
    prim = np.where((x < split)[:, None], np.asarray(left, float),
                    np.asarray(right, float))
    rho, u, p = prim[:, 0], prim[:, 1], prim[:, 2]
    so0 = np.stack([rho, rho * u, p / (gamma - 1.0) + 0.5 * rho * u * u],
                   axis=1)
    core.coord[...] = x
    core.gamma.fill(gamma)
    core.so0[...] = so0
    core.so1[...] = 0.0
    core.cfl.fill(0.0)
    core.setup_march()
    return core


def rect_mesh_def(nx, ny, extent=(4.0, 1.0), kind='quad', seed=0,
                  jitter=0.0):
    """(ndim, ndcrd, cltpn, clnds) of an nx-by-ny box grid on
    [0, extent[0]] x [0, extent[1]].

    ``kind`` 'quad' keeps one counter-clockwise quadrilateral per box;
    'tri' splits each box into two counter-clockwise triangles along a
    seeded diagonal.  Interior nodes move by up to ``jitter`` box widths
    per axis.  Boxes are numbered row by row from the lower left."""
    if kind not in ('quad', 'tri'):
        raise ValueError("kind must be 'quad' or 'tri'")
    rng = np.random.Generator(np.random.PCG64(seed))
    hx, hy = extent[0] / nx, extent[1] / ny
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing='ij')
    x, y = i * hx, j * hy
    if jitter:
        inner = (i > 0) & (i < nx) & (j > 0) & (j < ny)
        x = np.where(inner, x + rng.uniform(-jitter, jitter, x.shape) * hx,
                     x)
        y = np.where(inner, y + rng.uniform(-jitter, jitter, y.shape) * hy,
                     y)
    ndcrd = np.stack([x.ravel(), y.ravel()], axis=1)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing='ij')
    a = (jj * (nx + 1) + ii).ravel()
    b, c, d = a + 1, a + nx + 2, a + nx + 1
    if kind == 'quad':
        clnds = np.stack([np.full_like(a, 4), a, b, c, d], axis=1)
        cltpn = np.full(a.size, 3, dtype=np.int32)
    else:
        up = rng.integers(0, 2, size=a.size).astype(bool)
        t0 = np.where(up[:, None], np.stack([a, b, c], 1),
                      np.stack([a, b, d], 1))
        t1 = np.where(up[:, None], np.stack([a, c, d], 1),
                      np.stack([b, c, d], 1))
        tri = np.stack([t0, t1], axis=1).reshape(-1, 3)
        clnds = np.concatenate([np.full((tri.shape[0], 1), 3), tri], axis=1)
        cltpn = np.full(tri.shape[0], 4, dtype=np.int32)
    return 2, ndcrd, cltpn, clnds.astype(np.int32)


def edge_faces(bndfcs, fccnd, extent, tol=1e-9):
    """Boundary faces of a rectangle [0, X] x [0, Y] split by edge:
    (left, top, bottom, right), each in ascending face order.  ``bndfcs``
    holds face indices into ``fccnd`` (body rows)."""
    f = np.sort(np.asarray(bndfcs))
    c = np.asarray(fccnd)[f]
    left = f[np.abs(c[:, 0]) < tol]
    right = f[np.abs(c[:, 0] - extent[0]) < tol]
    bottom = f[np.abs(c[:, 1]) < tol]
    top = f[np.abs(c[:, 1] - extent[1]) < tol]
    return left, top, bottom, right
