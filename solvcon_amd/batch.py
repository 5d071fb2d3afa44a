"""Batched form: thousands of independent Euler solves in one device march.

The instances (2D or 3D meshes of any size, each with its own initial
state, time increment and step count) are joined into one disjoint-union
``StaticMesh`` built on the device, and one batched ``EulerCore`` marches
all of them.  Its cells are sorted by (cell type, instance rank, Hilbert
code), where ranks order the instances by descending step count.  Every
substep is one fused kernel launch per cell type over the cells of the
instances still marching, which form a prefix of each type group.  No
instance waits for another, and none sees another's data.  Each result is
bit-identical to marching that instance alone with the same dt, because a
cell's arithmetic does not depend on where its row sits.

There is no reference counterpart; the per-instance semantics are
``EulerCore.march`` (cpp/solvcon/multidim/euler_march.cpp:339-359).
"""

import numpy as np

from .gen import tri_mesh_def, riemann_box, quadrant_state
from .mesh import StaticMesh
from .multidim import EulerCore

__all__ = ['union_def', 'BatchedEuler', 'batch_config']


def union_def(defs):
    """Join mesh definitions (ndim, ndcrd, cltpn, clnds) into one; returns
    the joined definition, the instance of every cell, and the cell and
    node offsets of every instance."""
    ndim = defs[0][0]
    if any(d[0] != ndim for d in defs):
        raise ValueError('union_def: instances must share ndim')
    width = max(np.asarray(d[3]).shape[1] for d in defs)
    ncell = np.array([len(d[2]) for d in defs], dtype=np.int64)
    nnode = np.array([len(d[1]) for d in defs], dtype=np.int64)
    coff = np.concatenate([[0], np.cumsum(ncell)])
    noff = np.concatenate([[0], np.cumsum(nnode)])
    ndcrd = np.concatenate([np.asarray(d[1], dtype=np.float64) for d in defs])
    cltpn = np.concatenate([np.asarray(d[2], dtype=np.int32) for d in defs])
    if noff[-1] > np.iinfo(np.int32).max:
        raise ValueError('union_def: more nodes than int32 indexes')
    clnds = np.zeros((coff[-1], width), dtype=np.int32)
    for k, d in enumerate(defs):
        # The node ids of every cell (its first clnds[:, 0] entries) are
        # shifted by the instance's node offset.  When every cell fills
        # its row, the shift is one add fused with the copy.
        c = np.asarray(d[3], dtype=np.int32)
        blk = clnds[coff[k]:coff[k + 1], :c.shape[1]]
        full = (c[:, 0] == c.shape[1] - 1).all()
        if full and c.shape[1] == width:
            # Contiguous rows: shift every entry, then restore the counts.
            flat = clnds[coff[k]:coff[k + 1]].reshape(-1)
            np.add(c.reshape(-1), np.int32(noff[k]), out=flat)
            flat[0::width] = c[:, 0]
            continue
        blk[:, 0] = c[:, 0]
        if full:
            np.add(c[:, 1:], np.int32(noff[k]), out=blk[:, 1:])
        else:
            blk[:, 1:] = c[:, 1:]
            mask = np.arange(1, c.shape[1])[None, :] <= c[:, :1]
            blk[:, 1:][mask] += np.int32(noff[k])
    inst = np.repeat(np.arange(len(defs), dtype=np.int32), ncell)
    return (ndim, ndcrd, cltpn, clnds), inst, coff, noff


def build_mesh(defn, edge=False):
    """A device StaticMesh of a mesh definition (interior, boundary,
    ghosts)."""
    ndim, ndcrd, cltpn, clnds = defn
    ndcrd = np.asarray(ndcrd, dtype=np.float64)
    cltpn = np.asarray(cltpn, dtype=np.int32)
    clnds = np.asarray(clnds, dtype=np.int32)
    mh = StaticMesh(ndim=ndim, nnode=ndcrd.shape[0], nface=0,
                    ncell=cltpn.shape[0])
    mh.ndcrd.ndarray[...] = ndcrd
    mh.cltpn.ndarray[...] = cltpn
    mh.clnds.ndarray[:, :clnds.shape[1]] = clnds
    mh.build_interior(do_metric=True, build_edge=edge)
    mh.build_boundary()
    mh.build_ghost()
    return mh


class BatchedEuler(object):
    """Independent Euler solves marched together on the GPU.

    ``defs`` are mesh definitions; ``states`` the conserved body rows of
    each instance ([ncell_k, ndim + 2]), or a callable of the cell centroids
    [ncell_k, ndim] and the instance index; ``steps`` the step count of each
    instance; ``dts`` their time increments, or None for the dt that puts
    each instance's largest initial CFL number at ``cfl``.  Every instance
    gets non-reflective boundaries and gamma.  ``chunk_cells`` sets
    ``EulerCore.batch_chunk``.
    """

    def __init__(self, defs, states, steps, dts=None, gamma=1.4, cfl=0.8,
                 sfc=True, chunk_cells=0):
        self.ninst = len(defs)
        self.steps = np.asarray(steps, dtype=np.int64)
        if self.steps.size != self.ninst:
            raise ValueError('BatchedEuler: one step count per instance')
        defn, inst, self.coff, self.noff = union_def(defs)
        self.ndim = defn[0]
        self.mesh = build_mesh(defn)
        mh = self.mesh
        self.ngst = mh.ngstcell
        core = EulerCore(mh, 1.0, sfc=sfc, batch=(inst, self.steps))
        core.init_solution(gamma, 1.0, [0.0] * self.ndim, 1.0)
        so0n = core.so0n.ndarray
        body = so0n[self.ngst:]
        if callable(states):
            clcnd = mh.clcnd.ndarray[self.ngst:]
            for k in range(self.ninst):
                a, b = self.coff[k], self.coff[k + 1]
                body[a:b] = states(clcnd[a:b], k)
        else:
            for k in range(self.ninst):
                body[self.coff[k]:self.coff[k + 1]] = states[k]
        del so0n, body
        core.add_nonrefl(mh.bndfcs.ndarray[:, 0])
        core.bc_soln()
        core.bc_dsoln()
        if dts is None:
            self.dts = core.batch_cfl_dt(cfl)
        else:
            self.dts = np.asarray(dts, dtype=np.float64)
            core.batch_set_dt(self.dts)
        core.batch_chunk = chunk_cells
        self.core = core
        self.ncells = np.diff(self.coff)

    def march(self):
        """March every instance its own step count; returns device ms."""
        return self.core.march_batch()

    def done_ms(self):
        """Per instance, device ms from the start of the last march until
        its last step ran (instances with fewer steps finish earlier)."""
        return self.core.batch_done_ms()

    @property
    def cell_updates(self):
        return int((self.ncells * self.steps).sum())

    def body(self, name, k):
        """Body rows of table ``name`` (so0n, so1n, cflo, ...) of instance
        k, a copy."""
        a = getattr(self.core, name).ndarray
        return a[self.ngst + self.coff[k]:self.ngst + self.coff[k + 1]].copy()


def batch_config(ninst=4096, seed=4, lo=1e3, hi=2e5, smin=100, smax=1000):
    """The batched workload: per instance the quads per side m (mesh of
    2 m^2 triangles, sizes log-uniform in [lo, hi]), the mesh seed, the
    Riemann-box seed and the step count (uniform in [smin, smax])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = np.exp(rng.uniform(np.log(lo), np.log(hi), ninst))
    m = np.maximum(np.round(np.sqrt(n / 2)).astype(np.int64), 1)
    steps = rng.integers(smin, smax + 1, ninst)
    mseed = rng.integers(0, 2 ** 31, ninst)
    rseed = rng.integers(0, 2 ** 31, ninst)
    return m, mseed, rseed, steps


def config_defs(m, mseed):
    return [tri_mesh_def(int(mk), seed=int(sk)) for mk, sk in zip(m, mseed)]


def config_states(rseed, gamma=1.4):
    boxes = [riemann_box(int(s)) for s in rseed]

    def state(cen, k):
        rho, p = boxes[k]
        return quadrant_state(cen, rho, p, gamma)
    return state
