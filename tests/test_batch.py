"""Batched form: many independent solves in one device march.

CPU: the disjoint union of mesh definitions keeps every instance's cells
and nodes.  GPU: every instance marched in the batch equals the oracle's
march of that instance alone (the reference's EulerCore::march,
euler_march.cpp:339-359, restated) with the same dt and step count, value
for value, and is bit-identical, on every body and ghost row of every
table, to the same instance marched alone by the plain EulerCore.  That
holds for mixed sizes, step counts, cell types and dimensions, for the
committed tiny-mesh march fixtures taken as instances, and for a sample of
the batched workload's own generator.
"""

import numpy as np
import pytest

import meshdefs as MD


def _defs2d():
    from solvcon_amd.gen import tri_mesh_def
    d = [tri_mesh_def(m, seed=10 + m) for m in (3, 5, 8, 4)]
    d.append(MD.quad_channel(6, 3, 2.0, 1.0))
    d.append(tri_mesh_def(2, seed=7))
    return d


def test_union_def_keeps_instances():
    from solvcon_amd.batch import union_def
    defs = _defs2d()
    (ndim, ndcrd, cltpn, clnds), inst, coff, noff = union_def(defs)
    assert ndim == 2 and inst.size == coff[-1] == cltpn.size
    for k, d in enumerate(defs):
        c = np.asarray(d[3])
        u = clnds[coff[k]:coff[k + 1]]
        assert (inst[coff[k]:coff[k + 1]] == k).all()
        assert (u[:, 0] == c[:, 0]).all()
        for r in range(c.shape[0]):
            n = c[r, 0]
            got = ndcrd[u[r, 1:1 + n]]
            want = np.asarray(d[1], dtype=np.float64)[c[r, 1:1 + n]]
            assert (got == want).all()


# ----------------------------------------------------------------- GPU -----

TABLES = ('so0n', 'so1n', 'so0c', 'so1c', 'so0t', 'cflo', 'cflc')


def _alone(defn, state, dt, steps, gamma=1.4):
    from solvcon_amd.batch import build_mesh
    from solvcon_amd import EulerCore
    mh = build_mesh(defn)
    ec = EulerCore(mh, dt)
    ec.init_solution(gamma, 1.0, [0.0] * defn[0], 1.0)
    ec.so0n.ndarray[mh.ngstcell:] = state
    ec.add_nonrefl(mh.bndfcs.ndarray[:, 0])
    ec.bc_soln()
    ec.bc_dsoln()
    ec.march(steps)
    return mh, ec


def _oracle_alone(orc, defn, state, dt, steps, gamma=1.4):
    """The oracle's march of one instance: its own mesh, state, dt and
    step count, non-reflective boundaries."""
    o = orc.make_mesh(*defn)
    oe = orc.Euler(o, float(dt))
    oe.init_solution(gamma, 1.0, [0.0] * o.ndim, 1.0)
    oe.so0n[o.ngstcell:] = state
    oe.add_bc(orc.BC_NONREFL, o.bndfcs[:, 0].copy())
    oe.call('bc_soln')
    oe.call('bc_dsoln')
    oe.march(int(steps))
    return o, oe


def _check_oracle(orc, b, k, defn, state, steps, gamma=1.4):
    # Every body row of the batched instance equals the oracle's alone.
    _, oe = _oracle_alone(orc, defn, state, b.dts[k], steps, gamma)
    for name in ('so0n', 'so1n', 'cflo', 'cflc'):
        want = getattr(oe, name)[oe.ngstcell:]
        got = b.body(name, k)
        assert np.isfinite(want).all(), (k, name)
        assert np.array_equal(got, want), (k, name, 'oracle')


def _check_batch(defs, states, steps, chunk=0, orc=None):
    from solvcon_amd.batch import BatchedEuler
    b = BatchedEuler(defs, states, steps, chunk_cells=chunk)
    assert b.core.batch_chunk == chunk
    b.march()
    gofs = 0
    for k, d in enumerate(defs):
        print('  instance', k, 'cells', len(d[2]), 'steps', steps[k],
              flush=True)
        if orc is not None:
            _check_oracle(orc, b, k, d, states[k], steps[k])
        mh, ec = _alone(d, states[k], b.dts[k], steps[k])
        ng = mh.ngstcell
        for name in TABLES:
            a = getattr(ec, name).ndarray
            u = getattr(b.core, name).ndarray
            body = u[b.ngst + b.coff[k]:b.ngst + b.coff[k + 1]]
            assert np.array_equal(a[ng:], body), (k, name)
            # Ghost rows of instance k: a block of the union's ghost rows in
            # reverse order (ghost -1 is the last ghost row).
            ug = u[:b.ngst][::-1][gofs:gofs + ng]
            assert np.array_equal(a[:ng][::-1], ug), (k, name, 'ghost')
        gofs += ng
    assert gofs == b.ngst


@pytest.mark.gpu
@pytest.mark.parametrize('fans', [0, 1, 2])
@pytest.mark.parametrize('chunk', [0, 150])
def test_gpu_batch_equals_alone_2d(orc, chunk, fans):
    # chunk 150: rank chunks of about 150 cells, marched one after another.
    print('case: batch of 2D instances, chunk', chunk, 'fans', fans,
          flush=True)
    from solvcon_amd.gen import riemann_box, quadrant_state
    from solvcon_amd.batch import build_mesh
    # 3-triangle fans: with one the triangle count is odd and no range runs
    # the geometry prefetch; with two it is even, and with chunks some rank
    # ranges start at an odd cell and run without it.
    defs = _defs2d()
    states = []
    for k, d in enumerate(defs):
        mh = build_mesh(d)
        cen = mh.clcnd.ndarray[mh.ngstcell:].copy()
        rho, p = riemann_box(30 + k)
        states.append(quadrant_state(cen / cen.max(0), rho, p))
    for k in range(fans):
        defs.append(MD.TRIANGLES)
        states.append(np.tile([1.0, 0.1 * (k + 1), 0.0, 2.5], (3, 1)))
    _check_batch(defs, states, [4, 7, 2, 7, 5, 1, 6, 3][:len(defs)], chunk,
                 orc)


@pytest.mark.gpu
@pytest.mark.parametrize('chunk', [0, 200])
def test_gpu_batch_equals_alone_3d(orc, chunk):
    print('case: batch of 3 3D mixed instances, chunk', chunk, flush=True)
    from solvcon_amd.gen import mixed_mesh_def, blast_state
    from solvcon_amd.batch import build_mesh
    defs = [mixed_mesh_def(3, seed=3), mixed_mesh_def(4, seed=5),
            MD.hex_block(3, 2, 2, jitter=0.1, seed=2)]
    states = []
    for d in defs:
        mh = build_mesh(d)
        cen = mh.clcnd.ndarray[mh.ngstcell:].copy()
        states.append(blast_state(cen, r=0.4, center=cen.mean(0)))
    _check_batch(defs, states, [2, 3, 1], chunk, orc)


@pytest.mark.gpu
def test_gpu_batch_config_sample(orc):
    # Instances of the batched workload's own generator, sizes 50-3000
    # cells and 20-120 steps, against the oracle and the plain march.
    print('case: batch config sample', flush=True)
    from solvcon_amd.batch import (batch_config, config_defs, config_states,
                                   BatchedEuler, build_mesh)
    m, ms, rs, steps = batch_config(ninst=12, lo=50, hi=3000, smin=20,
                                    smax=120)
    defs = config_defs(m, ms)
    st = config_states(rs)
    b = BatchedEuler(defs, st, steps)
    b.march()
    for k in range(len(defs)):
        mh = build_mesh(defs[k])
        cen = mh.clcnd.ndarray[mh.ngstcell:].copy()
        _check_oracle(orc, b, k, defs[k], st(cen, k), steps[k])
        _, ec = _alone(defs[k], st(cen, k), b.dts[k], steps[k])
        assert np.array_equal(ec.so0n.ndarray[ec.ngstcell:],
                              b.body('so0n', k))


@pytest.mark.gpu
def test_gpu_batch_golden_instances(orc):
    # The committed tiny-mesh march fixtures (tests/golden/march_*.npz),
    # all 2D ones in one batch and all 3D ones in another, each with its
    # fixture's state, dt and step count but non-reflective boundaries
    # throughout: every instance equals the oracle's march alone.
    print('case: batch of the golden fixture meshes', flush=True)
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
    import make_golden as MG
    from solvcon_amd.batch import BatchedEuler, build_mesh
    cases = MG.march_cases()
    for nd in (2, 3):
        import pair
        names = sorted(n for n in cases if cases[n][0] == nd)
        defs = [pair.as_arrays(cases[n]) for n in names]
        states = []
        for d in defs:
            mh = build_mesh(d)
            states.append(MG.march_state(mh.clcnd.ndarray[mh.ngstcell:], nd))
        steps = [MG.MARCH_STEPS + i for i in range(len(defs))]
        b = BatchedEuler(defs, states, steps, dts=[MG.MARCH_DT] * len(defs),
                         gamma=MG.GAMMA)
        b.march()
        for k, d in enumerate(defs):
            print('  ', names[k], flush=True)
            _check_oracle(orc, b, k, d, states[k], steps[k])


@pytest.mark.gpu
def test_gpu_batch_guards():
    print('case: batch API guards', flush=True)
    from solvcon_amd.batch import BatchedEuler
    from solvcon_amd.gen import tri_mesh_def
    from solvcon_amd import EulerCore
    d = tri_mesh_def(3, seed=1)
    st = np.zeros((18, 4))
    st[:, 0] = 1.0
    st[:, 3] = 2.5
    b = BatchedEuler([d, d], [st, st], [2, 3], dts=[0.01, 0.02])
    with pytest.raises(RuntimeError):
        b.core.march(1)
    with pytest.raises(RuntimeError):
        b.core.calc_soln()
    b.march()
    ec = EulerCore(b.mesh, 0.01)
    with pytest.raises(RuntimeError):
        ec.batch_set_dt([0.1])
    with pytest.raises(ValueError):
        BatchedEuler([d], [st], [0])
    with pytest.raises(RuntimeError):
        ec.batch_done_ms()


@pytest.mark.gpu
@pytest.mark.parametrize('chunk', [0, 150])
def test_gpu_batch_completion_times(chunk):
    # Each instance's completion time is when its last step ran: within
    # the march, and (one chunk) ordered by step count, equal for equal
    # counts, the longest at the march's end.
    print('case: batch completion times, chunk', chunk, flush=True)
    from solvcon_amd.batch import BatchedEuler
    from solvcon_amd.gen import tri_mesh_def
    defs = [tri_mesh_def(m, seed=m) for m in (4, 6, 3, 5, 7, 4, 6, 5)]
    steps = [4, 7, 2, 7, 5, 1, 6, 3]
    states = []
    for d in defs:
        st = np.zeros((len(d[2]), 4))
        st[:, 0] = 1.0
        st[:, 3] = 2.5
        states.append(st)
    b = BatchedEuler(defs, states, steps, dts=[0.01] * 8, chunk_cells=chunk)
    total = b.march()
    done = b.done_ms()
    assert done.shape == (8,)
    assert (done > 0).all() and (done <= total * (1 + 1e-6)).all()
    if chunk == 0:
        for i in range(8):
            for j in range(8):
                if steps[i] < steps[j]:
                    assert done[i] <= done[j], (i, j)
                elif steps[i] == steps[j]:
                    assert done[i] == done[j], (i, j)
        assert done[1] >= 0.5 * total


@pytest.mark.gpu
def test_gpu_batch_graph_equals_eager():
    # The batched march replays one captured hipGraph per range of active
    # ranks (one step each); instances ending at different steps make
    # several ranges with several replays each.  Every table equals the
    # batch marched with its kernels launched one by one.
    print('case: batch graph vs eager', flush=True)
    from solvcon_amd.batch import (batch_config, config_defs, config_states,
                                   BatchedEuler)
    m, ms, rs, steps = batch_config(ninst=6, lo=50, hi=600, smin=3, smax=12)
    defs = config_defs(m, ms)
    st = config_states(rs)
    runs = []
    for graph in (True, False):
        b = BatchedEuler(defs, st, steps)
        b.core.use_graph = graph
        b.march()
        runs.append(b)
    for name in TABLES:
        assert np.array_equal(getattr(runs[0].core, name).ndarray,
                              getattr(runs[1].core, name).ndarray), name
