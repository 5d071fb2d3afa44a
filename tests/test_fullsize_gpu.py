"""Parity at the benchmark configurations' own meshes.

The 2D configuration (1024 x 1024 quads split into 2,097,152 jittered
triangles, Riemann box, non-reflective boundaries, dt from the CFL 0.8 rule)
is marched ten steps by the device and by the oracle on all host cores; every
row of the solution tables must be equal.  The 3D configuration's generator
and state (blast, slip walls on three sides) are checked the same way at 60^3
(525,600 cells, five steps): its full 19.3M-cell mesh takes minutes to build
on the CPU.
"""

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu



def _run(orc, defn, state, bcs, steps):
    import os
    from solvcon_amd.gen import build_static_mesh
    import solvcon_amd as sc
    mh = build_static_mesh(defn)
    ec = sc.EulerCore(mesh=mh, time_increment=1.0)
    ec.init_solution(1.4, 1.0, [0.0] * mh.ndim, 1.0)
    ngst = ec.ngstcell
    st = state(mh.clcnd.ndarray[ngst:])
    ec.so0n.ndarray[ngst:] = st
    groups = bcs(mh)
    for kind, faces in groups:
        getattr(ec, 'add_' + kind)(faces)
    ec.bc_soln()
    ec.bc_dsoln()
    dt = ec.cfl_dt(0.8)
    ec.set_time_increment(dt)
    print('case: device march', ec.ncell, 'cells', flush=True)
    ec.march(steps)
    m = orc.make_mesh(*defn)
    oe = orc.Euler(m, dt, nthreads=min(os.cpu_count() or 1, 16))
    oe.init_solution(1.4, 1.0, [0.0] * m.ndim, 1.0)
    oe.so0n[oe.ngstcell:] = st
    kinds = {'nonrefl': orc.BC_NONREFL, 'slipwall': orc.BC_SLIPWALL}
    for kind, faces in groups:
        oe.add_bc(kinds[kind], np.asarray(faces))
    oe.call('bc_soln')
    oe.call('bc_dsoln')
    print('case: oracle march', flush=True)
    oe.march(steps)
    for t in ('so0n', 'so1n', 'cflo', 'cflc'):
        a = getattr(ec, t).ndarray
        b = getattr(oe, t)
        assert a.shape == b.shape, t
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError('%s differs at %d places, first %s'
                                 % (t, len(bad), tuple(bad[0])))


def test_2d_config_mesh_bitwise(orc):
    defn, state, bcs, _ = bench.workload_2d(1024)
    _run(orc, defn, state, bcs, 10)


def test_3d_config_generator_bitwise(orc):
    defn, state, bcs, _ = bench.workload_3d(60)
    _run(orc, defn, state, bcs, 5)
