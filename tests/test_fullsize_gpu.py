"""Full-run parity at the benchmark configurations' own meshes.

- The 2D configuration (1024 x 1024 quads split into 2,097,152 jittered
  triangles, Riemann box, non-reflective boundaries, dt from the CFL 0.8
  rule) is marched its whole 500 steps by the device and by the oracle on
  all host cores.
- The 3D configuration's generator and state (mixed hex/tet/prism/pyramid
  block, blast, slip walls on three sides) are marched the configuration's
  whole 200 steps at 60^3 (525,600 cells): the full 19.3M-cell mesh takes
  minutes to build on the CPU, and its oracle march hours.

The bar is a relative error of at most 1e-12 in every column of so0n,
so1n, cflo and cflc (max |device - oracle| / max |oracle| per column); the
tests also report whether every value is equal, which they are.
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
    worst = 0.0
    for t in ('so0n', 'so1n', 'cflo', 'cflc'):
        a = getattr(ec, t).ndarray
        b = getattr(oe, t)
        assert a.shape == b.shape, t
        assert np.isfinite(b).all(), t
        a2 = a.reshape(a.shape[0], -1)
        b2 = b.reshape(b.shape[0], -1)
        scale = np.maximum(np.abs(b2).max(axis=0), 1e-300)
        rel = float((np.abs(a2 - b2).max(axis=0) / scale).max())
        nbad = int((a != b).sum())
        print('case: %s max relative error %.3g, %d values differ'
              % (t, rel, nbad), flush=True)
        assert rel <= 1e-12, t
        worst = max(worst, rel)
    return worst


@pytest.mark.timeout(900)
def test_2d_config_full_run(orc):
    defn, state, bcs, _ = bench.workload_2d(1024)
    assert _run(orc, defn, state, bcs, 500) == 0.0


@pytest.mark.timeout(900)
def test_3d_config_generator_full_run(orc):
    defn, state, bcs, _ = bench.workload_3d(60)
    assert _run(orc, defn, state, bcs, 200) == 0.0
