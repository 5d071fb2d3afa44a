"""Write the golden vectors of tests/golden from the oracle.

The reference library cannot be built or imported in this environment (it
needs C++23 <format>/<mdspan>, absent from GCC 11's libstdc++), so the
vectors come from the oracle, the CPU restatement under oracle/, whose
functions are pinned by the reference's own goldens and closed forms in
tests/test_oracle_*.py.  Run from the repository root:

    python tests/golden/make_golden.py

- sod_2001_600.npz: the Sod tube (2001 points) after 600 march_alpha2 steps.
- march_<case>.npz: tiny meshes (triangle fan, quad/triangle mix, quad
  channel, single tetrahedron, jittered hexahedra, mixed 3D block) with a
  Riemann-box state, alternating non-reflective and slip-wall boundaries,
  marched MARCH_STEPS steps of dt MARCH_DT: every row (ghosts first) of
  so0n, so1n, cflo and cflc.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

GAMMA = 1.4
MARCH_DT = 2e-3
MARCH_STEPS = 5
MARCH_SEED = 7
MARCH_TABLES = ('so0n', 'so1n', 'cflo', 'cflc')


def sod():
    from oracle import oracle as O
    o = O.sod_setup(ncoord=2001, time_increment=0.4 / 600)
    o.march(2, 600)
    np.savez_compressed(os.path.join(HERE, 'sod_2001_600.npz'),
                        so0=o.so0, so1=o.so1, cfl=o.cfl)


def march_cases():
    import meshdefs as MD
    from solvcon_amd.gen import mixed_mesh_def
    return {
        'triangles': MD.TRIANGLES,
        'mixed2d': MD.MIXED,
        'channel': MD.quad_channel(6, 3, 4.0, 2.0),
        'tet': MD.TETRA,
        'hexblk': MD.hex_block(3, 2, 2, jitter=0.2, seed=5),
        'mixed3d': mixed_mesh_def(4, seed=3),
    }


def march_state(clcnd_body, ndim):
    """The Riemann-box body state of the golden marches, from the mesh
    centroids of the body cells."""
    from solvcon_amd.gen import riemann_box, quadrant_state
    rho, p = riemann_box(MARCH_SEED)
    st = quadrant_state(np.asarray(clcnd_body)[:, :2], rho, p, GAMMA)
    if ndim == 3:
        st = np.insert(st, 3, 0.0, axis=1)
    return st


def march_oracle(defn):
    """Run one golden march on the oracle; returns {table: array}."""
    import pair
    from oracle import oracle as O
    m = pair.orc_mesh(O, defn)
    oe = O.Euler(m, MARCH_DT)
    oe.init_solution(GAMMA, 1.0, [0.0] * m.ndim, 1.0)
    oe.so0n[oe.ngstcell:] = march_state(m.clcnd[m.ngstcell:], m.ndim)
    faces = np.asarray(m.bndfcs[:, 0])
    if faces[0::2].size:
        oe.add_bc(O.BC_NONREFL, faces[0::2])
    if faces[1::2].size:
        oe.add_bc(O.BC_SLIPWALL, faces[1::2])
    oe.call('bc_soln')
    oe.call('bc_dsoln')
    oe.march(MARCH_STEPS)
    return {t: np.array(getattr(oe, t)) for t in MARCH_TABLES}


def marches():
    for name, defn in march_cases().items():
        out = march_oracle(defn)
        np.savez_compressed(os.path.join(HERE, 'march_%s.npz' % name), **out)


if __name__ == '__main__':
    sod()
    marches()
