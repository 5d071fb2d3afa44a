"""Committed march fixtures of tiny meshes (tests/golden/march_*.npz).

tests/golden/make_golden.py writes them with the oracle: a Riemann-box state
on a triangle fan, a quad/triangle mix, a quad channel, a single tetrahedron,
jittered hexahedra and a mixed 3D block, with alternating non-reflective and
slip-wall boundaries, marched five steps.  The oracle must still reproduce
them bit for bit (CPU), and so must the device march (GPU): every row,
ghosts included, of so0n, so1n, cflo and cflc.
"""

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))

import make_golden as MG  # noqa: E402

CASES = sorted(MG.march_cases())


def _fixture(name):
    with np.load(os.path.join(HERE, 'golden', 'march_%s.npz' % name)) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize('name', CASES)
def test_oracle_reproduces_fixture(orc, name):
    want = _fixture(name)
    got = MG.march_oracle(MG.march_cases()[name])
    for t in MG.MARCH_TABLES:
        assert np.array_equal(got[t], want[t]), t


@pytest.mark.gpu
@pytest.mark.parametrize('name', CASES)
def test_device_march_equals_fixture(name):
    print('case: golden march', name, flush=True)
    import pair
    import solvcon_amd as sc
    want = _fixture(name)
    g = pair.gpu_mesh(MG.march_cases()[name])
    ec = sc.EulerCore(mesh=g, time_increment=MG.MARCH_DT)
    ec.init_solution(MG.GAMMA, 1.0, [0.0] * g.ndim, 1.0)
    ec.so0n.ndarray[ec.ngstcell:] = MG.march_state(
        g.clcnd.ndarray[g.ngstcell:], g.ndim)
    faces = g.bndfcs.ndarray[:, 0].copy()
    if faces[0::2].size:
        ec.add_nonrefl(faces[0::2])
    if faces[1::2].size:
        ec.add_slipwall(faces[1::2])
    ec.bc_soln()
    ec.bc_dsoln()
    ec.march(MG.MARCH_STEPS)
    for t in MG.MARCH_TABLES:
        got = getattr(ec, t).ndarray
        assert got.shape == want[t].shape, t
        # IEEE equality (the sign of an exact zero is not compared).
        assert np.array_equal(got, want[t]), t
