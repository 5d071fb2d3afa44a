"""1D Euler CESE: oracle pinning (CPU) and HIP parity (GPU).

CPU:
- the exact Riemann solution (oracle/riemann.py) against the reference's
  ShockTube goldens (tests/test_onedim_euler.py:113-141), the jump
  conditions across its shock and the Riemann invariant across its fan;
- the oracle's derived-field accessors (Euler1DCore.hpp:106-168,
  Euler1DCore.cpp:37-68,174-205) and update_cfl (Euler1DCore.cpp:70-94)
  against closed forms on hand-chosen states and a non-uniform grid;
- the fine-grained stage API against march_alpha2 bit for bit (:53-71), an
  L1 bound against the exact Sod solution, and the committed golden vectors.
GPU: the HIP march against the oracle bit for bit on the Sod tube (2001
points, 600 steps) and on sizes that take the multi-kernel path.
"""

import math
import os

import numpy as np
import pytest

from oracle.riemann import RiemannSolution

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
SOD = ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1))


def test_riemann_sod_goldens():
    # The values the reference's ShockTube tests pin, rtol 1e-7.
    s = RiemannSolution(*SOD, gamma=1.4)
    assert (s.left_wave, s.right_wave) == ('rarefaction', 'shock')
    np.testing.assert_allclose(s.p_star / 0.1, 3.0313017805064697, rtol=1e-7)
    np.testing.assert_allclose([s.rho_star_left, s.rho_star_right],
                               [0.42631942817849494, 0.2655737117053072],
                               rtol=1e-7)
    np.testing.assert_allclose(s.p_star, 0.303130178050647, rtol=1e-7)
    np.testing.assert_allclose(s.u_star, 0.9274526200489503, rtol=1e-7)
    np.testing.assert_allclose(
        s.locations(t=0.1),
        [-0.11832159566199232, -0.007027281256118345, 0.09274526200489504,
         0.17521557320301784], rtol=1e-7)
    # The fan head moves at the left sound speed.
    assert s.locations(1.0)[0] == -math.sqrt(1.4)


def _flux(rho, u, p, g=1.4):
    e = p / (g - 1) + 0.5 * rho * u * u
    return np.array([rho * u, rho * u * u + p, u * (e + p)]), \
        np.array([rho, rho * u, e])


@pytest.mark.parametrize('left,right', [
    SOD, ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01)),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4)),       # two rarefactions
    ((1.0, 2.0, 0.4), (1.0, -2.0, 0.4)),       # two shocks
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.095))])
def test_riemann_jump_conditions(left, right):
    # Rankine-Hugoniot across every shock, u + 2c/(g-1) (left fan) or
    # u - 2c/(g-1) (right fan) constant through every rarefaction, and
    # continuity of p and u across the contact.
    g = 1.4
    s = RiemannSolution(left, right, gamma=g)
    lh, lt, us, rt, rh = s.speeds()
    assert lh <= lt <= us <= rt <= rh
    stars = ((s.rho_star_left, s.u_star, s.p_star),
             (s.rho_star_right, s.u_star, s.p_star))
    for side, wave, star, (a, b), sign in (
            (left, s.left_wave, stars[0], (lh, lt), -1),
            (right, s.right_wave, stars[1], (rt, rh), +1)):
        if wave == 'shock':
            fa, ua = _flux(*side)
            fb, ub = _flux(*star)
            np.testing.assert_allclose(fb - fa, a * (ub - ua), rtol=1e-10,
                                       atol=1e-10 * np.abs(fa).max())
        else:
            xi = np.linspace(a, b, 7)
            rho, u, p = s.sample(xi, 1.0)
            c = np.sqrt(g * p / rho)
            inv = u - sign * 2 * c / (g - 1)
            c0 = math.sqrt(g * side[2] / side[0])
            np.testing.assert_allclose(inv, side[1] - sign * 2 * c0 / (g - 1),
                                       rtol=1e-12)
            np.testing.assert_allclose(p / rho ** g, side[2] / side[0] ** g,
                                       rtol=1e-12)
    # Far from the origin the initial states persist.
    rho, u, p = s.sample([-1e3, 1e3], 1.0)
    assert (rho[0], u[0], p[0]) == tuple(map(float, left))
    assert (rho[1], u[1], p[1]) == tuple(map(float, right))


def test_riemann_mirror_symmetry():
    # Reflecting x -> -x swaps the states and negates the velocities.
    a = RiemannSolution((1.0, 0.3, 1.0), (0.125, -0.1, 0.1))
    b = RiemannSolution((0.125, 0.1, 0.1), (1.0, -0.3, 1.0))
    assert b.p_star == pytest.approx(a.p_star, rel=1e-14)
    assert b.u_star == pytest.approx(-a.u_star, rel=1e-14)
    x = np.linspace(-0.9, 0.9, 37)
    ra, ua, pa = a.sample(x, 0.3)
    rb, ub, pb = b.sample(-x, 0.3)
    np.testing.assert_allclose(rb, ra, rtol=1e-12)
    np.testing.assert_allclose(ub, -ua, rtol=1e-12, atol=1e-14)
    with pytest.raises(ValueError):
        RiemannSolution((1.0, -20.0, 1.0), (1.0, 20.0, 1.0))


# Hand-chosen conserved states (rho, m, E), exact in binary: their fields
# follow from the ideal-gas relations p = (g-1) rho e, T = p / (rho R),
# s = p / rho^g with e the specific internal energy E/rho - u^2/2.
_STATES = np.array([[2.0, 4.0, 10.0], [0.5, -1.0, 4.0], [1.0, 0.0, 2.5],
                    [4.0, 2.0, 0.125]])
_R = 8.31446261815324


def test_oracle_accessors_closed_form(orc):
    g = 1.4
    o = orc.Euler1D(5, 0.1)
    o.so0[:4] = _STATES
    o.so0[4] = [1.0, 0.0, 1.0]
    rho, m, e = _STATES.T
    u = m / rho
    eint = e / rho - 0.5 * u * u               # 3, 2, 2.5, -0.09375
    p = (g - 1) * rho * eint
    assert eint.tolist() == [3.0, 6.0, 2.5, -0.09375]
    np.testing.assert_allclose(o.derived('density')[:4], rho, rtol=0)
    np.testing.assert_allclose(o.derived('velocity')[:4], u, rtol=1e-15)
    np.testing.assert_allclose(o.derived('internal_energy')[:4], eint,
                               rtol=1e-15)
    np.testing.assert_allclose(o.derived('pressure')[:4], p, rtol=1e-15)
    np.testing.assert_allclose(o.derived('temperature')[:4],
                               p / (rho * _R), rtol=1e-14)
    # Entropy against p / rho^g; a negative pressure stays negative.
    ent = o.derived('entropy')[:4]
    np.testing.assert_allclose(ent, p / rho ** g, rtol=1e-14)
    assert ent[3] < 0
    # The reference's velocity divides by rho + 1e-100: zero density gives
    # a finite value instead of a division by zero.
    o.so0[0] = [0.0, 1e-300, 1.0]
    assert o.derived('velocity')[0] == pytest.approx(1e-200)


def test_oracle_update_cfl_closed_form(orc):
    # Non-uniform grid: each CFL number uses the shorter of its two
    # neighbouring intervals; negative pressure is clamped to 0, leaving
    # only the convective speed |u|.
    g, dt = 1.4, 0.2
    x = np.array([0.0, 1.0, 1.5, 3.5, 4.0, 4.25, 6.0, 7.0, 7.5])
    o = orc.Euler1D(9, dt)
    o.coord[:] = x
    o.so0[:] = [1.0, 0.0, 2.5]
    o.so0[2] = [2.0, 4.0, 10.0]      # u = 2, p = 2.4
    o.so0[3] = [0.5, -1.0, 4.0]      # u = -2, p = 1.2
    o.so0[4] = [4.0, 2.0, 0.125]     # p < 0 -> speed |u| = 0.5
    o.so0[5] = [1.0, 1.0, 0.5]       # p = 0 exactly -> speed 1
    rho, m, e = o.so0.T
    p = (g - 1) * (e - m * m / (2 * rho))
    speed = np.sqrt(g * np.maximum(p, 0) / rho) + np.abs(m) / rho
    dx = np.minimum(np.diff(x)[1:], np.diff(x)[:-1])     # rows 1..7
    want = np.full(9, -7.0)
    for odd, rows in ((False, [2, 4]), (True, [1, 3, 5])):
        o.cfl[:] = -7.0
        o.update_cfl(odd)
        got = o.cfl.copy()
        exp = want.copy()
        exp[rows] = dt / 2 * speed[rows] / dx[np.array(rows) - 1]
        np.testing.assert_allclose(got, exp, rtol=1e-15)
    assert exp[5] == dt / 2 * 1.0 / 0.25
    # The clamp: row 4's value uses |u| alone.
    o.cfl[:] = 0
    o.update_cfl(False)
    assert o.cfl[4] == pytest.approx(dt / 2 * 0.5 / 0.25, rel=1e-15)


def _periodic_solver(orc, resolution=8):
    # tests/test_onedim_euler.py:14-28
    time_stop = 2 * np.pi
    dx = 2 * np.pi / resolution * 2
    nstep = int(np.ceil(time_stop / dx))
    dt = time_stop / nstep
    o = orc.Euler1D(resolution + 1, dt)
    o.coord[:] = np.linspace(0, 2 * np.pi, num=resolution + 1)
    o.setup_march()
    return nstep, o


def test_oracle_fine_equals_march(orc):
    nstep, a = _periodic_solver(orc)
    _, b = _periodic_solver(orc)
    # A non-trivial state so the comparison is not of zeros.
    x = a.coord
    for o in (a, b):
        o.so0[:, 0] = 1.0 + 0.2 * np.sin(x)
        o.so0[:, 1] = 0.1 * np.cos(x)
        o.so0[:, 2] = 2.5 + 0.1 * np.sin(2 * x)
    for _ in range(nstep * 10):
        a.march_half_so0(False)
        a.treat_boundary_so0()
        a.update_cfl(True)
        a.march_half_so1(2, False)
        a.treat_boundary_so1()
        a.march_half_so0(True)
        a.update_cfl(False)
        a.march_half_so1(2, True)
        b.march(2, 1)
        assert a.so0.tolist() == b.so0.tolist()


def _l1(orc, ncoord, steps):
    o = orc.sod_setup(ncoord=ncoord, time_increment=0.4 / steps)
    o.march(2, steps)
    idx = np.arange(2, ncoord - 2, 2)
    rho_exact = RiemannSolution(*SOD).sample(o.coord[idx], 0.4)[0]
    return np.abs(o.so0[idx, 0] - rho_exact).mean()


def test_oracle_sod_against_analytic(orc):
    # CESE converges to the Riemann solution: the L1 density error at
    # t = 0.4 falls as the grid is refined.
    e1 = _l1(orc, 501, 150)
    e2 = _l1(orc, 2001, 600)
    assert e2 < 4e-3
    assert e2 < 0.5 * e1


def test_golden_sod_matches_oracle(orc):
    g = np.load(os.path.join(GOLDEN, 'sod_2001_600.npz'))
    o = orc.sod_setup(ncoord=2001, time_increment=0.4 / 600)
    o.march(2, 600)
    assert (o.so0 == g['so0']).all() and (o.so1 == g['so1']).all()
    assert (o.cfl == g['cfl']).all()
    # Expected CFL bound of the chosen dt (hdt*max(|u|+c)/dx < 0.8).
    assert 0.5 < o.cfl.max() < 0.8


def test_even_ncoord_raises_in_oracle(orc):
    with pytest.raises(ValueError):
        orc.Euler1D(10, 0.1)


# ----------------------------------------------------------------- GPU -----

def _gpu_sod(ncoord, steps):
    from solvcon_amd.gen import sod_tube
    return sod_tube(ncoord=ncoord, time_increment=0.4 / steps)


@pytest.mark.gpu
@pytest.mark.parametrize('alpha', [1, 2])
def test_gpu_sod_bitwise(orc, alpha):
    print('case: Sod 2001 points, 600 steps, alpha', alpha, flush=True)
    svr = _gpu_sod(2001, 600)
    o = orc.sod_setup(ncoord=2001, time_increment=0.4 / 600)
    assert (svr.coord == o.coord).all() and (svr.so0 == o.so0).all()
    assert (svr.cfl == o.cfl).all()
    getattr(svr, 'march_alpha%d' % alpha)(600)
    o.march(alpha, 600)
    assert (svr.so0 == o.so0).all()
    assert (svr.so1 == o.so1).all()
    assert (svr.cfl == o.cfl).all()
    if alpha == 2:
        g = np.load(os.path.join(GOLDEN, 'sod_2001_600.npz'))
        assert (svr.so0 == g['so0']).all()


@pytest.mark.gpu
@pytest.mark.parametrize('ncoord', [3401, 20001])
def test_gpu_large_bitwise(orc, ncoord):
    # n > 3400 takes the per-half-step kernels instead of the LDS kernel.
    # dt keeps the Sod tube's dt/dx, so the CFL number stays below 0.8.
    print('case: Sod', ncoord, 'points, 40 steps', flush=True)
    steps = 40
    dt = 0.4 / 600 * 2000 / (ncoord - 1)
    svr = _gpu_sod(ncoord, 0.4 / dt)
    o = orc.sod_setup(ncoord=ncoord, time_increment=svr.time_increment)
    svr.march_alpha2(steps)
    o.march(2, steps)
    assert np.isfinite(o.so0).all() and o.cfl.max() < 0.8
    assert (svr.so0 == o.so0).all()
    assert (svr.so1 == o.so1).all()
    assert (svr.cfl == o.cfl).all()


@pytest.mark.gpu
@pytest.mark.parametrize('ncoord,steps,alpha', [
    (5, 3, 2), (7, 65, 2), (129, 64, 1), (2001, 1, 2), (2001, 63, 1),
    (2001, 130, 2), (9001, 65, 2), (9001, 129, 1)])
@pytest.mark.parametrize('tiled', [1, 0])
def test_gpu_tiled_bitwise(orc, monkeypatch, ncoord, steps, alpha, tiled):
    # The tiled march (segments with halos, up to 64 steps per launch) and
    # the single-workgroup / per-half-step kernels both equal the oracle:
    # one and several launches, odd and even launch counts, the window
    # holding the whole tube or a segment of it.
    print('case: 1D tiled', tiled, ncoord, 'points', steps, 'steps, alpha',
          alpha, flush=True)
    monkeypatch.setenv('SC_1D_TILED', str(tiled))
    dt = 0.4 / 600 * 2000 / (ncoord - 1)
    svr = _gpu_sod(ncoord, 0.4 / dt)
    o = orc.sod_setup(ncoord=ncoord, time_increment=svr.time_increment)
    getattr(svr, 'march_alpha%d' % alpha)(steps)
    o.march(alpha, steps)
    assert np.isfinite(o.so0).all()
    assert (svr.so0 == o.so0).all()
    assert (svr.so1 == o.so1).all()
    assert (svr.cfl == o.cfl).all()


@pytest.mark.gpu
def test_gpu_fine_grained_api(orc):
    print('case: fine-grained 1D API', flush=True)
    from solvcon_amd.onedim import Euler1DCore
    nstep, o = _periodic_solver(orc)
    # Set up the way solvcon's Euler1DSolver.init_solver drives the core.
    svr = Euler1DCore(ncoord=9, time_increment=o.time_increment)
    svr.coord[...] = np.linspace(0.0, 2 * np.pi, num=9)
    svr.cfl.fill(0)
    svr.gamma.fill(1.4)
    svr.so0[...] = 0.0
    svr.so1[...] = 0.0
    x = svr.coord.copy()
    init = np.stack([1.0 + 0.2 * np.sin(x), 0.1 * np.cos(x),
                     2.5 + 0.1 * np.sin(2 * x)], axis=1)
    svr.so0[...] = init
    o.so0[...] = init
    svr.setup_march()
    o.setup_march()
    for _ in range(nstep * 10):
        svr.march_half_so0(odd_plane=False)
        svr.treat_boundary_so0()
        svr.update_cfl(odd_plane=True)
        svr.march_half_so1_alpha2(odd_plane=False)
        svr.treat_boundary_so1()
        svr.march_half_so0(odd_plane=True)
        svr.update_cfl(odd_plane=False)
        svr.march_half_so1_alpha2(odd_plane=True)
        o.march_half_so0(False)
        o.treat_boundary_so0()
        o.update_cfl(True)
        o.march_half_so1(2, False)
        o.treat_boundary_so1()
        o.march_half_so0(True)
        o.update_cfl(False)
        o.march_half_so1(2, True)
        assert svr.so0.tolist() == o.so0.tolist()
        assert svr.so1.tolist() == o.so1.tolist()
        assert svr.cfl.tolist() == o.cfl.tolist()
    # The half-step groups and the full march agree too.
    a = svr.clone()
    b = svr.clone()
    a.march_half1_alpha1()
    a.treat_boundary_so0()
    a.treat_boundary_so1()
    a.march_half2_alpha1()
    b.march_alpha1(1)
    assert (a.so0 == b.so0).all() and (a.so1 == b.so1).all()


@pytest.mark.gpu
def test_gpu_api_semantics(orc):
    print('case: 1D API semantics', flush=True)
    from solvcon_amd.onedim import Euler1DCore
    with pytest.raises(ValueError):
        Euler1DCore(ncoord=10, time_increment=0.1)
    c = Euler1DCore(ncoord=21, time_increment=0.01)
    assert c.nvar == 3 and c.ncoord == 21 and c.time_increment == 0.01
    assert c.so0.shape == (21, 3) and c.coord.shape == (21,)
    assert (c.gamma == 1.4).all() and (c.so0 == 0).all()
    held = c.so0
    c.so0[:, 0] = 1.0 + np.arange(21)
    c.so0[:, 2] = 2.5
    c.coord[:] = np.linspace(0, 1, 21)
    c.march_alpha2(3)
    # A held view keeps aliasing the solver state across device calls.
    assert (held == c.so0).all()
    o = orc.Euler1D(21, 0.01)
    o.coord[:] = np.linspace(0, 1, 21)
    o.so0[:, 0] = 1.0 + np.arange(21)
    o.so0[:, 2] = 2.5
    o.march(2, 3)
    assert (c.so0 == o.so0).all()
    for name in ('density', 'velocity', 'pressure', 'temperature',
                 'internal_energy'):
        assert (getattr(c, name) == o.derived(name)).all(), name
    np.testing.assert_allclose(c.entropy, o.derived('entropy'), rtol=1e-14)
    d = c.clone()
    assert (d.so0 == c.so0).all()
    d.so0[:] = 0
    d.march_alpha2(1)
    assert (c.so0 == o.so0).all()


@pytest.mark.gpu
def test_gpu_views_held_through_containers(orc):
    # A view of a table counts as held while any Python object references
    # it: in a list or dict, behind a memoryview, or as a sub-view.  Held
    # views see the state after a march, and writes through them reach the
    # device; a view that was held and released stops being refreshed.
    print('case: 1D views held through containers', flush=True)
    n, dt = 201, 0.4 / 60
    core = _gpu_sod(n, 60)
    o = orc.sod_setup(ncoord=n, time_increment=dt)
    box = {'list': [core.so0], 'dict': {'a': core.so1}}
    mv = memoryview(core.cfl)
    sub = core.so0[40:60, 1:]
    nested = [[core.so1[::2]]]
    core.march_alpha2(10)
    o.march(2, 10)
    assert (box['list'][0] == o.so0).all()
    assert (box['dict']['a'] == o.so1).all()
    assert (np.asarray(mv) == o.cfl).all()
    assert (sub == o.so0[40:60, 1:]).all()
    assert (nested[0][0] == o.so1[::2]).all()
    # A write through the held sub-view is uploaded before the next march.
    sub[0, 0] = 0.25
    o.so0[40, 1] = 0.25
    core.march_alpha2(3)
    o.march(2, 3)
    assert (box['list'][0] == o.so0).all()
    assert (core.so0 == o.so0).all()
    del box, mv, sub, nested
    core.march_alpha2(2)
    o.march(2, 2)
    assert (core.so0 == o.so0).all() and (core.so1 == o.so1).all()
