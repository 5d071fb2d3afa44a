"""Regular reflection of an oblique shock: the inlet / slip-wall /
non-reflective march of solvcon's obsrefl application
(solvcon/pilot/apps/obsrefl), on this build's own rectangle meshes.

- CPU: the oblique-shock relations (oracle/oblique.py) against the values
  the reference's tests/apps/obsrefl/test_driver.py pins (Anderson, Modern
  Compressible Flow, chapter 4), the jump conditions across both shocks,
  the rectangle meshes, and the oracle marched to a steady state whose
  three zones match the analytic ones.
- GPU: the device march with two inlets, a slip wall and a non-reflective
  edge against the oracle bit for bit, at the application's default 64 x 16
  boxes, for 10 steps and for the whole 600-step run to the steady state.
"""

import math

import numpy as np
import pytest

from oracle import oblique as OB
from solvcon_amd.gen import rect_mesh_def, edge_faces

EXTENT = (4.0, 1.0)
FLAVORS = {'quad': (4e-3, 3), 'tri': (3e-3, 4)}     # dt, cell type


# ------------------------------------------------------- relations (CPU) -----

def test_relation_values():
    beta = math.radians(37.8)
    assert OB.density_ratio(3, beta) == pytest.approx(2.4204302545,
                                                      abs=1e-10)
    assert OB.pressure_ratio(3, beta) == pytest.approx(3.7777114257,
                                                       abs=1e-10)
    assert OB.temperature_ratio(3, beta) == pytest.approx(1.5607602899,
                                                          abs=1e-10)
    assert OB.density_ratio(3, beta, g=1.2) == pytest.approx(2.7793244902,
                                                             abs=1e-10)
    assert OB.normal_mach_downstream(3) == pytest.approx(0.4751909633,
                                                         abs=1e-10)
    assert OB.downstream_mach(3, beta=beta) == pytest.approx(1.9924827009,
                                                             abs=1e-10)
    assert OB.downstream_mach(3, theta=math.radians(20)) == pytest.approx(
        1.9941316656, abs=1e-9)
    b = OB.shock_angle(4, math.radians(32))
    assert math.degrees(b) == pytest.approx(48.2584798722, abs=1e-8)
    assert math.degrees(OB.deflection(4, b)) == pytest.approx(32.0,
                                                              abs=1e-10)
    with pytest.raises(ValueError):
        OB.shock_angle(2, math.radians(40))
    with pytest.raises(ValueError):
        OB.downstream_mach(3)
    with pytest.raises(ValueError):
        OB.downstream_mach(3, beta=0.2, theta=0.1)


def _frame_flux(state, n):
    # Mass, normal-momentum, tangential-momentum and energy fluxes through a
    # surface of unit normal n.
    rho, (u, v), p = state
    un = u * n[0] + v * n[1]
    ut = -u * n[1] + v * n[0]
    h = 1.4 / 0.4 * p / rho + 0.5 * (u * u + v * v)
    return np.array([rho * un, rho * un * un + p, rho * un * ut,
                     rho * un * h])


def test_reflection_jump_conditions():
    # Rankine-Hugoniot across the incident and the reflected shock, computed
    # from the zone states and the shock directions alone.
    rr = OB.RegularReflection()
    b1 = rr.beta1
    n_inc = (math.sin(b1), math.cos(b1))           # normal of the incident
    np.testing.assert_allclose(_frame_flux(rr.zones[0], n_inc),
                               _frame_flux(rr.zones[1], n_inc), rtol=1e-12)
    a = rr.beta2 - rr.theta                        # reflected shock angle
    n_ref = (math.sin(a), -math.cos(a))
    np.testing.assert_allclose(_frame_flux(rr.zones[1], n_ref),
                               _frame_flux(rr.zones[2], n_ref), rtol=1e-12,
                               atol=1e-12)
    # Zone 3 flows along the wall again; pressures rise through each shock.
    assert rr.zones[2][1][1] == 0.0
    assert rr.zones[0][2] < rr.zones[1][2] < rr.zones[2][2]
    assert rr.wall_x == pytest.approx(1.0 / math.tan(b1))


# ---------------------------------------------------------- meshes (CPU) -----

@pytest.fixture(scope='module', params=sorted(FLAVORS))
def flavor(request, orc):
    name = request.param
    defn = rect_mesh_def(64, 16, extent=EXTENT, kind=name, seed=7)
    return name, defn, orc.make_mesh(*defn)


def test_rect_mesh(flavor):
    name, defn, o = flavor
    per_box = 1 if name == 'quad' else 2
    assert o.ncell == per_box * 64 * 16
    assert (o.cltpn[o.ngstcell:] == FLAVORS[name][1]).all()
    assert (o.clvol[o.ngstcell:] > 0).all()
    assert o.clvol[o.ngstcell:].sum() == pytest.approx(4.0, rel=1e-13)
    edges = edge_faces(o.bndfcs[:, 0], o.fccnd[o.ngstface:], EXTENT)
    assert [len(e) for e in edges] == [16, 64, 64, 16]
    allf = np.concatenate(edges)
    assert sorted(allf.tolist()) == sorted(o.bndfcs[:, 0].tolist())
    fn = o.fcnml[o.ngstface:]
    for faces, normal in zip(edges, ((-1, 0), (0, 1), (0, -1), (1, 0))):
        np.testing.assert_allclose(fn[faces], np.tile(normal,
                                                      (len(faces), 1)),
                                   atol=1e-12)
    with pytest.raises(ValueError):
        rect_mesh_def(2, 2, kind='hexagon')


def _setup(orc, defn, dt, gpu):
    rr = OB.RegularReflection()
    o = orc.make_mesh(*defn)
    edges = edge_faces(o.bndfcs[:, 0], o.fccnd[o.ngstface:], EXTENT)
    rho, (u, v), p = rr.zones[0]
    z2 = rr.zones[1]
    inl1 = [rho, u, v, p, 1.4]
    inl2 = [z2[0], z2[1][0], z2[1][1], z2[2], 1.4]
    oe = orc.Euler(o, dt)
    oe.init_solution(1.4, rho, [u, v], p)
    left, top, bottom, right = edges
    oe.add_bc(orc.BC_INLET, left, inl1)
    oe.add_bc(orc.BC_INLET, top, inl2)
    oe.add_bc(orc.BC_SLIPWALL, bottom)
    oe.add_bc(orc.BC_NONREFL, right)
    oe.call('bc_soln')
    oe.call('bc_dsoln')
    e = None
    if gpu:
        import solvcon_amd as sc
        import pair
        mh = pair.gpu_mesh(defn)
        e = sc.EulerCore(mesh=mh, time_increment=dt)
        e.init_solution(1.4, rho, [u, v], p)
        e.add_inlet(left, inl1)
        e.add_inlet(top, inl2)
        e.add_slipwall(bottom)
        e.add_nonrefl(right)
        e.bc_soln()
        e.bc_dsoln()
    return rr, o, oe, e


def _zone_errors(rr, o, so0n):
    ng = o.ngstcell
    zone = rr.zone_of(o.clcnd[ng:], margin=3 * EXTENT[0] / 64)
    out = []
    for z in range(3):
        k = zone == z
        assert k.sum() > 50
        out.append(so0n[ng:, 0][k].mean() / rr.zones[z][0] - 1.0)
    return out


def test_oracle_reaches_the_analytic_zones(orc, flavor):
    # 600 steps carry the free stream across the domain several times; the
    # mean density of each zone, away from the shocks, is then within 3%
    # of the analytic one.  This pins the oracle's inlet and slip-wall
    # conditions (euler_bc.cpp:143-320) against the analytic solution.
    name, defn, _ = flavor
    dt = FLAVORS[name][0]
    rr, o, oe, _ = _setup(orc, defn, dt, gpu=False)
    oe.march(600)
    ng = o.ngstcell
    assert np.isfinite(oe.so0n).all() and oe.cflo[ng:].max() < 0.6
    for err in _zone_errors(rr, o, oe.so0n):
        assert abs(err) < 0.03


# ------------------------------------------------------------------ GPU ------

@pytest.mark.gpu
@pytest.mark.parametrize('name', sorted(FLAVORS))
@pytest.mark.parametrize('steps', [10, 600])
def test_reflection_march_equals_oracle(orc, name, steps):
    # The device march with its two inlets, slip wall and non-reflective
    # edge equals the oracle's bit for bit, after 10 steps and after the
    # whole run to the steady state.
    print('case: oblique reflection', name, steps, 'steps', flush=True)
    import pair
    defn = rect_mesh_def(64, 16, extent=EXTENT, kind=name, seed=7)
    rr, o, oe, e = _setup(orc, defn, FLAVORS[name][0], gpu=True)
    e.march(steps)
    oe.march(steps)
    pair.same_euler(e, oe, names=('so0n', 'so1n', 'cflo', 'cflc'))
    if steps == 600:
        for err in _zone_errors(rr, o, e.so0n.ndarray):
            assert abs(err) < 0.03
