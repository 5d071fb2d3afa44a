"""One-dimensional Euler CESE solver on the GPU.

Drop-in for solvcon's ``solvcon.onedim.euler1d`` (solvcon/onedim/euler1d.py)
and the ``_solvcon.onedim.Euler1DCore`` binding it wraps
(cpp/solvcon/onedim/pymod/wrap_onedim.cpp:38-153).  The core's state lives
in HBM and every method runs a HIP kernel of libsolvcon_amd.so; the array
properties are host mirrors that keep aliasing the solver state the way the
reference's numpy views do.
"""

import ctypes

import numpy as np

from .. import _lib as L

__all__ = ['Euler1DCore', 'Euler1DSolver', 'ShockTube']

_ARRAYS = ('coord', 'cfl', 'so0', 'so1', 'gamma')
_DERIVED = ('density', 'velocity', 'pressure', 'temperature',
            'internal_energy', 'entropy')


class Euler1DCore(object):
    """GPU Euler1DCore (cpp/solvcon/onedim/Euler1DCore.hpp:39-142).

    ncoord must be odd (ValueError otherwise, Euler1DCore.cpp:25-28).
    """

    nvar = 3

    def __init__(self, ncoord, time_increment, _handle=None):
        L.require_gpu()
        lib = L.lib()
        if _handle is None:
            h = ctypes.c_void_p()
            L.check(lib.sc1d_create(int(ncoord), float(time_increment),
                                    ctypes.byref(h)))
            _handle = h
        self._h = _handle
        n = ctypes.c_int64()
        dt = ctypes.c_double()
        lib.sc1d_info(self._h, ctypes.byref(n), ctypes.byref(dt))
        self._n = n.value
        self._dt = dt.value
        self._mirrors = {}
        for i, name in enumerate(_ARRAYS):
            shape = (self._n, 3) if name in ('so0', 'so1') else (self._n,)
            self._mirrors[name] = L.Mirror(
                shape, np.float64,
                up=self._mk_up(i), down=self._mk_down(i))

    def _mk_up(self, which):
        def up(a):
            L.check(L.lib().sc1d_upload(self._h, which, L.dptr(a), a.size))
        return up

    def _mk_down(self, which):
        def down(a):
            L.check(L.lib().sc1d_download(self._h, which, L.dptr(a), a.size))
        return down

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and L._lib is not None:
            L._lib.sc1d_destroy(h)
            self._h = None

    def __str__(self):
        return 'Euler1DCore(ncoord=%d, time_increment=%g)' % (
            self._n, self._dt)

    # Run one device call with the host mirrors kept in step.
    def _call(self, fn, *args):
        for m in self._mirrors.values():
            m.before()
        try:
            L.check(fn(self._h, *args))
        finally:
            for m in self._mirrors.values():
                m.after()

    @property
    def ncoord(self):
        return self._n

    @property
    def time_increment(self):
        return self._dt

    @property
    def coord(self):
        return self._mirrors['coord'].view()

    @property
    def cfl(self):
        return self._mirrors['cfl'].view()

    @property
    def so0(self):
        return self._mirrors['so0'].view()

    @property
    def so1(self):
        return self._mirrors['so1'].view()

    @property
    def gamma(self):
        return self._mirrors['gamma'].view()

    def _derived(self, which):
        out = np.empty(self._n, dtype=np.float64)
        for m in self._mirrors.values():
            m.before()
        L.check(L.lib().sc1d_derived(self._h, which, L.dptr(out)))
        for m in self._mirrors.values():
            m.after(written=False)
        return out

    density = property(lambda self: self._derived(0))
    velocity = property(lambda self: self._derived(1))
    pressure = property(lambda self: self._derived(2))
    temperature = property(lambda self: self._derived(3))
    internal_energy = property(lambda self: self._derived(4))
    entropy = property(lambda self: self._derived(5))

    def clone(self):
        for m in self._mirrors.values():
            m.before()
        h = ctypes.c_void_p()
        L.check(L.lib().sc1d_clone(self._h, ctypes.byref(h)))
        return Euler1DCore(self._n, self._dt, _handle=h)

    def update_cfl(self, odd_plane):
        self._call(L.lib().sc1d_update_cfl, int(bool(odd_plane)))

    def march_half_so0(self, odd_plane):
        self._call(L.lib().sc1d_march_half_so0, int(bool(odd_plane)))

    def treat_boundary_so0(self):
        self._call(L.lib().sc1d_treat_boundary_so0)

    def treat_boundary_so1(self):
        self._call(L.lib().sc1d_treat_boundary_so1)

    def setup_march(self):
        self._call(L.lib().sc1d_setup_march)

    def march_half_so1_alpha1(self, odd_plane):
        self._call(L.lib().sc1d_march_half_so1, 1, int(bool(odd_plane)))

    def march_half_so1_alpha2(self, odd_plane):
        self._call(L.lib().sc1d_march_half_so1, 2, int(bool(odd_plane)))

    def march_half1_alpha1(self):
        self._call(L.lib().sc1d_march_half, 1, 0)

    def march_half1_alpha2(self):
        self._call(L.lib().sc1d_march_half, 2, 0)

    def march_half2_alpha1(self):
        self._call(L.lib().sc1d_march_half, 1, 1)

    def march_half2_alpha2(self):
        self._call(L.lib().sc1d_march_half, 2, 1)

    def march_alpha1(self, steps):
        self._call(L.lib().sc1d_march, 1, int(steps))

    def march_alpha2(self, steps):
        self._call(L.lib().sc1d_march, 2, int(steps))


class Euler1DSolver(object):
    """Numerical solver for the one-dimensional Euler equation by the CESE
    method (solvcon/onedim/euler1d.py:21-61)."""

    def __init__(self, xmin, xmax, ncoord, time_increment=0.05,
                 keep_edge=False):
        self._core = self.init_solver(xmin, xmax, ncoord, time_increment,
                                      gamma=1.4)
        _ = self.ncoord - 1
        start = 0 if keep_edge else 2
        stop = _ if keep_edge else (_ - 2)
        num = (stop - start) // 2 + 1
        self.xindices = np.linspace(start, stop, num, dtype='int32')

    def __getattr__(self, name):
        if name == '_core':
            raise AttributeError(name)
        return getattr(self._core, name)

    @staticmethod
    def init_solver(xmin, xmax, ncoord, time_increment, gamma):
        svr = Euler1DCore(ncoord=ncoord, time_increment=time_increment)
        svr.coord[...] = np.linspace(xmin, xmax, num=ncoord)
        svr.cfl.fill(0)
        svr.gamma.fill(gamma)
        svr.so0[...] = 0.0
        svr.so1[...] = 0.0
        return svr

    @staticmethod
    def calc_u2(gamma, density, velocity, pressure):
        ie = 1. / (gamma - 1) * pressure / density
        ke = velocity * velocity / 2
        return density * (ie + ke)


class ShockTube(object):
    """Sod shock tube: the analytic Riemann solution and the numerical setup
    (solvcon/onedim/euler1d.py:64-393).

    The analytic part follows J. D. Anderson, Modern Compressible Flow, 3/e:
    zones 1 and 5 are given; zone 4 from the normal-shock relations with the
    pressure ratio p4/p5 found by the secant method; zone 3 across the
    contact and the expansion fan in zone 2.
    """

    R = 8.31446261815324

    def __init__(self):
        self.gamma = None
        for z in ('1', '3', '4', '5'):
            for q in ('velocity', 'pressure', 'density', 'temperature',
                      'internal_energy', 'entropy', 'speedofsound'):
                setattr(self, q + z, None)
        self.velocity_shock = None
        self.velocity_con = None
        self.coord_field = None
        self.density_field = None
        self.velocity_field = None
        self.pressure_field = None
        self.temperature_field = None
        self.internal_energy_field = None
        self.entropy_field = None
        self.svr = None

    def build_numerical(self, xmin, xmax, ncoord, time_increment=0.05,
                        xdiaphragm=0.0, keep_edge=False):
        if self.gamma is None:
            raise ValueError("constants are not set; call build_constant()")
        self.svr = Euler1DSolver(xmin, xmax, ncoord,
                                 time_increment=time_increment,
                                 keep_edge=keep_edge)
        self.svr.gamma.fill(self.gamma)
        u0l, u0r = self.density1, self.density5
        u2l = self.svr.calc_u2(self.gamma, self.density1, 0.0,
                               self.pressure1)
        u2r = self.svr.calc_u2(self.gamma, self.density5, 0.0,
                               self.pressure5)
        left = self.svr.coord < xdiaphragm
        right = np.logical_not(left)
        so0 = self.svr.so0
        so0[left, 0] = u0l
        so0[right, 0] = u0r
        so0[:, 1] = 0.0
        so0[left, 2] = u2l
        so0[right, 2] = u2r
        del so0
        self.svr.so1.fill(0)
        self.svr.setup_march()

    def build_constant(self, gamma, pressure1, density1, pressure5,
                       density5):
        self.gamma = gamma
        for z, p, rho in (('1', pressure1, density1),
                          ('5', pressure5, density5)):
            setattr(self, 'velocity' + z, 0.0)
            setattr(self, 'pressure' + z, p)
            setattr(self, 'density' + z, rho)
            setattr(self, 'temperature' + z, p / (rho * self.R))
            setattr(self, 'speedofsound' + z,
                    self.calc_speedofsound(pressure=p, density=rho))
            setattr(self, 'internal_energy' + z,
                    self.calc_internal_energy(p, rho))
            setattr(self, 'entropy' + z, self.calc_entropy(p, rho))

        p45 = self.calc_pressure45()
        self.velocity4 = self.calc_velocity4(pressure45=p45)
        self.pressure4 = p45 * self.pressure5
        self.density4 = self.calc_density45(pressure45=p45) * self.density5
        self.temperature4 = self.calc_temperature45(p45) * self.temperature5
        self.speedofsound4 = self.calc_speedofsound(
            pressure=self.pressure4, density=self.density4)
        self.internal_energy4 = self.calc_internal_energy(self.pressure4,
                                                          self.density4)
        self.entropy4 = self.calc_entropy(self.pressure4, self.density4)

        mach = np.sqrt((gamma + 1) / (2 * gamma) * (p45 - 1) + 1)
        self.velocity_shock = self.speedofsound5 * mach
        self.velocity_con = self.velocity4

        self.velocity3 = self.velocity4
        self.pressure3 = self.pressure4
        self.density3 = self.calc_density2(x=0, t=0, velocity2=self.velocity3)
        self.temperature3 = self.calc_temperature2(x=0, t=0,
                                                   velocity2=self.velocity3)
        self.internal_energy3 = self.calc_internal_energy(self.pressure3,
                                                          self.density3)
        self.entropy3 = self.calc_entropy(self.pressure3, self.density3)
        self.speedofsound3 = self.calc_speedofsound(
            pressure=self.pressure3, density=self.density3)

    def calc_pressure45(self, tolerance=1.e-10, maxiter=50):
        """Shock strength p4/p5 by the secant method on the shock-tube
        relation p1/p5 = p4/p5 (1 - (g-1)(a5/a1)(p4/p5-1)
        / sqrt(2g(2g + (g+1)(p4/p5-1))))^(-2g/(g-1))."""
        p15 = self.pressure1 / self.pressure5
        a51 = self.speedofsound5 / self.speedofsound1
        g = self.gamma

        def resid(p45):
            v = p45 - 1
            nume = (g - 1) * a51 * v
            deno = np.sqrt(2 * g * (2 * g + (g + 1) * v))
            return p15 - p45 * (1 - nume / deno) ** (-2 * g / (g - 1))

        x0 = p15
        r0 = resid(x0)
        x1 = x0 / 2
        r1 = resid(x1)
        slope = (r1 - r0) / (x1 - x0)
        count = maxiter
        while np.abs(r1) > tolerance and count > 0:
            x0 = x1
            x1 -= r1 / slope
            r0 = r1
            r1 = resid(x1)
            slope = (r1 - r0) / (x1 - x0)
            count -= 1
        return x1

    def calc_density45(self, pressure45):
        gpn1 = (self.gamma + 1) / (self.gamma - 1)
        return (1 + gpn1 * pressure45) / (gpn1 + pressure45)

    def calc_temperature45(self, pressure45):
        gpn1 = (self.gamma + 1) / (self.gamma - 1)
        return pressure45 * ((gpn1 + pressure45) / (1 + gpn1 * pressure45))

    def calc_velocity4(self, pressure45):
        c = self.speedofsound5 / self.gamma
        nume = 2 * self.gamma / (self.gamma + 1)
        deno = pressure45 + (self.gamma - 1) / (self.gamma + 1)
        return c * (pressure45 - 1) * np.sqrt(nume / deno)

    def calc_velocity2(self, x, t):
        return 2 / (self.gamma + 1) * (self.speedofsound1 + x / t)

    def calc_speedofsound2_ratio(self, x, t, velocity2=None):
        if velocity2 is None:
            velocity2 = self.calc_velocity2(x, t)
        return 1 - (self.gamma - 1) / 2 * (velocity2 / self.speedofsound1)

    def calc_pressure2(self, x, t):
        c1 = self.calc_speedofsound2_ratio(x, t)
        return self.pressure1 * (c1 ** (2 * self.gamma / (self.gamma - 1)))

    def calc_density2(self, x, t, velocity2=None):
        c1 = self.calc_speedofsound2_ratio(x, t, velocity2=velocity2)
        return self.density1 * (c1 ** (2 / (self.gamma - 1)))

    def calc_temperature2(self, x, t, velocity2=None):
        return self.temperature1 * (self.calc_speedofsound2_ratio(
            x, t, velocity2=velocity2)) ** 2

    def calc_speedofsound(self, pressure, density):
        return np.sqrt(self.gamma * pressure / density)

    def calc_internal_energy(self, pressure, density):
        return pressure / (density * (self.gamma - 1))

    def calc_entropy(self, pressure, density):
        return pressure / (density ** self.gamma)

    def build_field(self, t, coord=None):
        """Fill the *_field arrays with the analytic solution at time t."""
        if coord is None:
            coord = self.svr.coord[self.svr.xindices]
        self.coord_field = coord.copy()
        loc12, loc23, loc34, loc45 = self.calc_locations(t=t)
        x = self.coord_field
        zones = [
            (x < loc12, '1'),
            (np.logical_and(x >= loc23, x < loc34), '3'),
            (np.logical_and(x >= loc34, x < loc45), '4'),
            (x >= loc45, '5'),
        ]
        fan = np.logical_and(x >= loc12, x < loc23)
        for q in ('density', 'velocity', 'pressure', 'temperature',
                  'internal_energy', 'entropy'):
            arr = np.zeros_like(x, dtype='float64')
            for slct, z in zones:
                arr[slct] = getattr(self, q + z)
            setattr(self, q + '_field', arr)
        for i in np.nonzero(fan)[0]:
            xi = coord[i]
            self.density_field[i] = self.calc_density2(xi, t)
            self.velocity_field[i] = self.calc_velocity2(xi, t)
            self.pressure_field[i] = self.calc_pressure2(xi, t)
            self.temperature_field[i] = self.calc_temperature2(
                xi, t, self.velocity_field[i])
            self.internal_energy_field[i] = self.calc_internal_energy(
                self.pressure_field[i], self.density_field[i])
            self.entropy_field[i] = self.calc_entropy(
                self.pressure_field[i], self.density_field[i])

    def calc_locations(self, t):
        """[x12, x23, x34, x45]: the expansion head and tail, the contact
        and the shock at time t."""
        return np.array([
            -self.speedofsound1 * t,
            (self.velocity3 - self.speedofsound3) * t,
            self.velocity3 * t,
            self.velocity_shock * t,
        ], dtype='float64')
