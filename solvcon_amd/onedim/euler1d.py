"""One-dimensional Euler CESE solver on the GPU.

Drop-in for the ``_solvcon.onedim.Euler1DCore`` binding
(cpp/solvcon/onedim/pymod/wrap_onedim.cpp:38-153) that solvcon's Python
solver classes (``Euler1DSolver`` and ``ShockTube`` in
solvcon/onedim/euler1d.py) construct and drive.  The core's state lives
in HBM and every method runs a HIP kernel of libsolvcon_amd.so; the array
properties are host mirrors that keep aliasing the solver state the way the
reference's numpy views do.
"""

import ctypes

import numpy as np

from .. import _lib as L

__all__ = ['Euler1DCore']

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
