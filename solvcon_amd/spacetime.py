"""One-dimensional space-time CESE scalar solvers on the MI355X.

Drop-in counterpart of ``solvcon.spacetime`` (solvcon/spacetime.py,
cpp/solvcon/spacetime/pymod/wrap_spacetime.{hpp,cpp}): ``Grid``,
``Kernel``, ``Solver`` (the default kernel hooks), ``LinearScalarSolver``,
``InviscidBurgersSolver`` and their conservation/solution elements.  The
marches and every element flux run in HIP kernels (libsolvcon_amd.so,
spacetime.hip); ``Grid`` builds its coordinates in the library's host code.

Python per-element hooks of the generic ``Solver`` (``Kernel.xn_calc = f``)
cannot run inside a kernel; assigning one raises ``NotImplementedError``.
"""

import ctypes
import sys

import numpy as np

from . import _lib
from ._lib import c_int, c_i64, c_dbl, c_vp, P_dbl, P_vp, check, dptr

__all__ = ['Grid', 'Kernel', 'Celm', 'Selm', 'Solver', 'LinearScalarSolver',
           'LinearScalarCelm', 'LinearScalarSelm', 'InviscidBurgersSolver',
           'InviscidBurgersCelm', 'InviscidBurgersSelm', 'SolverProxy']

_lib.register({
    'scst_grid': (c_int, [c_dbl, c_dbl, c_i64, P_dbl, c_i64]),
    'scst_grid_from': (c_int, [P_dbl, c_i64, P_dbl, c_i64]),
    'scst_create': (c_int, [c_int, P_dbl, c_i64, c_i64, c_dbl, P_vp]),
    'scst_clone': (c_int, [c_vp, P_vp]),
    'scst_destroy': (None, [c_vp]),
    'scst_set_time_increment': (c_int, [c_vp, c_dbl]),
    'scst_upload': (c_int, [c_vp, c_int, P_dbl, c_i64]),
    'scst_download': (c_int, [c_vp, c_int, P_dbl, c_i64]),
    'scst_stage': (c_int, [c_vp, c_int, c_int, c_int]),
    'scst_march': (c_int, [c_vp, c_int, c_i64]),
    'scst_march_async': (c_int, [c_vp, c_int, c_i64]),
    'scst_sync': (c_int, [c_vp]),
    'scst_eval': (c_int, [c_vp, c_int, c_int, c_i64, c_i64, c_int, P_dbl]),
    'scst_so0p': (c_int, [c_vp, c_i64, c_int, P_dbl, c_i64]),
})

_BOUND = 2
_K_DEFAULT, _K_LINEAR, _K_BURGERS = 0, 1, 2
_ST_SO0, _ST_SO1, _ST_CFL, _ST_BND0, _ST_BND1, _ST_HALF1, _ST_HALF2 = range(7)
_A_XC, _A_SO0, _A_SO1, _A_CFL = range(4)


def _num(v):
    """A double the way std::ostream prints it (precision 6, %g)."""
    return '%g' % v


class Array(object):
    """A SimpleArray-like holder of one numpy array (``ndarray``, ``shape``,
    indexing, numpy conversion)."""

    __slots__ = ('_a',)

    def __init__(self, a):
        self._a = a

    @property
    def ndarray(self):
        return self._a

    @property
    def shape(self):
        return self._a.shape

    @property
    def size(self):
        return self._a.size

    def __len__(self):
        return len(self._a)

    def __array__(self, dtype=None, copy=None):
        return self._a if dtype is None else self._a.astype(dtype)

    def __getitem__(self, k):
        return self._a[k]

    def __setitem__(self, k, v):
        self._a[k] = v

    def fill(self, v):
        self._a.fill(v)


class Grid(object):
    """Spatial grid (cpp/solvcon/spacetime/core.hpp:112-214, core.cpp:14-98):
    ``Grid(xmin, xmax, nelm)`` or ``Grid(xloc)``."""

    BOUND_COUNT = _BOUND

    def __init__(self, *args, **kw):
        if len(args) == 1 and not kw or (not args and set(kw) == {'xloc'}):
            xloc = np.ascontiguousarray(args[0] if args else kw['xloc'],
                                        dtype=np.float64).ravel()
            n = xloc.size
            xc = np.zeros(max(2 * (n - 1) + 5, 0))
            check(_lib.lib().scst_grid_from(dptr(xloc), n, dptr(xc),
                                            xc.size))
        else:
            names = ('xmin', 'xmax', 'nelm')
            vals = dict(zip(names, args))
            for k, v in kw.items():
                if k not in names or k in vals:
                    raise TypeError('Grid(): unexpected argument %r' % k)
                vals[k] = v
            if set(vals) != set(names):
                raise TypeError('Grid() takes (xmin, xmax, nelm) or (xloc)')
            ncelm = int(vals['nelm'])
            if ncelm < 0:
                raise TypeError('Grid(): nelm must be non-negative')
            xc = np.zeros(2 * max(ncelm, 0) + 5)
            check(_lib.lib().scst_grid(float(vals['xmin']),
                                       float(vals['xmax']), ncelm, dptr(xc),
                                       xc.size))
        self._xc = xc
        self._gen = 0
        self._ncelm = (xc.size - 5) // 2
        self._xmin = float(xc[_BOUND])
        self._xmax = float(xc[xc.size - _BOUND - 1])

    @property
    def xmin(self):
        return self._xmin

    @property
    def xmax(self):
        return self._xmax

    @property
    def ncelm(self):
        return self._ncelm

    @property
    def nselm(self):
        return self._ncelm + 1

    @property
    def xsize(self):
        return self._xc.size

    @property
    def xcoord(self):
        # Handing out the coordinates may change them: solvers re-upload.
        self._gen += 1
        return Array(self._xc)

    def _clone(self):
        g = Grid.__new__(Grid)
        g._xc = self._xc.copy()
        g._gen = 0
        g._ncelm = self._ncelm
        g._xmin = self._xmin
        g._xmax = self._xmax
        return g

    def __str__(self):
        return 'Grid(xmin=%s, xmax=%s, ncelm=%d)' % (
            _num(self.xmin), _num(self.xmax), self.ncelm)


class Kernel(object):
    """Solution-element hooks (core.hpp:222-275).  Only the defaults of
    ``reset`` (zero fluxes, so0p = so0, cfl = 0) run on the device."""

    _NAMES = ('xn_calc', 'xp_calc', 'tn_calc', 'tp_calc', 'so0p_calc',
              'cfl_updater')

    def __init__(self):
        object.__setattr__(self, '_hooks', {})
        self.reset()

    def reset(self):
        def zero(se, iv):
            return 0.0

        def so0p(se, iv):
            return se.get_so0(iv)

        def cfl(se):
            se.set_cfl(0.0)

        self._hooks.update(xn_calc=zero, xp_calc=zero, tn_calc=zero,
                           tp_calc=zero, so0p_calc=so0p, cfl_updater=cfl)

    def __getattr__(self, name):
        hooks = object.__getattribute__(self, '_hooks')
        if name in hooks:
            return hooks[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in self._NAMES:
            raise NotImplementedError(
                'Kernel.%s: Python solution-element hooks cannot run in a '
                'HIP kernel; use LinearScalarSolver or InviscidBurgersSolver'
                % name)
        object.__setattr__(self, name, value)


class _Elem(object):
    """Common part of Celm and Selm (ElementBase, core.hpp:24-94)."""

    __slots__ = ('_svr', '_xi')
    _LO = 1

    def __init__(self, svr, xindex):
        self._svr = svr
        self._xi = xindex

    @property
    def _xc(self):
        return self._svr._grid._xc

    @property
    def x(self):
        return float(self._xc[self._xi])

    @property
    def xneg(self):
        return float(self._xc[self._xi - 1])

    @property
    def xpos(self):
        return float(self._xc[self._xi + 1])

    @property
    def dx(self):
        return self.xpos - self.xneg

    @property
    def grid(self):
        return self._svr.grid

    @property
    def field(self):
        return self._svr

    @property
    def time_increment(self):
        return self._svr.time_increment

    @property
    def dt(self):
        return self._svr.dt

    @property
    def hdt(self):
        return self._svr.hdt

    @property
    def qdt(self):
        return self._svr.qdt

    def duplicate(self):
        return type(self)(self._svr, self._xi)

    @property
    def dup(self):
        return self.duplicate()

    def _move_at(self, offset):
        xi = self._xi + offset
        nx = self._svr._grid.xsize
        lo = self._LO
        if xi < lo or xi >= nx - lo:
            raise IndexError(
                '%s(xindex=%d)::move_at(offset=%d): xindex = %d outside the '
                'interval [%d, %d)' % (self._BASE, self._xi, offset, xi, lo,
                                       nx - lo))
        self._xi = xi
        return self

    def move(self, v):
        return self._move_at(int(v))

    def move_left(self):
        return self._move_at(-2)

    def move_right(self):
        return self._move_at(2)

    def move_neg(self):
        return self._move_at(-1)

    def move_pos(self):
        return self._move_at(1)

    @property
    def on_even_plane(self):
        return not self.on_odd_plane

    # Comparisons follow the element pointers of one solver
    # (core.hpp:76-81); elements of different solvers only differ.
    def _cmp(self, o, op):
        if not isinstance(o, _Elem) or self._BASE != o._BASE:
            return NotImplemented
        same = self._svr is o._svr
        if op == 'eq':
            return same and self._xi == o._xi
        if op == 'ne':
            return not same or self._xi != o._xi
        if not same:
            return False
        return {'lt': self._xi < o._xi, 'le': self._xi <= o._xi,
                'gt': self._xi > o._xi, 'ge': self._xi >= o._xi}[op]

    def __eq__(self, o):
        r = self._cmp(o, 'eq')
        return False if r is NotImplemented else r

    def __ne__(self, o):
        r = self._cmp(o, 'ne')
        return True if r is NotImplemented else r

    def __lt__(self, o):
        return self._cmp(o, 'lt')

    def __le__(self, o):
        return self._cmp(o, 'le')

    def __gt__(self, o):
        return self._cmp(o, 'gt')

    def __ge__(self, o):
        return self._cmp(o, 'ge')

    __hash__ = None

    def __str__(self):
        return '%s(%s, index=%d, x=%s, xneg=%s, xpos=%s)' % (
            self._NAME, 'odd' if self.on_odd_plane else 'even', self.index,
            _num(self.x), _num(self.xneg), _num(self.xpos))

    def _eval(self, what, iv=0, alpha=2, store=0):
        return self._svr._eval(what, self._xi, iv, alpha, store)


class Selm(_Elem):
    """Solution element (core.hpp:458-518)."""

    __slots__ = ()
    _BASE = 'Selm'
    _NAME = 'Selm'
    _LO = 1

    @property
    def index(self):
        return (self._xi >> 1) - 1

    @property
    def on_odd_plane(self):
        return bool((self._xi - _BOUND) & 1)

    @property
    def dxneg(self):
        return self.x - self.xneg

    @property
    def dxpos(self):
        return self.xpos - self.x

    @property
    def xctr(self):
        return (self.xneg + self.xpos) / 2

    def get_so0(self, iv):
        return float(self._svr._get(_A_SO0, self._xi, iv))

    def get_so1(self, iv):
        return float(self._svr._get(_A_SO1, self._xi, iv))

    def get_cfl(self):
        return float(self._svr._get(_A_CFL, self._xi))

    def set_so0(self, iv, val):
        self._svr._set(_A_SO0, self._xi, val, iv)

    def set_so1(self, iv, val):
        self._svr._set(_A_SO1, self._xi, val, iv)

    def set_cfl(self, val):
        self._svr._set(_A_CFL, self._xi, val)

    def xn(self, iv):
        return self._eval(0, iv)

    def xp(self, iv):
        return self._eval(1, iv)

    def tn(self, iv):
        return self._eval(2, iv)

    def tp(self, iv):
        return self._eval(3, iv)

    def so0p(self, iv):
        return self._eval(4, iv)

    def update_cfl(self):
        self._eval(5, 0, store=1)


class Celm(_Elem):
    """Compound conservation element (core.hpp:538-642)."""

    __slots__ = ()
    _BASE = 'Celm'
    _NAME = 'Celm'
    _LO = 2
    _SELM = Selm

    @property
    def index(self):
        return ((self._xi - 1) >> 1) - 1

    @property
    def on_odd_plane(self):
        return bool((self._xi - (1 + _BOUND)) & 1)

    @property
    def xctr(self):
        return self.x

    @property
    def selm_xn(self):
        return self._SELM(self._svr, self._xi - 1)

    @property
    def selm_xp(self):
        return self._SELM(self._svr, self._xi + 1)

    @property
    def selm_tn(self):
        return self._SELM(self._svr, self._xi)

    @property
    def selm_tp(self):
        return self._SELM(self._svr, self._xi)

    def calc_so0(self, iv):
        return self._eval(6, iv)

    def calc_so1_alpha0(self, iv):
        return self._eval(7, iv, 0)

    def calc_so1_alpha1(self, iv):
        return self._eval(7, iv, 1)

    def calc_so1_alpha2(self, iv):
        return self._eval(7, iv, 2)


class LinearScalarSelm(Selm):
    __slots__ = ()
    _NAME = 'LinearScalarSelm'


class LinearScalarCelm(Celm):
    __slots__ = ()
    _NAME = 'LinearScalarCelm'
    _SELM = LinearScalarSelm


class InviscidBurgersSelm(Selm):
    __slots__ = ()
    _NAME = 'InviscidBurgersSelm'


class InviscidBurgersCelm(Celm):
    __slots__ = ()
    _NAME = 'InviscidBurgersCelm'
    _SELM = InviscidBurgersSelm


class _ElemIter(object):
    """SolverElementIterator (wrap_spacetime.hpp:154-230)."""

    def __init__(self, svr, odd, selm):
        self._svr = svr
        self._odd = bool(odd)
        self._selm = selm
        self._cur = 0

    def nelem(self):
        g = self._svr._grid
        n = g.nselm if self._selm else g.ncelm
        return n - (1 if self._odd else 0)

    def __iter__(self):
        return self

    def __next__(self):
        if self._cur >= self.nelem():
            raise StopIteration
        i = self._cur
        self._cur += 1
        if self._selm:
            return self._svr.selm(i, self._odd)
        return self._svr.celm(i, self._odd)

    def __str__(self):
        return 'SolverElementIterator(%s, %s, current=%d, nelem=%d)' % (
            'selm' if self._selm else 'celm',
            'on_odd_plane' if self._odd else 'on_even_plane', self._cur,
            self.nelem())


class _SolverBase(object):
    """SolverBase (core.hpp:706-1036) over a device handle."""

    _KIND = _K_DEFAULT
    _NAME = 'Solver'
    _CELM = Celm
    _SELM = Selm

    def __init__(self, grid, time_increment, nvar=1):
        if not isinstance(grid, Grid):
            raise TypeError('grid must be a spacetime Grid')
        _lib.require_gpu()
        self._grid = grid
        self._nvar = int(nvar)
        self._dt = float(time_increment)
        h = c_vp()
        check(_lib.lib().scst_create(self._KIND, dptr(grid._xc),
                                     grid._xc.size, self._nvar, self._dt,
                                     ctypes.byref(h)))
        self._init_handle(h)

    def _init_handle(self, h):
        self._h = h
        self._xc_gen = self._grid._gen
        self._xc_copy = self._grid._xc.copy()
        nx = self._grid.xsize
        nv = self._nvar
        self._m = {
            _A_SO0: self._mirror_for(_A_SO0, (nx, nv)),
            _A_SO1: self._mirror_for(_A_SO1, (nx, nv)),
            _A_CFL: self._mirror_for(_A_CFL, (nx,)),
        }

    def _mirror_for(self, which, shape):
        L = _lib.lib()

        def up(a):
            check(L.scst_upload(self._h, which, dptr(a), a.size))

        def down(a):
            check(L.scst_download(self._h, which, dptr(a), a.size))

        return _lib.Mirror(shape, np.float64, up, down)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value is not None:
            _lib.lib().scst_destroy(h)
            self._h = None

    # Host <-> device bookkeeping around every device call.
    def _before(self):
        g = self._grid
        if g._gen != self._xc_gen or sys.getrefcount(g._xc) > 2:
            if not np.array_equal(g._xc, self._xc_copy):
                check(_lib.lib().scst_upload(self._h, _A_XC, dptr(g._xc),
                                             g._xc.size))
                self._xc_copy = g._xc.copy()
            self._xc_gen = g._gen
        for m in self._m.values():
            m.before()

    def _after(self):
        for m in self._m.values():
            m.after()

    def _get(self, which, xi, iv=None):
        a = self._m[which].view()
        return a[xi] if iv is None else a[xi, iv]

    def _set(self, which, xi, val, iv=None):
        a = self._m[which].view()
        if iv is None:
            a[xi] = val
        else:
            if iv >= self._nvar:
                raise IndexError('out of nvar range')
            a[xi, iv] = val

    def _eval(self, what, xi, iv, alpha, store):
        out = c_dbl()
        self._before()
        try:
            check(_lib.lib().scst_eval(self._h, what, alpha, xi, iv, store,
                                       ctypes.byref(out)))
        finally:
            self._after()
        return out.value

    def _call(self, fn, *args):
        self._before()
        try:
            check(fn(self._h, *args))
        finally:
            self._after()

    # -- construction and info ---------------------------------------------
    def clone(self, grid=False):
        g = self._grid._clone() if grid else self._grid
        o = type(self).__new__(type(self))
        o._grid = g
        o._nvar = self._nvar
        o._dt = self._dt
        self._before()
        h = c_vp()
        check(_lib.lib().scst_clone(self._h, ctypes.byref(h)))
        self._after()
        o._init_handle(h)
        return o

    @property
    def grid(self):
        return self._grid

    @property
    def nvar(self):
        return self._nvar

    @property
    def time_increment(self):
        return self._dt

    @time_increment.setter
    def time_increment(self, v):
        self._dt = float(v)
        check(_lib.lib().scst_set_time_increment(self._h, self._dt))

    @property
    def dt(self):
        return self._dt

    @property
    def hdt(self):
        return 0.5 * self._dt

    @property
    def qdt(self):
        return 0.25 * self._dt

    def __str__(self):
        return '%s(grid=%s)' % (self._NAME, self._grid)

    # -- elements ------------------------------------------------------------
    def _check_plane(self, ielm, odd, celm):
        xi = (3 if celm else 2) + 2 * int(ielm) + (1 if odd else 0)
        lo = 2 if celm else 1
        nx = self._grid.xsize
        if xi < lo or xi >= nx - lo:
            raise IndexError(
                'Field::%s_at(ielm=%d, odd_plane=%d): xindex = %d outside the '
                'interval [%d, %d)' % ('celm' if celm else 'selm', int(ielm),
                                       int(bool(odd)), xi, lo, nx - lo))
        return xi

    def celm(self, ielm, odd_plane=False):
        return self._CELM(self, self._check_plane(ielm, odd_plane, True))

    def selm(self, ielm, odd_plane=False):
        return self._SELM(self, self._check_plane(ielm, odd_plane, False))

    def celms(self, odd_plane=False):
        return _ElemIter(self, odd_plane, False)

    def selms(self, odd_plane=False):
        return _ElemIter(self, odd_plane, True)

    def _sel(self, odd):
        o = 1 if odd else 0
        return 2 + 2 * np.arange(self._grid.nselm - o) + o

    def x(self, odd_plane=False):
        return Array(self._grid._xc[self._sel(odd_plane)].copy())

    def xctr(self, odd_plane=False):
        y = self._sel(odd_plane)
        xc = self._grid._xc
        return (xc[y - 1] + xc[y + 1]) / 2

    # -- arrays --------------------------------------------------------------
    @property
    def so0(self):
        return Array(self._m[_A_SO0].view())

    @property
    def so1(self):
        return Array(self._m[_A_SO1].view())

    @property
    def cfl(self):
        return Array(self._m[_A_CFL].view())

    def _get_plane(self, which, iv, odd, name):
        if iv is not None and iv >= self._nvar:
            raise IndexError('%s(): out of nvar range' % name)
        a = self._m[which].host_copy()
        y = self._sel(odd)
        return Array(a[y] if iv is None else a[y, iv])

    def _set_plane(self, which, iv, arr, odd, name):
        if iv is not None and iv >= self._nvar:
            raise IndexError('%s(): out of nvar range' % name)
        arr = np.asarray(arr, dtype=np.float64)
        if arr.ndim != 1:
            raise IndexError('%s(): input not 1D' % name)
        y = self._sel(odd)
        if arr.size != y.size:
            raise IndexError('%s(): input wrong size' % name)
        a = self._m[which].view()
        if iv is None:
            a[y] = arr
        else:
            a[y, iv] = arr

    def get_so0(self, iv, odd_plane=False):
        return self._get_plane(_A_SO0, iv, odd_plane, 'get_so0')

    def get_so1(self, iv, odd_plane=False):
        return self._get_plane(_A_SO1, iv, odd_plane, 'get_so1')

    def get_cfl(self, odd_plane=False):
        return self._get_plane(_A_CFL, None, odd_plane, 'get_cfl')

    def set_so0(self, iv, arr, odd_plane=False):
        self._set_plane(_A_SO0, iv, arr, odd_plane, 'set_so0')

    def set_so1(self, iv, arr, odd_plane=False):
        self._set_plane(_A_SO1, iv, arr, odd_plane, 'set_so1')

    def set_cfl(self, arr, odd_plane=False):
        self._set_plane(_A_CFL, None, arr, odd_plane, 'set_cfl')

    def get_so0p(self, iv, odd_plane=False):
        if iv >= self._nvar:
            raise IndexError('get_so0p(): out of nvar range')
        out = np.empty(self._grid.nselm - (1 if odd_plane else 0))
        self._before()
        try:
            check(_lib.lib().scst_so0p(self._h, iv, int(bool(odd_plane)),
                                       dptr(out), out.size))
        finally:
            self._after()
        return Array(out)

    # -- march ----------------------------------------------------------------
    def _stage(self, stage, alpha=2, odd=False):
        self._call(_lib.lib().scst_stage, stage, alpha, int(bool(odd)))

    def update_cfl(self, odd_plane):
        self._stage(_ST_CFL, odd=odd_plane)

    def march_half_so0(self, odd_plane):
        self._stage(_ST_SO0, odd=odd_plane)

    def treat_boundary_so0(self):
        self._stage(_ST_BND0)

    def treat_boundary_so1(self):
        self._stage(_ST_BND1)

    def setup_march(self):
        self.update_cfl(False)

    def _march(self, alpha, steps):
        if steps < 0:
            raise TypeError('steps must be non-negative')
        self._call(_lib.lib().scst_march, alpha, int(steps))

    def march_half_so1_alpha0(self, odd_plane):
        self._stage(_ST_SO1, 0, odd_plane)

    def march_half_so1_alpha1(self, odd_plane):
        self._stage(_ST_SO1, 1, odd_plane)

    def march_half_so1_alpha2(self, odd_plane):
        self._stage(_ST_SO1, 2, odd_plane)

    def march_half1_alpha0(self):
        self._stage(_ST_HALF1, 0)

    def march_half1_alpha1(self):
        self._stage(_ST_HALF1, 1)

    def march_half1_alpha2(self):
        self._stage(_ST_HALF1, 2)

    def march_half2_alpha0(self):
        self._stage(_ST_HALF2, 0)

    def march_half2_alpha1(self):
        self._stage(_ST_HALF2, 1)

    def march_half2_alpha2(self):
        self._stage(_ST_HALF2, 2)

    def march_alpha0(self, steps):
        self._march(0, steps)

    def march_alpha1(self, steps):
        self._march(1, steps)

    def march_alpha2(self, steps):
        self._march(2, steps)

    # Device-resident march without host synchronisation (for batching
    # several solvers on their own streams).
    def march_async(self, alpha, steps):
        self._before()
        check(_lib.lib().scst_march_async(self._h, alpha, int(steps)))

    def sync(self):
        check(_lib.lib().scst_sync(self._h))
        self._after()


class Solver(_SolverBase):
    """Generic solver with the default kernel hooks (core.hpp:1038-1056)."""

    def __init__(self, grid, time_increment, nvar):
        super(Solver, self).__init__(grid, time_increment, nvar)
        self._kernel = Kernel()

    def _init_handle(self, h):
        super(Solver, self)._init_handle(h)
        if not hasattr(self, '_kernel'):
            self._kernel = Kernel()

    @property
    def kernel(self):
        return self._kernel


class LinearScalarSolver(_SolverBase):
    """Linear scalar advection u_t + u_x = 0 (kernel/linear_scalar.hpp)."""

    _KIND = _K_LINEAR
    _NAME = 'LinearScalarSolver'
    _CELM = LinearScalarCelm
    _SELM = LinearScalarSelm

    def __init__(self, grid, time_increment):
        super(LinearScalarSolver, self).__init__(grid, time_increment, 1)


class InviscidBurgersSolver(_SolverBase):
    """Inviscid Burgers u_t + (u^2/2)_x = 0 (kernel/inviscid_burgers.hpp)."""

    _KIND = _K_BURGERS
    _NAME = 'InviscidBurgersSolver'
    _CELM = InviscidBurgersCelm
    _SELM = InviscidBurgersSelm

    def __init__(self, grid, time_increment):
        super(InviscidBurgersSolver, self).__init__(grid, time_increment, 1)


class SolverProxy(object):
    """solvcon/spacetime.py:39-73 routes every element flux through Python
    callbacks; per-element Python code cannot run in a HIP kernel."""

    def __init__(self, *args, **kw):
        raise NotImplementedError(
            'SolverProxy: Python solution-element hooks cannot run in a HIP '
            'kernel; use LinearScalarSolver or InviscidBurgersSolver')
