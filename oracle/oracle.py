"""Oracle (test infrastructure only).

ctypes access to the CPU restatement of solvcon's CESE Euler march
(``liboracle.so``, built from the C++ sources beside this file).  Only
``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use it, always as the checker and never as the product path.

Tables are returned as numpy views over the oracle's own memory, with the
ghost rows first, the way the reference exposes ``SimpleArray.ndarray``.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, 'liboracle.so')
_lib = None

_I64P = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(['make', '-s', '-C', _HERE, '-B'], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            raise RuntimeError('oracle not built: run make -C oracle')
        L = ctypes.CDLL(_LIBPATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        L.orc_last_error.restype = ctypes.c_int
        L.orc_last_error.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.orc1d_stage.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int64, ctypes.c_double,
                                  dp, dp, dp, dp, dp]
        L.orc1d_march.argtypes = [ctypes.c_int, ctypes.c_int64,
                                  ctypes.c_int64, ctypes.c_double,
                                  dp, dp, dp, dp, dp]
        L.orc1d_derived.argtypes = [ctypes.c_int, ctypes.c_int64, dp, dp, dp]
        L.orcmesh_create.restype = vp
        L.orcmesh_create.argtypes = [ctypes.c_int, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_int32]
        L.orcmesh_destroy.argtypes = [vp]
        L.orcmesh_counts.argtypes = [vp, _I64P]
        L.orcmesh_array.argtypes = [vp, ctypes.c_char_p,
                                    ctypes.POINTER(ctypes.c_void_p), _I64P,
                                    ctypes.POINTER(ctypes.c_int), _I64P,
                                    ctypes.POINTER(ctypes.c_int)]
        L.orcmesh_build_interior.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orcmesh_build_boundary.argtypes = [vp]
        L.orcmesh_build_ghost.argtypes = [vp]
        L.orcmesh_add_bc.argtypes = [vp, ctypes.c_char_p, ip, ctypes.c_int64]
        L.orcmesh_bc.restype = ctypes.c_int64
        L.orcmesh_bc.argtypes = [vp, ctypes.c_int, ctypes.c_char_p,
                                 ctypes.c_int, ip, ip]
        L.orceuler_create.restype = vp
        L.orceuler_create.argtypes = [vp, ctypes.c_double]
        L.orceuler_destroy.argtypes = [vp]
        L.orceuler_set_threads.argtypes = [vp, ctypes.c_int]
        L.orceuler_counts.argtypes = [vp, _I64P]
        L.orceuler_get_param.restype = ctypes.c_double
        L.orceuler_get_param.argtypes = [vp, ctypes.c_int]
        L.orceuler_set_param.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.orceuler_array.argtypes = [vp, ctypes.c_char_p,
                                     ctypes.POINTER(ctypes.c_void_p), _I64P,
                                     ctypes.POINTER(ctypes.c_int), _I64P]
        L.orceuler_call.argtypes = [vp, ctypes.c_char_p]
        L.orceuler_march.argtypes = [vp, ctypes.c_int64]
        L.orceuler_init_solution.argtypes = [vp, ctypes.c_double,
                                             ctypes.c_double, dp,
                                             ctypes.c_double]
        L.orceuler_add_bc.argtypes = [vp, ctypes.c_int, ip, ctypes.c_int64,
                                      dp, ctypes.c_int64]
        L.orceuler_normal_matrix.argtypes = [vp, ctypes.c_int32, dp]
        L.orcge_build.argtypes = [vp, dp, ctypes.c_int64, ctypes.c_int64,
                                  ctypes.c_int64, ctypes.c_int32,
                                  ctypes.c_double, ip, ip, dp, dp, ip, dp]
        L.orcge_solve.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orcge_det.restype = ctypes.c_double
        L.orcge_det.argtypes = [ctypes.c_int, dp]
        L.orcge_adj.argtypes = [ctypes.c_int, dp, dp]
        st = [ctypes.c_int64, ctypes.c_int64, ctypes.c_double, dp, dp, dp,
              dp]
        L.orcst_grid.restype = ctypes.c_int
        L.orcst_grid.argtypes = [ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int64, dp]
        L.orcst_grid_from.restype = ctypes.c_int
        L.orcst_grid_from.argtypes = [dp, ctypes.c_int64, dp]
        L.orcst_stage.argtypes = [ctypes.c_int] * 4 + st
        L.orcst_march.argtypes = [ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int64] + st
        L.orcst_eval.restype = ctypes.c_double
        L.orcst_eval.argtypes = [ctypes.c_int] * 3 + [ctypes.c_int64] * 2 + st
        _lib = L
    return _lib


_EXC = {1: ValueError, 2: IndexError, 3: RuntimeError}


def _check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(512)
        kind = lib().orc_last_error(buf, 512)
        raise _EXC.get(kind or rc, RuntimeError)(buf.value.decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


# ---------------------------------------------------------------- 1D -------

class Euler1D:
    """Restated Euler1DCore state (Euler1DCore.cpp:22-35) with the fields of
    Euler1DSolver.init_solver (solvcon/onedim/euler1d.py:41-55)."""

    def __init__(self, ncoord, time_increment):
        if ncoord % 2 == 0:
            raise ValueError('ncoord cannot be even')
        self.ncoord = ncoord
        self.time_increment = time_increment
        self.coord = np.zeros(ncoord)
        self.cfl = np.zeros(ncoord)
        self.so0 = np.zeros((ncoord, 3))
        self.so1 = np.zeros((ncoord, 3))
        self.gamma = np.full(ncoord, 1.4)

    def clone(self):
        o = Euler1D(self.ncoord, self.time_increment)
        for k in ('coord', 'cfl', 'so0', 'so1', 'gamma'):
            getattr(o, k)[...] = getattr(self, k)
        return o

    def _args(self):
        return (self.ncoord, self.time_increment, _dp(self.coord),
                _dp(self.cfl), _dp(self.so0), _dp(self.so1), _dp(self.gamma))

    def stage(self, stage, alpha=2, odd=False):
        lib().orc1d_stage(stage, alpha, int(bool(odd)), *self._args())

    def update_cfl(self, odd_plane):
        self.stage(0, odd=odd_plane)

    def march_half_so0(self, odd_plane):
        self.stage(1, odd=odd_plane)

    def march_half_so1(self, alpha, odd_plane):
        self.stage(2, alpha=alpha, odd=odd_plane)

    def treat_boundary_so0(self):
        self.stage(3)

    def treat_boundary_so1(self):
        self.stage(4)

    def setup_march(self):
        self.update_cfl(False)

    def march_half1(self, alpha):
        self.stage(5, alpha=alpha)

    def march_half2(self, alpha):
        self.stage(6, alpha=alpha)

    def march(self, alpha, steps):
        lib().orc1d_march(alpha, steps, *self._args())

    def derived(self, which):
        names = ('density', 'velocity', 'pressure', 'temperature',
                 'internal_energy', 'entropy')
        out = np.zeros(self.ncoord)
        lib().orc1d_derived(names.index(which), self.ncoord, _dp(self.so0),
                            _dp(self.gamma), _dp(out))
        return out


def sod_setup(ncoord=2001, xmin=-1.0, xmax=1.0, time_increment=0.4 / 600,
              gamma=1.4, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1)):
    """The Sod tube the way ShockTube.build_numerical sets it up
    (solvcon/onedim/euler1d.py:124-168), on the restated core."""
    o = Euler1D(ncoord, time_increment)
    o.coord[:] = np.linspace(xmin, xmax, num=ncoord)
    o.gamma.fill(gamma)

    def u2(rho, v, p):
        ie = 1. / (gamma - 1) * p / rho
        ke = v * v / 2
        return rho * (ie + ke)

    sl = o.coord < 0.0
    sr = np.logical_not(sl)
    o.so0[sl, 0] = left[0]
    o.so0[sr, 0] = right[0]
    o.so0[:, 1] = 0.0
    o.so0[sl, 2] = u2(*left)
    o.so0[sr, 2] = u2(*right)
    o.so1.fill(0)
    o.setup_march()
    return o


# -------------------------------------------------------------- mesh -------

def _view(ptr, shape, dtype):
    n = int(np.prod(shape)) if len(shape) else 0
    if n == 0:
        return np.zeros(shape, dtype=dtype)
    ctype = ctypes.c_double if dtype == np.float64 else ctypes.c_int32
    buf = (ctype * n).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype).reshape(shape)


class Mesh:
    """Restated StaticMesh (cpp/solvcon/mesh/StaticMesh.hpp:264-456)."""

    ARRAYS = ('ndcrd', 'fccnd', 'fcnml', 'fcara', 'clcnd', 'clvol', 'fctpn',
              'cltpn', 'clgrp', 'fcnds', 'fccls', 'clnds', 'clfcs', 'ednds',
              'bndfcs')

    def __init__(self, ndim, nnode, nface=0, ncell=0):
        self._h = lib().orcmesh_create(ndim, nnode, nface, ncell)

    def __del__(self):
        if getattr(self, '_h', None) and _lib is not None:
            _lib.orcmesh_destroy(self._h)
            self._h = None

    def _counts(self):
        out = (ctypes.c_int64 * 10)()
        lib().orcmesh_counts(self._h, out)
        return list(out)

    ndim = property(lambda self: self._counts()[0])
    nnode = property(lambda self: self._counts()[1])
    nface = property(lambda self: self._counts()[2])
    ncell = property(lambda self: self._counts()[3])
    nbound = property(lambda self: self._counts()[4])
    ngstnode = property(lambda self: self._counts()[5])
    ngstface = property(lambda self: self._counts()[6])
    ngstcell = property(lambda self: self._counts()[7])
    nedge = property(lambda self: self._counts()[8])
    nbcs = property(lambda self: self._counts()[9])

    def array(self, name):
        """Full table (ghost rows first) as a numpy view and its nghost."""
        ptr = ctypes.c_void_p()
        shape = (ctypes.c_int64 * 3)()
        rank = ctypes.c_int()
        ngh = ctypes.c_int64()
        kind = ctypes.c_int()
        _check(lib().orcmesh_array(self._h, name.encode(), ctypes.byref(ptr),
                                   shape, ctypes.byref(rank),
                                   ctypes.byref(ngh), ctypes.byref(kind)))
        dt = np.float64 if kind.value == 0 else np.int32
        return _view(ptr.value or 0, tuple(shape[:rank.value]), dt), ngh.value

    def __getattr__(self, name):
        if name in Mesh.ARRAYS:
            return self.array(name)[0]
        raise AttributeError(name)

    def build_interior(self, do_metric=True, build_edge=True):
        _check(lib().orcmesh_build_interior(self._h, int(do_metric),
                                            int(build_edge)))

    def build_boundary(self):
        _check(lib().orcmesh_build_boundary(self._h))

    def build_ghost(self):
        _check(lib().orcmesh_build_ghost(self._h))

    def add_bc(self, name, faces):
        f = np.ascontiguousarray(faces, dtype=np.int32)
        _check(lib().orcmesh_add_bc(self._h, name.encode(), _ip(f), f.size))

    def bc(self, ibc):
        n = lib().orcmesh_bc(self._h, ibc, None, 0, None, None)
        name = ctypes.create_string_buffer(256)
        f0 = np.zeros(n, dtype=np.int32)
        f1 = np.zeros(n, dtype=np.int32)
        lib().orcmesh_bc(self._h, ibc, name, 256, _ip(f0), _ip(f1))
        return name.value.decode(), f0, f1


def make_mesh(ndim, ndcrd, cltpn, clnds, build=True, bcs=()):
    """Build a restated mesh from node coordinates and cell definitions.

    clnds rows are (nnode, n0, n1, ...) like StaticMesh.clnds."""
    ndcrd = np.asarray(ndcrd, dtype=np.float64)
    cltpn = np.asarray(cltpn, dtype=np.int32)
    m = Mesh(ndim, ndcrd.shape[0], 0, cltpn.shape[0])
    m.ndcrd[:, :] = ndcrd
    m.cltpn[:] = cltpn
    cl = m.clnds
    if isinstance(clnds, np.ndarray):
        cl[:, :clnds.shape[1]] = clnds
    else:
        for i, row in enumerate(clnds):
            cl[i, :len(row)] = row
    if build:
        m.build_interior(do_metric=True, build_edge=True)
        for name, faces in bcs:
            m.add_bc(name, faces)
        m.build_boundary()
        m.build_ghost()
    return m


# ------------------------------------------------------------- euler -------

BC_NONREFL, BC_SLIPWALL, BC_INLET = 0, 1, 2


class Euler:
    """Restated EulerCore (cpp/solvcon/multidim/euler.hpp:56-211)."""

    ARRAYS = ('cevol', 'cecnd', 'sfcnd', 'sfnml', 'so0c', 'so0n', 'so0t',
              'so1c', 'so1n', 'stm', 'cflo', 'cflc', 'gamma')

    def __init__(self, mesh, time_increment, nthreads=1):
        self.mesh = mesh
        self._h = lib().orceuler_create(mesh._h, time_increment)
        if not self._h:
            raise ValueError('EulerCore: ndim must be 2 or 3')
        lib().orceuler_set_threads(self._h, nthreads)

    def __del__(self):
        if getattr(self, '_h', None) and _lib is not None:
            _lib.orceuler_destroy(self._h)
            self._h = None

    def set_threads(self, n):
        lib().orceuler_set_threads(self._h, n)

    def _counts(self):
        out = (ctypes.c_int64 * 4)()
        lib().orceuler_counts(self._h, out)
        return list(out)

    ndim = property(lambda self: self._counts()[0])
    ncell = property(lambda self: self._counts()[1])
    ngstcell = property(lambda self: self._counts()[2])
    neq = property(lambda self: self._counts()[3])
    time_increment = property(
        lambda self: lib().orceuler_get_param(self._h, 3))

    def _param(i):
        return property(
            lambda self: lib().orceuler_get_param(self._h, i),
            lambda self, v: lib().orceuler_set_param(self._h, i, float(v)))

    sigma0 = _param(0)
    taumin = _param(1)
    tauscale = _param(2)
    del _param

    def array(self, name):
        ptr = ctypes.c_void_p()
        shape = (ctypes.c_int64 * 3)()
        rank = ctypes.c_int()
        ngh = ctypes.c_int64()
        _check(lib().orceuler_array(self._h, name.encode(), ctypes.byref(ptr),
                                    shape, ctypes.byref(rank),
                                    ctypes.byref(ngh)))
        return _view(ptr.value or 0, tuple(shape[:rank.value]),
                     np.float64), ngh.value

    def __getattr__(self, name):
        if name in Euler.ARRAYS:
            return self.array(name)[0]
        raise AttributeError(name)

    def call(self, what):
        _check(lib().orceuler_call(self._h, what.encode()))

    def march(self, steps):
        _check(lib().orceuler_march(self._h, steps))

    def init_solution(self, gamma, rho, v, p):
        if len(v) != self.ndim:
            raise ValueError('velocity size must equal ndim')
        vel = np.zeros(3)
        vel[:len(v)] = v
        _check(lib().orceuler_init_solution(self._h, gamma, rho, _dp(vel), p))

    def add_bc(self, kind, faces, value=()):
        f = np.ascontiguousarray(faces, dtype=np.int32)
        v = np.ascontiguousarray(value, dtype=np.float64)
        if v.size == 0:
            v = np.zeros(1)
            nv = 0
        else:
            nv = v.size
        _check(lib().orceuler_add_bc(self._h, kind, _ip(f), f.size, _dp(v),
                                     nv))

    def normal_matrix(self, ifc):
        out = np.zeros(9)
        lib().orceuler_normal_matrix(self._h, ifc, _dp(out))
        nd = self.ndim
        return out.reshape(3, 3)[:nd, :nd].copy()


def gradient_element(mesh, cecnd, nghost, icl, tau):
    """Restated GradientElement over cecnd (full table, ghost rows first)."""
    ce = np.ascontiguousarray(cecnd, dtype=np.float64)
    info = np.zeros(3, dtype=np.int32)
    rcl = np.zeros(6, dtype=np.int32)
    idis = np.zeros(18)
    jdis = np.zeros(18)
    faces = np.zeros(24, dtype=np.int32)
    ofg1 = ctypes.c_double()
    lib().orcge_build(mesh._h, _dp(ce), ce.shape[0], ce.shape[1], nghost,
                      icl, tau, _ip(info), _ip(rcl), _dp(idis), _dp(jdis),
                      _ip(faces), ctypes.byref(ofg1))
    nd, nfc, nfge = (int(x) for x in info)
    return dict(ndim=nd, clnfc=nfc, nfge=nfge, nfge_inverse=ofg1.value,
                rcl=rcl[:nfc].copy(), idis=idis.reshape(6, 3)[:nfc, :nd].copy(),
                jdis=jdis.reshape(6, 3)[:nfc, :nd].copy(),
                faces=faces.reshape(8, 3)[:nfge, :nd].copy())


def ge_solve(ndim, dst, udf):
    a = np.zeros((3, 3))
    a[:ndim, :ndim] = dst
    u = np.zeros(3)
    u[:ndim] = udf
    out = np.zeros(3)
    lib().orcge_solve(ndim, _dp(a), _dp(u), _dp(out))
    return out[:ndim]


def ge_det(ndim, a):
    m = np.zeros((3, 3))
    m[:ndim, :ndim] = a
    return lib().orcge_det(ndim, _dp(m))


def ge_adj(ndim, a):
    m = np.zeros((3, 3))
    m[:ndim, :ndim] = a
    out = np.zeros(9)
    lib().orcge_adj(ndim, _dp(m), _dp(out))
    return out.reshape(3, 3)[:ndim, :ndim].copy()


# --------------------------------------------------------- spacetime -------

ST_DEFAULT, ST_LINEAR, ST_BURGERS = 0, 1, 2


def st_grid(xmin=None, xmax=None, ncelm=None, xloc=None):
    """Coordinates of spacetime Grid(xmin, xmax, ncelm) or Grid(xloc)
    (cpp/solvcon/spacetime/core.cpp:14-98); ValueError on invalid
    arguments."""
    if xloc is not None:
        xloc = np.ascontiguousarray(xloc, dtype=np.float64)
        xc = np.zeros(max(2 * (xloc.size - 1) + 5, 0))
        rc = lib().orcst_grid_from(_dp(xloc), xloc.size, _dp(xc))
    else:
        xc = np.zeros(max(2 * int(ncelm) + 5, 0))
        rc = lib().orcst_grid(xmin, xmax, int(ncelm), _dp(xc))
    if rc != 0:
        raise ValueError('invalid grid')
    return xc


class StSolver:
    """Restated spacetime SolverBase (core.hpp:706-1036) over host arrays:
    kind 0 the default kernel hooks, 1 LinearScalar, 2 InviscidBurgers."""

    def __init__(self, kind, xc, time_increment, nvar=1):
        self.kind = kind
        self.xc = np.array(xc, dtype=np.float64)
        self.nx = self.xc.size
        self.ncelm = (self.nx - 5) // 2
        self.nselm = self.ncelm + 1
        self.nvar = nvar
        self.time_increment = time_increment
        self.so0 = np.zeros((self.nx, nvar))
        self.so1 = np.zeros((self.nx, nvar))
        self.cfl = np.zeros(self.nx)

    def clone(self):
        o = StSolver(self.kind, self.xc, self.time_increment, self.nvar)
        o.so0[...] = self.so0
        o.so1[...] = self.so1
        o.cfl[...] = self.cfl
        return o

    def _args(self):
        return (self.nx, self.nvar, self.time_increment, _dp(self.xc),
                _dp(self.so0), _dp(self.so1), _dp(self.cfl))

    def stage(self, stage, alpha=2, odd=False):
        lib().orcst_stage(self.kind, stage, alpha, int(bool(odd)),
                          *self._args())

    def march_half_so0(self, odd):
        self.stage(0, odd=odd)

    def march_half_so1(self, alpha, odd):
        self.stage(1, alpha, odd)

    def update_cfl(self, odd):
        self.stage(2, odd=odd)

    def treat_boundary_so0(self):
        self.stage(3)

    def treat_boundary_so1(self):
        self.stage(4)

    def setup_march(self):
        self.update_cfl(False)

    def march(self, alpha, steps):
        lib().orcst_march(self.kind, alpha, steps, *self._args())

    def sel(self, odd):
        """xindex of the solution elements of one plane."""
        return 2 + 2 * np.arange(self.nselm - int(bool(odd))) + int(bool(odd))

    def eval(self, what, y, iv=0, alpha=2):
        return lib().orcst_eval(self.kind, what, alpha, y, iv, *self._args())
