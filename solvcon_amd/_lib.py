"""Load libsolvcon_amd.so, the HIP library behind every solver class.

There is no CPU fallback: when the library is missing or no GPU is visible,
constructing a solver raises.
"""

import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, 'libsolvcon_amd.so')

_lib = None

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
P_dbl = ctypes.POINTER(c_dbl)
P_i32 = ctypes.POINTER(c_i32)
P_i64 = ctypes.POINTER(c_i64)
P_vp = ctypes.POINTER(c_vp)

# name: (restype, argtypes)
_SIGS = {
    'sc_last_error': (c_int, [ctypes.c_char_p, c_int]),
    'sc_device_count': (c_int, []),
    'sc_device_name': (c_int, [ctypes.c_char_p, c_int]),
    'sc_device_sync': (c_int, []),
    'sc_stream_copy': (c_int, [c_i64, c_int, P_dbl]),
    # 1D core
    'sc1d_create': (c_int, [c_i64, c_dbl, P_vp]),
    'sc1d_clone': (c_int, [c_vp, P_vp]),
    'sc1d_destroy': (None, [c_vp]),
    'sc1d_info': (c_int, [c_vp, P_i64, P_dbl]),
    'sc1d_device_ptr': (c_int, [c_vp, c_int, ctypes.POINTER(P_dbl), P_i64]),
    'sc1d_upload': (c_int, [c_vp, c_int, P_dbl, c_i64]),
    'sc1d_download': (c_int, [c_vp, c_int, P_dbl, c_i64]),
    'sc1d_update_cfl': (c_int, [c_vp, c_int]),
    'sc1d_march_half_so0': (c_int, [c_vp, c_int]),
    'sc1d_march_half_so1': (c_int, [c_vp, c_int, c_int]),
    'sc1d_treat_boundary_so0': (c_int, [c_vp]),
    'sc1d_treat_boundary_so1': (c_int, [c_vp]),
    'sc1d_setup_march': (c_int, [c_vp]),
    'sc1d_march_half': (c_int, [c_vp, c_int, c_int]),
    'sc1d_march': (c_int, [c_vp, c_int, c_i64]),
    'sc1d_march_async': (c_int, [c_vp, c_int, c_i64]),
    'sc1d_sync': (c_int, [c_vp]),
    'sc1d_derived': (c_int, [c_vp, c_int, P_dbl]),
}

_OPTIONAL = set()


def register(sigs, optional=False):
    """Add C signatures (used by the modules that own them)."""
    _SIGS.update(sigs)
    if optional:
        _OPTIONAL.update(sigs)
    if _lib is not None:
        _bind(_lib, sigs)


def _bind(L, sigs):
    for name, (res, args) in sigs.items():
        try:
            fn = getattr(L, name)
        except AttributeError:
            if name in _OPTIONAL:
                continue
            raise
        fn.restype = res
        fn.argtypes = args


def lib():
    """The loaded HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                'solvcon_amd: %s is missing; build it with '
                '`python -c "import __graft_entry__ as g; g.build()"`'
                % LIBPATH)
        L = ctypes.CDLL(LIBPATH)
        _bind(L, _SIGS)
        _lib = L
    return _lib


_checked_gpu = False


def require_gpu():
    """Raise unless a GPU is visible to the HIP runtime."""
    global _checked_gpu
    if not _checked_gpu:
        if lib().sc_device_count() < 1:
            raise RuntimeError(
                'solvcon_amd: no GPU visible; the solvers run only on the '
                'MI355X through libsolvcon_amd.so')
        _checked_gpu = True


_EXC = {1: ValueError, 2: IndexError, 3: RuntimeError, 4: RuntimeError}


def check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        kind = lib().sc_last_error(buf, 1024)
        raise _EXC.get(kind or rc, RuntimeError)(buf.value.decode())


def dptr(a):
    return a.ctypes.data_as(P_dbl)


def iptr(a):
    return a.ctypes.data_as(P_i32)


class Mirror(object):
    """Host mirror of one device table.

    The reference hands out numpy views that alias the solver's memory
    (cpp/solvcon/python/common.hpp:195-216).  A discrete GPU cannot alias
    host numpy memory at speed, so each table keeps one host buffer that
    stands in for it: handing out a view marks the buffer possibly written;
    the owner uploads such buffers before its next device call and refreshes
    the ones still referenced after it, so held views keep seeing the
    solver's state.
    """

    __slots__ = ('shape', 'dtype', 'nghost', 'base', 'valid', 'handed',
                 '_up', '_down')

    def __init__(self, shape, dtype, up, down, nghost=0):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nghost = int(nghost)
        self.base = None
        self.valid = False
        self.handed = False
        self._up = up
        self._down = down

    def _ensure(self):
        if self.base is None:
            self.base = np.zeros(self.shape, dtype=self.dtype)
            self.valid = False
        if not self.valid:
            if self.base.size:
                self._down(self.base)
            self.valid = True

    def view(self):
        self._ensure()
        self.handed = True
        return self.base[...]

    def held(self):
        return self.base is not None and sys.getrefcount(self.base) > 2

    def before(self):
        if self.handed and self.base is not None and self.base.size:
            self._up(self.base)
            # Device and host now agree; a view released since it was handed
            # out cannot write again, so later calls need no upload.
            if not self.held():
                self.handed = False

    def after(self, written=True):
        if not written:
            return
        if self.handed and self.held():
            if self.base.size:
                self._down(self.base)
            self.valid = True
        else:
            self.handed = False
            self.valid = False

    def host_copy(self):
        """Fresh copy of the device table (does not hand out a view)."""
        self._ensure()
        return self.base.copy()
