"""Ghost-aware table handle with the Python surface of solvcon's SimpleArray.

The reference exposes mesh and solver tables as SimpleArray objects whose
first axis holds ``nghost`` ghost rows ahead of the body rows; integer row
indices and slice bounds on that axis count from the first body row, so
row -1 is the last ghost (cpp/solvcon/buffer/SimpleArray.hpp:2241-2362,
cpp/solvcon/buffer/pymod/array_common.hpp:345-481).  ``ndarray`` is the
whole table, ghosts included.  Here a handle is bound to a named slot of
its owner and resolves the owner's current host mirror at every use, so a
held handle follows buffer swaps the way a held SimpleArray does.
"""

import numpy as np

__all__ = ['GhostArray']


class GhostArray(object):

    __slots__ = ('_owner', '_slot')

    def __init__(self, owner, slot):
        self._owner = owner
        self._slot = slot

    def _mirror(self):
        return self._owner._mirror(self._slot)

    @property
    def ndarray(self):
        return self._mirror().view()

    @property
    def nghost(self):
        return self._mirror().nghost

    @property
    def nbody(self):
        m = self._mirror()
        return m.shape[0] - m.nghost if m.shape else 0

    @property
    def shape(self):
        return self._mirror().shape

    @property
    def size(self):
        return int(np.prod(self._mirror().shape))

    @property
    def ndim(self):
        return len(self._mirror().shape)

    @property
    def dtype(self):
        return self._mirror().dtype

    def __len__(self):
        return self.size

    def __array__(self, dtype=None, copy=None):
        a = self.ndarray
        return a if dtype is None else a.astype(dtype)

    def _key(self, key):
        ngh = self._mirror().nghost
        rows = self._mirror().shape[0] if self._mirror().shape else 0

        def shift(k):
            if isinstance(k, slice):
                if ngh == 0:
                    return k
                start = None if k.start is None else k.start + ngh
                stop = None if k.stop is None else k.stop + ngh
                return slice(start, stop, k.step)
            if isinstance(k, (int, np.integer)):
                i = int(k) + ngh
                if i < 0 or i >= rows:
                    raise IndexError('row %d out of range [%d, %d)'
                                     % (int(k), -ngh, rows - ngh))
                return i
            return k

        if isinstance(key, tuple):
            if not key:
                return key
            if key[0] is Ellipsis:
                return key
            return (shift(key[0]),) + key[1:]
        if key is Ellipsis:
            return key
        return shift(key)

    def __getitem__(self, key):
        return self.ndarray[self._key(key)]

    def __setitem__(self, key, value):
        self.ndarray[self._key(key)] = value

    def fill(self, value):
        self.ndarray.fill(value)

    def copy(self):
        return self.ndarray.copy()

    def __repr__(self):
        return 'GhostArray(%s, shape=%s, nghost=%d)' % (
            self._slot, self.shape, self.nghost)
