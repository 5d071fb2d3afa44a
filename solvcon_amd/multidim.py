"""Multidimensional Euler CESE solver on the GPU.

Drop-in for solvcon's ``EulerCore`` and ``GradientElement`` bindings
(cpp/solvcon/multidim/pymod/wrap_multidim.cpp:17-414).  State and geometry
live in HBM; every stage runs a HIP kernel of euler.hip.  solvcon's Python
``EulerField`` (solvcon/multidim/euler/_field.py) reads these arrays
unchanged; ``EulerCore.reduce`` gives its mass and CFL range from one FP64
device reduction.
"""

import ctypes

import numpy as np

from . import _lib as L
from .array import GhostArray

__all__ = ['EulerCore', 'GradientElement']

L.register({
    'sceuler_create': (L.c_int, [L.c_vp, L.c_dbl, L.c_int, L.P_vp]),
    'sceuler_destroy': (None, [L.c_vp]),
    'sceuler_counts': (L.c_int, [L.c_vp, L.P_i64]),
    'sceuler_get_param': (L.c_dbl, [L.c_vp, L.c_int]),
    'sceuler_set_param': (L.c_int, [L.c_vp, L.c_int, L.c_dbl]),
    'sceuler_table_info': (L.c_int, [L.c_vp, L.c_int, L.P_i64]),
    'sceuler_download': (L.c_int, [L.c_vp, L.c_int, L.P_dbl, L.c_i64]),
    'sceuler_upload': (L.c_int, [L.c_vp, L.c_int, L.P_dbl, L.c_i64]),
    'sceuler_stage': (L.c_int, [L.c_vp, L.c_int]),
    'sceuler_march': (L.c_int, [L.c_vp, L.c_i64]),
    'sceuler_march_async': (L.c_int, [L.c_vp, L.c_i64]),
    'sceuler_sync': (L.c_int, [L.c_vp]),
    'sceuler_march_timed': (L.c_int, [L.c_vp, L.c_i64, L.P_dbl, L.P_dbl]),
    'sceuler_init_solution': (L.c_int, [L.c_vp, L.c_dbl, L.c_dbl, L.P_dbl,
                                        L.c_int, L.c_dbl]),
    'sceuler_add_bc': (L.c_int, [L.c_vp, L.c_int, L.P_i32, L.c_i64, L.P_dbl,
                                 L.c_i64]),
    'sceuler_normal_matrix': (L.c_int, [L.c_vp, L.c_i32, L.P_dbl]),
    'sceuler_reduce': (L.c_int, [L.c_vp, L.P_dbl]),
    'sceuler_create_batch': (L.c_int, [L.c_vp, L.c_i64, L.P_i32, L.P_i64,
                                       L.c_int, L.P_vp]),
    'sceuler_batch_set_dt': (L.c_int, [L.c_vp, L.P_dbl]),
    'sceuler_batch_cfl_dt': (L.c_int, [L.c_vp, L.c_dbl, L.P_dbl]),
    'sceuler_march_batch': (L.c_int, [L.c_vp, L.P_dbl]),
    'sceuler_batch_done_ms': (L.c_int, [L.c_vp, L.P_dbl, L.c_i64]),
    'sceuler_cfl_dt': (L.c_int, [L.c_vp, L.c_dbl, L.P_dbl]),
    'scge_build': (L.c_int, [L.c_vp, L.P_dbl, L.c_i64, L.c_i64, L.c_i64,
                             L.c_i32, L.c_dbl, L.P_i32, L.P_dbl]),
    'scge_solve': (L.c_int, [L.c_vp, L.c_int, L.P_dbl, L.P_dbl, L.P_dbl]),
})

# Table ids of the C ABI.
_IDS = {'cevol': 0, 'cecnd': 1, 'sfcnd': 2, 'sfnml': 3, 'so0c': 4,
        'so0n': 5, 'so0t': 6, 'so1c': 7, 'so1n': 8, 'stm': 9, 'cflo': 10,
        'cflc': 11, 'gamma': 12}
# Stage ids of sceuler_stage.
_ST = {'prepare_ce': 0, 'update': 1, 'calc_solt': 2, 'calc_soln': 3,
       'calc_cfl': 4, 'calc_dsoln': 5, 'bc_soln': 6, 'bc_dsoln': 7,
       'march_substep': 8, 'clear_bc': 9}
_NONREFL, _SLIPWALL, _INLET = 0, 1, 2


class EulerCore(object):
    """The CESE solver of the Euler equations on a StaticMesh.

    ``sfc`` (not in the reference) orders the body cells along a Hilbert
    curve inside each cell-type group; results are identical either way.
    """

    SUBSTEP_RUN = 2

    def __init__(self, mesh, time_increment, sfc=True, batch=None):
        L.require_gpu()
        mesh._before()
        h = ctypes.c_void_p()
        if batch is None:
            self._ninst = 0
            L.check(L.lib().sceuler_create(mesh._h, float(time_increment),
                                           int(bool(sfc)), ctypes.byref(h)))
        else:
            inst, steps = batch
            inst = np.ascontiguousarray(inst, dtype=np.int32)
            steps = np.ascontiguousarray(steps, dtype=np.int64)
            if inst.size != mesh.ncell:
                raise ValueError('EulerCore batch: one instance id per cell')
            self._ninst = steps.size
            L.check(L.lib().sceuler_create_batch(
                mesh._h, steps.size, L.iptr(inst),
                steps.ctypes.data_as(L.P_i64), int(bool(sfc)),
                ctypes.byref(h)))
        self._h = h
        self.mesh = mesh
        out = (ctypes.c_int64 * 4)()
        L.lib().sceuler_counts(self._h, out)
        self._ndim, self._ncell, self._ngstcell, self._neq = list(out)
        # Two physical buffers each for so0 and so1: slot *c is buffer cur.
        self._cur = 0
        self._mirrors = {}
        for name in ('cevol', 'cecnd', 'sfcnd', 'sfnml', 'so0t', 'stm',
                     'cflo', 'cflc', 'gamma'):
            self._mirrors[name] = self._mk(name)
        for b in (0, 1):
            self._mirrors[('so0', b)] = self._mk(('so0', b))
            self._mirrors[('so1', b)] = self._mk(('so1', b))

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and L._lib is not None:
            L._lib.sceuler_destroy(h)
            self._h = None

    # -- host mirrors -------------------------------------------------------

    def _tid(self, key):
        if isinstance(key, tuple):
            kind, b = key
            cur = b == self._cur
            return _IDS[kind + ('c' if cur else 'n')]
        return _IDS[key]

    def _mk(self, key):
        info = (ctypes.c_int64 * 4)()
        L.check(L.lib().sceuler_table_info(self._h, self._tid(key), info))
        rows, w, ngh, rank = list(info)
        if rank == 1:
            shape = (rows,)
        elif rank == 2:
            shape = (rows, w)
        else:
            shape = (rows, w // self._ndim, self._ndim)

        def up(a):
            L.check(L.lib().sceuler_upload(self._h, self._tid(key), L.dptr(a),
                                           a.size))

        def down(a):
            L.check(L.lib().sceuler_download(self._h, self._tid(key),
                                             L.dptr(a), a.size))
        return L.Mirror(shape, np.float64, up, down, nghost=ngh)

    def _mirror(self, slot):
        if slot in ('so0c', 'so0n', 'so1c', 'so1n'):
            b = self._cur if slot.endswith('c') else 1 - self._cur
            return self._mirrors[(slot[:3], b)]
        return self._mirrors[slot]

    def _before(self):
        self.mesh._before()
        for m in self._mirrors.values():
            m.before()

    def _after(self):
        for m in self._mirrors.values():
            m.after()

    def _stage(self, name):
        self._before()
        try:
            L.check(L.lib().sceuler_stage(self._h, _ST[name]))
        finally:
            self._after()

    # -- information ----------------------------------------------------------

    ndim = property(lambda self: self._ndim)
    ncell = property(lambda self: self._ncell)
    ngstcell = property(lambda self: self._ngstcell)
    neq = property(lambda self: self._neq)
    time_increment = property(
        lambda self: L.lib().sceuler_get_param(self._h, 3))

    def _param(i):
        def get(self):
            return L.lib().sceuler_get_param(self._h, i)

        def put(self, v):
            L.check(L.lib().sceuler_set_param(self._h, i, float(v)))
        return property(get, put)

    @property
    def kernel_forms(self):
        """{cell type id: form} chosen by the autotuner of the first substep
        (None before it; no reference counterpart).  A form is 'fV', the
        fused kernel in variant V, or 'sA/B', the split pair (soln + BC +
        CFL, then dsoln) in variants A and B.  A variant is the minimum
        wavefronts per SIMD it is compiled for, followed by 's' when the
        block's own rows are staged in LDS, 't' when the fused kernel
        reads the neighbours' so0t from memory and 'g' when it prefetches
        the sub-face geometry into LDS."""
        if L.lib().sceuler_get_param(self._h, 4) < 0:
            return None

        def var(c):
            return '%d%s%s%s' % (c % 10, 's' if (c // 10) % 10 else '',
                                 't' if (c // 100) % 10 else '',
                                 'g' if (c // 1000) % 10 else '')
        types = np.unique(self.mesh.cltpn.ndarray[self.mesh.ngstcell:])
        out = {}
        for t in types:
            c = int(L.lib().sceuler_get_param(self._h, 100 + int(t)))
            out[int(t)] = 'f' + var(c) if c < 10000 else 's%s/%s' % (
                var((c - 10000) // 100), var(c % 100))
        return out

    @property
    def exact_passes(self):
        """How many exact passes the kernel pairs have run (no reference
        counterpart).  Every kernel launch of a march or stage is a fast
        form (the Jacobian's structural zeros skipped, quotients by one
        denominator sharing a reciprocal) followed by the exact form, which
        rewrites the launch's cells only when a cell of the fast form met a
        value that is not finite or lies near the ends of the exponent
        range.  The results are the exact form's either way."""
        n = L.lib().sceuler_get_param(self._h, 7)
        if n < 0:
            raise RuntimeError('EulerCore: exact pass count unavailable')
        return int(n)

    sigma0 = _param(0)
    taumin = _param(1)
    tauscale = _param(2)
    del _param

    # -- preparation -----------------------------------------------------------

    def prepare_ce(self):
        self._stage('prepare_ce')

    def init_solution(self, gamma, rho, v, p):
        vel = np.zeros(3)
        v = list(v)
        if len(v) != self._ndim:
            raise ValueError(
                'EulerCore::init_solution: velocity size %d must equal ndim '
                '%d' % (len(v), self._ndim))
        vel[:len(v)] = v
        self._before()
        try:
            L.check(L.lib().sceuler_init_solution(
                self._h, float(gamma), float(rho), L.dptr(vel), len(v),
                float(p)))
        finally:
            self._after()

    # -- march -----------------------------------------------------------------

    def update(self):
        L.check(L.lib().sceuler_stage(self._h, _ST['update']))
        self._cur = 1 - self._cur

    def calc_solt(self):
        self._stage('calc_solt')

    def calc_soln(self):
        self._stage('calc_soln')

    def calc_cfl(self):
        self._stage('calc_cfl')

    def calc_dsoln(self):
        self._stage('calc_dsoln')

    def march_substep(self):
        self._before()
        try:
            L.check(L.lib().sceuler_stage(self._h, _ST['march_substep']))
            self._cur = 1 - self._cur
        finally:
            self._after()

    def march(self, steps):
        steps = int(steps)
        self._before()
        try:
            L.check(L.lib().sceuler_march(self._h, steps))
            if steps > 0:
                self._cur = (self._cur + 2 * steps) % 2
        finally:
            self._after()

    def march_timed(self, steps):
        """March and return (average fused-substep kernel time, total march
        time) in milliseconds from HIP events on the solver's stream."""
        self._before()
        sub = ctypes.c_double()
        tot = ctypes.c_double()
        try:
            L.check(L.lib().sceuler_march_timed(self._h, int(steps),
                                                ctypes.byref(sub),
                                                ctypes.byref(tot)))
        finally:
            self._after()
        return sub.value, tot.value

    # -- time step (no reference counterpart) ---------------------------------------

    def cfl_dt(self, cfl=0.8):
        """dt that puts the largest calc_cfl number of the current so0n at
        ``cfl``: cfl / max(wspd / (2 dist)), reduced on the device."""
        self._before()
        out = ctypes.c_double()
        try:
            L.check(L.lib().sceuler_cfl_dt(self._h, float(cfl),
                                           ctypes.byref(out)))
        finally:
            self._after()
        return out.value

    def set_time_increment(self, dt):
        """Change dt after construction (the reference fixes it in the
        constructor)."""
        L.check(L.lib().sceuler_set_param(self._h, 3, float(dt)))

    # -- batched solves (no reference counterpart) ---------------------------------

    def batch_set_dt(self, dt):
        """Time increment of every batched instance."""
        if self._ninst == 0:
            raise RuntimeError('EulerCore: not a batched solver')
        d = np.ascontiguousarray(dt, dtype=np.float64)
        if d.size != self._ninst:
            raise ValueError('EulerCore batch: one dt per instance')
        L.check(L.lib().sceuler_batch_set_dt(self._h, L.dptr(d)))

    def batch_cfl_dt(self, cfl=0.8):
        """Set and return the dt of every instance that puts its largest
        calc_cfl number of the current so0n at ``cfl``."""
        self._before()
        out = np.zeros(self._ninst)
        try:
            L.check(L.lib().sceuler_batch_cfl_dt(self._h, float(cfl),
                                                 L.dptr(out)))
        finally:
            self._after()
        return out

    @property
    def batch_chunk(self):
        """Cells per rank chunk of the batched march: consecutive ranks
        (instances by descending step count) holding about this many cells
        run all their steps before the next chunk, so a chunk can stay in the
        last-level cache; 0 marches every rank at once.  Results do not
        depend on it."""
        return int(L.lib().sceuler_get_param(self._h, 5))

    @batch_chunk.setter
    def batch_chunk(self, cells):
        L.check(L.lib().sceuler_set_param(self._h, 5, float(cells)))

    @property
    def use_graph(self):
        """Whether ``march`` and ``march_batch`` replay each whole step
        (the half-step pair) from a captured hipGraph instead of launching
        its kernels one by one (default False, or the environment's
        SC_EULER_GRAPH).  Results do not depend on it."""
        return L.lib().sceuler_get_param(self._h, 6) != 0.0

    @use_graph.setter
    def use_graph(self, on):
        L.check(L.lib().sceuler_set_param(self._h, 6, 1.0 if on else 0.0))

    def march_batch(self):
        """March every instance its own step count; returns the device
        time of the march in milliseconds."""
        self._before()
        ms = ctypes.c_double()
        try:
            L.check(L.lib().sceuler_march_batch(self._h, ctypes.byref(ms)))
        finally:
            self._after()
        return ms.value

    def batch_done_ms(self):
        """Per instance (by instance id), the milliseconds from the start
        of the last ``march_batch`` until that instance's last step ran, on
        the device clock."""
        out = np.zeros(self._ninst, dtype=np.float64)
        L.check(L.lib().sceuler_batch_done_ms(
            self._h, out.ctypes.data_as(L.P_dbl), out.size))
        return out

    # -- boundary conditions ------------------------------------------------------

    def _add_bc(self, kind, faces, value=()):
        f = np.ascontiguousarray(np.asarray(list(faces), dtype=np.int64)
                                 .astype(np.int32))
        v = np.ascontiguousarray(np.asarray(list(value), dtype=np.float64))
        vv = v if v.size else np.zeros(1)
        self._before()
        L.check(L.lib().sceuler_add_bc(self._h, kind, L.iptr(f), f.size,
                                       L.dptr(vv), v.size))

    def add_nonrefl(self, faces):
        self._add_bc(_NONREFL, faces)

    def add_slipwall(self, faces):
        self._add_bc(_SLIPWALL, faces)

    def add_inlet(self, faces, value):
        self._add_bc(_INLET, faces, value)

    def clear_bc(self):
        L.check(L.lib().sceuler_stage(self._h, _ST['clear_bc']))

    def get_normal_matrix(self, ifc):
        self._before()
        out = np.zeros(self._ndim * self._ndim)
        L.check(L.lib().sceuler_normal_matrix(self._h, int(ifc),
                                              L.dptr(out)))
        nd = self._ndim
        return [list(out[i * nd:(i + 1) * nd]) for i in range(nd)]

    def bc_soln(self):
        self._stage('bc_soln')

    def bc_dsoln(self):
        self._stage('bc_dsoln')

    # -- diagnostics ----------------------------------------------------------------

    def reduce(self):
        """(sum of so0n density times cell volume, min cflo, max cflo) over
        the body cells, reduced on the device in FP64."""
        self._before()
        out = np.zeros(3)
        L.check(L.lib().sceuler_reduce(self._h, L.dptr(out)))
        return float(out[0]), float(out[1]), float(out[2])

    # -- snapshots (no reference counterpart) --------------------------------

    #: Tables a snapshot holds: both solution levels, CFL numbers and gamma,
    #: every row (ghosts first).
    SNAPSHOT_TABLES = ('so0c', 'so0n', 'so1c', 'so1n', 'cflo', 'cflc',
                       'gamma')

    def snapshot(self):
        """Host copy of the solver state that a later march reads (the
        tables of SNAPSHOT_TABLES and time_increment).  Mesh, parameters and
        boundary conditions are configuration and are not included."""
        out = {t: getattr(self, t).ndarray.copy() for t in self.SNAPSHOT_TABLES}
        out['time_increment'] = np.float64(self.time_increment)
        return out

    def restore(self, snap):
        """Load a snapshot taken from a solver on an identical mesh; a march
        from here reproduces the original run bit for bit."""
        for t in self.SNAPSHOT_TABLES:
            dst = getattr(self, t).ndarray
            src = np.asarray(snap[t])
            if src.shape != dst.shape or src.dtype != dst.dtype:
                raise ValueError('EulerCore.restore: %s has shape %s (%s), '
                                 'expected %s (%s)' % (t, src.shape, src.dtype,
                                                       dst.shape, dst.dtype))
        for t in self.SNAPSHOT_TABLES:
            getattr(self, t).ndarray[...] = snap[t]
        if self._ninst == 0:
            # A batched solver keeps its per-instance dt.
            self.set_time_increment(float(snap['time_increment']))

    def save(self, path):
        """Write snapshot() to an .npz file."""
        np.savez(path, **self.snapshot())

    def load(self, path):
        """restore() from an .npz file written by save()."""
        with np.load(path, allow_pickle=False) as d:
            self.restore({k: d[k] for k in d.files})


for _name in ('cevol', 'cecnd', 'sfcnd', 'sfnml', 'so0c', 'so0n', 'so0t',
              'so1c', 'so1n', 'stm', 'cflo', 'cflc', 'gamma'):
    setattr(EulerCore, _name,
            property(lambda self, n=_name: GhostArray(self, n)))
del _name


class GradientElement(object):
    """Gradient element of one cell (GradientElement.hpp:86-164), evaluated
    on the GPU."""

    def __init__(self, mesh, cecnd, icl, tau):
        mesh._before()
        arr = cecnd.ndarray if isinstance(cecnd, GhostArray) else \
            np.asarray(cecnd)
        ngh = cecnd.nghost if isinstance(cecnd, GhostArray) else \
            mesh.ngstcell
        ce = np.ascontiguousarray(arr, dtype=np.float64)
        iout = np.zeros(34, dtype=np.int32)
        dout = np.zeros(37)
        L.check(L.lib().scge_build(mesh._h, L.dptr(ce), ce.shape[0],
                                   ce.shape[1], ngh, int(icl), float(tau),
                                   L.iptr(iout), L.dptr(dout)))
        self._mesh = mesh
        self._icl = int(icl)
        self._ndim = int(iout[0])
        self._clnfc = int(iout[1])
        self._nfge = int(iout[2])
        self._rcl = [int(x) for x in iout[4:4 + self._clnfc]]
        self._faces = iout[10:34].reshape(8, 3)
        self._ofg1 = float(dout[0])
        self._idis = dout[1:19].reshape(6, 3)
        self._jdis = dout[19:37].reshape(6, 3)

    icl = property(lambda self: self._icl)
    ndim = property(lambda self: self._ndim)
    clnfc = property(lambda self: self._clnfc)
    nfge = property(lambda self: self._nfge)
    nfge_inverse = property(lambda self: self._ofg1)

    def _ifl(self, ifl):
        if ifl < 0 or ifl >= self._clnfc:
            raise IndexError('GradientElement: ifl %d out of range [0, %d)'
                             % (ifl, self._clnfc))

    def _d(self, d):
        if d < 0 or d >= self._ndim:
            raise IndexError('GradientElement: d %d out of range [0, %d)'
                             % (d, self._ndim))

    def _ifge(self, ifge):
        if ifge < 0 or ifge >= self._nfge:
            raise IndexError('GradientElement: ifge %d out of range [0, %d)'
                             % (ifge, self._nfge))

    def rcl(self, ifl):
        self._ifl(ifl)
        return self._rcl[ifl]

    def idis(self, ifl, d):
        self._ifl(ifl)
        self._d(d)
        return float(self._idis[ifl, d])

    def jdis(self, ifl, d):
        self._ifl(ifl)
        self._d(d)
        return float(self._jdis[ifl, d])

    def faces(self, ifge):
        self._ifge(ifge)
        return [int(x) for x in self._faces[ifge, :self._ndim]]

    def displacement_matrix(self, ifge):
        self._ifge(ifge)
        nd = self._ndim
        return [[float(self._idis[f - 1, d]) for d in range(nd)]
                for f in self._faces[ifge, :nd]]

    def solve_gradient(self, ifge, udf):
        self._ifge(ifge)
        udf = list(udf)
        if len(udf) != self._ndim:
            raise ValueError('GradientElement: udf size %d must equal ndim %d'
                             % (len(udf), self._ndim))
        dst = np.zeros(9)
        dm = self.displacement_matrix(ifge)
        for i, row in enumerate(dm):
            dst[i * 3:i * 3 + len(row)] = row
        u = np.zeros(3)
        u[:len(udf)] = udf
        out = np.zeros(3)
        L.check(L.lib().scge_solve(self._mesh._h, self._ndim, L.dptr(dst),
                                   L.dptr(u), L.dptr(out)))
        return [float(x) for x in out[:self._ndim]]
