"""Serving form: continuous batching of 1D, 2D and 3D Euler solves on one
GPU.

Requests arrive at given times, each carrying its problem: kind, mesh
size, step count and seed.  The loop serves them as follows.
- Set-up: a worker thread builds each arrived request's solver on the
  device (mesh, CE geometry, initial state, boundary conditions, dt), in
  arrival order, while the device bytes of the solvers being set up fit
  the set-up budget.  The bytes are what the handles actually allocate
  (``device_bytes``); before a build they are estimated from the bytes per
  cell measured so far.
- Admission: a set-up solve joins its engine as soon as the engine has room
  for it.  2D and 3D solves join a resident arena (``Arena``, the
  ``scserve_*`` entries): their tables are copied into free whole-block
  ranges of the arena's per-type tables, at the arena's step boundary.
  1D tubes join the tube engine (``Tubes``, ``scserve1d_*``), which takes
  their device tables over.
- Stepping: each engine marches every solve it holds together, on its own
  stream: one fused launch per cell type per substep over the whole arena,
  and one tiled launch per round for all tubes, each solve with its own dt
  and step count.  The loop keeps at most two chunks of steps queued per
  engine, so arrivals wait at most that long to join.
- Completion: a solve that has run its steps stops at its last step, its
  state is drained to pinned host memory on the engine's stream, and the
  loop collects it when the copy has completed; its ranges are then free
  for the next admission.
Latency is arrival to result on the host.

The baseline serves the same requests one after another with the plain
blocking calls (each request a plain EulerCore / Euler1DCore marched
alone).  A FIFO single-server queue over the measured service times gives
its latencies at the same arrival times.

There is no reference counterpart; every solve is the reference's plain
march (Euler1DCore::march_alpha<2>, EulerCore::march), and its result is
the same alone and in the mix.
"""

import collections
import ctypes
import threading
import time

import numpy as np

from . import _lib as L
from .gen import (tri_mesh_def, riemann_box, quadrant_state, mixed_mesh_def,
                  blast_state, build_static_mesh)

__all__ = ['Request', 'make_requests', 'payload', 'Job', 'Arena', 'Tubes',
           'Server', 'serve', 'serve_serial', 'fifo_latency']

P_i64 = L.P_i64
L.register({
    'sc1d_query': (L.c_int, [L.c_vp, ctypes.POINTER(L.c_int)]),
    'sceuler_query': (L.c_int, [L.c_vp, ctypes.POINTER(L.c_int)]),
    'sceuler_device_bytes': (L.c_int, [L.c_vp, P_i64]),
    'scmesh_device_bytes': (L.c_int, [L.c_vp, P_i64]),
    'scserve_create': (L.c_int, [L.c_int, P_i64, L.c_i64, L.c_i64, L.P_vp]),
    'scserve_destroy': (None, [L.c_vp]),
    'scserve_admit': (L.c_int, [L.c_vp, L.c_vp, L.c_i64, P_i64]),
    'scserve_step': (L.c_int, [L.c_vp, L.c_i64]),
    'scserve_pending': (L.c_int, [L.c_vp, P_i64]),
    'scserve_query': (L.c_int, [L.c_vp, L.c_i64, ctypes.POINTER(L.c_int)]),
    'scserve_result': (L.c_int, [L.c_vp, L.c_i64, L.P_dbl, L.P_dbl]),
    'scserve_sync': (L.c_int, [L.c_vp]),
    'scserve_info': (L.c_int, [L.c_vp, P_i64]),
    'scserve1d_create': (L.c_int, [L.c_i64, L.P_vp]),
    'scserve1d_destroy': (None, [L.c_vp]),
    'scserve1d_admit': (L.c_int, [L.c_vp, L.c_vp, L.c_int, L.c_i64, P_i64]),
    'scserve1d_step': (L.c_int, [L.c_vp]),
    'scserve1d_pending': (L.c_int, [L.c_vp, P_i64]),
    'scserve1d_query': (L.c_int, [L.c_vp, L.c_i64,
                                  ctypes.POINTER(L.c_int)]),
    'scserve1d_result': (L.c_int, [L.c_vp, L.c_i64, L.P_dbl, L.P_dbl,
                                   L.P_dbl]),
})

Request = collections.namedtuple(
    'Request', 'rid kind size steps seed arrival')

# Cell type ids (solvcon's StaticMesh).
QUAD, TRI, HEX, TET, PRISM, PYRAMID = 3, 4, 5, 6, 7, 8


def make_requests(n, rate, mix=(0.5, 0.35, 0.15), lo=1e3, hi=2e6,
                  smin=100, smax=1000, seed=5):
    """n requests with Poisson arrivals at ``rate`` per second: kinds by
    ``mix`` (1D, 2D, 3D), sizes log-uniform in [lo, hi] cells, steps
    uniform in [smin, smax]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kinds = rng.choice(['1d', '2d', '3d'], size=n, p=list(mix))
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64)
    steps = rng.integers(smin, smax + 1, n)
    seeds = rng.integers(0, 2 ** 31, n)
    gaps = rng.exponential(1.0 / rate, n) if rate > 0 else np.zeros(n)
    arr = np.cumsum(gaps) - gaps[0]
    return [Request(i, str(kinds[i]), int(sizes[i]), int(steps[i]),
                    int(seeds[i]), float(arr[i])) for i in range(n)]


# Initial CFL number of the 2D and 3D requests.  The flow these problems
# develop raises the CFL number by up to about half again (a blast of
# pressure ratio 100, Riemann boxes of ratio up to 15); at 0.6 it stays
# below 1 and the solves stay finite.
SERVE_CFL = 0.6


def payload(req):
    """Host-side problem of a request (made by the client before it is
    sent): mesh definition, initial-state function and initial CFL number.
    1D: a Sod tube.  2D: a jittered triangle mesh of the unit square with a
    seeded Riemann box.  3D: the mixed block with a blast of radius
    max(0.1, 3 h) (at least three cells across, so it is resolved on the
    smallest blocks)."""
    if req.kind == '1d':
        return {'ncoord': max(5, req.size | 1)}
    if req.kind == '2d':
        m = max(2, int(round(np.sqrt(req.size / 2.0))))
        rho, p = riemann_box(req.seed % 100000)
        return {'defn': tri_mesh_def(m, seed=req.seed),
                'state': lambda cen: quadrant_state(cen, rho, p, 1.4),
                'cfl': SERVE_CFL}
    n = max(3, int(round((req.size / 2.41) ** (1.0 / 3.0))))
    r = max(0.1, 3.0 / n)
    return {'defn': mixed_mesh_def(n, seed=req.seed),
            'state': lambda cen: blast_state(cen, 1.4, r=r),
            'cfl': SERVE_CFL}


def device_bytes(solver):
    """Device bytes a set-up solver holds (its mesh's tables included)."""
    b = L.c_i64()
    if hasattr(solver, 'mesh'):
        L.check(L.lib().sceuler_device_bytes(solver._h, ctypes.byref(b)))
        m = L.c_i64()
        L.check(L.lib().scmesh_device_bytes(solver.mesh._h, ctypes.byref(m)))
        return b.value + m.value
    return 8 * 13 * solver.ncoord   # coord cfl gamma so0 so1 tso0 tso1


class Job(object):
    """One request: set-up of its plain solver, then either its own march
    (serially) or admission into an engine (served)."""

    def __init__(self, req, pl):
        self.req = req
        self.pl = pl
        self.solver = None
        self.result = None
        self.bytes = 0
        self.t_start = None
        self.t_ready = None
        self.t_end = None
        self.slot = None
        self.dt = None
        self.est = 0
        self.setup_s = 0.0

    def setup(self):
        req = self.req
        if req.kind == '1d':
            from .gen import sod_tube
            ncoord = self.pl['ncoord']
            # The Sod tube's dt/dx (CFL about 0.73).
            dt = 0.4 / 600 * 2000 / (ncoord - 1)
            self.solver = sod_tube(ncoord=ncoord, time_increment=dt)
            self.dt = dt
            L.check(L.lib().sc1d_sync(self.solver._h))
        else:
            from .multidim import EulerCore
            mh = build_static_mesh(self.pl['defn'])
            ngst = mh.ngstcell
            ec = EulerCore(mh, 1.0)
            ec.init_solution(1.4, 1.0, [0.0] * mh.ndim, 1.0)
            ec.so0n.ndarray[ngst:] = self.pl['state'](mh.clcnd.ndarray[ngst:])
            bf = mh.bndfcs.ndarray[:, 0].copy()
            if req.kind == '3d':
                c = mh.fccnd.ndarray[mh.ngstface + bf]
                wall = (np.abs(c) < 1e-12).any(axis=1)
                ec.add_slipwall(bf[wall])
                ec.add_nonrefl(bf[~wall])
            else:
                ec.add_nonrefl(bf)
            ec.bc_soln()
            ec.bc_dsoln()
            ec.set_time_increment(ec.cfl_dt(self.pl['cfl']))
            self.dt = ec.time_increment
            ec._before()
            L.check(L.lib().sceuler_sync(ec._h))
            self.solver = ec
        self.bytes = device_bytes(self.solver)

    def run_blocking(self):
        """Plain serial execution of this request: set-up, its own march,
        result on the host."""
        ts = time.perf_counter()
        self.setup()
        self.setup_s = time.perf_counter() - ts
        s = self.solver
        if self.req.kind == '1d':
            s.march_alpha2(self.req.steps)
            self.result = s.so0.copy()
        else:
            s.march(self.req.steps)
            self.result = s.so0n.ndarray[s.ngstcell:].copy()
        self.solver = None


class Arena(object):
    """A resident 2D or 3D serving arena (``scserve_*``): capacities in
    cells per cell type, ghost rows and solve slots, allocated once."""

    def __init__(self, ndim, type_cap, ghost_cap, max_slots=4096):
        L.require_gpu()
        caps = (ctypes.c_int64 * 9)()
        for t, c in type_cap.items():
            caps[int(t)] = int(c)
        h = ctypes.c_void_p()
        L.check(L.lib().scserve_create(int(ndim), caps, int(ghost_cap),
                                       int(max_slots), ctypes.byref(h)))
        self._h = h
        self.ndim = ndim
        self.neq = ndim + 2
        self.left = {}     # slot -> steps not yet enqueued
        self.ncell = {}

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and L._lib is not None:
            L._lib.scserve_destroy(h)
            self._h = None

    def info(self):
        out = (ctypes.c_int64 * 8)()
        L.check(L.lib().scserve_info(self._h, out))
        keys = ('used', 'marching', 'device_bytes', 'cell_cap', 'ghost_cap',
                'free_cells', 'free_ghosts', 'slot_cap')
        return dict(zip(keys, list(out)))

    def admit(self, core, steps):
        """Slot of the admitted solve, or None when it does not fit now.
        The EulerCore may be dropped afterwards."""
        core._before()
        slot = ctypes.c_int64()
        L.check(L.lib().scserve_admit(self._h, core._h, int(steps),
                                      ctypes.byref(slot)))
        if slot.value < 0:
            return None
        self.left[slot.value] = int(steps)
        self.ncell[slot.value] = core.ncell
        return slot.value

    def marching(self):
        return [s for s, k in self.left.items() if k > 0]

    def step(self, k):
        L.check(L.lib().scserve_step(self._h, int(k)))
        for s in self.left:
            self.left[s] = max(0, self.left[s] - k)

    def pending(self):
        n = ctypes.c_int64()
        L.check(L.lib().scserve_pending(self._h, ctypes.byref(n)))
        return n.value

    def done(self, slot):
        d = ctypes.c_int()
        L.check(L.lib().scserve_query(self._h, int(slot), ctypes.byref(d)))
        return d.value == 1

    def result(self, slot):
        n = self.ncell[slot]
        so0 = np.empty((n, self.neq))
        so1 = np.empty((n, self.neq, self.ndim))
        L.check(L.lib().scserve_result(self._h, int(slot), L.dptr(so0),
                                       L.dptr(so1)))
        del self.ncell[slot], self.left[slot]
        return so0, so1

    def sync(self):
        L.check(L.lib().scserve_sync(self._h))


class Tubes(object):
    """The 1D serving engine (``scserve1d_*``)."""

    def __init__(self, max_slots=4096):
        L.require_gpu()
        h = ctypes.c_void_p()
        L.check(L.lib().scserve1d_create(int(max_slots), ctypes.byref(h)))
        self._h = h
        self.left = {}
        self.n = {}

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and L._lib is not None:
            L._lib.scserve1d_destroy(h)
            self._h = None

    def admit(self, core, steps, alpha=2):
        """Slot of the admitted tube (the engine takes its device tables
        over), or None when every slot is in use."""
        for m in core._mirrors.values():
            m.before()
        L.check(L.lib().sc1d_sync(core._h))
        slot = ctypes.c_int64()
        L.check(L.lib().scserve1d_admit(self._h, core._h, int(alpha),
                                        int(steps), ctypes.byref(slot)))
        if slot.value < 0:
            return None
        core._h = None
        self.left[slot.value] = int(steps)
        self.n[slot.value] = core.ncoord
        return slot.value

    def marching(self):
        return [s for s, k in self.left.items() if k > 0]

    def step(self):
        L.check(L.lib().scserve1d_step(self._h))
        for s in self.left:
            seg = min(512, max(64, (self.n[s] + 127) // 128))
            kmax = min(64, (1024 - seg - 4) // 4)
            self.left[s] = max(0, self.left[s] - kmax)

    def pending(self):
        n = ctypes.c_int64()
        L.check(L.lib().scserve1d_pending(self._h, ctypes.byref(n)))
        return n.value

    def done(self, slot):
        d = ctypes.c_int()
        L.check(L.lib().scserve1d_query(self._h, int(slot), ctypes.byref(d)))
        return d.value == 1

    def result(self, slot):
        n = self.n[slot]
        so0 = np.empty((n, 3))
        so1 = np.empty((n, 3))
        cfl = np.empty(n)
        L.check(L.lib().scserve1d_result(self._h, int(slot), L.dptr(so0),
                                         L.dptr(so1), L.dptr(cfl)))
        del self.n[slot], self.left[slot]
        return so0, so1, cfl


# Arena capacities: cells per type.  The 3D split covers the mixed
# generator's shares with headroom: about 60% tetrahedra, 21% hexahedra,
# 17% prisms, and a pyramid layer whose share grows as the block shrinks
# (3% at 60^3, 10% at 20^3).
def arena_caps(dim, cells):
    if dim == 2:
        return {TRI: cells, QUAD: 0}
    return {TET: int(0.65 * cells), HEX: int(0.24 * cells),
            PRISM: int(0.20 * cells), PYRAMID: int(0.15 * cells)}


class Server(object):
    """The request loop over one GPU: a 1D tube engine and 2D / 3D arenas
    of fixed capacity, a set-up worker, and continuous admission."""

    def __init__(self, cells2=12_000_000, cells3=10_000_000,
                 ghosts2=2_000_000, ghosts3=3_000_000,
                 setup_budget=24 << 30, setup_workers=3, chunk_ms=8.0,
                 max_queued=2, poll_s=2e-4):
        self.engines = {'1d': Tubes(),
                        '2d': Arena(2, arena_caps(2, cells2), ghosts2),
                        '3d': Arena(3, arena_caps(3, cells3), ghosts3)}
        self.setup_budget = setup_budget
        self.setup_workers = setup_workers
        self.chunk_ms = chunk_ms
        self.max_queued = max_queued
        self.poll_s = poll_s
        # Set-up bytes per cell by kind, learned from the handles built.
        self.bpc = {'1d': 8 * 13.0, '2d': 2000.0, '3d': 5000.0}
        # Step time per cell by kind (ms), for the chunk size.
        self.ms_per_cell_step = {'2d': 3.2e-7, '3d': 1.2e-6}

    def device_bytes(self):
        return sum(e.info()['device_bytes'] for k, e in self.engines.items()
                   if k != '1d')

    def _chunk_steps(self, kind, eng):
        live = eng.marching()
        if not live:
            return 0
        least = min(eng.left[s] for s in live)
        cells = sum(eng.ncell[s] for s in live)
        est = max(1e-3, cells * self.ms_per_cell_step[kind])
        k = int(max(1, min(least, self.chunk_ms / est)))
        return k

    def run(self, requests, payloads):
        """Serve requests at their arrival times; returns the jobs (t_start
        set-up start, t_ready admitted, t_end result on host, relative to
        the first arrival) and the loop's wall time."""
        jobs = [Job(r, p) for r, p in zip(requests, payloads)]
        pending = collections.deque(sorted(jobs,
                                           key=lambda j: j.req.arrival))
        setup_q = collections.deque()
        ready = {k: collections.deque() for k in self.engines}
        owner = {}           # (kind, slot) -> job
        reap = collections.deque()
        lock = threading.Lock()
        cv = threading.Condition(lock)
        used = [0]
        stop = [False]
        errors = []
        t0 = time.perf_counter()

        def worker():
            while True:
                with cv:
                    while not setup_q and not reap and not stop[0]:
                        cv.wait()
                    if stop[0] and not setup_q and not reap:
                        return
                    if reap:
                        dead = reap.popleft()
                        job = None
                    else:
                        job = setup_q.popleft()
                        dead = None
                if dead is not None:
                    # Dropping an admitted solver waits for the copy out
                    # of it on the engine's stream: done here, off the loop.
                    del dead
                    continue
                try:
                    ts = time.perf_counter()
                    job.setup()
                    job.setup_s = time.perf_counter() - ts
                except Exception as e:  # reported by the loop
                    errors.append(e)
                    with cv:
                        stop[0] = True
                        cv.notify_all()
                    return
                with lock:
                    # The estimate reserved for this set-up becomes its
                    # actual bytes.
                    used[0] += job.bytes - job.est
                    n = max(1, job.req.size)
                    self.bpc[job.req.kind] = job.bytes / n
                    ready[job.req.kind].append(job)

        threads = [threading.Thread(target=worker, daemon=True)
                   for _ in range(self.setup_workers)]
        for t in threads:
            t.start()
        ndone = 0
        try:
            while ndone < len(jobs):
                if errors:
                    raise errors[0]
                now = time.perf_counter() - t0
                with cv:
                    # Set-up in arrival order while the set-up bytes fit;
                    # a request larger than the budget still runs, alone.
                    while pending and pending[0].req.arrival <= now:
                        job = pending[0]
                        job.est = int(self.bpc[job.req.kind]
                                      * max(1, job.req.size))
                        if used[0] + job.est > self.setup_budget and \
                                used[0] > 0:
                            break
                        pending.popleft()
                        used[0] += job.est
                        job.t_start = now
                        setup_q.append(job)
                        cv.notify()
                busy = False
                for kind, eng in self.engines.items():
                    # Admission in set-up order while the engine has room.
                    while True:
                        with lock:
                            job = ready[kind][0] if ready[kind] else None
                        if job is None:
                            break
                        if kind == '1d':
                            slot = eng.admit(job.solver, job.req.steps)
                        else:
                            slot = eng.admit(job.solver, job.req.steps)
                        if slot is None:
                            break
                        with cv:
                            ready[kind].popleft()
                            used[0] -= job.bytes
                            job.slot = slot
                            job.t_ready = time.perf_counter() - t0
                            owner[(kind, slot)] = job
                            reap.append(job.solver)
                            job.solver = None
                            cv.notify()
                    # Keep up to max_queued chunks of steps queued.
                    if eng.marching() and eng.pending() < self.max_queued:
                        if kind == '1d':
                            eng.step()
                        else:
                            eng.step(self._chunk_steps(kind, eng))
                        busy = True
                    # Collect finished solves.
                    for slot in [s for s, k in eng.left.items() if k == 0]:
                        if not eng.done(slot):
                            continue
                        job = owner.pop((kind, slot))
                        res = eng.result(slot)
                        job.result = res[0]
                        job.t_end = time.perf_counter() - t0
                        ndone += 1
                        busy = True
                if not busy:
                    time.sleep(self.poll_s)
        finally:
            with cv:
                stop[0] = True
                cv.notify_all()
            for t in threads:
                t.join()
        wall = time.perf_counter() - t0
        return jobs, wall


def serve(requests, payloads, **kw):
    """Serve the requests with a fresh Server (see Server.run)."""
    return Server(**kw).run(requests, payloads)


def serve_serial(requests, payloads):
    """The same requests one after another with blocking calls, each a plain
    solver marched alone; returns the jobs and their service times in
    seconds."""
    jobs = [Job(r, p) for r, p in zip(requests, payloads)]
    svc = []
    for job in jobs:
        t = time.perf_counter()
        job.run_blocking()
        svc.append(time.perf_counter() - t)
    return jobs, np.array(svc)


def fifo_latency(arrivals, service):
    """Latencies of a FIFO single server with the given arrival times and
    service times (both in seconds); also returns the makespan."""
    arrivals = np.asarray(arrivals, dtype=float)
    service = np.asarray(service, dtype=float)
    order = np.argsort(arrivals, kind='stable')
    free = 0.0
    lat = np.zeros(arrivals.size)
    for i in order:
        start = max(free, arrivals[i])
        free = start + service[i]
        lat[i] = free - arrivals[i]
    return lat, free - arrivals.min()
