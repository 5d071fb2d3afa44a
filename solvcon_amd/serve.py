"""Serving form: a request loop of 1D, 2D and 3D Euler solves on one GPU.

Requests arrive at given times, each carrying its problem: kind, mesh
size, step count and seed. The loop serves them as follows.
- It admits requests in arrival order while their estimated HBM footprint
  fits the budget.
- Set-up runs in a worker thread: the mesh is built on the device, then
  the solver, the initial state and the boundary conditions.
- Each admitted solve marches on its own HIP stream, so small solves
  overlap one another and the set-up of the next.
- Completed solves are detected by polling the streams, and their results
  are drained to host memory.
Latency is arrival to result on the host.

The baseline serves the same requests one after another with the plain
blocking calls. A FIFO single-server queue over the measured service
times gives its latencies at the same arrival times.

There is no reference counterpart; every solve is the reference's plain
march (Euler1DCore::march_alpha<2>, EulerCore::march).
"""

import collections
import ctypes
import threading
import time

import numpy as np

from . import _lib as L
from .gen import (tri_mesh_def, riemann_box, quadrant_state, mixed_mesh_def,
                  blast_state, build_static_mesh)

__all__ = ['Request', 'make_requests', 'Job', 'serve', 'serve_serial',
           'fifo_latency', 'footprint']

L.register({
    'sc1d_query': (L.c_int, [L.c_vp, ctypes.POINTER(L.c_int)]),
    'sceuler_query': (L.c_int, [L.c_vp, ctypes.POINTER(L.c_int)]),
})

# Device bytes per cell, for admission: solver state and geometry, mesh
# tables and the transient tables of the device mesh build.
BYTES_PER_CELL = {'1d': 200, '2d': 2500, '3d': 6000}

Request = collections.namedtuple(
    'Request', 'rid kind size steps seed arrival')


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


def footprint(req):
    return BYTES_PER_CELL[req.kind] * req.size


def payload(req):
    """Host-side problem of a request (made by the client before it is
    sent): mesh definition and initial-state function."""
    if req.kind == '1d':
        ncoord = max(5, req.size | 1)
        return {'ncoord': ncoord}
    if req.kind == '2d':
        m = max(2, int(round(np.sqrt(req.size / 2.0))))
        rho, p = riemann_box(req.seed % 100000)
        return {'defn': tri_mesh_def(m, seed=req.seed),
                'state': lambda cen: quadrant_state(cen, rho, p, 1.4)}
    n = max(3, int(round((req.size / 2.41) ** (1.0 / 3.0))))
    return {'defn': mixed_mesh_def(n, seed=req.seed),
            'state': lambda cen: blast_state(cen, 1.4)}


class Job(object):
    """One request: device set-up, asynchronous march, completion."""

    def __init__(self, req, pl):
        self.req = req
        self.pl = pl
        self.solver = None
        self.result = None
        self.t_start = None
        self.t_end = None

    def setup(self):
        req = self.req
        if req.kind == '1d':
            from .onedim import ShockTube
            ncoord = self.pl['ncoord']
            st = ShockTube()
            st.build_constant(gamma=1.4, pressure1=1.0, density1=1.0,
                              pressure5=0.1, density5=0.125)
            # The Sod tube's dt/dx (CFL about 0.73).
            dt = 0.4 / 600 * 2000 / (ncoord - 1)
            st.build_numerical(xmin=-1, xmax=1, ncoord=ncoord,
                               time_increment=dt)
            self.solver = st.svr._core
            return
        from .multidim import EulerCore
        defn = self.pl['defn']
        mh = build_static_mesh(defn)
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
        ec.set_time_increment(ec.cfl_dt(0.8))
        self.solver = ec

    def launch(self):
        s = self.solver
        if self.req.kind == '1d':
            for m in s._mirrors.values():
                m.before()
            L.check(L.lib().sc1d_march_async(s._h, 2, self.req.steps))
        else:
            s._before()
            L.check(L.lib().sceuler_march_async(s._h, self.req.steps))

    def done(self):
        d = L.c_int()
        fn = L.lib().sc1d_query if self.req.kind == '1d' else \
            L.lib().sceuler_query
        L.check(fn(self.solver._h, ctypes.byref(d)))
        return d.value == 1

    def finish(self):
        s = self.solver
        if self.req.kind == '1d':
            L.check(L.lib().sc1d_sync(s._h))
            for m in s._mirrors.values():
                m.after()
            self.result = s.so0.copy()
        else:
            L.check(L.lib().sceuler_sync(s._h))
            s._cur = (s._cur + 2 * self.req.steps) % 2
            s._after()
            self.result = s.so0n.ndarray[s.ngstcell:].copy()
        self.solver = None

    def run_blocking(self):
        """Plain serial execution of this request."""
        self.setup()
        self.launch()
        self.finish()


def serve(requests, payloads, budget_bytes=256 << 30, setup_workers=1,
          poll_s=2e-4):
    """Serve requests at their arrival times; returns the jobs with
    t_start (admission) and t_end (result on host) relative to the first
    arrival, and the loop's wall time."""
    jobs = [Job(r, p) for r, p in zip(requests, payloads)]
    pending = collections.deque(sorted(jobs, key=lambda j: j.req.arrival))
    setup_q = collections.deque()
    running = []
    lock = threading.Lock()
    cv = threading.Condition(lock)
    used = [0]
    stop = [False]
    errors = []
    t0 = time.perf_counter()

    def worker():
        while True:
            with cv:
                while not setup_q and not stop[0]:
                    cv.wait()
                if stop[0] and not setup_q:
                    return
                job = setup_q.popleft()
            try:
                job.setup()
                job.launch()
            except Exception as e:  # reported by the loop
                errors.append(e)
                with cv:
                    stop[0] = True
                    cv.notify_all()
                return
            with lock:
                running.append(job)

    threads = [threading.Thread(target=worker, daemon=True)
               for _ in range(setup_workers)]
    for t in threads:
        t.start()
    ndone = 0
    try:
        while ndone < len(jobs):
            if errors:
                raise errors[0]
            now = time.perf_counter() - t0
            with cv:
                # A request larger than the whole budget still runs, alone.
                while pending and pending[0].req.arrival <= now and \
                        (used[0] + footprint(pending[0].req) <= budget_bytes
                         or used[0] == 0):
                    job = pending.popleft()
                    used[0] += footprint(job.req)
                    job.t_start = now
                    setup_q.append(job)
                    cv.notify()
                snapshot = list(running)
            for job in snapshot:
                if job.done():
                    job.finish()
                    job.t_end = time.perf_counter() - t0
                    with lock:
                        running.remove(job)
                        used[0] -= footprint(job.req)
                    ndone += 1
            if not snapshot:
                wait = poll_s
                if pending and not setup_q:
                    wait = max(poll_s, min(0.01, pending[0].req.arrival
                                           - (time.perf_counter() - t0)))
                time.sleep(wait)
            else:
                time.sleep(poll_s)
    finally:
        with cv:
            stop[0] = True
            cv.notify_all()
        for t in threads:
            t.join()
    wall = time.perf_counter() - t0
    return jobs, wall


def serve_serial(requests, payloads):
    """The same requests one after another with blocking calls; returns the
    jobs and their service times in seconds."""
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
