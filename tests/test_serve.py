"""Serving form: continuous batching of 1D, 2D and 3D solves.

CPU: request generation, the FIFO baseline model, arena capacities.
GPU:
- every request served in the mix equals the same request solved alone by
  the plain solver, bit for bit, and equals the oracle's march of it;
- a small arena admits requests as earlier ones finish and free their
  ranges, and the results are unchanged;
- the 1D engine marches more tubes than one launch holds, each tube equal
  to its plain march;
- the arena's guards: wrong dimension, a solve larger than the arena,
  results of a marching slot, slots out of range.
"""

import numpy as np
import pytest


def test_make_requests_mix_and_arrivals():
    from solvcon_amd.serve import make_requests
    reqs = make_requests(4000, rate=50.0, seed=5)
    kinds = np.array([r.kind for r in reqs])
    frac = [(kinds == k).mean() for k in ('1d', '2d', '3d')]
    assert abs(frac[0] - 0.5) < 0.03 and abs(frac[1] - 0.35) < 0.03
    sizes = np.array([r.size for r in reqs])
    assert sizes.min() >= 1000 and sizes.max() <= 2e6
    # log-uniform: the median sits near the geometric mean of the bounds.
    assert abs(np.log(np.median(sizes)) - 0.5 * np.log(1e3 * 2e6)) < 0.2
    arr = np.array([r.arrival for r in reqs])
    assert arr[0] == 0.0 and (np.diff(arr) >= 0).all()
    # Poisson process: mean gap 1/rate.
    assert abs(np.diff(arr).mean() * 50.0 - 1.0) < 0.05
    again = make_requests(4000, rate=50.0, seed=5)
    assert again == reqs


def test_fifo_latency_closed_forms():
    from solvcon_amd.serve import fifo_latency
    # Arrivals slower than service: every latency is its service time.
    lat, span = fifo_latency([0.0, 10.0, 20.0], [1.0, 2.0, 3.0])
    assert lat.tolist() == [1.0, 2.0, 3.0] and span == 23.0
    # All at once: latencies are the prefix sums of the service times.
    lat, span = fifo_latency([0.0, 0.0, 0.0], [1.0, 2.0, 3.0])
    assert lat.tolist() == [1.0, 3.0, 6.0] and span == 6.0
    # Order by arrival, not by index.
    lat, _ = fifo_latency([5.0, 0.0], [1.0, 10.0])
    assert lat.tolist() == [6.0, 10.0]


def test_arena_caps_cover_the_generator_mix():
    # The 3D split leaves room for the mixed generator's type shares.
    from solvcon_amd.serve import arena_caps, TET, HEX, PRISM, PYRAMID
    from solvcon_amd.gen import mixed_mesh_def
    _, _, cltpn, _ = mixed_mesh_def(20, seed=3)
    caps = arena_caps(3, 1000 * len(cltpn))
    for t in (TET, HEX, PRISM, PYRAMID):
        assert caps[t] >= 1000 * (cltpn == t).sum(), t
    assert arena_caps(2, 500)[4] == 500


# ----------------------------------------------------------------- GPU -----

def _oracle_result(orc, job):
    """The oracle's march of one served request: same mesh, state,
    boundaries, dt and steps."""
    req, pl = job.req, job.pl
    if req.kind == '1d':
        o = orc.sod_setup(ncoord=pl['ncoord'], time_increment=job.dt)
        o.march(2, req.steps)
        return o.so0
    m = orc.make_mesh(*pl['defn'])
    oe = orc.Euler(m, job.dt)
    oe.init_solution(1.4, 1.0, [0.0] * m.ndim, 1.0)
    oe.so0n[oe.ngstcell:] = pl['state'](m.clcnd[m.ngstcell:])
    bf = m.bndfcs[:, 0].copy()
    if req.kind == '3d':
        c = m.fccnd[m.ngstface + bf]
        wall = (np.abs(c) < 1e-12).any(axis=1)
        oe.add_bc(orc.BC_SLIPWALL, bf[wall])
        oe.add_bc(orc.BC_NONREFL, bf[~wall])
    else:
        oe.add_bc(orc.BC_NONREFL, bf)
    oe.call('bc_soln')
    oe.call('bc_dsoln')
    oe.march(req.steps)
    return oe.so0n[oe.ngstcell:]


@pytest.mark.gpu
def test_gpu_served_equals_alone_and_oracle(orc):
    print('case: served mix equals alone and the oracle', flush=True)
    from solvcon_amd.serve import make_requests, payload, serve, serve_serial
    reqs = make_requests(16, rate=200.0, lo=300, hi=5000, smin=3, smax=60,
                         seed=9)
    pls = [payload(r) for r in reqs]
    jobs, wall = serve(reqs, pls, cells2=200_000, cells3=200_000,
                       ghosts2=50_000, ghosts3=100_000)
    sj, svc = serve_serial(reqs, pls)
    kinds = set()
    for a, b in zip(jobs, sj):
        kinds.add(a.req.kind)
        assert a.t_end >= a.t_ready >= a.t_start >= a.req.arrival - 1e-9
        assert np.isfinite(a.result).all()
        assert a.dt == b.dt
        assert np.array_equal(a.result, b.result), a.req
        want = _oracle_result(orc, a)
        assert np.array_equal(a.result, want), ('oracle', a.req)
    assert kinds == {'1d', '2d', '3d'}


@pytest.mark.gpu
def test_gpu_arena_admits_as_ranges_free(orc):
    # An arena of 5120 triangles holds two of these meshes at a time:
    # later requests join as earlier ones finish, at the step boundaries of
    # the others, and every result still equals its plain march.
    print('case: continuous admission into a small arena', flush=True)
    from solvcon_amd.serve import Request, payload, serve, serve_serial
    reqs = [Request(i, '2d', 2 * 40 * 40, 20 + 7 * i, 100 + i, 0.0)
            for i in range(6)]
    pls = [payload(r) for r in reqs]
    jobs, _ = serve(reqs, pls, cells2=5120, cells3=256, ghosts2=4096,
                    ghosts3=256)
    sj, _ = serve_serial(reqs, pls)
    ends = sorted(j.t_end for j in jobs)
    late = [j for j in jobs if j.t_ready > ends[0]]
    assert len(late) >= 3
    for a, b in zip(jobs, sj):
        assert np.array_equal(a.result, b.result), a.req
    assert np.array_equal(jobs[5].result, _oracle_result(orc, jobs[5]))


@pytest.mark.gpu
def test_gpu_tubes_many_per_launch():
    # 21 tubes (more than the 16 one launch holds), sizes 5 to 70001 and
    # 1 to 200 steps, alpha 1 and 2: each equals its plain march.
    print('case: 1D engine, 21 tubes', flush=True)
    from solvcon_amd.gen import sod_tube
    from solvcon_amd.serve import Tubes
    rng = np.random.default_rng(3)
    sizes = [5, 7, 129, 2001, 70001] + [int(x) | 1 for x in
                                        rng.integers(11, 20000, 16)]
    steps = [1, 200, 64, 65, 130] + [int(x) for x in rng.integers(1, 150, 16)]
    alphas = [2, 1, 2, 1, 2] + [int(x) for x in rng.integers(1, 3, 16)]
    eng = Tubes(max_slots=32)
    slots = []
    for n, k, a in zip(sizes, steps, alphas):
        dt = 0.4 / 600 * 2000 / (n - 1)
        slots.append(eng.admit(sod_tube(ncoord=n, time_increment=dt), k,
                               alpha=a))
    while eng.marching():
        eng.step()
    for n, k, a, s in zip(sizes, steps, alphas, slots):
        while not eng.done(s):
            pass
        so0, so1, cfl = eng.result(s)
        dt = 0.4 / 600 * 2000 / (n - 1)
        ref = sod_tube(ncoord=n, time_increment=dt)
        getattr(ref, 'march_alpha%d' % a)(k)
        assert np.array_equal(so0, ref.so0), (n, k, a)
        assert np.array_equal(so1, ref.so1), (n, k, a)
        assert np.array_equal(cfl, ref.cfl), (n, k, a)


@pytest.mark.gpu
def test_gpu_arena_guards():
    print('case: arena guards', flush=True)
    from solvcon_amd.serve import Arena, Request, Job, payload, TRI
    a2 = Arena(2, {TRI: 1000}, 1000, max_slots=2)
    j3 = Job(Request(0, '3d', 200, 3, 1, 0.0), None)
    j3.pl = payload(j3.req)
    j3.setup()
    with pytest.raises(ValueError):
        a2.admit(j3.solver, 3)
    big = Job(Request(1, '2d', 2 * 30 * 30, 3, 1, 0.0), None)
    big.pl = payload(big.req)
    big.setup()
    assert a2.admit(big.solver, 3) is None       # 1800 cells > 1024
    small = Job(Request(2, '2d', 2 * 10 * 10, 3, 2, 0.0), None)
    small.pl = payload(small.req)
    small.setup()
    s = a2.admit(small.solver, 3)
    assert s is not None
    assert a2.info()['used'] == 1
    with pytest.raises(ValueError):
        a2.admit(small.solver, 0)
    with pytest.raises(RuntimeError):
        a2.result(s)                            # still marching
    with pytest.raises(IndexError):
        a2.done(7)
    a2.step(3)
    a2.sync()
    assert a2.done(s)
    so0, so1 = a2.result(s)
    assert so0.shape == (200, 4) and so1.shape == (200, 4, 2)
    assert a2.info()['used'] == 0 and a2.info()['free_cells'] == 1024
