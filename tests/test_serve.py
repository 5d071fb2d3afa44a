"""Serving form: request generation, the FIFO baseline model, and (GPU)
every request served in the mix equals the same request solved alone.
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


def test_footprint_scales_with_size():
    from solvcon_amd.serve import Request, footprint
    a = Request(0, '2d', 1000, 10, 1, 0.0)
    b = Request(1, '2d', 2000, 10, 1, 0.0)
    assert footprint(b) == 2 * footprint(a)


# ----------------------------------------------------------------- GPU -----

@pytest.mark.gpu
def test_gpu_served_equals_alone():
    print('case: served mix equals alone', flush=True)
    from solvcon_amd.serve import (make_requests, payload, serve,
                                   serve_serial)
    reqs = make_requests(12, rate=200.0, lo=300, hi=5000, smin=3, smax=12,
                         seed=9)
    pls = [payload(r) for r in reqs]
    jobs, wall = serve(reqs, pls, budget_bytes=64 << 20)
    sj, svc = serve_serial(reqs, pls)
    for a, b in zip(jobs, sj):
        assert a.t_end >= a.t_start >= a.req.arrival - 1e-9
        assert np.array_equal(a.result, b.result), a.req
        assert np.isfinite(a.result).all()
