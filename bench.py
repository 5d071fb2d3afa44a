"""Benchmark of the CESE Euler march on one MI355X.

Default workload: the 2D Euler CESE march on the jittered triangle mesh of
the unit square (1024 x 1024 quads split into 2,097,152 triangles, interior
nodes jittered by up to 0.2h, seed 2), a four-quadrant Riemann box with
rho and p drawn from U(0.1, 1.5), u = v = 0, gamma = 1.4, non-reflective
boundaries and dt set so the largest initial CFL number is 0.8.  A step is
one CESE step (two substeps) over the whole mesh; value is cell-updates/s
with the state resident in HBM.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K]
[--warmup W] [--size M] [--no-cpu]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ('cell-updates/s FP64 Euler CESE; batched solves/s + p99 latency '
          'vs plain')
PEAK_HBM_GBS = 8000.0
# Algorithmic bytes per cell per substep of a triangle at packed widths:
# state in 8q + 8qd + 8 (gamma), state out 8q + 8qd + 16 (cflo, cflc),
# CE geometry 8(f+1) + 8d(f+1) + 16 S d, neighbour list 4(f+1);
# q = 4, d = 2, f = 3, S = 6.
B_SUBSTEP_TRI = (8 * 4 + 8 * 8 + 8) + (8 * 4 + 8 * 8 + 16) \
    + (8 * 4 + 8 * 2 * 4 + 16 * 6 * 2) + 4 * 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup_2d(m, seed=2, cfl=0.8):
    import numpy as np
    import solvcon_amd as sc
    from solvcon_amd.gen import (tri_mesh_def, riemann_box, quadrant_state,
                                 build_static_mesh, cfl_dt)
    defn = tri_mesh_def(m, seed=seed)
    mh = build_static_mesh(defn)
    ncell = mh.ncell
    ngst = mh.ngstcell
    clcnd = mh.clcnd.ndarray[ngst:].copy()
    rho, p = riemann_box(seed)
    state = quadrant_state(clcnd, rho, p, 1.4)
    # dt from the CFL rule needs the CE geometry: a provisional core gives it.
    probe = sc.EulerCore(mesh=mh, time_increment=1.0)
    dt = cfl_dt(probe.cecnd.ndarray, ncell, ngst, 2, 3, state, 1.4, cfl)
    del probe
    ec = sc.EulerCore(mesh=mh, time_increment=dt)
    ec.init_solution(1.4, 1.0, [0.0, 0.0], 1.0)
    ec.so0n.ndarray[ngst:] = state
    ec.add_nonrefl(mh.bndfcs.ndarray[:, 0])
    ec.bc_soln()
    ec.bc_dsoln()
    return defn, mh, ec, state, dt


def traffic_from_profiles():
    """HBM bytes per fused launch from the committed rocprofv3 counter
    summary, when one exists."""
    path = os.path.join(ROOT, 'profiles', 'hbm_traffic_tri2d.md')
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith('traffic_bytes_per_launch:'):
            return float(line.split(':', 1)[1])
    return None


def cpu_baseline(defn, state, dt, steps, threads):
    """The oracle (the CPU restatement of the reference march) on the same
    mesh and state, timed over a few steps."""
    from oracle import oracle as O
    m = O.make_mesh(*defn)
    oe = O.Euler(m, dt, nthreads=threads)
    oe.init_solution(1.4, 1.0, [0.0, 0.0], 1.0)
    oe.so0n[oe.ngstcell:] = state
    oe.add_bc(O.BC_NONREFL, m.bndfcs[:, 0])
    oe.call('bc_soln')
    oe.call('bc_dsoln')
    oe.march(1)
    t0 = time.perf_counter()
    oe.march(steps)
    el = time.perf_counter() - t0
    return oe.ncell * steps / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--size', type=int, default=1024)
    ap.add_argument('--cpu-steps', type=int, default=60)
    ap.add_argument('--copy-probe', type=int, default=1024,
                    help='MiB moved by the stream-copy probe (0: skip)')
    ap.add_argument('--no-cpu', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build measures one GPU; --gpus %d is not '
              'supported (every GPU box here has one GPU)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import solvcon_amd as sc
    from solvcon_amd import _lib
    log('bench: building the %dx%dx2 triangle mesh on the GPU' %
        (args.size, args.size))
    t0 = time.perf_counter()
    defn, mh, ec, state, dt = setup_2d(args.size)
    log('bench: set-up %.1f s, %d cells, dt %.6g' %
        (time.perf_counter() - t0, ec.ncell, dt))
    ec.march(args.warmup)
    _lib.check(_lib.lib().sc_device_sync())
    t0 = time.perf_counter()
    ec.march(args.steps)
    _lib.check(_lib.lib().sc_device_sync())
    el = time.perf_counter() - t0
    ncell = ec.ncell
    value = ncell * args.steps / el
    log('bench: %d steps in %.3f s, %.4g cell-updates/s' %
        (args.steps, el, value))
    # Kernel time per launch from HIP events on the solver's stream.
    nprof = max(10, min(args.steps, 100))
    sub_ms, tot_ms = ec.march_timed(nprof)
    achieved = B_SUBSTEP_TRI * ncell / (sub_ms * 1e-3) / 1e9
    log('bench: fused substep kernel %.4f ms avg (%d launches), %.1f GB/s '
        'algorithmic' % (sub_ms, 2 * nprof, achieved))
    copy_gbs = None
    if args.copy_probe > 0:
        import ctypes
        cms = ctypes.c_double()
        nb = args.copy_probe * 1024 * 1024
        _lib.check(_lib.lib().sc_stream_copy(nb, 10, ctypes.byref(cms)))
        copy_gbs = 2 * nb / (cms.value * 1e-3) / 1e9
        log('bench: stream copy of %d MiB: %.4f ms, %.1f GB/s (read+write)'
            % (args.copy_probe, cms.value, copy_gbs))
    mass, cmin, cmax = ec.reduce()
    traffic = traffic_from_profiles()
    roof = {'bound': 'hbm', 'achieved': round(achieved, 1),
            'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
            'frac': round(achieved / PEAK_HBM_GBS, 4),
            'traffic': traffic}
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get('OMP_NUM_THREADS', '0') or 0) or \
            min(os.cpu_count() or 1, 16)
        log('bench: timing the oracle on %d threads, %d steps' %
            (threads, args.cpu_steps))
        cval, cel = cpu_baseline(defn, state, dt, args.cpu_steps, threads)
        cpu = {'value': round(cval, 1), 'unit': 'cell-updates/s',
               'cores': threads, 'kind': 'port',
               'sample': '%d steps of the same %d-cell mesh and state '
                         '(%.1f s), OpenMP over cells' %
                         (args.cpu_steps, ncell, cel)}
    out = {
        'metric': METRIC,
        'value': round(value, 1),
        'unit': 'cell-updates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(el / args.steps * 1e3, 5),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic',
        'config': {
            'workload': '2D Euler CESE march, jittered triangle mesh of the '
                        'unit square (%dx%d quads x 2 = %d cells), '
                        'Riemann box, non-reflective BCs, CFL 0.8'
                        % (args.size, args.size, ncell),
            'cells': ncell,
            'dt': dt,
            'fused_substep_ms': round(sub_ms, 5),
            'b_alg_per_cell_update': 2 * B_SUBSTEP_TRI,
            'mass': mass,
            'stream_copy_gbs': None if copy_gbs is None else round(copy_gbs, 1),
            'cfl_range': [cmin, cmax],
        },
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
