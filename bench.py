"""Benchmark of the CESE Euler march on one MI355X.

Default workload (``--config 3d``): the 3D Euler CESE march on the mixed
hex/tet/prism/pyramid block of the unit cube.
- Mesh: 200^3 hexes, 30% in tetrahedral layers and 20% in prismatic ones,
  with a pyramid transition layer: 19,280,000 cells.
- Initial state: a spherical blast (rho = 1, p = 10 inside r = 0.1 of the
  centre, p = 0.1 outside), u = 0, gamma = 1.4.
- Slip walls on x = 0, y = 0 and z = 0, non-reflective boundaries
  elsewhere; dt set so that the largest initial CFL number is 0.8.
A step is one CESE step (two substeps) over the whole mesh.  ``value`` is
cell-updates/s with the state resident in HBM.  The same JSON line carries
the batched form (``batched``: 4096 independent 2D solves in one batched
march against the same solves marched one by one with plain solvers) unless
``--no-batched``.

Other modes:
- ``--config 2d``: the 2D march on the jittered triangle mesh of the unit
  square (1024 x 1024 quads split into 2,097,152 triangles, Riemann box,
  non-reflective boundaries, 500 steps).
- ``--config 1d``: the Sod tube, 2001 points, 600 steps of march_alpha2.
- ``--batched``: the batched form alone, as its own JSON line.
- ``--served``: a Poisson request stream of 1D/2D/3D solves through the
  request loop (continuous batching), with p50/p99 latency, against serial
  plain execution.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K]
[--warmup W] [--config 3d|2d|1d] [--batched] [--served] [--no-cpu]
[--no-batched]
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

# Cell type ids (solvcon's StaticMesh) -> (ndim, faces, sub-faces).
CELL_SHAPE = {3: (2, 4, 8), 4: (2, 3, 6), 5: (3, 6, 24), 6: (3, 4, 12),
              7: (3, 5, 18), 8: (3, 5, 16)}
CELL_NAME = {3: 'quad', 4: 'triangle', 5: 'hexahedron', 6: 'tetrahedron',
             7: 'prism', 8: 'pyramid'}


def b_substep(ctype):
    """Algorithmic bytes per cell and substep at packed widths:
    - state in: so0c 8q, so1c 8qd, gamma 8;
    - state out: so0n 8q, so1n 8qd, cflo and cflc 16;
    - CE geometry: volumes 8(f+1), BCE centroids 8d(f+1), sub-face
      centroids and normals 16 S d;
    - neighbour list 4(f+1).
    Here q = d + 2."""
    d, f, s = CELL_SHAPE[ctype]
    q = d + 2
    state = (8 * q + 8 * q * d + 8) + (8 * q + 8 * q * d + 16)
    geom = 8 * (f + 1) + 8 * d * (f + 1) + 16 * s * d + 4 * (f + 1)
    return state + geom


B_SUBSTEP_TRI = b_substep(4)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(name):
    """HBM bytes per fused launch from a committed rocprofv3 counter
    summary, when one exists."""
    path = os.path.join(ROOT, 'profiles', name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith('traffic_bytes_per_launch:'):
            return float(line.split(':', 1)[1])
    return None


def workload_2d(m, seed=2):
    from solvcon_amd.gen import tri_mesh_def, riemann_box, quadrant_state
    defn = tri_mesh_def(m, seed=seed)
    rho, p = riemann_box(seed)

    def state(cen):
        return quadrant_state(cen, rho, p, 1.4)

    def bcs(mh):
        return [('nonrefl', mh.bndfcs.ndarray[:, 0].copy())]
    desc = ('2D Euler CESE march, jittered triangle mesh of the unit square '
            '(%dx%d quads x 2 = %d cells), Riemann box, non-reflective BCs, '
            'CFL 0.8' % (m, m, 2 * m * m))
    return defn, state, bcs, desc


def workload_3d(n, seed=3):
    import numpy as np
    from solvcon_amd.gen import mixed_mesh_def, blast_state
    defn = mixed_mesh_def(n, seed=seed)

    def state(cen):
        return blast_state(cen, 1.4)

    def bcs(mh):
        bf = mh.bndfcs.ndarray[:, 0].copy()
        c = mh.fccnd.ndarray[mh.ngstface + bf]
        wall = (np.abs(c) < 1e-12).any(axis=1)
        return [('slipwall', bf[wall]), ('nonrefl', bf[~wall])]
    desc = ('3D Euler CESE march, mixed hex/tet/prism/pyramid block of the '
            'unit cube (%d^3 hexes, 30%% tetrahedral and 20%% prismatic '
            'layers, %d cells), spherical blast, slip walls on x=0/y=0/z=0, '
            'non-reflective elsewhere, CFL 0.8' % (n, len(defn[2])))
    return defn, state, bcs, desc


def setup_mesh_core(defn, state, bcs, cfl=0.8):
    """Device mesh, EulerCore, initial state, BCs and the CFL-rule dt."""
    import solvcon_amd as sc
    from solvcon_amd.gen import build_static_mesh
    mh = build_static_mesh(defn)
    ngst = mh.ngstcell
    ec = sc.EulerCore(mesh=mh, time_increment=1.0)
    ec.init_solution(1.4, 1.0, [0.0] * mh.ndim, 1.0)
    st = state(mh.clcnd.ndarray[ngst:])
    ec.so0n.ndarray[ngst:] = st
    for kind, faces in bcs(mh):
        getattr(ec, 'add_' + kind)(faces)
    ec.bc_soln()
    ec.bc_dsoln()
    dt = ec.cfl_dt(cfl)
    ec.set_time_increment(dt)
    return mh, ec, st, dt


class _View(object):
    def __init__(self, a):
        self.ndarray = a


def host_threads():
    """Host threads this process may use: OMP_NUM_THREADS when set (the
    GPU box sets it to the CPU share of its one GPU), else the CPUs of the
    process's affinity mask."""
    env = int(os.environ.get('OMP_NUM_THREADS', '0') or 0)
    if env > 0:
        return env
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline_mesh(defn, state, bcs_of, dt, plan):
    """The oracle (the CPU restatement of the reference march) on the given
    mesh, state and boundaries: for each (threads, steps) of ``plan`` one
    untimed step, then the steps timed.  Returns [(cell-updates/s,
    seconds)], and the cell count."""
    import numpy as np
    from oracle import oracle as O
    m = O.make_mesh(*defn)
    oe = O.Euler(m, dt, nthreads=plan[0][0])
    oe.init_solution(1.4, 1.0, [0.0] * m.ndim, 1.0)
    ngst = oe.ngstcell
    oe.so0n[ngst:] = state(m.clcnd[ngst:])
    mv = type('MeshView', (), {})()
    mv.bndfcs = _View(m.bndfcs)
    mv.fccnd = _View(m.fccnd)
    mv.ngstface = m.ngstface
    kinds = {'nonrefl': O.BC_NONREFL, 'slipwall': O.BC_SLIPWALL}
    for kind, faces in bcs_of(mv):
        oe.add_bc(kinds[kind], np.asarray(faces))
    oe.call('bc_soln')
    oe.call('bc_dsoln')
    out = []
    for threads, steps in plan:
        oe.set_threads(threads)
        oe.march(1)
        t0 = time.perf_counter()
        oe.march(steps)
        el = time.perf_counter() - t0
        out.append((oe.ncell * steps / el, el))
    return out, oe.ncell


def group_mix(ec):
    """Body cells per cell type of an EulerCore, from its mesh."""
    import numpy as np
    cl = ec.mesh.cltpn.ndarray[ec.mesh.ngstcell:]
    t, c = np.unique(cl, return_counts=True)
    return {int(a): int(b) for a, b in zip(t, c)}


def run_mesh(args, which):
    from solvcon_amd import _lib
    if which == '2d':
        defn, state, bcs, desc = workload_2d(args.size)
        traffic_name = 'hbm_traffic_tri2d.md'
    else:
        defn, state, bcs, desc = workload_3d(args.size3)
        traffic_name = 'hbm_traffic_mixed3d.md'
    log('bench: %s mesh of %d cells generated' % (which, len(defn[2])))
    t0 = time.perf_counter()
    mh, ec, st, dt = setup_mesh_core(defn, state, bcs)
    _lib.check(_lib.lib().sc_device_sync())
    log('bench: device set-up %.1f s, %d cells, dt %.6g' %
        (time.perf_counter() - t0, ec.ncell, dt))
    ec.march(args.warmup)
    _lib.check(_lib.lib().sc_device_sync())
    # The timed region: wall clock around the march, and HIP events on the
    # solver's stream around every substep of the same march for the
    # kernel time per substep.
    t0 = time.perf_counter()
    sub_ms, tot_ms = ec.march_timed(args.steps)
    _lib.check(_lib.lib().sc_device_sync())
    el = time.perf_counter() - t0
    ncell = ec.ncell
    value = ncell * args.steps / el
    log('bench: %d steps in %.3f s, %.4g cell-updates/s' %
        (args.steps, el, value))
    mix = group_mix(ec)
    b_sub = sum(b_substep(t) * n for t, n in mix.items())  # bytes/substep
    achieved = b_sub / (sub_ms * 1e-3) / 1e9
    log('bench: substep kernels %.4f ms avg over the %d substeps of the '
        'timed region, %.1f GB/s algorithmic; whole march by events %.4f ms '
        'per step' % (sub_ms, 2 * args.steps, achieved, tot_ms / args.steps))
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
    roof = {'bound': 'hbm', 'achieved': round(achieved, 1),
            'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
            'frac': round(achieved / PEAK_HBM_GBS, 4),
            'traffic': traffic_from_profiles(traffic_name)}
    cpu = None
    if not args.no_cpu:
        nt = host_threads()
        if which == '2d':
            cdefn = defn
            plan = [(nt, args.cpu_steps), (1, max(1, args.cpu_steps // 12))]
            what = 'the same %d-cell mesh and state'
        else:
            # The full 3D mesh takes minutes to build on the CPU: the
            # sample is the same generator, state and boundaries at a
            # smaller size.
            cdefn = workload_3d(args.cpu_size3)[0]
            plan = [(nt, args.cpu_steps3), (1, args.cpu_steps3_1t)]
            what = 'the same generator, state and boundaries at %d cells'
        log('bench: timing the oracle on %d threads and on 1' % nt)
        res, ccells = cpu_baseline_mesh(cdefn, state, bcs, dt, plan)
        (vn, en), (v1, e1) = res
        cpu = {'value': round(vn, 1), 'unit': 'cell-updates/s',
               'cores': nt, 'kind': 'port',
               'sample': ('%d steps of ' + what + ' (%.1f s), OpenMP over '
                          'cells on %d threads') % (plan[0][1], ccells, en,
                                                    nt),
               'cpu_model': cpu_model(),
               'single_thread': {
                   'value': round(v1, 1), 'cores': 1,
                   'sample': ('%d steps of ' + what + ' (%.1f s)')
                   % (plan[1][1], ccells, e1)}}
        log('bench: oracle %.4g cell-updates/s on %d threads, %.4g on 1 '
            '(%s)' % (vn, nt, v1, cpu['cpu_model']))
    batched = None
    if which == '3d' and not args.no_batched:
        batched = run_batched(args, line=False)
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
            'workload': desc,
            'cells': ncell,
            'cell_mix': {CELL_NAME[t]: n for t, n in mix.items()},
            'dt': dt,
            'fused_substep_ms': round(sub_ms, 5),
            'b_alg_per_cell_update': round(2 * b_sub / ncell, 1),
            'mass': mass,
            'stream_copy_gbs': None if copy_gbs is None else round(copy_gbs, 1),
            'cfl_range': [cmin, cmax],
            'kernel_forms': {CELL_NAME[t]: f for t, f in
                             (ec.kernel_forms or {}).items()},
            'exact_passes': ec.exact_passes,
        },
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    if batched is not None:
        out['batched'] = batched
    print(json.dumps(out), flush=True)


def run_1d(args):
    """Sod tube, 2001 points, 600 steps of march_alpha2 per run
    (cell-updates per run = 600 x (2001 - 4))."""
    from solvcon_amd.gen import sod_tube
    from solvcon_amd import _lib
    ncoord, steps = 2001, 600
    base = sod_tube(ncoord=ncoord, time_increment=0.4 / steps)
    runs = []
    for _ in range(args.warmup + args.steps):
        s = base.clone()
        _lib.check(_lib.lib().sc_device_sync())
        t0 = time.perf_counter()
        s.march_alpha2(steps)
        runs.append(time.perf_counter() - t0)
    runs = sorted(runs[args.warmup:])
    med = runs[len(runs) // 2]
    upd = steps * (ncoord - 4)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        ts = []
        for _ in range(5):
            o = O.sod_setup(ncoord=ncoord, time_increment=0.4 / steps)
            t0 = time.perf_counter()
            o.march(2, steps)
            ts.append(time.perf_counter() - t0)
        ts.sort()
        cpu = {'value': round(upd / ts[2], 1), 'unit': 'cell-updates/s',
               'cores': 1, 'kind': 'port',
               'sample': 'the whole Sod run (2001 points, 600 steps), '
                         'median of 5'}
    out = {
        'metric': 'cell-updates/s FP64 1D Euler CESE (Sod tube)',
        'value': round(upd / med, 1), 'unit': 'cell-updates/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup,
        'ms_per_run': round(med * 1e3, 4), 'higher_is_better': True,
        'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': 'Sod shock tube, 2001 points, 600 steps of '
                               'march_alpha2 per run; one run per timed step',
                   'run_ms_min': round(runs[0] * 1e3, 4),
                   'run_ms_max': round(runs[-1] * 1e3, 4)},
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


def run_batched(args, line=True):
    """The batched form: ``--batch-n`` independent 2D triangle solves.
    - Mesh sizes are log-uniform in [1e3, 2e5] cells.
    - Step counts are uniform in [100, 1000].
    - Each instance has its own seeded Riemann box and its own dt, from its
      initial CFL number ``--batch-cfl``.
    All instances run in one batched march.  The plain form is measured on
    a stratified sample: the instances are ordered by work (cells x steps)
    and cut into ``--plain-sample`` strata of equal size, and the middle
    instance of each stratum is set up and marched alone by a plain
    EulerCore (kernel form fixed, so no autotuning is timed).  The serial
    plain time of the whole batch is the stratified estimate: each
    stratum's size times its sample's time.  Instances whose state is not
    finite at the end are counted and left out of solves/s.
    Returns the result dict; with ``line`` also prints it as a JSON line."""
    import numpy as np
    from solvcon_amd import _lib
    from solvcon_amd.batch import (batch_config, config_defs, config_states,
                                   BatchedEuler, build_mesh)
    from solvcon_amd import EulerCore
    m, mseed, rseed, steps = batch_config(ninst=args.batch_n)
    t0 = time.perf_counter()
    defs = config_defs(m, mseed)
    states = config_states(rseed)
    t1 = time.perf_counter()
    b = BatchedEuler(defs, states, steps, cfl=args.batch_cfl)
    _lib.check(_lib.lib().sc_device_sync())
    t2 = time.perf_counter()
    log('bench: batch of %d instances, %d cells, %.3g cell-updates; '
        'generate %.1f s, device set-up %.1f s' %
        (b.ninst, b.coff[-1], b.cell_updates, t1 - t0, t2 - t1))
    t3 = time.perf_counter()
    dev_ms = b.march()
    t4 = time.perf_counter()
    done = b.done_ms()
    wall = t4 - t3
    so0 = b.core.so0n.ndarray[b.ngst:]
    fin = np.array([np.isfinite(so0[b.coff[k]:b.coff[k + 1]]).all()
                    for k in range(b.ninst)])
    del so0
    nbad = int((~fin).sum())
    log('bench: batched march %.3f s wall (%.1f ms device), %.4g '
        'cell-updates/s, %.1f solves/s, %d of %d instances not finite' %
        (wall, dev_ms, b.cell_updates / wall, fin.sum() / wall, nbad,
         b.ninst))
    # The plain form on a stratified sample.
    work = b.ncells * steps
    order = np.argsort(work, kind='stable')
    strata = np.array_split(order, args.plain_sample)
    prev_form = os.environ.get('SC_EULER_FORM')
    os.environ['SC_EULER_FORM'] = args.plain_form
    rows = []
    try:
        for st in strata:
            k = int(st[len(st) // 2])
            ta = time.perf_counter()
            mh = build_mesh(defs[k])
            ec = EulerCore(mh, float(b.dts[k]))
            ec.init_solution(1.4, 1.0, [0.0, 0.0], 1.0)
            cen = mh.clcnd.ndarray[mh.ngstcell:]
            ec.so0n.ndarray[mh.ngstcell:] = states(cen, k)
            ec.add_nonrefl(mh.bndfcs.ndarray[:, 0])
            ec.bc_soln()
            ec.bc_dsoln()
            _lib.check(_lib.lib().sc_device_sync())
            tb = time.perf_counter()
            ec.march(int(steps[k]))
            tc = time.perf_counter()
            # Bit for bit; a non-finite instance carries its NaNs to the
            # same places alone and in the batch.
            same = np.array_equal(ec.so0n.ndarray[mh.ngstcell:],
                                  b.body('so0n', k),
                                  equal_nan=not bool(fin[k]))
            rows.append((len(st), b.ncells[k], steps[k], tb - ta, tc - tb,
                         same, fin[k]))
    finally:
        if prev_form is None:
            os.environ.pop('SC_EULER_FORM', None)
        else:
            os.environ['SC_EULER_FORM'] = prev_form
    tm = np.array(rows, dtype=float)
    plain_march = float((tm[:, 0] * tm[:, 4]).sum())
    plain_setup = float((tm[:, 0] * tm[:, 3]).sum())
    # A two-term model of the sample (per-step cost plus per-cell-update
    # cost), to report how well such a fit would have done.
    A = np.stack([tm[:, 2], tm[:, 1] * tm[:, 2]], axis=1)
    coef, *_ = np.linalg.lstsq(A, tm[:, 4], rcond=None)
    resid = np.abs(A @ coef - tm[:, 4]) / tm[:, 4]
    fit_total = float(coef[0] * steps.sum() + coef[1] * work.sum())
    log('bench: plain sample of %d (one per stratum of %d): %.2f s '
        'measured; whole batch serially %.1f s march + %.1f s set-up '
        '(stratified); a two-term fit gives %.1f s, residual median %.1f%% '
        'max %.1f%%; sample equal to the batch: %s' %
        (len(rows), len(strata[0]), tm[:, 4].sum(), plain_march,
         plain_setup, fit_total, 100 * np.median(resid), 100 * resid.max(),
         bool(tm[:, 5].all())))
    # Latency of each solve, all submitted at the march's start (set-up
    # excluded on both sides): batched, the device time until the
    # instance's last step ran; serially, the fitted march times of the
    # instances before it and its own, in instance order.
    lat = done[fin] / 1e3
    slat = np.cumsum(coef[0] * steps + coef[1] * work)[fin]
    pct = {'p50_s': float(np.percentile(lat, 50)),
           'p99_s': float(np.percentile(lat, 99)),
           'serial_p50_s_est': float(np.percentile(slat, 50)),
           'serial_p99_s_est': float(np.percentile(slat, 99))}
    log('bench: solve latency p50/p99 %.2f/%.2f s batched, %.2f/%.2f s '
        'serially (est.)' % (pct['p50_s'], pct['p99_s'],
                             pct['serial_p50_s_est'],
                             pct['serial_p99_s_est']))
    res = {
        'value': round(float(fin.sum()) / wall, 2),
        'unit': 'solves/s',
        'workload': '%d independent 2D jittered-triangle Euler solves, '
                    'sizes log-uniform 1e3-2e5 cells, 100-1000 steps, '
                    'Riemann boxes, CFL %g' % (b.ninst, args.batch_cfl),
        'cells': int(b.coff[-1]),
        'cell_updates': b.cell_updates,
        'batched_wall_s': round(wall, 4),
        'batched_device_ms': round(dev_ms, 2),
        'batched_cell_updates_per_s': round(b.cell_updates / wall, 1),
        'batched_setup_s': round(t2 - t1, 2),
        'not_finite': nbad,
        'plain_sample': int(len(rows)),
        'plain_sample_march_s': round(float(tm[:, 4].sum()), 3),
        'plain_serial_march_s': round(plain_march, 2),
        'plain_serial_setup_s': round(plain_setup, 2),
        'plain_form': args.plain_form,
        'plain_fit_residual_median': round(float(np.median(resid)), 4),
        'sample_equal_bitwise': bool(tm[:, 5].all()),
        'sample_not_finite': int((tm[:, 6] == 0).sum()),
        'latency_s': {k: round(v, 3) for k, v in pct.items()},
        'latency_note': 'set-up excluded on both sides; serial latencies '
                        'from the two-term fit of the plain sample',
        'p99_s': round(pct['p99_s'], 3),
        'vs_plain': round(plain_march / wall, 3),
        'vs_plain_end_to_end': round((plain_march + plain_setup)
                                     / (wall + t2 - t1), 3),
    }
    if line:
        out = {
            'metric': 'batched solves/s (FP64 Euler CESE, 2D triangles)',
            'value': res['value'], 'unit': 'solves/s', 'n_gpus': 1,
            'higher_is_better': True, 'dtype': 'f64', 'data': 'synthetic',
            'config': {k: v for k, v in res.items()
                       if k not in ('value', 'unit', 'p99_s', 'vs_plain',
                                    'vs_plain_end_to_end')},
            'p99_s': res['p99_s'], 'vs_plain': res['vs_plain'],
            'vs_plain_end_to_end': res['vs_plain_end_to_end'],
        }
        print(json.dumps(out), flush=True)
    return res


def run_served(args):
    """The serving form: ``--serve-n`` requests of 1D/2D/3D solves in a
    50/35/15 mix, served by the request loop's continuous batching (a
    resident 2D and 3D arena and a 1D tube engine).
    - Sizes are log-uniform in 1e3-2e6 cells; steps are uniform in
      100-1000.
    - Arrivals follow a Poisson process at several loads, given as
      multiples of the serial plain capacity.
    Each load is served by the request loop, and the same stream is
    modelled as a FIFO queue over the measured serial plain service
    times."""
    import numpy as np
    from solvcon_amd.serve import (make_requests, payload, serve,
                                   serve_serial, fifo_latency)
    base = make_requests(args.serve_n, rate=1.0, seed=5)
    t0 = time.perf_counter()
    payloads = [payload(r) for r in base]
    log('bench: %d request payloads generated in %.1f s (client side)' %
        (len(base), time.perf_counter() - t0))
    sj, svc = serve_serial(base, payloads)
    mu = 1.0 / svc.mean()
    log('bench: serial set-up %.4f s mean of the %.4f s mean service' %
        (np.mean([j.setup_s for j in sj]), svc.mean()))
    log('bench: serial plain service %.4f s mean, capacity %.2f solves/s' %
        (svc.mean(), mu))
    loads = []
    equal = True
    for frac in args.serve_loads:
        reqs = make_requests(args.serve_n, rate=frac * mu, seed=5)
        jobs, wall = serve(reqs, payloads,
                           setup_budget=int(args.serve_budget_gb * 2 ** 30),
                           setup_workers=args.serve_setup_workers)
        lat = np.array([j.t_end - j.req.arrival for j in jobs])
        span = max(j.t_end for j in jobs)
        arr = np.array([r.arrival for r in reqs])
        slat, sspan = fifo_latency(arr, svc)
        nonfin = 0
        for a, b in zip(jobs, sj):
            fin = bool(np.isfinite(b.result).all())
            nonfin += not fin
            equal = equal and np.array_equal(a.result, b.result,
                                             equal_nan=not fin)
        row = {'load': frac, 'rate': round(frac * mu, 3),
               'solves_s': round(len(jobs) / span, 3),
               'p50_s': round(float(np.percentile(lat, 50)), 4),
               'p99_s': round(float(np.percentile(lat, 99)), 4),
               'serial_solves_s': round(len(jobs) / sspan, 3),
               'serial_p50_s': round(float(np.percentile(slat, 50)), 4),
               'serial_p99_s': round(float(np.percentile(slat, 99)), 4),
               'not_finite': nonfin,
               'setup_s_mean': round(float(np.mean([j.setup_s
                                                   for j in jobs])), 4),
               'admit_wait_s_mean': round(float(np.mean(
                   [j.t_ready - j.t_start - j.setup_s for j in jobs])), 4)}
        loads.append(row)
        log('bench: load %.2f: %s' % (frac, row))
    top = loads[-1]
    out = {
        'metric': 'served solves/s and p99 latency (FP64 Euler CESE, '
                  '1D/2D/3D mix)',
        'value': top['solves_s'], 'unit': 'solves/s', 'n_gpus': 1,
        'higher_is_better': True, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': '%d requests, 50/35/15 1D/2D/3D, sizes '
                               'log-uniform 1e3-2e6 cells, 100-1000 steps, '
                               'Poisson arrivals' % args.serve_n,
                   'serial_service_mean_s': round(float(svc.mean()), 4),
                   'serial_capacity_solves_s': round(mu, 3),
                   'budget_gb': args.serve_budget_gb,
                   'results_equal_serial_bitwise': bool(equal),
                   'loads': loads},
        'p99_s': top['p99_s'],
        'vs_plain': round(top['solves_s'] / top['serial_solves_s'], 3),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None)
    ap.add_argument('--warmup', type=int, default=None)
    ap.add_argument('--config', choices=('2d', '3d', '1d'), default='3d')
    ap.add_argument('--size', type=int, default=1024)
    ap.add_argument('--form', default=None,
                    help='fix the multidim kernel forms instead of the '
                         'autotuner (SC_EULER_FORM syntax, e.g. '
                         '"6:f2,5:s1s/1")')
    ap.add_argument('--size3', type=int, default=200)
    ap.add_argument('--cpu-steps', type=int, default=60)
    ap.add_argument('--cpu-size3', type=int, default=60)
    ap.add_argument('--cpu-steps3', type=int, default=20)
    ap.add_argument('--cpu-steps3-1t', type=int, default=3)
    ap.add_argument('--no-batched', action='store_true',
                    help='leave the batched form out of the 3D line')
    ap.add_argument('--batch-cfl', type=float, default=0.8)
    ap.add_argument('--plain-form', default='4:f1',
                    help='kernel form of the plain sample (no autotuning): '
                    'the form the autotuner keeps for the 2D march')
    ap.add_argument('--copy-probe', type=int, default=1024,
                    help='MiB moved by the stream-copy probe (0: skip)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--graph', action='store_true',
                    help='replay a captured hipGraph per step instead of '
                         'launching its kernels one by one')
    ap.add_argument('--batched', action='store_true',
                    help='measure the batched form instead')
    ap.add_argument('--batch-n', type=int, default=4096)
    ap.add_argument('--plain-sample', type=int, default=410)
    ap.add_argument('--served', action='store_true',
                    help='measure the serving form instead')
    ap.add_argument('--serve-n', type=int, default=60)
    ap.add_argument('--serve-loads', type=float, nargs='+',
                    default=[0.5, 1.0, 2.0, 4.0])
    ap.add_argument('--serve-setup-workers', type=int, default=3)
    ap.add_argument('--serve-budget-gb', type=float, default=24.0,
                    help='device bytes of solves being set up at once')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build measures one GPU; --gpus %d is not '
              'supported (no multi-GPU path is built: the sharded design is '
              'described, not implemented)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.form:
        os.environ['SC_EULER_FORM'] = args.form
    if args.graph:
        os.environ['SC_EULER_GRAPH'] = '1'
    defaults = {'2d': (500, 20), '3d': (20, 3), '1d': (21, 3)}[args.config]
    if args.steps is None:
        args.steps = defaults[0]
    if args.warmup is None:
        args.warmup = defaults[1]
    if args.batched:
        run_batched(args)
    elif args.served:
        run_served(args)
    elif args.config == '1d':
        run_1d(args)
    else:
        run_mesh(args, args.config)


if __name__ == '__main__':
    main()
