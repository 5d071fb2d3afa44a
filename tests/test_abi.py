"""The C ABI: libsolvcon_amd.so loads and exports every function that
include/solvcon_amd.h declares (no device calls; runs without a GPU)."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, 'include', 'solvcon_amd.h')).read()
    txt = re.sub(r'/\*.*?\*/', ' ', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(sc\w+)\s*\(', txt)))


def test_header_declares_entry_points():
    names = _declared()
    for pre in ('sc_', 'sc1d_', 'scmesh_', 'sceuler_', 'scge_', 'scst_',
                'scserve_', 'scserve1d_'):
        assert any(n.startswith(pre) for n in names), pre


def test_library_exports_every_declared_symbol():
    from solvcon_amd import _lib
    L = ctypes.CDLL(_lib.LIBPATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_device_count_without_gpu_is_safe():
    # sc_device_count never fails; the solvers refuse to start without one.
    from solvcon_amd import _lib
    assert _lib.lib().sc_device_count() >= 0


def test_bench_refuses_more_than_one_gpu():
    # bench.py measures one GPU; asked for more it says so and exits 2
    # before it touches a device.
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'),
                        '--gpus', '2'], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr + r.stdout
    assert r.stdout.strip() == ''
