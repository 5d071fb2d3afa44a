"""Write the golden vectors of tests/golden from the oracle.

The reference library cannot be built or imported in this environment (it
needs C++23 <format>/<mdspan>, absent from GCC 11's libstdc++), so the
vectors come from the oracle, the CPU restatement under oracle/, whose
functions are pinned by the reference's own goldens and closed forms in
tests/test_oracle_*.py.  Run from the repository root:

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def sod():
    from oracle import oracle as O
    o = O.sod_setup(ncoord=2001, time_increment=0.4 / 600)
    o.march(2, 600)
    np.savez_compressed(os.path.join(HERE, 'sod_2001_600.npz'),
                        so0=o.so0, so1=o.so1, cfl=o.cfl)


if __name__ == '__main__':
    sod()
