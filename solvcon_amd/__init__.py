"""MI355X-native CESE Euler march with solvcon's solver API.

The solver classes keep the Python surface of solvcon (Euler1DSolver,
ShockTube, StaticMesh, EulerCore, GradientElement, EulerField and the
spacetime scalar solvers) while every march, mesh build and reduction runs as a hand-written HIP kernel of
libsolvcon_amd.so on the GPU.
"""

from . import onedim  # noqa: F401
from . import obsrefl  # noqa: F401
from . import spacetime  # noqa: F401
from ._lib import lib as _load_lib  # noqa: F401
from .mesh import StaticMesh, StaticMeshBc  # noqa: F401
from .gmsh import Gmsh  # noqa: F401
from .multidim import (EulerCore, GradientElement, EulerField,  # noqa: F401
                       RunHistory)

__all__ = ['onedim', 'spacetime', 'StaticMesh', 'StaticMeshBc', 'Gmsh',
           'EulerCore', 'GradientElement', 'EulerField', 'RunHistory']
