"""MI355X-native CESE Euler march with solvcon's solver API.

The classes stand in for the compiled ``_solvcon`` module that solvcon's
Python layer loads (``onedim.Euler1DCore``, ``StaticMesh``, ``EulerCore``,
``GradientElement``, ``Gmsh`` and the spacetime solvers): every march, mesh
build and reduction runs as a hand-written HIP kernel of libsolvcon_amd.so
on the GPU.  The batched and served forms (``batch``, ``serve``) add many
independent solves per launch.
"""

from . import onedim  # noqa: F401
from . import spacetime  # noqa: F401
from ._lib import lib as _load_lib  # noqa: F401
from .mesh import StaticMesh, StaticMeshBc  # noqa: F401
from .gmsh import Gmsh  # noqa: F401
from .multidim import EulerCore, GradientElement  # noqa: F401

__all__ = ['onedim', 'spacetime', 'StaticMesh', 'StaticMeshBc', 'Gmsh',
           'EulerCore', 'GradientElement']
