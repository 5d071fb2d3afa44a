"""Oblique-shock reflection driven end to end on the device march.

The problem of solvcon's obsrefl application (solvcon/pilot/apps/obsrefl:
_driver.py, _analytic.py, _session.py), without its Qt panel:

- A uniform supersonic stream enters a rectangle from the left (inlet).
- The top boundary holds the state behind an incident oblique shock
  (inlet).
- The bottom slip wall reflects the shock, and the flow leaves through a
  non-reflective outflow on the right.

Oblique-shock relations give the three zone states. A ``Reflection``
measures a run against them: zone averages, the fitted incident-shock angle,
the reflection point and horizontal profiles. A ``ReflectionSession`` marches
a run chunk by chunk and records mass and CFL range.  Every march is
``EulerCore.march`` on the GPU with inlet, slip-wall and non-reflective
boundaries; the analysis is host numpy over the solver's body-cell views.
"""

import collections
import dataclasses
import math

import numpy as np

from .mesh import StaticMesh
from .multidim import EulerCore, EulerField, RunHistory

__all__ = ['ObliqueShockMesher', 'ObliqueShockRelation', 'ObliqueShock',
           'Segment', 'SegmentPad', 'ZoneInfo', 'AngleFit', 'WallPoint',
           'HorizontalProfile', 'Reflection', 'ReflectionSession']


# ----------------------------------------------------------------- mesh ------

def _circumcircle(x, y, a, b, c):
    """Circumcentre and squared radius of the triangle (a, b, c) of the
    point arrays x, y (the textbook determinant formula)."""
    ax, ay, bx, by, cx, cy = x[a], y[a], x[b], y[b], x[c], y[c]
    d = 2.0 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
    a2, b2, c2 = ax * ax + ay * ay, bx * bx + by * by, cx * cx + cy * cy
    ux = (a2 * (by - cy) + b2 * (cy - ay) + c2 * (ay - by)) / d
    uy = (a2 * (cx - bx) + b2 * (ax - cx) + c2 * (bx - ax)) / d
    return ux, uy, (ax - ux) ** 2 + (ay - uy) ** 2


def delaunay(points):
    """Bowyer-Watson Delaunay triangulation of 2D points; returns
    counter-clockwise index triples.

    The published incremental method: a super-triangle far outside the
    points, then each point in turn carves out the triangles whose
    circumcircle strictly contains it and fans the cavity rim to it; the
    triangles touching a super-triangle vertex are dropped at the end.  A
    point on a circumcircle (within a relative 1e-12) counts as outside, so
    the cocircular points of a regular boundary give a well-defined cavity
    (the choices of solvcon/pilot/apps/obsrefl/_driver.py:168-235, so both
    triangulate the same point cloud alike).

    A. Bowyer, The Computer Journal 24(2):162-166, 1981;
    D. F. Watson, The Computer Journal 24(2):167-172, 1981.
    """
    pts = np.asarray(points, dtype=np.float64)
    npt = pts.shape[0]
    xmid = 0.5 * (pts[:, 0].min() + pts[:, 0].max())
    ymid = 0.5 * (pts[:, 1].min() + pts[:, 1].max())
    span = max(float(np.ptp(pts[:, 0])), float(np.ptp(pts[:, 1])), 1.0)
    x = np.concatenate([pts[:, 0], [xmid - 64.0 * span, xmid + 64.0 * span,
                                    xmid]])
    y = np.concatenate([pts[:, 1], [ymid - 32.0 * span, ymid - 32.0 * span,
                                    ymid + 64.0 * span]])
    tri = [(npt, npt + 1, npt + 2)]
    circ = [_circumcircle(x, y, npt, npt + 1, npt + 2)]
    alive = [True]
    cx = np.array([circ[0][0]])
    cy = np.array([circ[0][1]])
    cr = np.array([circ[0][2]])
    live = np.array([True])
    for ip in range(npt):
        px, py = x[ip], y[ip]
        bad = np.nonzero(live & ((px - cx) ** 2 + (py - cy) ** 2
                                 < cr * (1.0 - 1e-12)))[0]
        edges = set()
        for it in bad:
            a, b, c = tri[it]
            edges.update(((a, b), (b, c), (c, a)))
            alive[it] = False
        live[bad] = False
        new = [(ea, eb, ip) for ea, eb in sorted(edges)
               if (eb, ea) not in edges]
        ncirc = [_circumcircle(x, y, *t) for t in new]
        tri.extend(new)
        alive.extend([True] * len(new))
        cx = np.concatenate([cx, [c[0] for c in ncirc]])
        cy = np.concatenate([cy, [c[1] for c in ncirc]])
        cr = np.concatenate([cr, [c[2] for c in ncirc]])
        live = np.concatenate([live, np.ones(len(new), dtype=bool)])
    out = []
    for t, ok in zip(tri, alive):
        if not ok or max(t) >= npt:
            continue
        a, b, c = t
        if (x[b] - x[a]) * (y[c] - y[a]) - (x[c] - x[a]) * (y[b] - y[a]) < 0.0:
            t = (a, c, b)
        out.append(t)
    return out


class ObliqueShockMesher(object):
    """The rectangle from ``ll`` to ``ur`` in ``nx`` by ``ny`` grid boxes
    (ObliqueShockMesher, solvcon/pilot/apps/obsrefl/_driver.py:21-319).

    ``make_mesh`` builds one of three flavors, all with the same boundary
    nodes (nx segments on the bottom and top, ny on the left and right):
    ``'quad'`` one quadrilateral per box; ``'triangle'`` two triangles per
    box cut along the lower-left-to-upper-right diagonal, the other diagonal
    at the upper-left and lower-right corner boxes so no triangle has two
    boundary faces; ``'unstructured'`` (the default) a Delaunay
    triangulation of the boundary nodes and deterministically jittered
    interior points, with Steiner points until no cell has two boundary
    faces.  The mesh is built on the device with one named boundary group
    per domain edge.
    """

    BOUNDARY_NAMES = ('left', 'top', 'bottom', 'right')

    def __init__(self, nx=64, ny=16, ll=(0.0, 0.0), ur=(4.0, 1.0)):
        self.nx = nx
        self.ny = ny
        (self.x0, self.y0), (self.x1, self.y1) = ll, ur

    @property
    def cell_extent(self):
        """(width, height) of one grid box."""
        return ((self.x1 - self.x0) / self.nx, (self.y1 - self.y0) / self.ny)

    def node(self, it, jt):
        """Coordinates of the logical grid point (it, jt) (fractional
        indices allowed)."""
        return (self.x0 + it * (self.x1 - self.x0) / self.nx,
                self.y0 + jt * (self.y1 - self.y0) / self.ny)

    def _nid(self, it, jt):
        return jt * (self.nx + 1) + it

    def _box(self, it, jt):
        return (self._nid(it, jt), self._nid(it + 1, jt),
                self._nid(it + 1, jt + 1), self._nid(it, jt + 1))

    def _grid_nodes(self):
        return [self.node(it, jt) for jt in range(self.ny + 1)
                for it in range(self.nx + 1)]

    def point_cloud(self):
        """Boundary nodes in a counter-clockwise walk from the lower-left
        corner, then the interior nodes displaced in logical space by a
        fixed phase of at most 0.3 boxes per axis (no random generator, so
        the mesh is reproducible)."""
        nx, ny = self.nx, self.ny
        pts = [self.node(it, 0) for it in range(nx + 1)]
        pts += [self.node(nx, jt) for jt in range(1, ny + 1)]
        pts += [self.node(it, ny) for it in range(nx - 1, -1, -1)]
        pts += [self.node(0, jt) for jt in range(ny - 1, 0, -1)]
        for jt in range(1, ny):
            for it in range(1, nx):
                pts.append(self.node(
                    it + 0.3 * math.sin(12.9898 * it + 78.233 * jt),
                    jt + 0.3 * math.cos(26.651 * it + 41.347 * jt)))
        return pts

    @staticmethod
    def _boundary_edge_count(tris):
        count = collections.Counter()
        for t in tris:
            for k in range(3):
                count[frozenset((t[k], t[(k + 1) % 3]))] += 1
        return count

    def _unstructured(self):
        pts = self.point_cloud()
        for _ in range(16):
            tris = delaunay(pts)
            shared = self._boundary_edge_count(tris)
            ears = [t for t in tris
                    if sum(shared[frozenset((t[k], t[(k + 1) % 3]))] == 1
                           for k in range(3)) >= 2]
            if not ears:
                return pts, tris
            for t in ears:
                pts.append((sum(pts[i][0] for i in t) / 3.0,
                            sum(pts[i][1] for i in t) / 3.0))
        raise RuntimeError('boundary-cell refinement did not converge')

    def definition(self, cell_type='unstructured'):
        """(ndim, ndcrd, cltpn, clnds) of the flavor ``cell_type``, cells
        counter-clockwise."""
        nx, ny = self.nx, self.ny
        if cell_type == 'quad':
            nodes = self._grid_nodes()
            tpn = StaticMesh.QUADRILATERAL
            cells = [(4,) + self._box(it, jt)
                     for jt in range(ny) for it in range(nx)]
        elif cell_type == 'triangle':
            nodes = self._grid_nodes()
            tpn = StaticMesh.TRIANGLE
            cells = []
            for jt in range(ny):
                for it in range(nx):
                    ll, lr, ur, ul = self._box(it, jt)
                    if (it, jt) in ((0, ny - 1), (nx - 1, 0)):
                        cells += [(3, ll, lr, ul), (3, lr, ur, ul)]
                    else:
                        cells += [(3, ll, lr, ur), (3, ll, ur, ul)]
        elif cell_type == 'unstructured':
            tpn = StaticMesh.TRIANGLE
            nodes, tris = self._unstructured()
            cells = [(3,) + t for t in tris]
        else:
            raise ValueError("unknown cell_type '%s'" % cell_type)
        return (2, np.asarray(nodes, dtype=np.float64),
                np.full(len(cells), tpn, dtype=np.int32),
                np.asarray(cells, dtype=np.int32))

    def make_mesh(self, cell_type='unstructured'):
        """The device StaticMesh of the flavor ``cell_type``, with interior
        (metrics and edges), the four named boundary groups, boundary and
        ghosts built."""
        _, ndcrd, cltpn, clnds = self.definition(cell_type)
        mh = StaticMesh(ndim=2, nnode=ndcrd.shape[0], nface=0,
                        ncell=cltpn.shape[0])
        mh.ndcrd.ndarray[...] = ndcrd
        mh.cltpn.ndarray[...] = cltpn
        mh.clnds.ndarray[:, :clnds.shape[1]] = clnds
        mh.build_interior(do_metric=True)
        for name, faces in zip(self.BOUNDARY_NAMES,
                               self.classify_boundary(mh)):
            mh.add_bc(name, faces)
        mh.build_boundary()
        mh.build_ghost()
        return mh

    @staticmethod
    def classify_boundary(mh, tol=1e-9):
        """Boundary faces of ``mh`` (no neighbour cell) sorted into the
        (left, top, bottom, right) domain edges by face centre and outward
        normal; each list ascending.  The edge positions come from the face
        centres, since ghost nodes lie outside the domain.  ``mh`` is a
        StaticMesh or any object with the body-face tables as arrays."""
        def body(name):
            a = getattr(mh, name)
            a = getattr(a, 'ndarray', a)
            return np.asarray(a)[mh.ngstface:]
        fccls, fccnd, fcnml = body('fccls'), body('fccnd'), body('fcnml')
        bf = np.nonzero(fccls[:, 1] < 0)[0]
        xc, yc = fccnd[bf, 0], fccnd[bf, 1]
        nxv, nyv = fcnml[bf, 0], fcnml[bf, 1]
        xmin, xmax, ymax = xc.min(), xc.max(), yc.max()
        left = (np.abs(xc - xmin) <= tol) & (nxv < 0.0)
        right = ~left & (np.abs(xc - xmax) <= tol) & (nxv > 0.0)
        top = ~left & ~right & (np.abs(yc - ymax) <= tol) & (nyv > 0.0)
        bottom = ~left & ~right & ~top
        return tuple(sorted(bf[m].tolist()) for m in (left, top, bottom,
                                                      right))


# ------------------------------------------------------------- relations -----

class ObliqueShockRelation(object):
    """Jumps across an oblique shock of angle ``beta`` with deflection
    ``theta`` (radians, from the upstream flow), after chapter 4 of
    J. D. Anderson, Modern Compressible Flow, 3rd ed. (the relations of
    solvcon/pilot/apps/obsrefl/_driver.py:238-396)."""

    def __init__(self, gamma):
        self.gamma = gamma

    def _mn1sq(self, mach1, beta):
        return (mach1 * math.sin(beta)) ** 2

    def calc_density_ratio(self, mach1, beta):
        g = self.gamma
        m = self._mn1sq(mach1, beta)
        return (g + 1) * m / ((g - 1) * m + 2)

    def calc_pressure_ratio(self, mach1, beta):
        g = self.gamma
        return 1 + 2 * g / (g + 1) * (self._mn1sq(mach1, beta) - 1)

    def calc_temperature_ratio(self, mach1, beta):
        return (self.calc_pressure_ratio(mach1, beta)
                / self.calc_density_ratio(mach1, beta))

    def calc_normal_dmach(self, mach_n1):
        """Downstream Mach number normal to the shock."""
        g = self.gamma
        m = mach_n1 * mach_n1
        return math.sqrt(((g - 1) * m + 2) / (2 * g * m - (g - 1)))

    def calc_dmach(self, mach1, beta=None, theta=None, delta=1):
        """Downstream Mach number from exactly one of the shock angle and
        the deflection angle (``delta`` 1 weak, 0 strong branch)."""
        if (beta is None) == (theta is None):
            raise ValueError('got (beta=%s, theta=%s), but I need to take '
                             'either beta or theta' % (beta, theta))
        if beta is None:
            beta = self.calc_shock_angle(mach1, theta, delta=delta)
        if theta is None:
            theta = self.calc_flow_angle(mach1, beta)
        return (self.calc_normal_dmach(mach1 * math.sin(beta))
                / math.sin(beta - theta))

    def calc_flow_tangent(self, mach1, beta):
        """tan(theta) of the theta-beta-M relation."""
        g = self.gamma
        m1sq = mach1 * mach1
        return (2 / math.tan(beta) * (m1sq * math.sin(beta) ** 2 - 1)
                / (m1sq * (g + math.cos(2 * beta)) + 2))

    def calc_flow_angle(self, mach1, beta):
        return math.atan(self.calc_flow_tangent(mach1, beta))

    def calc_shock_tangent_aux(self, mach1, theta):
        """(lambda, chi) of the closed-form inversion of theta-beta-M;
        ValueError when no attached shock exists."""
        g = self.gamma
        m1sq = mach1 * mach1
        t2 = math.tan(theta) ** 2
        a = 1 + (g - 1) / 2 * m1sq
        disc = (m1sq - 1) ** 2 - 3 * a * (1 + (g + 1) / 2 * m1sq) * t2
        if disc <= 0.0:
            raise ValueError('no attached shock for mach1=%g and theta=%g deg'
                             % (mach1, math.degrees(theta)))
        lmbd = math.sqrt(disc)
        chi = ((m1sq - 1) ** 3
               - 9 * a * (a + (g + 1) / 4 * m1sq * m1sq) * t2) / lmbd ** 3
        return lmbd, chi

    def calc_shock_tangent(self, mach1, theta, delta):
        g = self.gamma
        m1sq = mach1 * mach1
        lmbd, chi = self.calc_shock_tangent_aux(mach1, theta)
        num = (m1sq - 1
               + 2 * lmbd * math.cos((4 * math.pi * delta + math.acos(chi))
                                     / 3))
        return num / (3 * (1 + (g - 1) / 2 * m1sq) * math.tan(theta))

    def calc_shock_angle(self, mach1, theta, delta=1):
        return math.atan(self.calc_shock_tangent(mach1, theta, delta))


# ---------------------------------------------------------------- driver -----

class ObliqueShock(object):
    """The reflection problem and its solver (ObliqueShock,
    solvcon/pilot/apps/obsrefl/_driver.py:399-553)."""

    def __init__(self):
        self.gamma = self.density = self.pressure = self.mach = None
        self.speedofsound = self.velocity = self.relation = None
        self.theta = self.shock_angle = None
        self.density2 = self.pressure2 = self.mach2 = None
        self.velocity2 = self.shock_angle2 = None
        self.density3 = self.pressure3 = self.mach3 = self.velocity3 = None
        self.mesher = self.mesh = self.svr = None

    def build_constant(self, gamma=1.4, density=1.0, pressure=1.0, mach=3.0,
                       angle=10.0):
        """Zone states: the free stream (1) entering horizontally at
        ``mach``; behind the incident shock of deflection ``angle`` degrees
        (2, imposed at the top boundary, velocity ``angle`` below
        horizontal); behind the reflected shock (3, horizontal again)."""
        rel = ObliqueShockRelation(gamma=gamma)
        self.gamma, self.density, self.pressure, self.mach = (
            gamma, density, pressure, mach)
        self.speedofsound = math.sqrt(gamma * pressure / density)
        self.velocity = mach * self.speedofsound
        self.relation = rel
        self.theta = theta = math.radians(angle)
        self.shock_angle = beta = rel.calc_shock_angle(mach, theta)
        self.density2 = density * rel.calc_density_ratio(mach, beta)
        self.pressure2 = pressure * rel.calc_pressure_ratio(mach, beta)
        self.mach2 = rel.calc_dmach(mach, beta=beta)
        c2 = math.sqrt(gamma * self.pressure2 / self.density2)
        self.velocity2 = (self.mach2 * c2 * math.cos(theta),
                          -self.mach2 * c2 * math.sin(theta))
        self.shock_angle2 = beta2 = rel.calc_shock_angle(self.mach2, theta)
        self.density3 = self.density2 * rel.calc_density_ratio(self.mach2,
                                                               beta2)
        self.pressure3 = self.pressure2 * rel.calc_pressure_ratio(self.mach2,
                                                                  beta2)
        self.mach3 = rel.calc_dmach(self.mach2, beta=beta2)
        c3 = math.sqrt(gamma * self.pressure3 / self.density3)
        self.velocity3 = (self.mach3 * c3, 0.0)

    def build_numerical(self, cell_type='unstructured', time_increment=2.e-3,
                        sigma0=3.0, taumin=0.0, tauscale=1.0, **mesher_kw):
        """The mesh and the device EulerCore: free stream everywhere, inlet
        on the left (zone 1) and the top (zone 2), slip wall at the bottom,
        non-reflective outflow on the right, ghosts primed."""
        if self.gamma is None:
            raise ValueError('constants are not set; call build_constant()')
        self.mesher = ObliqueShockMesher(**mesher_kw)
        self.mesh = self.mesher.make_mesh(cell_type=cell_type)
        svr = EulerCore(mesh=self.mesh, time_increment=time_increment)
        svr.sigma0 = sigma0
        svr.taumin = taumin
        svr.tauscale = tauscale
        svr.init_solution(self.gamma, self.density, [self.velocity, 0.0],
                          self.pressure)
        left, top, bottom, right = (
            self.mesh.bc(name).facn.ndarray[:, 0].tolist()
            for name in self.mesher.BOUNDARY_NAMES)
        svr.add_inlet(left, value=[self.density, self.velocity, 0.0,
                                   self.pressure, self.gamma])
        svr.add_inlet(top, value=[self.density2, self.velocity2[0],
                                  self.velocity2[1], self.pressure2,
                                  self.gamma])
        svr.add_slipwall(bottom)
        svr.add_nonrefl(right)
        svr.bc_soln()
        svr.bc_dsoln()
        self.svr = svr

    def march(self, steps):
        self.svr.march(steps)

    def zone_states(self):
        """(rho, vx, vy, p) of zones 1, 2 and 3."""
        return [(self.density, self.velocity, 0.0, self.pressure),
                (self.density2, self.velocity2[0], self.velocity2[1],
                 self.pressure2),
                (self.density3, self.velocity3[0], self.velocity3[1],
                 self.pressure3)]

    def shock_path(self):
        """Corners of the analytic shock path: the upper-left corner, the
        reflection point on the wall and where the reflected shock leaves;
        two corners when the incident shock leaves through the outflow."""
        if self.mesher is None:
            raise ValueError('mesh is not built; call build_numerical()')
        m = self.mesher
        xhit = m.x0 + (m.y1 - m.y0) / math.tan(self.shock_angle)
        if xhit >= m.x1:
            return [(m.x0, m.y1),
                    (m.x1, m.y1 - (m.x1 - m.x0) * math.tan(self.shock_angle))]
        slope = math.tan(self.shock_angle2 - self.theta)
        xend = min(m.x1, xhit + (m.y1 - m.y0) / slope)
        return [(m.x0, m.y1), (xhit, m.y0),
                (xend, m.y0 + (xend - xhit) * slope)]


# -------------------------------------------------------------- analysis -----

Segment = collections.namedtuple('Segment', 'x0 y0 x1 y1')


class SegmentPad(object):
    """Plane segments (the analytic shock arms) with a packed [n, 4] view."""

    def __init__(self):
        self._segs = []

    def append(self, x0, y0, x1, y1):
        self._segs.append(Segment(float(x0), float(y0), float(x1), float(y1)))

    def __len__(self):
        return len(self._segs)

    def __getitem__(self, i):
        return self._segs[i]

    def pack_array(self):
        class _Packed(object):
            pass
        out = _Packed()
        out.ndarray = np.array(self._segs, dtype=np.float64).reshape(-1, 4)
        return out


@dataclasses.dataclass
class ZoneInfo(object):
    """One zone's mean of a computed field against its analytic value."""

    ZONES = (None, 1, 2, 3)
    _zone: int
    count: int
    computed: float
    analytic: float
    error: float

    def __post_init__(self):
        self.zone = self._zone

    @property
    def zone(self):
        return self._zone

    @zone.setter
    def zone(self, value):
        if value not in self.ZONES or isinstance(value, str):
            raise ValueError('zone must be one of %s, not %r'
                             % (self.ZONES, value))
        self._zone = value


@dataclasses.dataclass
class AngleFit(object):
    degree: float
    analytic: float
    error: float
    npoint: int


@dataclasses.dataclass
class WallPoint(object):
    x: float
    analytic: float
    error: float


@dataclasses.dataclass
class HorizontalProfile(object):
    x: np.ndarray
    computed: np.ndarray
    analytic: np.ndarray


class Reflection(object):
    """A run measured against the analytic solution (Reflection,
    solvcon/pilot/apps/obsrefl/_analytic.py:127-436).  No measurement raises
    on a field whose shocks have not formed; it reports nan instead.
    ``margin`` is the clearance kept from the analytic shocks when averaging
    a zone, as a fraction of the domain height."""

    MARGIN_FRACTION = 0.1

    def __init__(self, shock, margin=MARGIN_FRACTION):
        if shock.svr is None:
            raise ValueError('solver is not built; call build_numerical()')
        self.shock = shock
        self.field = EulerField(shock.svr, shock.mesh)
        self.margin = margin
        self.arms = SegmentPad()
        path = shock.shock_path()
        for (xa, ya), (xb, yb) in zip(path[:-1], path[1:]):
            self.arms.append(xa, ya, xb, yb)

    @property
    def has_reflection(self):
        return len(self.arms) == 2

    @property
    def incident(self):
        return self.arms[0]

    @property
    def reflected(self):
        return self.arms[1] if self.has_reflection else None

    @staticmethod
    def _relative(value, target):
        return (value - target) / target if target else float('nan')

    @staticmethod
    def _slope(seg):
        return (seg.y1 - seg.y0) / (seg.x1 - seg.x0)

    @classmethod
    def _offset(cls, seg, xs, ys):
        """Signed distance from the line of ``seg``, positive above."""
        k = cls._slope(seg)
        return (ys - seg.y0 - k * (xs - seg.x0)) / math.hypot(1.0, k)

    @classmethod
    def _abscissa(cls, seg, height):
        return seg.x0 + (height - seg.y0) / cls._slope(seg)

    def zone_masks(self, margin=None):
        """Body-cell masks of zones 1, 2 and 3: a cell's centroid on the
        zone's side of both shocks, at least ``margin`` away from them."""
        margin = self.margin if margin is None else margin
        m = self.shock.mesher
        gap = margin * (m.y1 - m.y0)
        cen = self.field.centroid
        xs, ys = cen[:, 0], cen[:, 1]
        inc = self._offset(self.incident, xs, ys)
        ref = (self._offset(self.reflected, xs, ys) if self.has_reflection
               else np.full(xs.shape, np.inf))
        return [inc < -gap, (inc > gap) & (ref > gap), ref < -gap]

    def zone_fields(self):
        """Analytic value of zones 1, 2, 3 by field name (every EulerField
        field but the CFL number)."""
        g = self.shock.gamma
        rho, vx, vy, p = np.array(self.shock.zone_states(),
                                  dtype=np.float64).T
        v2 = vx * vx + vy * vy
        return {'density': rho, 'velx': vx, 'vely': vy, 'speed': np.sqrt(v2),
                'pressure': p, 'mach': np.sqrt(v2 / (g * p / rho)),
                'total_energy': p / (g - 1.0) + 0.5 * rho * v2}

    def zone_field(self, name):
        fields = self.zone_fields()
        if name not in fields:
            raise ValueError("unknown field '%s'" % name)
        return fields[name]

    def color_range(self, name, vmin, vmax):
        """(vmin, vmax) widened to the analytic values of the field."""
        fields = self.zone_fields()
        if name not in fields:
            return vmin, vmax
        z = fields[name]
        return min(vmin, float(z.min())), max(vmax, float(z.max()))

    def zone_conserved(self):
        """Analytic zone states as conserved rows [3, 4]."""
        f = self.zone_fields()
        rho = f['density']
        return np.column_stack([rho, rho * f['velx'], rho * f['vely'],
                                f['total_energy']])

    def zone_info(self, name='density'):
        """ZoneInfo of zones 1-3 for the named field ([] for a field with
        no analytic value; ValueError for an unknown name)."""
        values = self.field.field(name)
        fields = self.zone_fields()
        if name not in fields:
            return []
        out = []
        for iz, mask in enumerate(self.zone_masks()):
            count = int(mask.sum())
            comp = float(values[mask].mean()) if count else float('nan')
            target = float(fields[name][iz])
            out.append(ZoneInfo(iz + 1, count, comp, target,
                                self._relative(comp, target)))
        return out

    def crossings(self, nbin=None):
        """Where density crosses (rho1 + rho2) / 2 under the incident
        shock: the domain up to the end of the incident arm in ``nbin``
        column bins (one per box column by default), each walked upward from
        the wall to the first cell at or above the mid value, interpolated
        with the cell below; [npoint, 2]."""
        m = self.shock.mesher
        xend = self.incident.x1
        if nbin is None:
            nbin = max(1, round((xend - m.x0) / m.cell_extent[0]))
        cen = self.field.centroid
        rho = self.field.density
        mid = 0.5 * (self.shock.density + self.shock.density2)
        edges = np.linspace(m.x0, xend, nbin + 1)
        found = []
        for ib in range(nbin):
            sel = (cen[:, 0] >= edges[ib]) & (cen[:, 0] < edges[ib + 1])
            order = np.argsort(cen[sel, 1])
            ys, vs = cen[sel, 1][order], rho[sel][order]
            hi = np.nonzero(vs >= mid)[0]
            if hi.size == 0 or hi[0] == 0:
                continue
            k = hi[0]
            found.append((0.5 * (edges[ib] + edges[ib + 1]),
                          ys[k - 1] + (mid - vs[k - 1]) * (ys[k] - ys[k - 1])
                          / (vs[k] - vs[k - 1])))
        return np.array(found, dtype=np.float64).reshape(-1, 2)

    def fit_line(self, nbin=None):
        """(slope, intercept, npoint) of the least-squares line through the
        crossings (nan with fewer than two)."""
        pts = self.crossings(nbin)
        if len(pts) < 2:
            return float('nan'), float('nan'), len(pts)
        k, c = np.polyfit(pts[:, 0], pts[:, 1], 1)
        return float(k), float(c), len(pts)

    def fit_incident_angle(self, nbin=None):
        k, _, n = self.fit_line(nbin)
        analytic = math.degrees(self.shock.shock_angle)
        deg = math.degrees(math.atan(-k)) if k < 0.0 else float('nan')
        return AngleFit(deg, analytic, self._relative(deg, analytic), n)

    def reflection_point(self, nbin=None):
        k, c, _ = self.fit_line(nbin)
        xfit = (self.shock.mesher.y0 - c) / k if k < 0.0 else float('nan')
        analytic = self.reflected.x0 if self.has_reflection else float('nan')
        return WallPoint(xfit, analytic, self._relative(xfit, analytic))

    def profile(self, height, name='density', halfwidth=None):
        """The field along y = ``height``: cells within ``halfwidth`` (half
        a box by default) sorted by abscissa, with the analytic step."""
        if halfwidth is None:
            halfwidth = 0.5 * self.shock.mesher.cell_extent[1]
        cen = self.field.centroid
        sel = np.abs(cen[:, 1] - height) <= halfwidth
        order = np.argsort(cen[sel, 0])
        xs = cen[sel, 0][order]
        return HorizontalProfile(xs, self.field.field(name)[sel][order],
                                 self.analytic_profile(xs, height, name))

    def analytic_profile(self, xs, height, name='density'):
        z = self.zone_field(name)
        xs = np.asarray(xs)
        out = np.where(xs < self._abscissa(self.incident, height), z[0], z[1])
        if self.has_reflection:
            out = np.where(xs >= self._abscissa(self.reflected, height), z[2],
                           out)
        return out


# --------------------------------------------------------------- session -----

class ReflectionSession(object):
    """One reflection run marched chunk by chunk (ReflectionSession,
    solvcon/pilot/apps/obsrefl/_session.py:73-169): each ``advance``
    marches ``steps_per_chunk`` steps (the last chunk trimmed to
    ``max_steps``) and records step, mass and CFL range in ``history``."""

    STEPS_PER_CHUNK = 5
    MAX_STEPS = 2000

    def __init__(self, gamma=1.4, density=1.0, pressure=1.0, mach=3.0,
                 angle=10.0, cell_type='unstructured', time_increment=2.e-3,
                 nx=64, ny=16, steps_per_chunk=STEPS_PER_CHUNK,
                 max_steps=MAX_STEPS):
        self.shock = ObliqueShock()
        self.shock.build_constant(gamma=gamma, density=density,
                                  pressure=pressure, mach=mach, angle=angle)
        self.shock.build_numerical(cell_type=cell_type,
                                   time_increment=time_increment, nx=nx,
                                   ny=ny)
        self.analysis = Reflection(self.shock)
        self.history = RunHistory()
        self.steps_per_chunk = steps_per_chunk
        self.max_steps = max_steps
        self.step = 0
        self.stop_reason = None

    @property
    def field(self):
        return self.analysis.field

    @property
    def finished(self):
        return self.stop_reason is not None

    def advance(self):
        if self.finished:
            return None
        steps = min(self.steps_per_chunk, self.max_steps - self.step)
        self.shock.march(steps)
        self.step += steps
        cmin, cmax = self.field.calc_cfl_range()
        rec = self.history.append(self.step, self.field.calc_overall_mass(),
                                  cmin, cmax)
        if self.step >= self.max_steps:
            self.stop_reason = 'cap'
        return rec

    def run(self):
        while not self.finished:
            self.advance()
        return self.stop_reason

    def stop(self):
        if not self.finished:
            self.stop_reason = 'stopped'

    def zone_info(self, name='density'):
        return self.analysis.zone_info(name)

    def fit_incident_angle(self, nbin=None):
        return self.analysis.fit_incident_angle(nbin)

    def reflection_point(self, nbin=None):
        return self.analysis.reflection_point(nbin)

    def profile(self, height, name='density', halfwidth=None):
        return self.analysis.profile(height, name, halfwidth)
