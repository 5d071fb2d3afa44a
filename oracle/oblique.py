"""Oblique-shock relations and the regular reflection of an oblique shock
from a wall, for a calorically perfect gas.

Test infrastructure: the analytic checker of the inlet / slip-wall march
that solvcon's oblique-shock reflection application runs
(solvcon/pilot/apps/obsrefl).  Never imported by the product package.

Written from the textbook relations (J. D. Anderson, *Modern Compressible
Flow*, 3rd ed., sections 3.6 and 4.3-4.5), with M_n1 = M_1 sin(beta):
- rho2/rho1 = (g+1) M_n1^2 / ((g-1) M_n1^2 + 2);
- p2/p1 = 1 + 2g/(g+1) (M_n1^2 - 1);
- M_n2^2 = (M_n1^2 + 2/(g-1)) / (2g/(g-1) M_n1^2 - 1);
- M_2 = M_n2 / sin(beta - theta);
- tan(theta) = 2 cot(beta) (M_1^2 sin^2 beta - 1) / (M_1^2 (g + cos 2 beta)
  + 2), the theta-beta-M relation, inverted here for the weak shock by
  bisection on [asin(1/M_1), beta at the largest deflection].
"""

import math

import numpy as np

__all__ = ['density_ratio', 'pressure_ratio', 'temperature_ratio',
           'normal_mach_downstream', 'downstream_mach', 'deflection',
           'shock_angle', 'RegularReflection']


def density_ratio(m1, beta, g=1.4):
    mn2 = (m1 * math.sin(beta)) ** 2
    return (g + 1.0) * mn2 / ((g - 1.0) * mn2 + 2.0)


def pressure_ratio(m1, beta, g=1.4):
    mn2 = (m1 * math.sin(beta)) ** 2
    return 1.0 + 2.0 * g / (g + 1.0) * (mn2 - 1.0)


def temperature_ratio(m1, beta, g=1.4):
    # Ideal gas: T2/T1 = (p2/p1) / (rho2/rho1).
    return pressure_ratio(m1, beta, g) / density_ratio(m1, beta, g)


def normal_mach_downstream(mn1, g=1.4):
    return math.sqrt((mn1 * mn1 + 2.0 / (g - 1.0))
                     / (2.0 * g / (g - 1.0) * mn1 * mn1 - 1.0))


def deflection(m1, beta, g=1.4):
    """Flow deflection theta behind a shock of angle beta."""
    s2 = math.sin(beta) ** 2
    num = 2.0 / math.tan(beta) * (m1 * m1 * s2 - 1.0)
    den = m1 * m1 * (g + math.cos(2.0 * beta)) + 2.0
    return math.atan(num / den)


def _beta_max(m1, g):
    # The largest deflection: maximise deflection() by golden section.
    lo, hi = math.asin(1.0 / m1), 0.5 * math.pi
    r = 0.5 * (math.sqrt(5.0) - 1.0)
    a, b = hi - r * (hi - lo), lo + r * (hi - lo)
    for _ in range(200):
        if deflection(m1, a, g) > deflection(m1, b, g):
            hi, b = b, a
            a = hi - r * (hi - lo)
        else:
            lo, a = a, b
            b = lo + r * (hi - lo)
    return 0.5 * (lo + hi)


def shock_angle(m1, theta, g=1.4):
    """Weak-shock angle beta for Mach m1 and deflection theta (radians).
    ValueError when theta exceeds the largest attached deflection."""
    lo = math.asin(1.0 / m1)
    hi = _beta_max(m1, g)
    if theta > deflection(m1, hi, g):
        raise ValueError('deflection %g rad exceeds the attached-shock '
                         'limit at Mach %g' % (theta, m1))
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if deflection(m1, mid, g) < theta:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def downstream_mach(m1, beta=None, theta=None, g=1.4):
    """M_2 behind a shock given either its angle or the deflection."""
    if (beta is None) == (theta is None):
        raise ValueError('give exactly one of beta and theta')
    if beta is None:
        beta = shock_angle(m1, theta, g)
    th = deflection(m1, beta, g)
    return normal_mach_downstream(m1 * math.sin(beta), g) \
        / math.sin(beta - th)


class RegularReflection(object):
    """Incident shock from the upper-left corner (0, height) of a channel,
    deflecting a free stream (rho1, p1, Mach m1, along +x) by ``theta``
    towards the wall y = 0, and its regular reflection from that wall.

    Zone 1 is the free stream, zone 2 the flow behind the incident shock
    (turned by -theta), zone 3 the flow behind the reflected shock (turned
    back to +x)."""

    def __init__(self, m1=3.0, theta_deg=10.0, rho1=1.0, p1=1.0, g=1.4,
                 height=1.0):
        self.g = g
        self.height = height
        th = math.radians(theta_deg)
        self.theta = th
        self.beta1 = shock_angle(m1, th, g)
        c1 = math.sqrt(g * p1 / rho1)
        m2 = downstream_mach(m1, beta=self.beta1, g=g)
        rho2 = rho1 * density_ratio(m1, self.beta1, g)
        p2 = p1 * pressure_ratio(m1, self.beta1, g)
        c2 = math.sqrt(g * p2 / rho2)
        self.beta2 = shock_angle(m2, th, g)
        rho3 = rho2 * density_ratio(m2, self.beta2, g)
        p3 = p2 * pressure_ratio(m2, self.beta2, g)
        m3 = downstream_mach(m2, beta=self.beta2, g=g)
        c3 = math.sqrt(g * p3 / rho3)
        self.zones = [
            (rho1, (m1 * c1, 0.0), p1),
            (rho2, (m2 * c2 * math.cos(th), -m2 * c2 * math.sin(th)), p2),
            (rho3, (m3 * c3, 0.0), p3)]
        self.machs = (m1, m2, m3)
        self.wall_x = height / math.tan(self.beta1)
        # The reflected shock leaves the wall at beta2 - theta to it.
        self.reflect_slope = math.tan(self.beta2 - th)

    def conserved(self, zone):
        rho, (u, v), p = self.zones[zone]
        return np.array([rho, rho * u, rho * v,
                         p / (self.g - 1.0) + 0.5 * rho * (u * u + v * v)])

    def zone_of(self, xy, margin=0.0):
        """Zone index (0, 1, 2) of each point, or -1 within ``margin`` of
        a shock."""
        x, y = xy[:, 0], xy[:, 1]
        tb = math.tan(self.beta1)
        # Signed distances: positive downstream of each shock.
        d_inc = (x * tb + y - self.height) / math.hypot(tb, 1.0)
        s = self.reflect_slope
        d_ref = (s * (x - self.wall_x) - y) / math.hypot(s, 1.0)
        z = np.where(d_inc < 0, 0, np.where(d_ref < 0, 1, 2))
        near = (np.abs(d_inc) < margin) | \
            ((np.abs(d_ref) < margin) & (x > self.wall_x - margin))
        return np.where(near, -1, z)
