"""Exact solution of the 1D Euler Riemann problem for an ideal gas.

Test infrastructure: a checker for the marched Sod tube, never imported by
the product package.

It is written from the textbook construction (E. F. Toro, *Riemann Solvers
and Numerical Methods for Fluid Dynamics*, 3rd ed., chapter 4):
- the star-region pressure p* is the root of
  f(p) = f_L(p) + f_R(p) + (u_R - u_L), where f_K is the shock
  (Rankine-Hugoniot) branch for p > p_K and the isentropic rarefaction
  branch for p <= p_K;
- Newton's method from the two-rarefaction estimate finds the root;
- the solution is then sampled self-similarly at xi = (x - x0) / t.

The reference gets the same solution another way: its ``ShockTube``
(solvcon/onedim/euler1d.py:64-393) runs a secant iteration on the
shock-tube pressure relation.  The values its tests pin
(tests/test_onedim_euler.py:113-141: p4/p5, the zone densities, the contact
velocity and the four wave positions at t = 0.1) pin this solver; the jump
conditions and the Riemann invariant are checked separately.
"""

import math

import numpy as np

__all__ = ['RiemannSolution', 'conserved']


def conserved(rho, u, p, gamma):
    """(rho, rho u, E) of a primitive state, E = p/(gamma-1) + rho u^2/2."""
    return rho, rho * u, p / (gamma - 1.0) + 0.5 * rho * u * u


class _Side(object):
    """One initial state and its branch of the pressure function."""

    def __init__(self, rho, u, p, gamma, sign):
        self.rho, self.u, self.p = float(rho), float(u), float(p)
        self.g = gamma
        self.sign = sign            # -1 for the left state, +1 for the right
        self.c = math.sqrt(gamma * p / rho)
        self.a = 2.0 / ((gamma + 1.0) * rho)
        self.b = (gamma - 1.0) / (gamma + 1.0) * p

    def f(self, p):
        """Velocity jump across this side's wave at star pressure p, and
        its derivative."""
        g = self.g
        if p > self.p:              # shock
            q = math.sqrt(self.a / (p + self.b))
            return (p - self.p) * q, q * (1.0 - 0.5 * (p - self.p)
                                          / (p + self.b))
        r = (p / self.p) ** ((g - 1.0) / (2.0 * g))     # rarefaction
        return (2.0 * self.c / (g - 1.0) * (r - 1.0),
                1.0 / (self.rho * self.c) * (p / self.p)
                ** (-(g + 1.0) / (2.0 * g)))

    def star_density(self, ps):
        g = self.g
        t = ps / self.p
        if ps > self.p:
            k = (g - 1.0) / (g + 1.0)
            return self.rho * (t + k) / (k * t + 1.0)
        return self.rho * t ** (1.0 / g)

    def shock_speed(self, ps):
        g = self.g
        return self.u + self.sign * self.c * math.sqrt(
            (g + 1.0) / (2.0 * g) * ps / self.p + (g - 1.0) / (2.0 * g))

    def fan(self, xi):
        """State inside this side's rarefaction fan at xi = x/t."""
        g = self.g
        s = self.sign
        base = 2.0 / (g + 1.0) - s * (g - 1.0) / ((g + 1.0) * self.c) \
            * (self.u - xi)
        rho = self.rho * base ** (2.0 / (g - 1.0))
        u = 2.0 / (g + 1.0) * (-s * self.c + 0.5 * (g - 1.0) * self.u + xi)
        p = self.p * base ** (2.0 * g / (g - 1.0))
        return rho, u, p


class RiemannSolution(object):
    """The self-similar solution of one Riemann problem.

    ``left`` and ``right`` are primitive states (rho, u, p); the membrane
    sits at ``x0``.  After construction:
    - ``p_star`` and ``u_star`` are the star-region pressure and velocity;
    - ``rho_star_left`` / ``rho_star_right`` the densities on either side
      of the contact;
    - ``left_wave`` / ``right_wave`` are 'shock' or 'rarefaction'.
    """

    def __init__(self, left, right, gamma=1.4, x0=0.0, tol=1e-15,
                 maxiter=100):
        self.gamma = float(gamma)
        self.x0 = float(x0)
        self.L = _Side(*left, gamma=self.gamma, sign=-1)
        self.R = _Side(*right, gamma=self.gamma, sign=+1)
        du = self.R.u - self.L.u
        g = self.gamma
        # A vacuum forms when the rarefactions alone cannot meet.
        if 2.0 * (self.L.c + self.R.c) / (g - 1.0) <= du:
            raise ValueError('the initial states generate a vacuum')
        # Two-rarefaction estimate as the starting point.
        z = (g - 1.0) / (2.0 * g)
        p = ((self.L.c + self.R.c - 0.5 * (g - 1.0) * du)
             / (self.L.c / self.L.p ** z + self.R.c / self.R.p ** z)) \
            ** (1.0 / z)
        p = max(p, 1e-14)
        for it in range(maxiter):
            fl, dl = self.L.f(p)
            fr, dr = self.R.f(p)
            step = (fl + fr + du) / (dl + dr)
            pn = max(p - step, 1e-14 * p)
            if abs(pn - p) <= tol * 0.5 * (pn + p):
                p = pn
                break
            p = pn
        else:
            raise RuntimeError('Newton iteration for p* did not converge')
        self.iterations = it + 1
        self.p_star = p
        fl, _ = self.L.f(p)
        fr, _ = self.R.f(p)
        self.u_star = 0.5 * (self.L.u + self.R.u) + 0.5 * (fr - fl)
        self.rho_star_left = self.L.star_density(p)
        self.rho_star_right = self.R.star_density(p)
        self.left_wave = 'shock' if p > self.L.p else 'rarefaction'
        self.right_wave = 'shock' if p > self.R.p else 'rarefaction'

    def speeds(self):
        """Wave speeds from left to right: (left head, left tail, contact,
        right tail, right head); a shock has equal head and tail."""
        g = self.gamma
        ps, us = self.p_star, self.u_star
        out = []
        for side in (self.L, self.R):
            if ps > side.p:
                s = side.shock_speed(ps)
                out.append((s, s))
            else:
                cs = side.c * (ps / side.p) ** ((g - 1.0) / (2.0 * g))
                out.append((side.u + side.sign * side.c,
                            us + side.sign * cs))
        (lh, lt), (rh, rt) = out
        return lh, lt, us, rt, rh

    def locations(self, t):
        """Positions at time t of the four fronts of a left rarefaction and
        a right shock: fan head, fan tail, contact, shock."""
        lh, lt, us, rt, rh = self.speeds()
        return [self.x0 + t * s for s in (lh, lt, us, rh)]

    def sample(self, x, t):
        """Primitive fields (rho, u, p) at positions x and time t > 0."""
        x = np.asarray(x, dtype=np.float64)
        rho = np.empty_like(x)
        u = np.empty_like(x)
        p = np.empty_like(x)
        lh, lt, us, rt, rh = self.speeds()
        for k, xk in enumerate(x.ravel()):
            xi = (xk - self.x0) / t
            if xi < us:
                side, head, tail = self.L, lh, lt
                star = self.rho_star_left
                if xi < head:
                    st = (side.rho, side.u, side.p)
                elif xi >= tail:
                    st = (star, self.u_star, self.p_star)
                else:
                    st = side.fan(xi)
            else:
                side, head, tail = self.R, rh, rt
                star = self.rho_star_right
                if xi > head:
                    st = (side.rho, side.u, side.p)
                elif xi <= tail:
                    st = (star, self.u_star, self.p_star)
                else:
                    st = side.fan(xi)
            rho.flat[k], u.flat[k], p.flat[k] = st
        return rho, u, p
