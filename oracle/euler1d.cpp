/*
 * Oracle (test infrastructure only): CPU restatement of the 1D Euler CESE
 * core, following cpp/solvcon/onedim/Euler1DCore.hpp:106-349 and
 * cpp/solvcon/onedim/Euler1DCore.cpp:22-205 operation for operation.  It acts
 * on caller-owned arrays: coord[n], cfl[n], so0[n,3], so1[n,3], gamma[n].
 */

#include "common.hpp"

#include <cmath>
#include <cstdint>

namespace orc
{
namespace
{

constexpr double TINY1D = 1.e-100; // Euler1DCore.hpp:47,182
constexpr double RGAS = 8.31446261815324; // Euler1DCore.hpp:48

struct Arrays1D
{
    int64_t n;
    double dt;
    double * coord;
    double * cfl;
    double * so0;
    double * so1;
    double const * gamma;
};

// One solution element: Euler1DKernel (Euler1DCore.hpp:180-282).
struct SolElem
{
    double ga, hdt, qdt, x, xctr, xneg, xpos;
    double u[3], ux[3], jac[3][3], f[3], ut[3], ft[3], up[3];

    void load(Arrays1D const & a, int64_t ic)
    {
        hdt = a.dt / 2.0;
        qdt = hdt / 2.0;
        ga = a.gamma[ic];
        x = a.coord[ic];
        xneg = a.coord[ic - 1];
        xpos = a.coord[ic + 1];
        xctr = (xpos + xneg) * 0.5;
        for (int k = 0; k < 3; ++k)
        {
            u[k] = a.so0[ic * 3 + k];
            ux[k] = a.so1[ic * 3 + k];
        }
    }

    // Euler1DCore.hpp:210-244
    void derive()
    {
        double const t = TINY1D;
        jac[0][0] = 0.0;
        jac[0][1] = 1.0;
        jac[0][2] = 0.0;
        jac[1][0] = (ga - 3.0) / 2.0 * u[1] * u[1] / (u[0] * u[0] + t);
        jac[1][1] = -(ga - 3.0) * u[1] / (u[0] + t);
        jac[1][2] = ga - 1.0;
        jac[2][0] = (ga - 1.0) * u[1] * u[1] * u[1] / (u[0] * u[0] * u[0] + t) - ga * u[1] * u[2] / (u[0] * u[0] + t);
        jac[2][1] = ga * u[2] / (u[0] + t) - 3.0 / 2.0 * (ga - 1.0) * u[1] * u[1] / (u[0] * u[0] + t);
        jac[2][2] = ga * u[1] / (u[0] + t);

        f[0] = u[1];
        f[1] = (ga - 1.0) * u[2] + (3.0 - ga) / 2.0 * u[1] * u[1] / (u[0] + t);
        f[2] = ga * u[1] * u[2] / (u[0] + t) - (ga - 1.0) / 2.0 * u[1] * u[1] * u[1] / (u[0] * u[0] + t);

        for (int i = 0; i < 3; ++i)
        {
            ut[i] = -jac[i][0] * ux[0] - jac[i][1] * ux[1] - jac[i][2] * ux[2];
        }
        for (int i = 0; i < 3; ++i)
        {
            ft[i] = jac[i][0] * ut[0] + jac[i][1] * ut[1] + jac[i][2] * ut[2];
        }
        for (int i = 0; i < 3; ++i)
        {
            up[i] = u[i] + (x - xctr) * ux[i] + hdt * ut[i];
        }
    }

    // calc_flux_ll, Euler1DCore.hpp:246-255
    void flux_ll(double r[3]) const
    {
        double const deltax = xpos - x;
        double const dxmid = 0.5 * (x + xpos) - xctr;
        double const dxctr = x - xctr;
        for (int i = 0; i < 3; ++i)
        {
            r[i] = deltax * (u[i] + dxmid * ux[i]) + hdt * (f[i] - (dxctr * ut[i]) + (qdt * ft[i]));
        }
    }

    // calc_flux_lr, Euler1DCore.hpp:257-266
    void flux_lr(double r[3]) const
    {
        double const deltax = x - xneg;
        double const dxmid = 0.5 * (x + xneg) - xctr;
        double const dxctr = x - xctr;
        for (int i = 0; i < 3; ++i)
        {
            r[i] = deltax * (u[i] + dxmid * ux[i]) - hdt * (f[i] - (dxctr * ut[i]) + (qdt * ft[i]));
        }
    }
};

inline int64_t plane_start(bool odd) { return 2 - (odd ? 1 : 0); }
inline int64_t plane_stop(int64_t n, bool odd) { return n - 2 - (odd ? 0 : 1); }

// Euler1DCore.cpp:70-94
void update_cfl(Arrays1D const & a, bool odd)
{
    double const hdt = a.dt / 2;
    for (int64_t it = plane_start(odd); it < plane_stop(a.n, odd); it += 2)
    {
        double const ga = a.gamma[it];
        double wspd = a.so0[it * 3 + 1];
        wspd *= wspd;
        double const ke = wspd / (2.0 * a.so0[it * 3 + 0]);
        double pr = (ga - 1.0) * (a.so0[it * 3 + 2] - ke);
        pr = (pr + std::fabs(pr)) / 2.0;
        wspd = std::sqrt(ga * pr / a.so0[it * 3 + 0]) + std::sqrt(wspd) / a.so0[it * 3 + 0];
        double const dxpos = a.coord[it + 1] - a.coord[it];
        double const dxneg = a.coord[it] - a.coord[it - 1];
        a.cfl[it] = hdt * wspd / (dxpos < dxneg ? dxpos : dxneg);
    }
}

// Euler1DCore.cpp:96-130
void half_so0(Arrays1D const & a, bool odd)
{
    int64_t const start = plane_start(odd);
    int64_t const stop = plane_stop(a.n, odd);
    SolElem xn;
    SolElem xp;
    xp.load(a, start);
    xp.derive();
    for (int64_t ic = start; ic < stop; ic += 2)
    {
        xn = xp;
        xp.load(a, ic + 2);
        xp.derive();
        double fll[3];
        double flr[3];
        xn.flux_ll(fll);
        xp.flux_lr(flr);
        double const dx = a.coord[ic + 2] - a.coord[ic];
        for (int i = 0; i < 3; ++i)
        {
            a.so0[(ic + 1) * 3 + i] = (fll[i] + flr[i]) / dx;
        }
    }
}

template <int ALPHA>
inline double powa(double v)
{
    // math.hpp:21-46: pow<N> by repeated multiplication.
    if constexpr (ALPHA == 1)
    {
        return v;
    }
    else
    {
        return v * v;
    }
}

// Euler1DCore.hpp:284-320
template <int ALPHA>
void half_so1(Arrays1D const & a, bool odd)
{
    int64_t const start = plane_start(odd);
    int64_t const stop = plane_stop(a.n, odd);
    SolElem xn;
    SolElem xp;
    xp.load(a, start);
    xp.derive();
    for (int64_t ic = start; ic < stop; ic += 2)
    {
        xn = xp;
        xp.load(a, ic + 2);
        xp.derive();
        for (int iv = 0; iv < 3; ++iv)
        {
            double const utp = a.so0[(ic + 1) * 3 + iv];
            double const duxn = (utp - xn.up[iv]) / (xn.xpos - xn.x);
            double const duxp = (xp.up[iv] - utp) / (xp.x - xp.xneg);
            double const fan = powa<ALPHA>(std::fabs(duxn));
            double const fap = powa<ALPHA>(std::fabs(duxp));
            a.so1[(ic + 1) * 3 + iv] = (fap * duxn + fan * duxp) / (fap + fan + TINY1D);
        }
    }
}

void half_so1(Arrays1D const & a, int alpha, bool odd)
{
    if (alpha == 1)
    {
        half_so1<1>(a, odd);
    }
    else
    {
        half_so1<2>(a, odd);
    }
}

// Euler1DCore.cpp:132-172
void boundary(double * arr, int64_t n)
{
    for (int i = 0; i < 3; ++i)
    {
        arr[1 * 3 + i] = arr[2 * 3 + i];
    }
    for (int i = 0; i < 3; ++i)
    {
        arr[(n - 2) * 3 + i] = arr[(n - 3) * 3 + i];
    }
}

} /* end namespace */
} /* end namespace orc */

using orc::Arrays1D;

extern "C"
{

// stage: 0 update_cfl, 1 march_half_so0, 2 march_half_so1 (alpha), 3 treat_boundary_so0,
// 4 treat_boundary_so1, 5 march_half1 (alpha), 6 march_half2 (alpha)
void orc1d_stage(int stage, int alpha, int odd, int64_t n, double dt,
                 double * coord, double * cfl, double * so0, double * so1, double const * gamma)
{
    Arrays1D const a{n, dt, coord, cfl, so0, so1, gamma};
    bool const op = odd != 0;
    switch (stage)
    {
    case 0: orc::update_cfl(a, op); break;
    case 1: orc::half_so0(a, op); break;
    case 2: orc::half_so1(a, alpha, op); break;
    case 3: orc::boundary(so0, n); break;
    case 4: orc::boundary(so1, n); break;
    case 5: // Euler1DCore.hpp:322-328
        orc::half_so0(a, false);
        orc::update_cfl(a, false);
        orc::half_so1(a, alpha, false);
        break;
    case 6: // Euler1DCore.hpp:330-337
        orc::half_so0(a, true);
        orc::update_cfl(a, true);
        orc::half_so1(a, alpha, true);
        break;
    default: break;
    }
}

// Euler1DCore.hpp:339-349
void orc1d_march(int alpha, int64_t steps, int64_t n, double dt,
                 double * coord, double * cfl, double * so0, double * so1, double const * gamma)
{
    Arrays1D const a{n, dt, coord, cfl, so0, so1, gamma};
    for (int64_t it = 0; it < steps; ++it)
    {
        orc::half_so0(a, false);
        orc::update_cfl(a, false);
        orc::half_so1(a, alpha, false);
        orc::boundary(so0, n);
        orc::boundary(so1, n);
        orc::half_so0(a, true);
        orc::update_cfl(a, true);
        orc::half_so1(a, alpha, true);
    }
}

// Derived fields, Euler1DCore.hpp:106-168 and Euler1DCore.cpp:37-68,174-205.
// which: 0 density, 1 velocity, 2 pressure, 3 temperature, 4 internal energy, 5 entropy
void orc1d_derived(int which, int64_t n, double const * so0, double const * gamma, double * out)
{
    for (int64_t it = 0; it < n; ++it)
    {
        double const u0 = so0[it * 3 + 0];
        double const u1 = so0[it * 3 + 1];
        double const u2 = so0[it * 3 + 2];
        double const ga = gamma[it];
        auto pressure = [&]()
        {
            double r = u1;
            r *= r;
            r /= 2.0 * u0 + orc::TINY1D;
            r = u2 - r;
            r *= ga - 1.0;
            return r;
        };
        double v = 0.0;
        switch (which)
        {
        case 0: v = u0; break;
        case 1: v = u1 / (u0 + orc::TINY1D); break;
        case 2: v = pressure(); break;
        case 3:
            v = (ga - 1.0) / orc::RGAS;
            v *= u2 / u0 - 0.5 * (u1 * u1) / (u0 * u0);
            break;
        case 4:
            v = u2 / u0;
            v -= 0.5 * ((u1 * u1) / (u0 * u0));
            break;
        case 5: v = pressure() / std::pow(u0, ga); break;
        default: break;
        }
        out[it] = v;
    }
}

} /* end extern "C" */
