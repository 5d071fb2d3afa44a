/*
 * Oracle (test infrastructure only): CPU restatement of solvcon's 1D
 * space-time CESE scalar engine, cpp/solvcon/spacetime/core.hpp and
 * core.cpp, with the element kernels of kernel/linear_scalar.hpp and
 * kernel/inviscid_burgers.hpp and the default hooks of Kernel::reset
 * (core.hpp:520-536).  Used by the tests as the checker of the HIP engine;
 * never part of the product path.
 *
 * Coordinates: xc[nx] with nx = 2*ncelm + 5 (Grid::init_from_array,
 * core.cpp:48-98).  Solution element j sits at xindex 2 + 2j + odd, the
 * conservation element j at 3 + 2j + odd (core.hpp:168-182).  Arrays so0 and
 * so1 are [nx][nvar] row-major; only variable 0 is marched
 * (core.hpp:954,978).
 */

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace
{

enum Kind : int
{
    K_DEFAULT = 0,
    K_LINEAR = 1,
    K_BURGERS = 2,
};

struct St
{
    int kind;
    int64_t nx;
    int64_t nvar;
    double dt, hdt, qdt;
    double const * xc;
    double * so0;
    double * so1;
    double * cfl;

    int64_t ncelm() const { return (nx - 5) / 2; }
    int64_t nselm() const { return ncelm() + 1; }

    // Selm geometry (core.hpp:478-484, ElementBase :44-56).
    double x(int64_t y) const { return xc[y]; }
    double xneg(int64_t y) const { return xc[y - 1]; }
    double xpos(int64_t y) const { return xc[y + 1]; }
    double xctr(int64_t y) const { return (xneg(y) + xpos(y)) / 2; }
    double dxneg(int64_t y) const { return x(y) - xneg(y); }
    double dxpos(int64_t y) const { return xpos(y) - x(y); }
    double dx(int64_t y) const { return xpos(y) - xneg(y); }
    double u(int64_t y, int64_t iv) const { return so0[y * nvar + iv]; }
    double ux(int64_t y, int64_t iv) const { return so1[y * nvar + iv]; }

    // linear_scalar.hpp:45-87 and inviscid_burgers.hpp:51-109.
    double xn(int64_t y, int64_t iv) const
    {
        if (kind == K_DEFAULT)
        {
            return 0.0;
        }
        double const displacement = 0.5 * (x(y) + xneg(y)) - xctr(y);
        return dxneg(y) * (u(y, iv) + displacement * ux(y, iv));
    }
    double xp(int64_t y, int64_t iv) const
    {
        if (kind == K_DEFAULT)
        {
            return 0.0;
        }
        double const displacement = 0.5 * (x(y) + xpos(y)) - xctr(y);
        return dxpos(y) * (u(y, iv) + displacement * ux(y, iv));
    }
    double tn(int64_t y, int64_t iv) const
    {
        if (kind == K_DEFAULT)
        {
            return 0.0;
        }
        double const displacement = x(y) - xctr(y);
        if (kind == K_LINEAR)
        {
            double ret = u(y, iv);
            ret += displacement * ux(y, iv);
            ret += qdt * ux(y, iv);
            return hdt * ret;
        }
        double const uu = u(y, iv);
        double const u_2 = uu * uu;
        double ret = 0.5 * u_2;
        ret += displacement * uu * ux(y, iv);
        ret += qdt * u_2 * ux(y, iv);
        return hdt * ret;
    }
    double tp(int64_t y, int64_t iv) const
    {
        if (kind == K_DEFAULT)
        {
            return 0.0;
        }
        double const displacement = x(y) - xctr(y);
        if (kind == K_LINEAR)
        {
            double ret = u(y, iv);
            ret += displacement * ux(y, iv);
            ret -= qdt * ux(y, iv);
            return hdt * ret;
        }
        double const uu = u(y, iv);
        double const u_2 = uu * uu;
        double ret = 0.5 * u_2;
        ret += displacement * uu * ux(y, iv);
        ret -= qdt * u_2 * ux(y, iv);
        return hdt * ret;
    }
    double so0p(int64_t y, int64_t iv) const
    {
        if (kind == K_DEFAULT)
        {
            return u(y, iv);
        }
        double ret = u(y, iv);
        ret += (x(y) - xctr(y)) * ux(y, iv);
        ret -= hdt * ux(y, iv);
        return ret;
    }
    double cfl_of(int64_t y) const
    {
        if (kind == K_DEFAULT)
        {
            return 0.0;
        }
        double const hdx = std::min(dxneg(y), dxpos(y));
        if (kind == K_LINEAR)
        {
            return hdt / hdx;
        }
        return std::fabs(u(y, 0)) * hdt / hdx;
    }

    // Celm::calc_so0 and calc_so1_alpha (core.hpp:645-672); X is the CE
    // centre, its bounding SEs sit at X - 1, X + 1 and X (top).
    double calc_so0(int64_t X, int64_t iv) const
    {
        double const flux_ll = xp(X - 1, iv) + tp(X - 1, iv);
        double const flux_ur = xn(X + 1, iv) - tp(X + 1, iv);
        return (flux_ll + flux_ur) / dx(X);
    }
    double calc_so1(int64_t X, int64_t iv, int alpha) const
    {
        double const upn = so0p(X - 1, iv);
        double const upp = so0p(X + 1, iv);
        double const utp = u(X, iv);
        double const duxn = (utp - upn) / dxpos(X - 1);
        double const duxp = (upp - utp) / dxneg(X + 1);
        double fan = 1.0;
        double fap = 1.0;
        if (alpha >= 1)
        {
            fan = std::fabs(duxn);
            fap = std::fabs(duxp);
        }
        if (alpha == 2)
        {
            fan = fan * std::fabs(duxn);
            fap = fap * std::fabs(duxp);
        }
        constexpr double tiny = std::numeric_limits<double>::min();
        return (fap * duxn + fan * duxp) / (fap + fan + tiny);
    }

    // SolverBase passes (core.hpp:946-1036).
    void half_so0(bool odd)
    {
        int64_t const start = odd ? -1 : 0;
        for (int64_t ic = start; ic < ncelm(); ++ic)
        {
            int64_t const X = 3 + 2 * ic + (odd ? 1 : 0);
            so0[X * nvar] = calc_so0(X, 0);
        }
    }
    void half_so1(int alpha, bool odd)
    {
        int64_t const start = odd ? -1 : 0;
        for (int64_t ic = start; ic < ncelm(); ++ic)
        {
            int64_t const X = 3 + 2 * ic + (odd ? 1 : 0);
            so1[X * nvar] = calc_so1(X, 0, alpha);
        }
    }
    void update_cfl(bool odd)
    {
        int64_t const start = odd ? -1 : 0;
        for (int64_t ic = start; ic < nselm(); ++ic)
        {
            int64_t const y = 2 + 2 * ic + (odd ? 1 : 0);
            cfl[y] = cfl_of(y);
        }
    }
    // Periodic: the outer odd-plane SEs take the opposite inner ones.
    void boundary(double * a)
    {
        int64_t const n = ncelm();
        double const left_in = a[3 * nvar];
        double const right_in = a[(2 * n + 1) * nvar];
        a[1 * nvar] = right_in;
        a[(2 * n + 3) * nvar] = left_in;
    }
    void half1(int alpha)
    {
        half_so0(false);
        boundary(so0);
        update_cfl(true);
        half_so1(alpha, false);
        boundary(so1);
    }
    void half2(int alpha)
    {
        half_so0(true);
        update_cfl(false);
        half_so1(alpha, true);
    }
};

St make(int kind, int64_t nx, int64_t nvar, double dt, double const * xc, double * so0, double * so1, double * cfl)
{
    St s;
    s.kind = kind;
    s.nx = nx;
    s.nvar = nvar;
    s.dt = dt;
    s.hdt = 0.5 * dt; // Field::set_time_increment, core.cpp:109-114
    s.qdt = 0.25 * dt;
    s.xc = xc;
    s.so0 = so0;
    s.so1 = so1;
    s.cfl = cfl;
    return s;
}

} /* end namespace */

extern "C"
{

// Grid::init_from_array (core.cpp:48-98) on xloc[n]: returns 0, or 1 for
// invalid arguments; xc receives 2*(n-1)+5 values.
int orcst_grid_from(double const * xloc, int64_t n, double * xc)
{
    if (n < 2)
    {
        return 1;
    }
    for (int64_t it = 0; it + 1 < n; ++it)
    {
        if (xloc[it] >= xloc[it + 1])
        {
            return 1;
        }
    }
    constexpr int64_t B = 2;
    int64_t const ncelm = n - 1;
    int64_t const nx = ncelm * 2 + (1 + B * 2);
    for (int64_t it = 0; it < n; ++it)
    {
        xc[it * 2 + B] = xloc[it];
    }
    for (int64_t it = 0; it < ncelm; ++it)
    {
        int64_t const ref = it * 2 + B + 1;
        xc[ref] = (xc[ref - 1] + xc[ref + 1]) / 2;
    }
    for (int64_t it = 1; it <= B; ++it)
    {
        xc[B - it] = xc[B] + xc[B] - xc[B + it];
        int64_t const ref = nx - B - 1;
        xc[ref + it] = xc[ref] + xc[ref] - xc[ref - it];
    }
    return 0;
}

// Grid::Grid(xmin, xmax, ncelm) (core.cpp:14-45): returns 0, or 1 for
// invalid arguments; xc receives 2*ncelm+5 values.
int orcst_grid(double xmin, double xmax, int64_t ncelm, double * xc)
{
    if (ncelm < 1 || xmin >= xmax)
    {
        return 1;
    }
    double const xspace = (xmax - xmin) / static_cast<double>(ncelm);
    double * xloc = new double[ncelm + 1];
    xloc[0] = xmin;
    for (int64_t it = 1; it < ncelm; ++it)
    {
        xloc[it] = xloc[it - 1] + xspace;
    }
    xloc[ncelm] = xmax;
    int const rc = orcst_grid_from(xloc, ncelm + 1, xc);
    delete[] xloc;
    return rc;
}

// One pass on host arrays: 0 march_half_so0, 1 march_half_so1_alpha,
// 2 update_cfl, 3 treat_boundary_so0, 4 treat_boundary_so1,
// 5 march_half1_alpha, 6 march_half2_alpha.
void orcst_stage(int kind, int stage, int alpha, int odd, int64_t nx, int64_t nvar, double dt, double const * xc,
                 double * so0, double * so1, double * cfl)
{
    St s = make(kind, nx, nvar, dt, xc, so0, so1, cfl);
    switch (stage)
    {
    case 0: s.half_so0(odd != 0); break;
    case 1: s.half_so1(alpha, odd != 0); break;
    case 2: s.update_cfl(odd != 0); break;
    case 3: s.boundary(so0); break;
    case 4: s.boundary(so1); break;
    case 5: s.half1(alpha); break;
    case 6: s.half2(alpha); break;
    default: break;
    }
}

// march_alpha<alpha>(steps), core.hpp:1026-1036.
void orcst_march(int kind, int alpha, int64_t steps, int64_t nx, int64_t nvar, double dt, double const * xc,
                 double * so0, double * so1, double * cfl)
{
    St s = make(kind, nx, nvar, dt, xc, so0, so1, cfl);
    for (int64_t it = 0; it < steps; ++it)
    {
        s.half1(alpha);
        s.half2(alpha);
    }
}

// One element value at xindex y (variable iv): 0 xn, 1 xp, 2 tn, 3 tp,
// 4 so0p, 5 the CFL number update_cfl would store, 6 calc_so0,
// 7 calc_so1_alpha.
double orcst_eval(int kind, int what, int alpha, int64_t y, int64_t iv, int64_t nx, int64_t nvar, double dt,
                  double const * xc, double * so0, double * so1, double * cfl)
{
    St s = make(kind, nx, nvar, dt, xc, so0, so1, cfl);
    switch (what)
    {
    case 0: return s.xn(y, iv);
    case 1: return s.xp(y, iv);
    case 2: return s.tn(y, iv);
    case 3: return s.tp(y, iv);
    case 4: return s.so0p(y, iv);
    case 5: return s.cfl_of(y);
    case 6: return s.calc_so0(y, iv);
    case 7: return s.calc_so1(y, iv, alpha);
    default: return 0.0;
    }
}

} /* end extern "C" */
