/*
 * Oracle (test infrastructure only): CPU restatement of solvcon's EulerCore.
 * Arithmetic follows the cited reference lines operation for operation so
 * that results are reproducible bit for bit (compiled with
 * -ffp-contract=off and without fast-math).  The per-cell loops may run on
 * several OpenMP threads; every cell writes only its own rows, so the result
 * does not depend on the thread count.
 */

#include "euler.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace orc
{

// GradientElement.cpp:16-79
GeType const & getype(int32_t tpn)
{
    static std::array<GeType, 9> const tab = []()
    {
        std::array<GeType, 9> t{};
        auto set = [&](int32_t id, int32_t clnfc, int32_t nfge, std::vector<std::array<int32_t, 3>> f)
        {
            t[id].clnfc = clnfc;
            t[id].nfge = nfge;
            t[id].nfge_inverse = 1.0 / nfge;
            for (size_t i = 0; i < f.size(); ++i)
            {
                t[id].faces[i] = f[i];
            }
        };
        set(QUADRILATERAL, 4, 4, {{1, 2, -1}, {2, 3, -1}, {3, 4, -1}, {4, 1, -1}});
        set(TRIANGLE, 3, 3, {{1, 2, -1}, {2, 3, -1}, {3, 1, -1}});
        set(HEXAHEDRON, 6, 8, {{2, 3, 5}, {6, 3, 2}, {4, 3, 6}, {5, 3, 4}, {5, 1, 2}, {2, 1, 6}, {6, 1, 4}, {4, 1, 5}});
        set(TETRAHEDRON, 4, 4, {{3, 1, 2}, {2, 1, 4}, {4, 1, 3}, {2, 4, 3}});
        set(PRISM, 5, 6, {{5, 2, 4}, {3, 2, 5}, {4, 2, 3}, {4, 1, 5}, {5, 1, 3}, {3, 1, 4}});
        set(PYRAMID, 5, 6, {{1, 5, 2}, {2, 5, 3}, {3, 5, 4}, {4, 5, 1}, {1, 3, 4}, {3, 1, 2}});
        return t;
    }();
    return tab[static_cast<size_t>(tpn)];
}

namespace
{

using vec3 = std::array<double, 3>;
using mat3 = std::array<vec3, 3>;

// calc_gge_centroid, GradientElement.cpp:81-133
vec3 gge_centroid(GeType const & ge, vec3 const & avg, std::array<vec3, CLMFC> const & gp, int ndim)
{
    vec3 cnd = {0, 0, 0};
    double voc = 0.0;
    for (int32_t isub = 0; isub < ge.nfge; ++isub)
    {
        vec3 subcnd = avg;
        mat3 dst = {};
        for (int ivx = 0; ivx < ndim; ++ivx)
        {
            int32_t const fl = ge.faces[isub][ivx] - 1;
            for (int d = 0; d < ndim; ++d)
            {
                subcnd[d] += gp[fl][d];
                dst[ivx][d] = gp[fl][d] - avg[d];
            }
        }
        for (int d = 0; d < ndim; ++d)
        {
            subcnd[d] /= static_cast<double>(ndim + 1);
        }
        double vob = 0.0;
        if (ndim == 2)
        {
            vob = dst[0][0] * dst[1][1] - dst[0][1] * dst[1][0];
        }
        else
        {
            vob = dst[0][0] * (dst[1][1] * dst[2][2] - dst[1][2] * dst[2][1])
                - dst[0][1] * (dst[1][0] * dst[2][2] - dst[1][2] * dst[2][0])
                + dst[0][2] * (dst[1][0] * dst[2][1] - dst[1][1] * dst[2][0]);
        }
        voc += vob;
        for (int d = 0; d < ndim; ++d)
        {
            cnd[d] += subcnd[d] * vob;
        }
    }
    for (int d = 0; d < ndim; ++d)
    {
        cnd[d] /= voc;
    }
    return cnd;
}

} /* end namespace */

// GradientElement.cpp:137-213
GradElem::GradElem(Mesh const & m, Table<double> const & cecnd, int32_t icl_in, double tau)
    : icl(icl_in)
    , ndim(m.ndim)
    , clnfc(m.clfcs(icl_in, 0))
{
    ge = &getype(m.cltpn(icl));
    vec3 icecnd = {0, 0, 0};
    for (int d = 0; d < ndim; ++d)
    {
        icecnd[d] = cecnd(icl, d);
    }
    std::array<vec3, CLMFC> gp{};
    std::array<vec3, CLMFC> jd{};
    for (int32_t ifl = 0; ifl < clnfc; ++ifl)
    {
        int32_t const ifc = m.clfcs(icl, ifl + 1);
        int32_t const jcl = m.fcrcl(ifc, icl);
        rcls[ifl] = jcl;
        for (int d = 0; d < ndim; ++d)
        {
            double const mid = cecnd(icl, (ifl + 1) * ndim + d);
            double const jce = (jcl >= 0) ? cecnd(jcl, d) : m.clcnd(jcl, d);
            gp[ifl][d] = mid + tau * (jce - mid);
            jd[ifl][d] = gp[ifl][d] - jce;
        }
    }
    vec3 avg = {0, 0, 0};
    for (int32_t ifl = 0; ifl < clnfc; ++ifl)
    {
        for (int d = 0; d < ndim; ++d)
        {
            avg[d] += gp[ifl][d];
        }
    }
    for (int d = 0; d < ndim; ++d)
    {
        avg[d] /= clnfc;
    }
    vec3 const cnd = gge_centroid(*ge, avg, gp, ndim);
    for (int32_t ifl = 0; ifl < clnfc; ++ifl)
    {
        for (int d = 0; d < ndim; ++d)
        {
            idis[ifl][d] = gp[ifl][d] - cnd[d];
            jdis[ifl][d] = jd[ifl][d] + icecnd[d] - cnd[d];
        }
    }
}

// GradientElement.cpp:383-394
double ge_determinant(mat3 const & a, int ndim)
{
    if (ndim == 2)
    {
        return a[0][0] * a[1][1] - a[0][1] * a[1][0];
    }
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
         - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
         + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

// GradientElement.cpp:396-417
mat3 ge_adjugate(mat3 const & a, int ndim)
{
    mat3 r = {};
    if (ndim == 2)
    {
        r[0][0] = a[1][1];
        r[0][1] = -a[0][1];
        r[1][0] = -a[1][0];
        r[1][1] = a[0][0];
        return r;
    }
    r[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    r[0][1] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
    r[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    r[1][0] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    r[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    r[1][2] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
    r[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    r[2][1] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
    r[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    return r;
}

// GradientElement.cpp:432-445
mat3 GradElem::displacement_matrix(int32_t ifge) const
{
    mat3 dst = {};
    for (int ivx = 0; ivx < ndim; ++ivx)
    {
        int32_t const ifl = ge->faces[ifge][ivx] - 1;
        for (int d = 0; d < ndim; ++d)
        {
            dst[ivx][d] = idis[ifl][d];
        }
    }
    return dst;
}

// GradientElement.cpp:419-430,447-458
vec3 GradElem::solve_gradient(int32_t ifge, vec3 const & udf) const
{
    mat3 const dst = displacement_matrix(ifge);
    mat3 const adj = ge_adjugate(dst, ndim);
    vec3 adjudf = {0, 0, 0};
    for (int i = 0; i < ndim; ++i)
    {
        for (int j = 0; j < ndim; ++j)
        {
            adjudf[i] += adj[i][j] * udf[j];
        }
    }
    double const det = ge_determinant(dst, ndim);
    vec3 grad = {0, 0, 0};
    for (int d = 0; d < ndim; ++d)
    {
        grad[d] = adjudf[d] / det;
    }
    return grad;
}

namespace
{

constexpr double TINYJ = 1.e-60; // euler_march.cpp:30,89

// EulerJacobian<2>::update, euler_march.cpp:26-83
struct Jac2
{
    double jacos[4][4][2];
    double fcn[4][2];

    void update(double gamma, double const * sol)
    {
        double const ga = gamma;
        double const ga1 = ga - 1;
        double const ga3 = ga - 3;
        double const ga1h = ga1 / 2;
        double const u1 = sol[0] + TINYJ;
        double const u2 = sol[1];
        double const u3 = sol[2];
        double const u4 = sol[3];
        double const rho2 = u1 * u1;
        double const v1 = u2 / u1;
        double const v1o2 = v1 * v1;
        double const v2 = u3 / u1;
        double const v2o2 = v2 * v2;
        double const ke2 = (u2 * u2 + u3 * u3) / u1;
        double const g1ke2 = ga1 * ke2;
        double const vs = ke2 / u1;
        double const gretot = ga * u4;
        double const getot = gretot / u1;
        double const pr = ga1 * u4 - ga1h * ke2;

        fcn[0][0] = u2;
        fcn[0][1] = u3;
        fcn[1][0] = pr + u2 * v1;
        fcn[1][1] = u2 * v2;
        fcn[2][0] = u3 * v1;
        fcn[2][1] = pr + u3 * v2;
        fcn[3][0] = (pr + u4) * v1;
        fcn[3][1] = (pr + u4) * v2;

        auto put = [this](int i, int j, double a, double b)
        {
            jacos[i][j][0] = a;
            jacos[i][j][1] = b;
        };
        put(0, 0, 0, 0);
        put(0, 1, 1, 0);
        put(0, 2, 0, 1);
        put(0, 3, 0, 0);
        put(1, 0, -v1o2 + ga1h * vs, -v1 * v2);
        put(1, 1, -ga3 * v1, v2);
        put(1, 2, -ga1 * v2, v1);
        put(1, 3, ga1, 0);
        put(2, 0, -v2 * v1, -v2o2 + ga1h * vs);
        put(2, 1, v2, -ga1 * v1);
        put(2, 2, v1, -ga3 * v2);
        put(2, 3, 0, ga1);
        put(3, 0, (-gretot + g1ke2) * u2 / rho2, (-gretot + g1ke2) * u3 / rho2);
        put(3, 1, getot - ga1h * (vs + 2 * v1o2), -ga1 * v1 * v2);
        put(3, 2, -ga1 * v2 * v1, getot - ga1h * (vs + 2 * v2o2));
        put(3, 3, ga * v1, ga * v2);
    }
};

// EulerJacobian<3>::update, euler_march.cpp:85-158
struct Jac3
{
    double jacos[5][5][3];
    double fcn[5][3];

    void update(double gamma, double const * sol)
    {
        double const ga = gamma;
        double const ga1 = ga - 1;
        double const ga3 = ga - 3;
        double const ga1h = ga1 / 2;
        double const u1 = sol[0] + TINYJ;
        double const u2 = sol[1];
        double const u3 = sol[2];
        double const u4 = sol[3];
        double const u5 = sol[4];
        double const rho2 = u1 * u1;
        double const v1 = u2 / u1;
        double const v1o2 = v1 * v1;
        double const v2 = u3 / u1;
        double const v2o2 = v2 * v2;
        double const v3 = u4 / u1;
        double const v3o2 = v3 * v3;
        double const ke2 = (u2 * u2 + u3 * u3 + u4 * u4) / u1;
        double const g1ke2 = ga1 * ke2;
        double const vs = ke2 / u1;
        double const gretot = ga * u5;
        double const getot = gretot / u1;
        double const pr = ga1 * u5 - ga1h * ke2;

        auto putf = [this](int i, double a, double b, double c)
        {
            fcn[i][0] = a;
            fcn[i][1] = b;
            fcn[i][2] = c;
        };
        putf(0, u2, u3, u4);
        putf(1, pr + u2 * v1, u2 * v2, u2 * v3);
        putf(2, u3 * v1, pr + u3 * v2, u3 * v3);
        putf(3, u4 * v1, u4 * v2, pr + u4 * v3);
        putf(4, (pr + u5) * v1, (pr + u5) * v2, (pr + u5) * v3);

        auto put = [this](int i, int j, double a, double b, double c)
        {
            jacos[i][j][0] = a;
            jacos[i][j][1] = b;
            jacos[i][j][2] = c;
        };
        put(0, 0, 0, 0, 0);
        put(0, 1, 1, 0, 0);
        put(0, 2, 0, 1, 0);
        put(0, 3, 0, 0, 1);
        put(0, 4, 0, 0, 0);
        put(1, 0, -v1o2 + ga1h * vs, -v1 * v2, -v1 * v3);
        put(1, 1, -ga3 * v1, v2, v3);
        put(1, 2, -ga1 * v2, v1, 0);
        put(1, 3, -ga1 * v3, 0, v1);
        put(1, 4, ga1, 0, 0);
        put(2, 0, -v2 * v1, -v2o2 + ga1h * vs, -v2 * v3);
        put(2, 1, v2, -ga1 * v1, 0);
        put(2, 2, v1, -ga3 * v2, v3);
        put(2, 3, 0, -ga1 * v3, v2);
        put(2, 4, 0, ga1, 0);
        put(3, 0, -v3 * v1, -v3 * v2, -v3o2 + ga1h * vs);
        put(3, 1, v3, 0, -ga1 * v1);
        put(3, 2, 0, v3, -ga1 * v2);
        put(3, 3, v1, v2, -ga3 * v3);
        put(3, 4, 0, 0, ga1);
        put(4, 0, (-gretot + g1ke2) * u2 / rho2, (-gretot + g1ke2) * u3 / rho2, (-gretot + g1ke2) * u4 / rho2);
        put(4, 1, getot - ga1h * (vs + 2 * v1o2), -ga1 * v1 * v2, -ga1 * v1 * v3);
        put(4, 2, -ga1 * v2 * v1, getot - ga1h * (vs + 2 * v2o2), -ga1 * v2 * v3);
        put(4, 3, -ga1 * v3 * v1, -ga1 * v3 * v2, getot - ga1h * (vs + 2 * v3o2));
        put(4, 4, ga * v1, ga * v2, ga * v3);
    }
};

template <int ND>
struct JacOf;
template <>
struct JacOf<2>
{
    using type = Jac2;
};
template <>
struct JacOf<3>
{
    using type = Jac3;
};

// calc_solt_impl, euler_march.cpp:160-191
template <int ND>
void solt_impl(Euler & ec)
{
    constexpr int NEQ = ND + 2;
    int64_t const lo = -static_cast<int64_t>(ec.ngstcell);
    int64_t const hi = ec.ncell;
#pragma omp parallel for num_threads(ec.nthreads) schedule(static)
    for (int64_t icl = lo; icl < hi; ++icl)
    {
        typename JacOf<ND>::type jaco;
        double sol[NEQ];
        for (int ieq = 0; ieq < NEQ; ++ieq)
        {
            sol[ieq] = ec.so0c(icl, ieq);
        }
        jaco.update(ec.gamma(icl), sol);
        for (int ieq = 0; ieq < NEQ; ++ieq)
        {
            double val = 0.0;
            for (int idm = 0; idm < ND; ++idm)
            {
                for (int jeq = 0; jeq < NEQ; ++jeq)
                {
                    val += jaco.jacos[ieq][jeq][idm] * ec.so1c(icl, jeq, idm);
                }
            }
            ec.so0t(icl, ieq) = -val;
        }
    }
}

// calc_soln_impl, euler_march.cpp:193-303
template <int ND>
void soln_impl(Euler & ec)
{
    constexpr int NEQ = ND + 2;
    Mesh const & msh = *ec.mesh;
    double const dt = ec.dt;
    double const qdt = dt * 0.25;
    double const hdt = dt * 0.5;
#pragma omp parallel for num_threads(ec.nthreads) schedule(static)
    for (int64_t icl64 = 0; icl64 < ec.ncell; ++icl64)
    {
        int32_t const icl = static_cast<int32_t>(icl64);
        typename JacOf<ND>::type jaco;
        int32_t const clnfc = msh.clfcs(icl, 0);
        double acc[NEQ] = {};
        for (int32_t ifl = 1; ifl <= clnfc; ++ifl)
        {
            int32_t const ifc = msh.clfcs(icl, ifl);
            int32_t const jcl = msh.fcrcl(ifc, icl);
            double jcecnd[ND];
            for (int d = 0; d < ND; ++d)
            {
                jcecnd[d] = (jcl >= 0) ? ec.cecnd(jcl, d) : msh.clcnd(jcl, d);
            }
            double const bvol = ec.cevol(icl, ifl);
            double bcnd[ND];
            for (int d = 0; d < ND; ++d)
            {
                bcnd[d] = ec.cecnd(icl, ifl * ND + d);
            }
            for (int ieq = 0; ieq < NEQ; ++ieq)
            {
                double fusp = ec.so0c(jcl, ieq);
                for (int d = 0; d < ND; ++d)
                {
                    fusp += (bcnd[d] - jcecnd[d]) * ec.so1c(jcl, ieq, d);
                }
                acc[ieq] += fusp * bvol;
            }
            double jsol[NEQ];
            for (int ieq = 0; ieq < NEQ; ++ieq)
            {
                jsol[ieq] = ec.so0c(jcl, ieq);
            }
            jaco.update(ec.gamma(jcl), jsol);
            int32_t const fcnnd = msh.fcnds(ifc, 0);
            int32_t const sf_base = (ifl - 1) * FCMND;
            for (int32_t inf = 0; inf < fcnnd; ++inf)
            {
                int32_t const sfi = sf_base + inf;
                double usfc[NEQ];
                for (int ieq = 0; ieq < NEQ; ++ieq)
                {
                    usfc[ieq] = qdt * ec.so0t(jcl, ieq);
                    for (int d = 0; d < ND; ++d)
                    {
                        usfc[ieq] += (ec.sfcnd(icl, sfi, d) - jcecnd[d]) * ec.so1c(jcl, ieq, d);
                    }
                }
                for (int ieq = 0; ieq < NEQ; ++ieq)
                {
                    double dot = 0.0;
                    for (int d2 = 0; d2 < ND; ++d2)
                    {
                        double dfcn = jaco.fcn[ieq][d2];
                        for (int jeq = 0; jeq < NEQ; ++jeq)
                        {
                            dfcn += jaco.jacos[ieq][jeq][d2] * usfc[jeq];
                        }
                        dot += dfcn * ec.sfnml(icl, sfi, d2);
                    }
                    acc[ieq] -= hdt * dot;
                }
            }
        }
        double const cvol = ec.cevol(icl, 0);
        for (int ieq = 0; ieq < NEQ; ++ieq)
        {
            ec.so0n(icl, ieq) = acc[ieq] / cvol;
        }
    }
}

// calc_dsoln, euler_gradient.cpp:23-130
template <int ND>
void dsoln_impl(Euler & ec)
{
    constexpr int NEQ = ND + 2;
    constexpr double ALMOST_ZERO = 1.e-200;
    double const hdt = ec.dt * 0.5;
    Mesh const & msh = *ec.mesh;
#pragma omp parallel for num_threads(ec.nthreads) schedule(static)
    for (int64_t icl64 = 0; icl64 < ec.ncell; ++icl64)
    {
        int32_t const icl = static_cast<int32_t>(icl64);
        double const acfl = std::fabs(ec.cflc(icl));
        double const sgm0 = ec.sigma0 / acfl;
        double const tau = ec.taumin + acfl * ec.tauscale;
        GradElem const gelem(msh, ec.cecnd, icl, tau);
        int32_t const nfge = gelem.ge->nfge;
        double const ofg1 = gelem.ge->nfge_inverse;

        std::array<std::array<vec3, 5>, 8> grad = {};
        std::array<std::array<double, 5>, 8> widv = {};
        std::array<double, 5> wacc = {};
        std::array<double, 5> sigma_max = {};
        for (int32_t ifge = 0; ifge < nfge; ++ifge)
        {
            std::array<vec3, 5> udf = {};
            for (int ivx = 0; ivx < ND; ++ivx)
            {
                int32_t const ifl = gelem.ge->faces[ifge][ivx] - 1;
                int32_t const jcl = gelem.rcls[ifl];
                for (int ieq = 0; ieq < NEQ; ++ieq)
                {
                    double val = ec.so0c(jcl, ieq) + hdt * ec.so0t(jcl, ieq) - ec.so0n(icl, ieq);
                    for (int d = 0; d < ND; ++d)
                    {
                        val += gelem.jdis[ifl][d] * ec.so1c(jcl, ieq, d);
                    }
                    udf[ieq][ivx] = val;
                }
            }
            for (int ieq = 0; ieq < NEQ; ++ieq)
            {
                vec3 const g = gelem.solve_gradient(ifge, udf[ieq]);
                grad[ifge][ieq] = g;
                double sq = 0.0;
                for (int d = 0; d < ND; ++d)
                {
                    sq += g[d] * g[d];
                }
                double const wgt = 1.0 / std::sqrt(sq + ALMOST_ZERO);
                wacc[ieq] += wgt;
                widv[ifge][ieq] = wgt;
            }
        }
        std::array<std::array<double, 2>, 5> wpa = {};
        for (int32_t ifge = 0; ifge < nfge; ++ifge)
        {
            for (int ieq = 0; ieq < NEQ; ++ieq)
            {
                double const wgt = widv[ifge][ieq] / wacc[ieq] - ofg1;
                widv[ifge][ieq] = wgt;
                wpa[ieq][0] = std::fmax(wpa[ieq][0], wgt);
                wpa[ieq][1] = std::fmin(wpa[ieq][1], wgt);
            }
        }
        for (int ieq = 0; ieq < NEQ; ++ieq)
        {
            double const sm = std::fmin((1.0 - ofg1) / (wpa[ieq][0] + ALMOST_ZERO),
                                        -ofg1 / (wpa[ieq][1] - ALMOST_ZERO));
            sigma_max[ieq] = std::fmin(sm, sgm0);
        }
        for (int ieq = 0; ieq < NEQ; ++ieq)
        {
            for (int d = 0; d < ND; ++d)
            {
                ec.so1n(icl, ieq, d) = 0.0;
            }
        }
        for (int32_t isub = 0; isub < nfge; ++isub)
        {
            for (int ieq = 0; ieq < NEQ; ++ieq)
            {
                double const wgt = ofg1 + sigma_max[ieq] * widv[isub][ieq];
                for (int d = 0; d < ND; ++d)
                {
                    ec.so1n(icl, ieq, d) += wgt * grad[isub][ieq][d];
                }
            }
        }
    }
}

// normal_matrix<2>, normal_matrix<3>, euler_bc.cpp:87-128
mat3 normal_matrix_of(int nd, vec3 const & n)
{
    mat3 r = {};
    if (nd == 2)
    {
        r[0][0] = n[0];
        r[0][1] = n[1];
        r[1][0] = -n[1];
        r[1][1] = n[0];
        return r;
    }
    vec3 axis = {0.0, 0.0, 0.0};
    double const ax = std::fabs(n[0]);
    double const ay = std::fabs(n[1]);
    double const az = std::fabs(n[2]);
    if (ax <= ay && ax <= az)
    {
        axis = {1.0, 0.0, 0.0};
    }
    else if (ay <= az)
    {
        axis = {0.0, 1.0, 0.0};
    }
    else
    {
        axis = {0.0, 0.0, 1.0};
    }
    vec3 t1 = {n[1] * axis[2] - n[2] * axis[1], n[2] * axis[0] - n[0] * axis[2], n[0] * axis[1] - n[1] * axis[0]};
    double const len = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    t1 = {t1[0] / len, t1[1] / len, t1[2] / len};
    vec3 const t2 = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
    r[0] = n;
    r[1] = t1;
    r[2] = t2;
    return r;
}

// bc_matvec, bc_transpose, bc_matmat (euler_bc.cpp:36-81)
vec3 matvec(int nd, mat3 const & m, vec3 const & v)
{
    vec3 r = {0, 0, 0};
    for (int i = 0; i < nd; ++i)
    {
        for (int j = 0; j < nd; ++j)
        {
            r[i] += m[i][j] * v[j];
        }
    }
    return r;
}

mat3 transpose(int nd, mat3 const & m)
{
    mat3 r = {};
    for (int i = 0; i < nd; ++i)
    {
        for (int j = 0; j < nd; ++j)
        {
            r[i][j] = m[j][i];
        }
    }
    return r;
}

mat3 matmat(int nd, mat3 const & a, mat3 const & b)
{
    mat3 r = {};
    for (int i = 0; i < nd; ++i)
    {
        for (int j = 0; j < nd; ++j)
        {
            for (int k = 0; k < nd; ++k)
            {
                r[i][j] += a[i][k] * b[k][j];
            }
        }
    }
    return r;
}

} /* end namespace */

Euler::Euler(std::shared_ptr<Mesh> m, double time_increment)
    : mesh(std::move(m))
    , dt(time_increment)
{
    // euler_ce.cpp:15-60, euler_solution.cpp:19-41
    ndim = mesh->ndim;
    ncell = mesh->ncell;
    ngstcell = mesh->ngstcell;
    neq = ndim + 2;
    int64_t const total = static_cast<int64_t>(ngstcell) + ncell;
    cevol = make2<double>(total, CLMFC + 1, 0.0, ngstcell);
    cecnd = make2<double>(total, (CLMFC + 1) * ndim, 0.0, ngstcell);
    sfcnd = make3<double>(ncell, CLMFC * FCMND, ndim, 0.0);
    sfnml = make3<double>(ncell, CLMFC * FCMND, ndim, 0.0);
    so0c = make2<double>(total, neq, 0.0, ngstcell);
    so0n = make2<double>(total, neq, 0.0, ngstcell);
    so0t = make2<double>(total, neq, 0.0, ngstcell);
    so1c = make3<double>(total, neq, ndim, 0.0, ngstcell);
    so1n = make3<double>(total, neq, ndim, 0.0, ngstcell);
    stm = make2<double>(total, neq, 0.0, ngstcell);
    cflo = make1<double>(total, 0.0, ngstcell);
    cflc = make1<double>(total, 0.0, ngstcell);
    gamma = make1<double>(total, 0.0, ngstcell);
    prepare_ce();
}

// euler_ce.cpp:62-347
void Euler::prepare_ce()
{
    Mesh const & msh = *mesh;
    if (ndim == 2)
    {
        for (int32_t icl = 0; icl < ncell; ++icl)
        {
            int32_t const clnfc = msh.clfcs(icl, 0);
            double volc = 0.0;
            double cecnd0 = 0.0;
            double cecnd1 = 0.0;
            double const crdi0 = msh.clcnd(icl, 0);
            double const crdi1 = msh.clcnd(icl, 1);
            for (int32_t ifl = 1; ifl <= clnfc; ++ifl)
            {
                int32_t const ifc = msh.clfcs(icl, ifl);
                int32_t const jcl = msh.fcrcl(ifc, icl);
                double const crde0 = msh.clcnd(jcl, 0);
                double const crde1 = msh.clcnd(jcl, 1);
                int32_t const ind0 = msh.fcnds(ifc, 1);
                int32_t const ind1 = msh.fcnds(ifc, 2);
                double const crd10 = msh.ndcrd(ind0, 0);
                double const crd11 = msh.ndcrd(ind0, 1);
                double const crd20 = msh.ndcrd(ind1, 0);
                double const crd21 = msh.ndcrd(ind1, 1);
                double const cndi0 = (crd10 + crd20 + crdi0) / 3.0;
                double const cndi1 = (crd11 + crd21 + crdi1) / 3.0;
                double const voli = std::fabs((crd10 - crdi0) * (crd21 - crdi1) - (crd11 - crdi1) * (crd20 - crdi0)) / 2.0;
                double const cnde0 = (crd10 + crd20 + crde0) / 3.0;
                double const cnde1 = (crd11 + crd21 + crde1) / 3.0;
                double const vole = std::fabs((crd10 - crde0) * (crd21 - crde1) - (crd11 - crde1) * (crd20 - crde0)) / 2.0;
                double const volb = voli + vole;
                cevol(icl, ifl) = volb;
                cecnd(icl, ifl * 2 + 0) = (cndi0 * voli + cnde0 * vole) / volb;
                cecnd(icl, ifl * 2 + 1) = (cndi1 * voli + cnde1 * vole) / volb;
                volc += volb;
                cecnd0 += (cndi0 * voli + cnde0 * vole);
                cecnd1 += (cndi1 * voli + cnde1 * vole);
                int32_t const sf = (ifl - 1) * FCMND;
                double const cross = (crd10 - crde0) * (crd21 - crde1) - (crd11 - crde1) * (crd20 - crde0);
                double const sgn = (cross > 0.0) ? 1.0 : -1.0;
                sfcnd(icl, sf + 0, 0) = (crde0 + crd10) / 2.0;
                sfcnd(icl, sf + 0, 1) = (crde1 + crd11) / 2.0;
                double const s0dx = crde0 - crd10;
                double const s0dy = crde1 - crd11;
                sfnml(icl, sf + 0, 0) = -s0dy * sgn;
                sfnml(icl, sf + 0, 1) = s0dx * sgn;
                sfcnd(icl, sf + 1, 0) = (crd20 + crde0) / 2.0;
                sfcnd(icl, sf + 1, 1) = (crd21 + crde1) / 2.0;
                double const s1dx = crd20 - crde0;
                double const s1dy = crd21 - crde1;
                sfnml(icl, sf + 1, 0) = -s1dy * sgn;
                sfnml(icl, sf + 1, 1) = s1dx * sgn;
            }
            cevol(icl, 0) = volc;
            cecnd(icl, 0) = cecnd0 / volc;
            cecnd(icl, 1) = cecnd1 / volc;
        }
    }
    else if (ndim == 3)
    {
        for (int32_t icl = 0; icl < ncell; ++icl)
        {
            int32_t const clnfc = msh.clfcs(icl, 0);
            double volc = 0.0;
            double cc[3] = {0.0, 0.0, 0.0};
            double crdi[3];
            for (int d = 0; d < 3; ++d)
            {
                crdi[d] = msh.clcnd(icl, d);
            }
            for (int32_t ifl = 1; ifl <= clnfc; ++ifl)
            {
                int32_t const ifc = msh.clfcs(icl, ifl);
                int32_t const jcl = msh.fcrcl(ifc, icl);
                int32_t const fcnnd = msh.fcnds(ifc, 0);
                double crde[3];
                double pfc[3];
                for (int d = 0; d < 3; ++d)
                {
                    crde[d] = msh.clcnd(jcl, d);
                    pfc[d] = msh.fccnd(ifc, d);
                }
                std::array<std::array<double, 3>, FCMND + 2> crd{};
                for (int32_t inf = 1; inf <= fcnnd; ++inf)
                {
                    int32_t const ind = msh.fcnds(ifc, inf);
                    for (int d = 0; d < 3; ++d)
                    {
                        crd[inf][d] = msh.ndcrd(ind, d);
                    }
                }
                crd[fcnnd + 1] = crd[1];
                double volb = 0.0;
                double bc[3] = {0.0, 0.0, 0.0};
                double const voe = (msh.fccls(ifc, 0) == icl) ? 0.5 : -0.5;
                int32_t const sf = (ifl - 1) * FCMND;
                for (int32_t inf = 1; inf <= fcnnd; ++inf)
                {
                    double const u0 = crd[inf][0] - pfc[0];
                    double const u1 = crd[inf][1] - pfc[1];
                    double const u2 = crd[inf][2] - pfc[2];
                    double const v0 = crd[inf + 1][0] - pfc[0];
                    double const v1 = crd[inf + 1][1] - pfc[1];
                    double const v2 = crd[inf + 1][2] - pfc[2];
                    double const w0 = u1 * v2 - u2 * v1;
                    double const w1 = u2 * v0 - u0 * v2;
                    double const w2 = u0 * v1 - u1 * v0;
                    double const voli = std::fabs(w0 * (crdi[0] - pfc[0]) + w1 * (crdi[1] - pfc[1]) + w2 * (crdi[2] - pfc[2])) / 6.0;
                    double cndi[3];
                    for (int d = 0; d < 3; ++d)
                    {
                        cndi[d] = (crd[inf][d] + crd[inf + 1][d] + pfc[d] + crdi[d]) / 4.0;
                    }
                    double const vole = std::fabs(w0 * (crde[0] - pfc[0]) + w1 * (crde[1] - pfc[1]) + w2 * (crde[2] - pfc[2])) / 6.0;
                    double cnde[3];
                    for (int d = 0; d < 3; ++d)
                    {
                        cnde[d] = (crd[inf][d] + crd[inf + 1][d] + pfc[d] + crde[d]) / 4.0;
                    }
                    volb += voli + vole;
                    for (int d = 0; d < 3; ++d)
                    {
                        bc[d] += cndi[d] * voli + cnde[d] * vole;
                    }
                    int32_t const si = sf + (inf - 1);
                    for (int d = 0; d < 3; ++d)
                    {
                        sfcnd(icl, si, d) = (crde[d] + crd[inf][d] + crd[inf + 1][d]) / 3.0;
                    }
                    double const su0 = crd[inf][0] - crde[0];
                    double const su1 = crd[inf][1] - crde[1];
                    double const su2 = crd[inf][2] - crde[2];
                    double const sv0 = crd[inf + 1][0] - crde[0];
                    double const sv1 = crd[inf + 1][1] - crde[1];
                    double const sv2 = crd[inf + 1][2] - crde[2];
                    sfnml(icl, si, 0) = (su1 * sv2 - su2 * sv1) * voe;
                    sfnml(icl, si, 1) = (su2 * sv0 - su0 * sv2) * voe;
                    sfnml(icl, si, 2) = (su0 * sv1 - su1 * sv0) * voe;
                }
                cevol(icl, ifl) = volb;
                for (int d = 0; d < 3; ++d)
                {
                    cecnd(icl, ifl * 3 + d) = bc[d] / volb;
                }
                volc += volb;
                for (int d = 0; d < 3; ++d)
                {
                    cc[d] += bc[d];
                }
            }
            cevol(icl, 0) = volc;
            for (int d = 0; d < 3; ++d)
            {
                cecnd(icl, d) = cc[d] / volc;
            }
        }
    }
    else
    {
        set_error(ERR_VALUE, "EulerCore: ndim must be 2 or 3");
        throw std::invalid_argument("prepare_ce");
    }
}

// euler_solution.cpp:54-99
void Euler::init_solution(double ga, double rho, std::array<double, 3> const & v, double p)
{
    if (ga <= 1.0)
    {
        set_error(ERR_VALUE, "EulerCore::init_solution: gamma must be > 1");
        throw std::invalid_argument("init_solution");
    }
    if (rho <= 0.0)
    {
        set_error(ERR_VALUE, "EulerCore::init_solution: rho must be > 0");
        throw std::invalid_argument("init_solution");
    }
    if (p < 0.0)
    {
        set_error(ERR_VALUE, "EulerCore::init_solution: pressure must be >= 0");
        throw std::invalid_argument("init_solution");
    }
    double vsq = 0.0;
    for (int d = 0; d < ndim; ++d)
    {
        vsq += v[d] * v[d];
    }
    double const energy = p / (ga - 1.0) + 0.5 * rho * vsq;
    std::fill(gamma.data.begin(), gamma.data.end(), ga);
    for (int32_t icl = 0; icl < ncell; ++icl)
    {
        so0n(icl, 0) = rho;
        for (int d = 0; d < ndim; ++d)
        {
            so0n(icl, 1 + d) = rho * v[d];
        }
        so0n(icl, neq - 1) = energy;
    }
}

// euler.hpp:123-127
void Euler::update()
{
    std::swap(so0c, so0n);
    std::swap(so1c, so1n);
}

void Euler::calc_solt()
{
    if (ndim == 2)
    {
        solt_impl<2>(*this);
    }
    else
    {
        solt_impl<3>(*this);
    }
}

void Euler::calc_soln()
{
    if (ndim == 2)
    {
        soln_impl<2>(*this);
    }
    else
    {
        soln_impl<3>(*this);
    }
}

void Euler::calc_dsoln()
{
    if (ndim == 2)
    {
        dsoln_impl<2>(*this);
    }
    else
    {
        dsoln_impl<3>(*this);
    }
}

// euler_solution.cpp:106-160
void Euler::calc_cfl()
{
    constexpr double TINY = 1.e-60;
    double const hdt = dt / 2.0;
    Mesh const & msh = *mesh;
    int const nd = ndim;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i64 = 0; i64 < ncell; ++i64)
    {
        int32_t const icl = static_cast<int32_t>(i64);
        int32_t const clnfc = msh.clfcs(icl, 0);
        double dist = std::numeric_limits<double>::max();
        for (int32_t ifl = 1; ifl <= clnfc; ++ifl)
        {
            double d2 = 0.0;
            for (int d = 0; d < nd; ++d)
            {
                double const diff = cecnd(icl, ifl * nd + d) - cecnd(icl, d);
                d2 += diff * diff;
            }
            dist = std::min(dist, std::sqrt(d2));
        }
        double const ga = gamma(icl);
        double const ga1 = ga - 1.0;
        double const density = so0n(icl, 0);
        double momsq = 0.0;
        for (int d = 0; d < nd; ++d)
        {
            double const mom = so0n(icl, 1 + d);
            momsq += mom * mom;
        }
        double const ke = momsq / (2.0 * density);
        double const energy = so0n(icl, neq - 1);
        double const pr = ga1 * (energy - ke);
        double const pr_adj = (pr + std::fabs(pr)) / 2.0;
        double const wspd = std::sqrt(ga * pr_adj / density) + std::sqrt(momsq) / density;
        double const co = hdt * wspd / dist;
        cflo(icl) = co;
        cflc(icl) = (co - 1.0) * pr_adj / (pr_adj + TINY) + 1.0;
        so0n(icl, neq - 1) = pr_adj / ga1 + ke + TINY;
    }
}

mat3 Euler::normal_matrix(int32_t ifc) const
{
    vec3 n = {0, 0, 0};
    for (int d = 0; d < ndim; ++d)
    {
        n[d] = mesh->fcnml(ifc, d);
    }
    return normal_matrix_of(ndim, n);
}

// bc_soln_impl, euler_bc.cpp:141-215
void Euler::bc_soln()
{
    Mesh const & msh = *mesh;
    int const nd = ndim;
    for (EulerBoundary const & bnd : boundaries)
    {
        for (int32_t ifc : bnd.faces)
        {
            int32_t const icl = msh.fccls(ifc, 0);
            int32_t const jcl = msh.fccls(ifc, 1);
            if (bnd.kind == BC_NONREFL)
            {
                for (int ieq = 0; ieq < neq; ++ieq)
                {
                    so0n(jcl, ieq) = so0n(icl, ieq);
                }
            }
            else if (bnd.kind == BC_SLIPWALL)
            {
                mat3 const mat = normal_matrix(ifc);
                mat3 const inv = transpose(nd, mat);
                so0n(jcl, 0) = so0n(icl, 0);
                so0n(jcl, neq - 1) = so0n(icl, neq - 1);
                vec3 mom = {0, 0, 0};
                for (int d = 0; d < nd; ++d)
                {
                    mom[d] = so0n(icl, 1 + d);
                }
                vec3 rot = matvec(nd, mat, mom);
                rot[0] = -rot[0];
                vec3 const out = matvec(nd, inv, rot);
                for (int d = 0; d < nd; ++d)
                {
                    so0n(jcl, 1 + d) = out[d];
                }
            }
            else
            {
                double const rho = bnd.value[0];
                double const p = bnd.value[1 + nd];
                double const ga = bnd.value[2 + nd];
                double vsq = 0.0;
                for (int d = 0; d < nd; ++d)
                {
                    vsq += bnd.value[1 + d] * bnd.value[1 + d];
                }
                so0n(jcl, 0) = rho;
                for (int d = 0; d < nd; ++d)
                {
                    so0n(jcl, 1 + d) = rho * bnd.value[1 + d];
                }
                so0n(jcl, neq - 1) = p / (ga - 1.0) + 0.5 * rho * vsq;
            }
        }
    }
}

// bc_dsoln_impl, euler_bc.cpp:217-320
void Euler::bc_dsoln()
{
    Mesh const & msh = *mesh;
    int const nd = ndim;
    for (EulerBoundary const & bnd : boundaries)
    {
        for (int32_t ifc : bnd.faces)
        {
            int32_t const icl = msh.fccls(ifc, 0);
            int32_t const jcl = msh.fccls(ifc, 1);
            if (bnd.kind == BC_INLET)
            {
                for (int ieq = 0; ieq < neq; ++ieq)
                {
                    for (int d = 0; d < nd; ++d)
                    {
                        so1n(jcl, ieq, d) = 0.0;
                    }
                }
                continue;
            }
            mat3 const mat = normal_matrix(ifc);
            mat3 const inv = transpose(nd, mat);
            if (bnd.kind == BC_NONREFL)
            {
                for (int ieq = 0; ieq < neq; ++ieq)
                {
                    vec3 g = {0, 0, 0};
                    for (int d = 0; d < nd; ++d)
                    {
                        g[d] = so1c(icl, ieq, d);
                    }
                    vec3 rot = matvec(nd, mat, g);
                    rot[0] = 0.0;
                    vec3 const out = matvec(nd, inv, rot);
                    for (int d = 0; d < nd; ++d)
                    {
                        so1n(jcl, ieq, d) = out[d];
                    }
                }
                continue;
            }
            int const scal[2] = {0, neq - 1};
            for (int ieq : scal)
            {
                vec3 g = {0, 0, 0};
                for (int d = 0; d < nd; ++d)
                {
                    g[d] = so1c(icl, ieq, d);
                }
                vec3 rot = matvec(nd, mat, g);
                rot[0] = -rot[0];
                vec3 const out = matvec(nd, inv, rot);
                for (int d = 0; d < nd; ++d)
                {
                    so1n(jcl, ieq, d) = out[d];
                }
            }
            mat3 mg = {};
            for (int a = 0; a < nd; ++a)
            {
                for (int b = 0; b < nd; ++b)
                {
                    mg[a][b] = so1c(icl, 1 + a, b);
                }
            }
            mat3 uv = matmat(nd, matmat(nd, mat, mg), inv);
            for (int it = 1; it < nd; ++it)
            {
                uv[0][it] = -uv[0][it];
                uv[it][0] = -uv[it][0];
            }
            mat3 const out = matmat(nd, matmat(nd, inv, uv), mat);
            for (int a = 0; a < nd; ++a)
            {
                for (int b = 0; b < nd; ++b)
                {
                    so1n(jcl, 1 + a, b) = out[a][b];
                }
            }
        }
    }
}

// euler_march.cpp:339-359
void Euler::march_substep()
{
    update();
    calc_solt();
    calc_soln();
    bc_soln();
    calc_cfl();
    calc_dsoln();
    bc_dsoln();
}

void Euler::march(int64_t steps)
{
    for (int64_t istep = 0; istep < steps; ++istep)
    {
        for (int isub = 0; isub < 2; ++isub)
        {
            march_substep();
        }
    }
}

// euler_bc.cpp:324-391
void Euler::add_bc(int32_t kind, std::vector<int32_t> const & faces, std::vector<double> const & value)
{
    int32_t const nface = mesh->nface;
    for (int32_t ifc : faces)
    {
        if (ifc < 0 || ifc >= nface)
        {
            set_error(ERR_VALUE, "EulerCore::add_bc: face out of range");
            throw std::invalid_argument("add_bc");
        }
        if (mesh->fccls(ifc, 1) >= 0)
        {
            set_error(ERR_VALUE, "EulerCore::add_bc: face is not a boundary face");
            throw std::invalid_argument("add_bc");
        }
    }
    if (kind == BC_INLET)
    {
        if (value.size() != static_cast<size_t>(ndim) + 3)
        {
            set_error(ERR_VALUE, "EulerCore::add_bc: inlet value size must equal ndim+3");
            throw std::invalid_argument("add_bc");
        }
        double const rho = value[0];
        double const p = value[1 + ndim];
        double const ga = value[2 + ndim];
        if (ga <= 1.0 || rho <= 0.0 || p < 0.0)
        {
            set_error(ERR_VALUE, "EulerCore::add_bc: inlet state out of bounds");
            throw std::invalid_argument("add_bc");
        }
    }
    EulerBoundary b;
    b.kind = kind;
    b.faces = faces;
    b.value = value;
    boundaries.push_back(std::move(b));
}

} /* end namespace orc */
