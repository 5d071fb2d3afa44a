/*
 * Oracle (test infrastructure only): C entry points over the CPU restatement
 * so that Python tests can drive it through ctypes.  Tables are exposed as
 * raw host pointers plus shape, with ghost rows stored first.
 */

#include "euler.hpp"

#include <cstring>
#include <memory>
#include <string>

namespace orc
{
namespace
{
thread_local int g_err_kind = 0;
thread_local std::string g_err_msg;
} /* end namespace */

void set_error(int kind, char const * msg)
{
    g_err_kind = kind;
    g_err_msg = msg;
}

} /* end namespace orc */

using orc::Euler;
using orc::Mesh;

namespace
{

struct MeshHandle
{
    std::shared_ptr<Mesh> mesh;
};

template <typename F>
int guarded(F && f)
{
    try
    {
        f();
        return 0;
    }
    catch (std::exception const & e)
    {
        if (orc::g_err_kind == 0)
        {
            orc::set_error(orc::ERR_RUNTIME, e.what());
        }
        return orc::g_err_kind;
    }
}

template <typename T>
int expose(orc::Table<T> & t, void ** ptr, int64_t * shape, int * rank, int64_t * nghost)
{
    *ptr = t.data.data();
    *rank = static_cast<int>(t.shape.size());
    for (size_t i = 0; i < t.shape.size(); ++i)
    {
        shape[i] = t.shape[i];
    }
    *nghost = t.nghost;
    return 0;
}

} /* end namespace */

extern "C"
{

int orc_last_error(char * buf, int n)
{
    if (buf != nullptr && n > 0)
    {
        std::strncpy(buf, orc::g_err_msg.c_str(), static_cast<size_t>(n) - 1);
        buf[n - 1] = '\0';
    }
    int const k = orc::g_err_kind;
    orc::g_err_kind = 0;
    return k;
}

void * orcmesh_create(int ndim, int32_t nnode, int32_t nface, int32_t ncell)
{
    auto * h = new MeshHandle();
    h->mesh = std::make_shared<Mesh>(ndim, nnode, nface, ncell);
    return h;
}

void orcmesh_destroy(void * h) { delete static_cast<MeshHandle *>(h); }

// ndim nnode nface ncell nbound ngstnode ngstface ngstcell nedge nbcs
void orcmesh_counts(void * h, int64_t * out)
{
    Mesh const & m = *static_cast<MeshHandle *>(h)->mesh;
    out[0] = m.ndim;
    out[1] = m.nnode;
    out[2] = m.nface;
    out[3] = m.ncell;
    out[4] = m.nbound;
    out[5] = m.ngstnode;
    out[6] = m.ngstface;
    out[7] = m.ngstcell;
    out[8] = m.ednds.shape[0];
    out[9] = static_cast<int64_t>(m.bcs.size());
}

// kind: 0 double, 1 int32
int orcmesh_array(void * h, char const * name, void ** ptr, int64_t * shape, int * rank, int64_t * nghost, int * kind)
{
    Mesh & m = *static_cast<MeshHandle *>(h)->mesh;
    std::string const s(name);
    *kind = 0;
#define ORC_D(N) \
    if (s == #N) \
        return expose(m.N, ptr, shape, rank, nghost);
    ORC_D(ndcrd)
    ORC_D(fccnd)
    ORC_D(fcnml)
    ORC_D(fcara)
    ORC_D(clcnd)
    ORC_D(clvol)
#undef ORC_D
    *kind = 1;
#define ORC_I(N) \
    if (s == #N) \
        return expose(m.N, ptr, shape, rank, nghost);
    ORC_I(fctpn)
    ORC_I(cltpn)
    ORC_I(clgrp)
    ORC_I(fcnds)
    ORC_I(fccls)
    ORC_I(clnds)
    ORC_I(clfcs)
    ORC_I(ednds)
    ORC_I(bndfcs)
#undef ORC_I
    orc::set_error(orc::ERR_VALUE, "unknown mesh array");
    return orc::ERR_VALUE;
}

int orcmesh_build_interior(void * h, int do_metric, int do_edge)
{
    return guarded([&]() { static_cast<MeshHandle *>(h)->mesh->build_interior(do_metric != 0, do_edge != 0); });
}

int orcmesh_build_boundary(void * h)
{
    return guarded([&]() { static_cast<MeshHandle *>(h)->mesh->build_boundary(); });
}

int orcmesh_build_ghost(void * h)
{
    return guarded([&]() { static_cast<MeshHandle *>(h)->mesh->build_ghost(); });
}

int orcmesh_add_bc(void * h, char const * name, int32_t const * faces, int64_t n)
{
    std::vector<int32_t> f(faces, faces + n);
    return guarded([&]() { static_cast<MeshHandle *>(h)->mesh->add_bc(name, f); });
}

// Copies BC ibc's name and its facn columns 0 and 1; returns the face count.
int64_t orcmesh_bc(void * h, int ibc, char * name, int nname, int32_t * facn0, int32_t * facn1)
{
    Mesh const & m = *static_cast<MeshHandle *>(h)->mesh;
    orc::MeshBc const & b = m.bcs.at(static_cast<size_t>(ibc));
    if (name != nullptr && nname > 0)
    {
        std::strncpy(name, b.name.c_str(), static_cast<size_t>(nname) - 1);
        name[nname - 1] = '\0';
    }
    for (size_t i = 0; i < b.facn0.size(); ++i)
    {
        if (facn0 != nullptr)
        {
            facn0[i] = b.facn0[i];
        }
        if (facn1 != nullptr)
        {
            facn1[i] = b.facn1[i];
        }
    }
    return static_cast<int64_t>(b.facn0.size());
}

void * orceuler_create(void * mh, double dt)
{
    try
    {
        return new Euler(static_cast<MeshHandle *>(mh)->mesh, dt);
    }
    catch (std::exception const &)
    {
        return nullptr;
    }
}

void orceuler_destroy(void * h) { delete static_cast<Euler *>(h); }

void orceuler_set_threads(void * h, int n) { static_cast<Euler *>(h)->nthreads = n < 1 ? 1 : n; }

// ndim ncell ngstcell neq
void orceuler_counts(void * h, int64_t * out)
{
    Euler const & e = *static_cast<Euler *>(h);
    out[0] = e.ndim;
    out[1] = e.ncell;
    out[2] = e.ngstcell;
    out[3] = e.neq;
}

// which: 0 sigma0, 1 taumin, 2 tauscale, 3 time_increment (read only)
double orceuler_get_param(void * h, int which)
{
    Euler const & e = *static_cast<Euler *>(h);
    switch (which)
    {
    case 0: return e.sigma0;
    case 1: return e.taumin;
    case 2: return e.tauscale;
    default: return e.dt;
    }
}

void orceuler_set_param(void * h, int which, double v)
{
    Euler & e = *static_cast<Euler *>(h);
    switch (which)
    {
    case 0: e.sigma0 = v; break;
    case 1: e.taumin = v; break;
    case 2: e.tauscale = v; break;
    default: break;
    }
}

int orceuler_array(void * h, char const * name, void ** ptr, int64_t * shape, int * rank, int64_t * nghost)
{
    Euler & e = *static_cast<Euler *>(h);
    std::string const s(name);
#define ORC_E(N) \
    if (s == #N) \
        return expose(e.N, ptr, shape, rank, nghost);
    ORC_E(cevol)
    ORC_E(cecnd)
    ORC_E(sfcnd)
    ORC_E(sfnml)
    ORC_E(so0c)
    ORC_E(so0n)
    ORC_E(so0t)
    ORC_E(so1c)
    ORC_E(so1n)
    ORC_E(stm)
    ORC_E(cflo)
    ORC_E(cflc)
    ORC_E(gamma)
#undef ORC_E
    orc::set_error(orc::ERR_VALUE, "unknown euler array");
    return orc::ERR_VALUE;
}

int orceuler_call(void * h, char const * what)
{
    Euler & e = *static_cast<Euler *>(h);
    std::string const s(what);
    return guarded(
        [&]()
        {
            if (s == "prepare_ce")
                e.prepare_ce();
            else if (s == "update")
                e.update();
            else if (s == "calc_solt")
                e.calc_solt();
            else if (s == "calc_soln")
                e.calc_soln();
            else if (s == "calc_cfl")
                e.calc_cfl();
            else if (s == "calc_dsoln")
                e.calc_dsoln();
            else if (s == "bc_soln")
                e.bc_soln();
            else if (s == "bc_dsoln")
                e.bc_dsoln();
            else if (s == "march_substep")
                e.march_substep();
            else if (s == "clear_bc")
                e.boundaries.clear();
            else
            {
                orc::set_error(orc::ERR_VALUE, "unknown euler call");
                throw std::invalid_argument(s);
            }
        });
}

int orceuler_march(void * h, int64_t steps)
{
    return guarded([&]() { static_cast<Euler *>(h)->march(steps); });
}

int orceuler_init_solution(void * h, double gamma, double rho, double const * v, double p)
{
    std::array<double, 3> vel = {v[0], v[1], v[2]};
    return guarded([&]() { static_cast<Euler *>(h)->init_solution(gamma, rho, vel, p); });
}

int orceuler_add_bc(void * h, int kind, int32_t const * faces, int64_t n, double const * value, int64_t nv)
{
    std::vector<int32_t> f(faces, faces + n);
    std::vector<double> v(value, value + nv);
    return guarded([&]() { static_cast<Euler *>(h)->add_bc(kind, f, v); });
}

void orceuler_normal_matrix(void * h, int32_t ifc, double * out)
{
    auto const m = static_cast<Euler *>(h)->normal_matrix(ifc);
    for (int i = 0; i < 3; ++i)
    {
        for (int j = 0; j < 3; ++j)
        {
            out[i * 3 + j] = m[i][j];
        }
    }
}

// GradientElement over a caller-supplied cecnd table [ngst + ncell, ncol].
// Outputs: info = {ndim, clnfc, nfge}, rcl[6], idis[6*3], jdis[6*3],
// faces[8*3], and nfge_inverse.
void orcge_build(void * mh, double const * cecnd, int64_t nrow, int64_t ncol, int64_t nghost,
                 int32_t icl, double tau, int32_t * info, int32_t * rcl, double * idis, double * jdis,
                 int32_t * faces, double * ofg1)
{
    Mesh const & m = *static_cast<MeshHandle *>(mh)->mesh;
    orc::Table<double> t = orc::make2<double>(nrow, ncol, 0.0, nghost);
    std::memcpy(t.data.data(), cecnd, sizeof(double) * static_cast<size_t>(nrow * ncol));
    orc::GradElem const g(m, t, icl, tau);
    info[0] = g.ndim;
    info[1] = g.clnfc;
    info[2] = g.ge->nfge;
    *ofg1 = g.ge->nfge_inverse;
    for (int i = 0; i < orc::CLMFC; ++i)
    {
        rcl[i] = g.rcls[i];
        for (int d = 0; d < 3; ++d)
        {
            idis[i * 3 + d] = g.idis[i][d];
            jdis[i * 3 + d] = g.jdis[i][d];
        }
    }
    for (int i = 0; i < 8; ++i)
    {
        for (int k = 0; k < 3; ++k)
        {
            faces[i * 3 + k] = g.ge->faces[i][k];
        }
    }
}

// solve_gradient: dst rows are idis of the FGE faces (GradientElement.cpp:447-458).
void orcge_solve(int ndim, double const * dst9, double const * udf3, double * out3)
{
    std::array<std::array<double, 3>, 3> a{};
    for (int i = 0; i < 3; ++i)
    {
        for (int j = 0; j < 3; ++j)
        {
            a[i][j] = dst9[i * 3 + j];
        }
    }
    auto const adj = orc::ge_adjugate(a, ndim);
    double r[3] = {0, 0, 0};
    for (int i = 0; i < ndim; ++i)
    {
        for (int j = 0; j < ndim; ++j)
        {
            r[i] += adj[i][j] * udf3[j];
        }
    }
    double const det = orc::ge_determinant(a, ndim);
    for (int d = 0; d < 3; ++d)
    {
        out3[d] = d < ndim ? r[d] / det : 0.0;
    }
}

double orcge_det(int ndim, double const * a9)
{
    std::array<std::array<double, 3>, 3> a{};
    for (int i = 0; i < 3; ++i)
    {
        for (int j = 0; j < 3; ++j)
        {
            a[i][j] = a9[i * 3 + j];
        }
    }
    return orc::ge_determinant(a, ndim);
}

void orcge_adj(int ndim, double const * a9, double * out9)
{
    std::array<std::array<double, 3>, 3> a{};
    for (int i = 0; i < 3; ++i)
    {
        for (int j = 0; j < 3; ++j)
        {
            a[i][j] = a9[i * 3 + j];
        }
    }
    auto const r = orc::ge_adjugate(a, ndim);
    for (int i = 0; i < 3; ++i)
    {
        for (int j = 0; j < 3; ++j)
        {
            out9[i * 3 + j] = r[i][j];
        }
    }
}

} /* end extern "C" */
