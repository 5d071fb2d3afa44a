/*
 * GradientElement of one cell on the GPU, for the drop-in Python class of
 * the same name (cpp/solvcon/multidim/GradientElement.cpp:137-458 and the
 * binding cpp/solvcon/multidim/pymod/wrap_multidim.cpp:79-208): the
 * neighbour list, the displacements idis/jdis, and the per-FGE gradient
 * solve.  The march itself inlines the same arithmetic per cell type.
 */

#include "mesh.hpp"

#include <cstring>
#include <vector>

namespace scamd
{
namespace
{

struct GeTab
{
    int32_t clnfc;
    int32_t nfge;
    int32_t faces[8][3];
};

// GradientElement.cpp:16-79, indexed by cell type id.
__constant__ GeTab c_ge[9] = {
    {-1, -1, {}},
    {-1, -1, {}},
    {-1, -1, {}},
    {4, 4, {{1, 2, -1}, {2, 3, -1}, {3, 4, -1}, {4, 1, -1}}},
    {3, 3, {{1, 2, -1}, {2, 3, -1}, {3, 1, -1}}},
    {6, 8, {{2, 3, 5}, {6, 3, 2}, {4, 3, 6}, {5, 3, 4}, {5, 1, 2}, {2, 1, 6}, {6, 1, 4}, {4, 1, 5}}},
    {4, 4, {{3, 1, 2}, {2, 1, 4}, {4, 1, 3}, {2, 4, 3}}},
    {5, 6, {{5, 2, 4}, {3, 2, 5}, {4, 2, 3}, {4, 1, 5}, {5, 1, 3}, {3, 1, 4}}},
    {5, 6, {{1, 5, 2}, {2, 5, 3}, {3, 5, 4}, {4, 5, 1}, {1, 3, 4}, {3, 1, 2}}},
};

__device__ double det_n(double const (&a)[3][3], int nd)
{
    if (nd == 2)
    {
        return a[0][0] * a[1][1] - a[0][1] * a[1][0];
    }
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
         + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

// out: [0] ndim, [1] clnfc, [2] nfge, [3] valid, then rcl[6], then faces[24]
// (ints); dbl: [0] nfge_inverse, idis[18], jdis[18].
__global__ void k_ge_build(int nd, int32_t icl, double tau, int32_t const * cltpn_b, int32_t const * clfcs_b,
                           int32_t const * fccls_b, double const * clcnd_b, double const * cecnd_b, int ncol,
                           int32_t * iout, double * dout)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
    {
        return;
    }
    int32_t const tpn = cltpn_b[icl];
    int32_t const * cf = clfcs_b + static_cast<int64_t>(icl) * (CLMFC + 1);
    int32_t const clnfc = cf[0];
    GeTab const & ge = c_ge[(tpn >= 0 && tpn <= 8) ? tpn : 0];
    iout[0] = nd;
    iout[1] = clnfc;
    iout[2] = ge.nfge;
    iout[3] = (ge.nfge > 0 && clnfc > 0 && clnfc <= CLMFC) ? 1 : 0;
    if (!iout[3])
    {
        return;
    }
    double ice[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d)
    {
        ice[d] = cecnd_b[static_cast<int64_t>(icl) * ncol + d];
    }
    double gp[CLMFC][3] = {};
    double jd[CLMFC][3] = {};
    for (int32_t ifl = 0; ifl < clnfc; ++ifl)
    {
        int32_t const ifc = cf[ifl + 1];
        int32_t const jcl = fccls_b[static_cast<int64_t>(ifc) * FCREL] + fccls_b[static_cast<int64_t>(ifc) * FCREL + 1] - icl;
        iout[4 + ifl] = jcl;
        for (int d = 0; d < nd; ++d)
        {
            double const mid = cecnd_b[static_cast<int64_t>(icl) * ncol + (ifl + 1) * nd + d];
            double const jce = (jcl >= 0) ? cecnd_b[static_cast<int64_t>(jcl) * ncol + d]
                                          : clcnd_b[static_cast<int64_t>(jcl) * nd + d];
            gp[ifl][d] = mid + tau * (jce - mid);
            jd[ifl][d] = gp[ifl][d] - jce;
        }
    }
    double avg[3] = {0, 0, 0};
    for (int32_t ifl = 0; ifl < clnfc; ++ifl)
    {
        for (int d = 0; d < nd; ++d)
        {
            avg[d] += gp[ifl][d];
        }
    }
    for (int d = 0; d < nd; ++d)
    {
        avg[d] /= clnfc;
    }
    double cnd[3] = {0, 0, 0};
    double voc = 0.0;
    for (int32_t e = 0; e < ge.nfge; ++e)
    {
        double sub[3] = {avg[0], avg[1], avg[2]};
        double dst[3][3] = {};
        for (int v = 0; v < nd; ++v)
        {
            int32_t const fl = ge.faces[e][v] - 1;
            for (int d = 0; d < nd; ++d)
            {
                sub[d] += gp[fl][d];
                dst[v][d] = gp[fl][d] - avg[d];
            }
        }
        for (int d = 0; d < nd; ++d)
        {
            sub[d] /= static_cast<double>(nd + 1);
        }
        double const vob = det_n(dst, nd);
        voc += vob;
        for (int d = 0; d < nd; ++d)
        {
            cnd[d] += sub[d] * vob;
        }
    }
    for (int d = 0; d < nd; ++d)
    {
        cnd[d] /= voc;
    }
    dout[0] = 1.0 / ge.nfge;
    for (int32_t ifl = 0; ifl < clnfc; ++ifl)
    {
        for (int d = 0; d < nd; ++d)
        {
            dout[1 + ifl * 3 + d] = gp[ifl][d] - cnd[d];
            dout[19 + ifl * 3 + d] = jd[ifl][d] + ice[d] - cnd[d];
        }
    }
    for (int e = 0; e < 8; ++e)
    {
        for (int v = 0; v < 3; ++v)
        {
            iout[10 + e * 3 + v] = ge.faces[e][v];
        }
    }
}

// solve_gradient (GradientElement.cpp:419-458) for a row-major dst[3][3].
__global__ void k_ge_solve(int nd, double const * dst9, double const * udf3, double * out3)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
    {
        return;
    }
    double a[3][3];
    for (int i = 0; i < 3; ++i)
    {
        for (int j = 0; j < 3; ++j)
        {
            a[i][j] = dst9[i * 3 + j];
        }
    }
    double r[3][3] = {};
    if (nd == 2)
    {
        r[0][0] = a[1][1];
        r[0][1] = -a[0][1];
        r[1][0] = -a[1][0];
        r[1][1] = a[0][0];
    }
    else
    {
        r[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
        r[0][1] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
        r[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
        r[1][0] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
        r[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
        r[1][2] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
        r[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
        r[2][1] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
        r[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    }
    double adjudf[3] = {0, 0, 0};
    for (int i = 0; i < nd; ++i)
    {
        for (int j = 0; j < nd; ++j)
        {
            adjudf[i] += r[i][j] * udf3[j];
        }
    }
    double const det = det_n(a, nd);
    for (int d = 0; d < 3; ++d)
    {
        out3[d] = d < nd ? adjudf[d] / det : 0.0;
    }
}

} /* end namespace */
} /* end namespace scamd */

using scamd::DMesh;
using scamd::guard;

extern "C"
{

// Build the GradientElement of cell icl over a host cecnd table
// [nrow][ncol] (ghost rows first).  iout[34]: ndim, clnfc, nfge, valid,
// rcl[6], faces[8][3]; dout[37]: nfge_inverse, idis[6][3], jdis[6][3].
int scge_build(void * mesh, double const * cecnd, int64_t nrow, int64_t ncol, int64_t nghost, int32_t icl, double tau,
               int32_t * iout, double * dout)
{
    return guard(
        [&]()
        {
            DMesh & m = *static_cast<DMesh *>(mesh);
            if (icl < 0 || icl >= m.ncell)
            {
                scamd::fail(scamd::SC_EINDEX, "GradientElement: icl out of range");
            }
            if (m.ndim != 2 && m.ndim != 3)
            {
                scamd::fail(scamd::SC_EVALUE, "GradientElement: ndim must be 2 or 3");
            }
            if (nghost != m.ngstcell || nrow != m.ngstcell + m.ncell || ncol != (scamd::CLMFC + 1) * m.ndim)
            {
                scamd::fail(scamd::SC_EVALUE, "GradientElement: cecnd shape does not match the mesh");
            }
            scamd::DevBuf<double> dce, dd;
            scamd::DevBuf<int32_t> di;
            dce.alloc(nrow * ncol);
            dd.alloc(37);
            di.alloc(34);
            SC_HIP(hipMemsetAsync(dd.ptr, 0, dd.bytes(), m.stream));
            SC_HIP(hipMemsetAsync(di.ptr, 0, di.bytes(), m.stream));
            SC_HIP(hipMemcpyAsync(dce.ptr, cecnd, dce.bytes(), hipMemcpyHostToDevice, m.stream));
            scamd::k_ge_build<<<1, 64, 0, m.stream>>>(m.ndim, icl, tau, m.cltpn_b(), m.clfcs_b(), m.fccls_b(),
                                                      m.clcnd_b(), dce.ptr + nghost * ncol, static_cast<int>(ncol),
                                                      di.ptr, dd.ptr);
            SC_HIP(hipGetLastError());
            SC_HIP(hipMemcpyAsync(iout, di.ptr, di.bytes(), hipMemcpyDeviceToHost, m.stream));
            SC_HIP(hipMemcpyAsync(dout, dd.ptr, dd.bytes(), hipMemcpyDeviceToHost, m.stream));
            m.sync();
        });
}

int scge_solve(void * mesh, int ndim, double const * dst9, double const * udf3, double * out3)
{
    return guard(
        [&]()
        {
            DMesh & m = *static_cast<DMesh *>(mesh);
            scamd::DevBuf<double> buf;
            buf.alloc(15);
            SC_HIP(hipMemcpyAsync(buf.ptr, dst9, 9 * sizeof(double), hipMemcpyHostToDevice, m.stream));
            SC_HIP(hipMemcpyAsync(buf.ptr + 9, udf3, 3 * sizeof(double), hipMemcpyHostToDevice, m.stream));
            scamd::k_ge_solve<<<1, 64, 0, m.stream>>>(ndim, buf.ptr, buf.ptr + 9, buf.ptr + 12);
            SC_HIP(hipGetLastError());
            SC_HIP(hipMemcpyAsync(out3, buf.ptr + 12, 3 * sizeof(double), hipMemcpyDeviceToHost, m.stream));
            m.sync();
        });
}

} /* end extern "C" */
