/*
 * GPU builders of the unstructured mesh, producing solvcon's StaticMesh
 * tables bit for bit in the reference numbering:
 *
 *  - build_faces_from_cells (StaticMesh_interior.cpp:16-564): faces are
 *    emitted per cell in (cell, local face) order; duplicates (same face
 *    type and node set) collapse onto their first occurrence, found by a
 *    minimum over the faces that share the face's first node (node-to-face
 *    table built with atomics and a prefix sum); surviving faces are
 *    renumbered by a prefix sum, and fccls takes the first two claimants in
 *    cell order.
 *  - calc_metric (:623-976): per-face centroid/normal/area, per-cell
 *    centroid, then the reorder_face flips, which the reference applies
 *    while sweeping cells in order.  Each face is visited by its two cells
 *    and the lower-numbered cell fccls[0] goes first, so the flip decisions
 *    are evaluated per face in that order; cell volumes use |vol|, which
 *    does not depend on the normal's sign.
 *  - build_edge (:566-610), build_boundary and build_ghost
 *    (StaticMesh_boundary.cpp:103-627): edge, boundary-face and ghost
 *    numbering by prefix sums in the reference's loop order.
 *
 * The arithmetic of every metric follows the reference in order (built with
 * -ffp-contract=off).  Node indices in clnds are validated on the device
 * before use.
 */

#include "mesh.hpp"
#include "scan.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

namespace scamd
{
namespace
{

constexpr unsigned BLK = 256;

struct FaceDef
{
    int32_t type;
    int32_t nnd;
    int32_t nd[4];
};

struct CellDef
{
    int32_t nface;
    int32_t nnode;
    FaceDef face[6];
};

// Face tables of FaceBuilder::add_* (StaticMesh_interior.cpp:41-308) with
// 0-based local node positions, indexed by cell type id.
__constant__ CellDef c_cells[9] = {
    {0, 0, {}}, // non-cell type
    {0, 1, {}}, // point (no faces built)
    {2, 2, {{CT_POINT, 1, {0}}, {CT_POINT, 1, {1}}}},
    {4, 4, {{CT_LINE, 2, {0, 1}}, {CT_LINE, 2, {1, 2}}, {CT_LINE, 2, {2, 3}}, {CT_LINE, 2, {3, 0}}}},
    {3, 3, {{CT_LINE, 2, {0, 1}}, {CT_LINE, 2, {1, 2}}, {CT_LINE, 2, {2, 0}}}},
    {6,
     8,
     {{CT_QUAD, 4, {0, 3, 2, 1}},
      {CT_QUAD, 4, {1, 2, 6, 5}},
      {CT_QUAD, 4, {4, 5, 6, 7}},
      {CT_QUAD, 4, {0, 4, 7, 3}},
      {CT_QUAD, 4, {0, 1, 5, 4}},
      {CT_QUAD, 4, {2, 3, 7, 6}}}},
    {4, 4, {{CT_TRI, 3, {0, 2, 1}}, {CT_TRI, 3, {0, 1, 3}}, {CT_TRI, 3, {0, 3, 2}}, {CT_TRI, 3, {1, 2, 3}}}},
    {5,
     6,
     {{CT_TRI, 3, {0, 1, 2}},
      {CT_TRI, 3, {3, 5, 4}},
      {CT_QUAD, 4, {0, 3, 4, 1}},
      {CT_QUAD, 4, {0, 2, 5, 3}},
      {CT_QUAD, 4, {1, 4, 5, 2}}}},
    {5,
     5,
     {{CT_TRI, 3, {0, 4, 3}},
      {CT_TRI, 3, {1, 4, 0}},
      {CT_TRI, 3, {2, 4, 1}},
      {CT_TRI, 3, {3, 4, 2}},
      {CT_QUAD, 4, {0, 3, 2, 1}}}},
};

__device__ inline int32_t cell_nface(int32_t tpn) { return (tpn >= 0 && tpn <= 8) ? c_cells[tpn].nface : 0; }


template <typename T>
__global__ void k_fill(T * a, int64_t n, T v)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
    {
        a[i] = v;
    }
}

template <typename T>
void fill(T * a, int64_t n, T v, hipStream_t s)
{
    if (n > 0)
    {
        k_fill<T><<<grid_for(n, BLK), BLK, 0, s>>>(a, n, v);
        SC_HIP(hipGetLastError());
    }
}

// Count faces per cell and validate the node ids the face tables will read.
__global__ void k_count_faces(int32_t ncell, int32_t nnode, int32_t const * cltpn, int32_t const * clnds, int32_t * cnt,
                              int32_t * bad)
{
    int32_t const icl = blockIdx.x * blockDim.x + threadIdx.x;
    if (icl >= ncell)
    {
        return;
    }
    int32_t const tpn = cltpn[icl];
    int32_t const nf = cell_nface(tpn);
    cnt[icl] = nf;
    if (nf > 0)
    {
        int32_t const need = c_cells[tpn].nnode;
        for (int32_t k = 1; k <= need; ++k)
        {
            int32_t const ind = clnds[static_cast<int64_t>(icl) * (CLMND + 1) + k];
            if (ind < 0 || ind >= nnode)
            {
                atomicOr(bad, 1);
            }
        }
    }
    int32_t const nnd = clnds[static_cast<int64_t>(icl) * (CLMND + 1)];
    if (nnd < 0 || nnd > CLMND)
    {
        atomicOr(bad, 2);
        return;
    }
    for (int32_t k = 1; k <= nnd; ++k)
    {
        int32_t const ind = clnds[static_cast<int64_t>(icl) * (CLMND + 1) + k];
        if (ind < 0 || ind >= nnode)
        {
            atomicOr(bad, 2);
        }
    }
}

// FaceBuilder::populate (:351-386).
__global__ void k_populate(int32_t ncell, int32_t const * cltpn, int32_t const * clnds, int32_t const * foff,
                           int32_t * bfctpn, int32_t * bfcnds, int32_t * bclfcs, int32_t * fcell)
{
    int32_t const icl = blockIdx.x * blockDim.x + threadIdx.x;
    if (icl >= ncell)
    {
        return;
    }
    int32_t * cf = bclfcs + static_cast<int64_t>(icl) * (CLMFC + 1);
    for (int k = 0; k <= CLMFC; ++k)
    {
        cf[k] = -1;
    }
    int32_t const tpn = cltpn[icl];
    int32_t const nf = cell_nface(tpn);
    if (nf == 0)
    {
        return;
    }
    cf[0] = nf;
    int32_t const * nds = clnds + static_cast<int64_t>(icl) * (CLMND + 1);
    for (int32_t k = 0; k < nf; ++k)
    {
        FaceDef const & fd = c_cells[tpn].face[k];
        int32_t const f = foff[icl] + k;
        bfctpn[f] = fd.type;
        int32_t * fn = bfcnds + static_cast<int64_t>(f) * (FCMND + 1);
        fn[0] = fd.nnd;
        for (int m = 0; m < FCMND; ++m)
        {
            fn[m + 1] = m < fd.nnd ? nds[fd.nd[m] + 1] : -1;
        }
        cf[k + 1] = f;
        fcell[f] = icl;
    }
}

__global__ void k_node_count(int32_t mface, int32_t const * bfcnds, int32_t * cnt)
{
    int32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= mface)
    {
        return;
    }
    int32_t const * fn = bfcnds + static_cast<int64_t>(f) * (FCMND + 1);
    for (int m = 1; m <= fn[0]; ++m)
    {
        atomicAdd(&cnt[fn[m]], 1);
    }
}

__global__ void k_node_fill(int32_t mface, int32_t const * bfcnds, int32_t const * off, int32_t * cur, int32_t * lst)
{
    int32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= mface)
    {
        return;
    }
    int32_t const * fn = bfcnds + static_cast<int64_t>(f) * (FCMND + 1);
    for (int m = 1; m <= fn[0]; ++m)
    {
        int32_t const nd = fn[m];
        int32_t const pos = atomicAdd(&cur[nd], 1);
        lst[off[nd] + pos] = f;
    }
}

// make_dedupmap (:428-477): the representative of a face is the first face
// with the same type whose nodes are all among its nodes.
__global__ void k_dedup(int32_t mface, int32_t const * bfctpn, int32_t const * bfcnds, int32_t const * off,
                        int32_t const * cnt, int32_t const * lst, int32_t * dup)
{
    int32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= mface)
    {
        return;
    }
    int32_t const * fn = bfcnds + static_cast<int64_t>(f) * (FCMND + 1);
    int32_t const nnd = fn[0];
    int32_t const tp = bfctpn[f];
    int32_t best = f;
    int32_t const nd0 = fn[1];
    int32_t const b = off[nd0];
    int32_t const e = b + cnt[nd0];
    for (int32_t p = b; p < e; ++p)
    {
        int32_t const g = lst[p];
        if (g >= best || bfctpn[g] != tp)
        {
            continue;
        }
        int32_t const * gn = bfcnds + static_cast<int64_t>(g) * (FCMND + 1);
        if (gn[0] != nnd)
        {
            continue;
        }
        int match = 0;
        for (int j = 1; j <= gn[0]; ++j)
        {
            for (int i = 1; i <= nnd; ++i)
            {
                if (fn[i] == gn[j])
                {
                    ++match;
                    break;
                }
            }
        }
        if (match == gn[0])
        {
            best = g;
        }
    }
    dup[f] = best;
}

__global__ void k_canon(int32_t mface, int32_t const * dup, int32_t * flag)
{
    int32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < mface)
    {
        flag[f] = dup[f] == f ? 1 : 0;
    }
}

// remap_face (:479-523), first half: compaction and the first claimant.
__global__ void k_compact(int32_t mface, int32_t const * dup, int32_t const * newidx, int32_t const * bfctpn,
                          int32_t const * bfcnds, int32_t const * fcell, int32_t * fctpn, int32_t * fcnds,
                          int32_t * fccls, int32_t * second)
{
    int32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= mface)
    {
        return;
    }
    int32_t const d = dup[f];
    int32_t const j = newidx[d];
    if (d == f)
    {
        fctpn[j] = bfctpn[f];
        for (int m = 0; m <= FCMND; ++m)
        {
            fcnds[static_cast<int64_t>(j) * (FCMND + 1) + m] = bfcnds[static_cast<int64_t>(f) * (FCMND + 1) + m];
        }
        int32_t * fc = fccls + static_cast<int64_t>(j) * FCREL;
        fc[0] = fcell[f];
        fc[2] = -1;
        fc[3] = -1;
    }
    else
    {
        atomicMin(&second[j], f);
    }
}

// Second claimant in cell order: the next-lowest original face of the group.
__global__ void k_second(int32_t nface, int32_t const * second, int32_t const * fcell, int32_t * fccls)
{
    int32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nface)
    {
        int32_t const s = second[j];
        fccls[static_cast<int64_t>(j) * FCREL + 1] = s == INT_MAX ? -1 : fcell[s];
    }
}

__global__ void k_remap_clfcs(int32_t ncell, int32_t const * bclfcs, int32_t const * dup, int32_t const * newidx,
                              int32_t * clfcs)
{
    int32_t const icl = blockIdx.x * blockDim.x + threadIdx.x;
    if (icl >= ncell)
    {
        return;
    }
    int32_t const * src = bclfcs + static_cast<int64_t>(icl) * (CLMFC + 1);
    int32_t * dst = clfcs + static_cast<int64_t>(icl) * (CLMFC + 1);
    dst[0] = src[0];
    for (int k = 1; k <= CLMFC; ++k)
    {
        dst[k] = (src[0] > 0 && k <= src[0]) ? newidx[dup[src[k]]] : -1;
    }
}

// ------------------------------------------------------------- metrics ----

// Face centroid, normal and area (StaticMesh_interior.cpp:625-768,
// StaticMesh_boundary.cpp:374-503); f may be negative (ghost face).
__device__ void face_metric(int ndim, int32_t f, int32_t const * fcnds, double const * ndcrd, double * fccnd,
                            double * fcnml, double * fcara)
{
    int32_t const * fn = fcnds + static_cast<int64_t>(f) * (FCMND + 1);
    if (ndim == 2)
    {
        int32_t ind = fn[1];
        double c0 = ndcrd[static_cast<int64_t>(ind) * 2];
        double c1 = ndcrd[static_cast<int64_t>(ind) * 2 + 1];
        ind = fn[2];
        c0 += ndcrd[static_cast<int64_t>(ind) * 2];
        c1 += ndcrd[static_cast<int64_t>(ind) * 2 + 1];
        c0 /= 2;
        c1 /= 2;
        fccnd[static_cast<int64_t>(f) * 2] = c0;
        fccnd[static_cast<int64_t>(f) * 2 + 1] = c1;
        int32_t const i1 = fn[1];
        int32_t const i2 = fn[2];
        double n0 = ndcrd[static_cast<int64_t>(i2) * 2 + 1] - ndcrd[static_cast<int64_t>(i1) * 2 + 1];
        double n1 = ndcrd[static_cast<int64_t>(i1) * 2] - ndcrd[static_cast<int64_t>(i2) * 2];
        double const ar = sqrt(n0 * n0 + n1 * n1);
        n0 /= ar;
        n1 /= ar;
        fcnml[static_cast<int64_t>(f) * 2] = n0;
        fcnml[static_cast<int64_t>(f) * 2 + 1] = n1;
        fcara[f] = ar;
        return;
    }
    if (ndim != 3)
    {
        return;
    }
    int32_t const nnd = fn[0];
    double cfd[FCMND + 2][3];
    cfd[0][0] = cfd[0][1] = cfd[0][2] = 0.0;
    for (int32_t inf = 1; inf <= nnd; ++inf)
    {
        int32_t const ind = fn[inf];
        for (int d = 0; d < 3; ++d)
        {
            double const x = ndcrd[static_cast<int64_t>(ind) * 3 + d];
            cfd[inf][d] = x;
            cfd[0][d] += x;
        }
    }
    for (int d = 0; d < 3; ++d)
    {
        cfd[nnd + 1][d] = cfd[1][d];
    }
    for (int d = 0; d < 3; ++d)
    {
        cfd[0][d] /= nnd;
    }
    double voc = 0.0;
    double cc[3] = {0.0, 0.0, 0.0};
    for (int32_t inf = 1; inf <= nnd; ++inf)
    {
        double crd[3];
        for (int d = 0; d < 3; ++d)
        {
            crd[d] = (cfd[0][d] + cfd[inf][d] + cfd[inf + 1][d]) / 3;
        }
        double const du0 = cfd[inf][0] - cfd[0][0];
        double const du1 = cfd[inf][1] - cfd[0][1];
        double const du2 = cfd[inf][2] - cfd[0][2];
        double const dv0 = cfd[inf + 1][0] - cfd[0][0];
        double const dv1 = cfd[inf + 1][1] - cfd[0][1];
        double const dv2 = cfd[inf + 1][2] - cfd[0][2];
        double const dw0 = du1 * dv2 - du2 * dv1;
        double const dw1 = du2 * dv0 - du0 * dv2;
        double const dw2 = du0 * dv1 - du1 * dv0;
        double const vob = sqrt(dw0 * dw0 + dw1 * dw1 + dw2 * dw2);
        for (int d = 0; d < 3; ++d)
        {
            cc[d] += crd[d] * vob;
        }
        voc += vob;
    }
    for (int d = 0; d < 3; ++d)
    {
        cc[d] /= voc;
        fccnd[static_cast<int64_t>(f) * 3 + d] = cc[d];
    }
    double rv[FCMND][3];
    for (int32_t inf = 0; inf < nnd; ++inf)
    {
        int32_t const ind = fn[inf + 1];
        for (int d = 0; d < 3; ++d)
        {
            rv[inf][d] = ndcrd[static_cast<int64_t>(ind) * 3 + d] - cc[d];
        }
    }
    double n0 = rv[nnd - 1][1] * rv[0][2] - rv[nnd - 1][2] * rv[0][1];
    double n1 = rv[nnd - 1][2] * rv[0][0] - rv[nnd - 1][0] * rv[0][2];
    double n2 = rv[nnd - 1][0] * rv[0][1] - rv[nnd - 1][1] * rv[0][0];
    for (int32_t k = 1; k < nnd; ++k)
    {
        n0 += rv[k - 1][1] * rv[k][2] - rv[k - 1][2] * rv[k][1];
        n1 += rv[k - 1][2] * rv[k][0] - rv[k - 1][0] * rv[k][2];
        n2 += rv[k - 1][0] * rv[k][1] - rv[k - 1][1] * rv[k][0];
    }
    double ar = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    n0 /= ar;
    n1 /= ar;
    n2 /= ar;
    ar /= 2.0;
    fcnml[static_cast<int64_t>(f) * 3] = n0;
    fcnml[static_cast<int64_t>(f) * 3 + 1] = n1;
    fcnml[static_cast<int64_t>(f) * 3 + 2] = n2;
    fcara[f] = ar;
}

__global__ void k_face_metric(int ndim, int32_t lo, int32_t hi, int32_t const * fcnds, double const * ndcrd,
                              double * fccnd, double * fcnml, double * fcara)
{
    int32_t const f = lo + static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (f < hi)
    {
        face_metric(ndim, f, fcnds, ndcrd, fccnd, fcnml, fcara);
    }
}

// Cell centroid from the pyramids on each face about the node average
// (StaticMesh_interior.cpp:802-836,882-922; StaticMesh_boundary.cpp:506-592).
__global__ void k_cell_centroid(int ndim, int32_t lo, int32_t hi, int32_t const * clnds, int32_t const * clfcs,
                                double const * ndcrd, double const * fccnd, double const * fcnml,
                                double const * fcara, double * clcnd)
{
    int32_t const icl = lo + static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (icl >= hi || (ndim != 2 && ndim != 3))
    {
        return;
    }
    int32_t const * cn = clnds + static_cast<int64_t>(icl) * (CLMND + 1);
    int32_t const * cf = clfcs + static_cast<int64_t>(icl) * (CLMFC + 1);
    double crd[3] = {0.0, 0.0, 0.0};
    int32_t const nnd = cn[0];
    for (int32_t k = 1; k <= nnd; ++k)
    {
        for (int d = 0; d < ndim; ++d)
        {
            crd[d] += ndcrd[static_cast<int64_t>(cn[k]) * ndim + d];
        }
    }
    for (int d = 0; d < ndim; ++d)
    {
        crd[d] /= nnd;
    }
    double voc = 0.0;
    double cc[3] = {0.0, 0.0, 0.0};
    int32_t const nfc = cf[0];
    for (int32_t ifl = 1; ifl <= nfc; ++ifl)
    {
        int32_t const f = cf[ifl];
        double du[3];
        for (int d = 0; d < ndim; ++d)
        {
            du[d] = crd[d] - fccnd[static_cast<int64_t>(f) * ndim + d];
        }
        double dot;
        if (ndim == 2)
        {
            dot = du[0] * fcnml[static_cast<int64_t>(f) * 2] + du[1] * fcnml[static_cast<int64_t>(f) * 2 + 1];
        }
        else
        {
            dot = du[0] * fcnml[static_cast<int64_t>(f) * 3] + du[1] * fcnml[static_cast<int64_t>(f) * 3 + 1]
                + du[2] * fcnml[static_cast<int64_t>(f) * 3 + 2];
        }
        double const vob = fabs(dot * fcara[f]);
        voc += vob;
        for (int d = 0; d < ndim; ++d)
        {
            double const dv = fccnd[static_cast<int64_t>(f) * ndim + d] + du[d] / (ndim + 1);
            cc[d] += dv * vob;
        }
    }
    for (int d = 0; d < ndim; ++d)
    {
        clcnd[static_cast<int64_t>(icl) * ndim + d] = cc[d] / voc;
    }
}

__device__ inline double face_vol(int ndim, int32_t f, int32_t icl, double const * fccnd, double const * fcnml,
                                  double const * fcara, double const * clcnd, double sgn)
{
    double vol = 0.0;
    for (int d = 0; d < ndim; ++d)
    {
        vol += (fccnd[static_cast<int64_t>(f) * ndim + d] - clcnd[static_cast<int64_t>(icl) * ndim + d])
             * (sgn * fcnml[static_cast<int64_t>(f) * ndim + d]);
    }
    vol *= fcara[f];
    return vol;
}

// reorder_face decisions (:944-975) evaluated per face in visiting order.
__global__ void k_face_flip(int ndim, int32_t nface, int32_t const * fccls, double const * fccnd, double * fcnml,
                            double const * fcara, double const * clcnd, int32_t * fcnds)
{
    int32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface)
    {
        return;
    }
    int32_t const a = fccls[static_cast<int64_t>(f) * FCREL];
    int32_t const b = fccls[static_cast<int64_t>(f) * FCREL + 1];
    if (a < 0)
    {
        return;
    }
    double sgn = 1.0;
    int flips = 0;
    if (face_vol(ndim, f, a, fccnd, fcnml, fcara, clcnd, sgn) < 0.0)
    {
        ++flips;
        sgn = -sgn;
    }
    if (b >= 0 && b != a)
    {
        if (!(face_vol(ndim, f, b, fccnd, fcnml, fcara, clcnd, sgn) < 0.0))
        {
            ++flips;
        }
    }
    if (flips & 1)
    {
        int32_t * fn = fcnds + static_cast<int64_t>(f) * (FCMND + 1);
        int32_t const nnd = fn[0];
        int32_t rev[FCMND];
        for (int32_t j = 0; j < nnd; ++j)
        {
            rev[j] = fn[nnd - j];
        }
        for (int32_t j = 0; j < nnd; ++j)
        {
            fn[j + 1] = rev[j];
        }
        for (int d = 0; d < ndim; ++d)
        {
            fcnml[static_cast<int64_t>(f) * ndim + d] = -fcnml[static_cast<int64_t>(f) * ndim + d];
        }
    }
}

__global__ void k_cell_volume(int ndim, int32_t ncell, int32_t const * clfcs, double const * fccnd,
                              double const * fcnml, double const * fcara, double const * clcnd, double * clvol)
{
    int32_t const icl = blockIdx.x * blockDim.x + threadIdx.x;
    if (icl >= ncell)
    {
        return;
    }
    int32_t const * cf = clfcs + static_cast<int64_t>(icl) * (CLMFC + 1);
    double v = 0.0;
    for (int32_t it = 1; it <= cf[0]; ++it)
    {
        double vol = face_vol(ndim, cf[it], icl, fccnd, fcnml, fcara, clcnd, 1.0);
        if (vol < 0.0)
        {
            vol = -vol;
        }
        v += vol;
    }
    clvol[icl] = v / ndim;
}

// ---------------------------------------------------------------- edges ---

__global__ void k_col0(int32_t n, int32_t const * fcnds, int32_t * out)
{
    int32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        out[i] = fcnds[static_cast<int64_t>(i) * (FCMND + 1)];
    }
}

__global__ void k_edge_cands(int32_t nface, int32_t const * fcnds, int32_t const * coff, int32_t * ea, int32_t * eb,
                             int32_t * cnt)
{
    int32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface)
    {
        return;
    }
    int32_t const * fn = fcnds + static_cast<int64_t>(f) * (FCMND + 1);
    int32_t const nnd = fn[0];
    for (int32_t inf = 1; inf <= nnd; ++inf)
    {
        int32_t const inx = (nnd == inf) ? 1 : inf + 1;
        int32_t a = fn[inf];
        int32_t b = fn[inx];
        if (a > b)
        {
            int32_t const t = a;
            a = b;
            b = t;
        }
        int32_t const c = coff[f] + inf - 1;
        ea[c] = a;
        eb[c] = b;
        atomicAdd(&cnt[a], 1);
    }
}

__global__ void k_edge_fill(int32_t ncand, int32_t const * ea, int32_t const * off, int32_t * cur, int32_t * lst)
{
    int32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncand)
    {
        int32_t const a = ea[c];
        lst[off[a] + atomicAdd(&cur[a], 1)] = c;
    }
}

__global__ void k_edge_first(int32_t ncand, int32_t const * ea, int32_t const * eb, int32_t const * off,
                             int32_t const * cnt, int32_t const * lst, int32_t * flag)
{
    int32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand)
    {
        return;
    }
    int32_t const a = ea[c];
    int32_t const b = eb[c];
    int first = 1;
    for (int32_t p = off[a]; p < off[a] + cnt[a]; ++p)
    {
        int32_t const d = lst[p];
        if (d < c && eb[d] == b)
        {
            first = 0;
            break;
        }
    }
    flag[c] = first;
}

__global__ void k_edge_write(int32_t ncand, int32_t const * ea, int32_t const * eb, int32_t const * flag,
                             int32_t const * pos, int32_t * ednds)
{
    int32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncand && flag[c])
    {
        ednds[static_cast<int64_t>(pos[c]) * 2] = ea[c];
        ednds[static_cast<int64_t>(pos[c]) * 2 + 1] = eb[c];
    }
}

// -------------------------------------------------------------- boundary --

__global__ void k_bnd_flag(int32_t nface, int32_t const * fccls, int32_t * flag)
{
    int32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nface)
    {
        flag[f] = fccls[static_cast<int64_t>(f) * FCREL + 1] < 0 ? 1 : 0;
    }
}

__global__ void k_unflag(int32_t n, int32_t const * faces, int32_t * flag)
{
    int32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        flag[faces[i]] = 0;
    }
}

__global__ void k_bnd_rest(int32_t nface, int32_t const * flag, int32_t const * pos, int32_t base, int32_t ibc,
                           int32_t * bndfcs, int32_t * rest)
{
    int32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nface && flag[f])
    {
        int32_t const row = base + pos[f];
        bndfcs[static_cast<int64_t>(row) * BFREL] = f;
        bndfcs[static_cast<int64_t>(row) * BFREL + 1] = ibc;
        rest[pos[f]] = f;
    }
}

__global__ void k_gather_fcjcl(int32_t n, int32_t const * faces, int32_t const * fccls, int32_t nface, int32_t * out)
{
    int32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        int32_t const f = faces[i];
        out[i] = (f >= 0 && f < nface) ? fccls[static_cast<int64_t>(f) * FCREL + 1] : 0;
    }
}

// ----------------------------------------------------------------- ghost --

__global__ void k_ghost_count(int32_t nbound, int32_t const * bndfcs, int32_t const * fccls, int32_t const * cltpn,
                              int32_t const * clnds, int32_t const * fcnds, int32_t * gface, int32_t * gnode)
{
    int32_t const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbound)
    {
        return;
    }
    int32_t const f = bndfcs[static_cast<int64_t>(k) * BFREL];
    int32_t const icl = fccls[static_cast<int64_t>(f) * FCREL];
    gface[k] = cell_nface(cltpn[icl]) - 1;
    gnode[k] = clnds[static_cast<int64_t>(icl) * (CLMND + 1)] - fcnds[static_cast<int64_t>(f) * (FCMND + 1)];
}

// fill_ghost connectivity and node mirroring (StaticMesh_boundary.cpp:283-372)
// for ghost cell igcl = -(k+1).  All pointers are body pointers.
__global__ void k_ghost_fill(int ndim, int32_t nbound, int32_t const * bndfcs, int32_t const * noff,
                             int32_t const * foff, int32_t * fccls, int32_t * cltpn, int32_t * clgrp, int32_t * clnds,
                             int32_t * clfcs, int32_t * fctpn, int32_t * fcnds, double * ndcrd, double const * fccnd,
                             double const * fcnml)
{
    int32_t const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbound)
    {
        return;
    }
    int32_t const igcl = -(k + 1);
    int32_t const ibfc = bndfcs[static_cast<int64_t>(k) * BFREL];
    int32_t const icl = fccls[static_cast<int64_t>(ibfc) * FCREL];
    cltpn[igcl] = cltpn[icl];
    clgrp[igcl] = clgrp[icl];
    int32_t const * cn = clnds + static_cast<int64_t>(icl) * (CLMND + 1);
    int32_t * gn = clnds + static_cast<int64_t>(igcl) * (CLMND + 1);
    for (int i = 0; i <= CLMND; ++i)
    {
        gn[i] = cn[i];
    }
    int32_t const * bfn = fcnds + static_cast<int64_t>(ibfc) * (FCMND + 1);
    int32_t mapfrom[CLMND];
    int32_t mapto[CLMND];
    int nmap = 0;
    int32_t ignd = -(noff[k] + 1);
    for (int32_t inl = 1; inl <= cn[0]; ++inl)
    {
        int32_t const ind = cn[inl];
        bool found = false;
        for (int32_t inf = 1; inf <= bfn[0]; ++inf)
        {
            if (ind == bfn[inf])
            {
                found = true;
                break;
            }
        }
        if (!found)
        {
            mapfrom[nmap] = ind;
            mapto[nmap] = ignd;
            ++nmap;
            gn[inl] = ignd;
            double dist = 0.0;
            for (int d = 0; d < ndim; ++d)
            {
                dist += (fccnd[static_cast<int64_t>(ibfc) * ndim + d] - ndcrd[static_cast<int64_t>(ind) * ndim + d])
                      * fcnml[static_cast<int64_t>(ibfc) * ndim + d];
            }
            for (int d = 0; d < ndim; ++d)
            {
                ndcrd[static_cast<int64_t>(ignd) * ndim + d] = ndcrd[static_cast<int64_t>(ind) * ndim + d]
                                                             + 2 * dist * fcnml[static_cast<int64_t>(ibfc) * ndim + d];
            }
            ignd -= 1;
        }
    }
    fccls[static_cast<int64_t>(ibfc) * FCREL + 1] = igcl;
    int32_t const * cf = clfcs + static_cast<int64_t>(icl) * (CLMFC + 1);
    int32_t * gf = clfcs + static_cast<int64_t>(igcl) * (CLMFC + 1);
    for (int i = 0; i <= CLMFC; ++i)
    {
        gf[i] = cf[i];
    }
    int32_t igfc = -(foff[k] + 1);
    for (int32_t ifl = 1; ifl <= cf[0]; ++ifl)
    {
        int32_t const ifc = cf[ifl];
        if (ifc == ibfc)
        {
            continue;
        }
        fctpn[igfc] = fctpn[ifc];
        fccls[static_cast<int64_t>(igfc) * FCREL] = igcl;
        gf[ifl] = igfc;
        int32_t const * sn = fcnds + static_cast<int64_t>(ifc) * (FCMND + 1);
        int32_t * dn = fcnds + static_cast<int64_t>(igfc) * (FCMND + 1);
        for (int i = 0; i <= FCMND; ++i)
        {
            dn[i] = sn[i];
        }
        for (int32_t inf = 1; inf <= dn[0]; ++inf)
        {
            for (int m = 0; m < nmap; ++m)
            {
                if (dn[inf] == mapfrom[m])
                {
                    dn[inf] = mapto[m];
                    break;
                }
            }
        }
        igfc -= 1;
    }
}

// Ghost cell volumes and the outward flip of ghost-owned faces
// (StaticMesh_boundary.cpp:594-626).
__global__ void k_ghost_volume(int ndim, int32_t nbound, int32_t const * clfcs, int32_t const * fccls,
                               double const * fccnd, double * fcnml, double const * fcara, double const * clcnd,
                               double * clvol)
{
    int32_t const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbound)
    {
        return;
    }
    int32_t const icl = -(k + 1);
    int32_t const * cf = clfcs + static_cast<int64_t>(icl) * (CLMFC + 1);
    double v = 0.0;
    for (int32_t it = 1; it <= cf[0]; ++it)
    {
        int32_t const f = cf[it];
        double vol = face_vol(ndim, f, icl, fccnd, fcnml, fcara, clcnd, 1.0);
        if (vol < 0.0)
        {
            if (fccls[static_cast<int64_t>(f) * FCREL] == icl)
            {
                for (int d = 0; d < ndim; ++d)
                {
                    fcnml[static_cast<int64_t>(f) * ndim + d] = -fcnml[static_cast<int64_t>(f) * ndim + d];
                }
            }
            vol = -vol;
        }
        v += vol;
    }
    clvol[icl] = v / ndim;
}

template <typename T>
void grow(DevBuf<T> & b, int64_t nbody_rows, int64_t ngh, int64_t width, T init, hipStream_t s)
{
    DevBuf<T> n;
    n.alloc((nbody_rows + ngh) * width);
    fill<T>(n.ptr, ngh * width, init, s);
    if (nbody_rows * width > 0)
    {
        SC_HIP(hipMemcpyAsync(n.ptr + ngh * width, b.ptr, sizeof(T) * static_cast<size_t>(nbody_rows * width),
                              hipMemcpyDeviceToDevice, s));
    }
    SC_HIP(hipStreamSynchronize(s));
    b = std::move(n);
}

template <typename T>
T to_host(T const * dptr, hipStream_t s)
{
    T v;
    SC_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    return v;
}

} /* end namespace */

DMesh::DMesh(int nd, int32_t nn, int32_t nf, int32_t nc)
    : ndim(nd)
    , nnode(nn)
    , nface(nf)
    , ncell(nc)
{
    if (nd < 1 || nd > 3)
    {
        fail(SC_EVALUE, "StaticMesh: ndim must be 1, 2 or 3");
    }
    if (nn < 0 || nf < 0 || nc < 0)
    {
        fail(SC_EVALUE, "StaticMesh: negative size");
    }
    AllocScope const scope(stream);
    // StaticMesh.hpp:290-311; tables the reference leaves uninitialised are
    // zero here.
    ndcrd.alloc(static_cast<int64_t>(nnode) * ndim);
    fccnd.alloc(static_cast<int64_t>(nface) * ndim);
    fcnml.alloc(static_cast<int64_t>(nface) * ndim);
    fcara.alloc(nface);
    clcnd.alloc(static_cast<int64_t>(ncell) * ndim);
    clvol.alloc(ncell);
    fctpn.alloc(nface);
    cltpn.alloc(ncell);
    clgrp.alloc(ncell);
    fcnds.alloc(static_cast<int64_t>(nface) * (FCMND + 1));
    fccls.alloc(static_cast<int64_t>(nface) * FCREL);
    clnds.alloc(static_cast<int64_t>(ncell) * (CLMND + 1));
    clfcs.alloc(static_cast<int64_t>(ncell) * (CLMFC + 1));
    for (auto * b : {&ndcrd, &fccnd, &fcnml, &fcara, &clcnd, &clvol})
    {
        if (b->count)
        {
            SC_HIP(hipMemsetAsync(b->ptr, 0, b->bytes(), stream));
        }
    }
    for (auto * b : {&fctpn, &cltpn, &clgrp, &fcnds, &fccls, &clnds, &clfcs})
    {
        if (b->count)
        {
            SC_HIP(hipMemsetAsync(b->ptr, 0, b->bytes(), stream));
        }
    }
    sync();
}

DMesh::~DMesh() = default;

void DMesh::table(int id, int64_t & rows, int64_t & cols, int64_t & ngh, bool & is_int, void *& ptr) const
{
    is_int = id >= MA_FCTPN;
    switch (id)
    {
    case MA_NDCRD: rows = ngstnode + nnode; cols = ndim; ngh = ngstnode; ptr = ndcrd.ptr; break;
    case MA_FCCND: rows = ngstface + nface; cols = ndim; ngh = ngstface; ptr = fccnd.ptr; break;
    case MA_FCNML: rows = ngstface + nface; cols = ndim; ngh = ngstface; ptr = fcnml.ptr; break;
    case MA_FCARA: rows = ngstface + nface; cols = 0; ngh = ngstface; ptr = fcara.ptr; break;
    case MA_CLCND: rows = ngstcell + ncell; cols = ndim; ngh = ngstcell; ptr = clcnd.ptr; break;
    case MA_CLVOL: rows = ngstcell + ncell; cols = 0; ngh = ngstcell; ptr = clvol.ptr; break;
    case MA_FCTPN: rows = ngstface + nface; cols = 0; ngh = ngstface; ptr = fctpn.ptr; break;
    case MA_CLTPN: rows = ngstcell + ncell; cols = 0; ngh = ngstcell; ptr = cltpn.ptr; break;
    case MA_CLGRP: rows = ngstcell + ncell; cols = 0; ngh = ngstcell; ptr = clgrp.ptr; break;
    case MA_FCNDS: rows = ngstface + nface; cols = FCMND + 1; ngh = ngstface; ptr = fcnds.ptr; break;
    case MA_FCCLS: rows = ngstface + nface; cols = FCREL; ngh = ngstface; ptr = fccls.ptr; break;
    case MA_CLNDS: rows = ngstcell + ncell; cols = CLMND + 1; ngh = ngstcell; ptr = clnds.ptr; break;
    case MA_CLFCS: rows = ngstcell + ncell; cols = CLMFC + 1; ngh = ngstcell; ptr = clfcs.ptr; break;
    case MA_EDNDS: rows = nedge; cols = 2; ngh = 0; ptr = ednds.ptr; break;
    case MA_BNDFCS: rows = nbound; cols = BFREL; ngh = 0; ptr = bndfcs.ptr; break;
    default: fail(SC_EVALUE, "StaticMesh: unknown table id");
    }
}

void DMesh::build_interior(bool do_metric, bool do_edge)
{
    AllocScope const scope(stream);
    build_faces_from_cells();
    if (do_metric)
    {
        calc_metric();
    }
    if (do_edge)
    {
        build_edge();
    }
}

void DMesh::build_faces_from_cells()
{
    AllocScope const scope(stream);
    hipStream_t const s = stream;
    if (ngstcell != 0 || ngstface != 0 || ngstnode != 0)
    {
        fail(SC_ERUNTIME, "StaticMesh::build_interior: ghost entities already built");
    }
    int32_t const nc = ncell;
    DevBuf<int32_t> cnt;
    DevBuf<int32_t> bad;
    cnt.alloc(nc + 1);
    bad.alloc(1);
    SC_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(int32_t), s));
    if (nc > 0)
    {
        k_count_faces<<<grid_for(nc, BLK), BLK, 0, s>>>(nc, nnode, cltpn.ptr, clnds.ptr, cnt.ptr, bad.ptr);
        SC_HIP(hipGetLastError());
    }
    if (to_host(bad.ptr, s) != 0)
    {
        fail(SC_EVALUE, "StaticMesh::build_interior: clnds refers to a node outside [0, nnode)");
    }
    int64_t const mface64 = exclusive_scan<int32_t>(cnt.ptr, cnt.ptr, nc, s);
    if (mface64 >= INT_MAX / 2)
    {
        fail(SC_EVALUE, "StaticMesh::build_interior: too many faces");
    }
    int32_t const mface = static_cast<int32_t>(mface64);
    DevBuf<int32_t> bfctpn, bfcnds, bclfcs, fcell;
    bfctpn.alloc(mface);
    bfcnds.alloc(static_cast<int64_t>(mface) * (FCMND + 1));
    bclfcs.alloc(static_cast<int64_t>(nc) * (CLMFC + 1));
    fcell.alloc(mface);
    if (nc > 0)
    {
        k_populate<<<grid_for(nc, BLK), BLK, 0, s>>>(nc, cltpn.ptr, clnds.ptr, cnt.ptr, bfctpn.ptr, bfcnds.ptr,
                                                      bclfcs.ptr, fcell.ptr);
        SC_HIP(hipGetLastError());
    }

    // Node-to-face table.
    DevBuf<int32_t> ncnt, noff, ncur, nlst;
    ncnt.alloc(nnode + 1);
    noff.alloc(nnode + 1);
    ncur.alloc(nnode + 1);
    SC_HIP(hipMemsetAsync(ncnt.ptr, 0, ncnt.bytes(), s));
    SC_HIP(hipMemsetAsync(ncur.ptr, 0, ncur.bytes(), s));
    if (mface > 0)
    {
        k_node_count<<<grid_for(mface, BLK), BLK, 0, s>>>(mface, bfcnds.ptr, ncnt.ptr);
        SC_HIP(hipGetLastError());
    }
    int64_t const nent = exclusive_scan<int32_t>(ncnt.ptr, noff.ptr, nnode, s);
    nlst.alloc(nent);
    DevBuf<int32_t> dup, flag;
    dup.alloc(mface);
    flag.alloc(mface + 1);
    if (mface > 0)
    {
        k_node_fill<<<grid_for(mface, BLK), BLK, 0, s>>>(mface, bfcnds.ptr, noff.ptr, ncur.ptr, nlst.ptr);
        SC_HIP(hipGetLastError());
        k_dedup<<<grid_for(mface, BLK), BLK, 0, s>>>(mface, bfctpn.ptr, bfcnds.ptr, noff.ptr, ncnt.ptr, nlst.ptr,
                                                     dup.ptr);
        SC_HIP(hipGetLastError());
        k_canon<<<grid_for(mface, BLK), BLK, 0, s>>>(mface, dup.ptr, flag.ptr);
        SC_HIP(hipGetLastError());
    }
    nlst.release();
    int32_t const nf = static_cast<int32_t>(exclusive_scan<int32_t>(flag.ptr, flag.ptr, mface, s));
    nface = nf;
    fctpn.alloc(nf);
    fcnds.alloc(static_cast<int64_t>(nf) * (FCMND + 1));
    fccls.alloc(static_cast<int64_t>(nf) * FCREL);
    DevBuf<int32_t> second;
    second.alloc(nf);
    fill<int32_t>(second.ptr, nf, INT_MAX, s);
    if (mface > 0)
    {
        k_compact<<<grid_for(mface, BLK), BLK, 0, s>>>(mface, dup.ptr, flag.ptr, bfctpn.ptr, bfcnds.ptr, fcell.ptr,
                                                       fctpn.ptr, fcnds.ptr, fccls.ptr, second.ptr);
        SC_HIP(hipGetLastError());
    }
    if (nf > 0)
    {
        k_second<<<grid_for(nf, BLK), BLK, 0, s>>>(nf, second.ptr, fcell.ptr, fccls.ptr);
        SC_HIP(hipGetLastError());
    }
    if (nc > 0)
    {
        k_remap_clfcs<<<grid_for(nc, BLK), BLK, 0, s>>>(nc, bclfcs.ptr, dup.ptr, flag.ptr, clfcs.ptr);
        SC_HIP(hipGetLastError());
    }
    // The metric tables are remade zero (StaticMesh_interior.cpp:560-562).
    fccnd.alloc(static_cast<int64_t>(nf) * ndim);
    fcnml.alloc(static_cast<int64_t>(nf) * ndim);
    fcara.alloc(nf);
    for (auto * b : {&fccnd, &fcnml, &fcara})
    {
        if (b->count)
        {
            SC_HIP(hipMemsetAsync(b->ptr, 0, b->bytes(), s));
        }
    }
    sync();
}

void DMesh::calc_metric()
{
    AllocScope const scope(stream);
    hipStream_t const s = stream;
    if (nface > 0 && (ndim == 2 || ndim == 3))
    {
        k_face_metric<<<grid_for(nface, BLK), BLK, 0, s>>>(ndim, 0, nface, fcnds_b(), ndcrd_b(), fccnd_b(), fcnml_b(),
                                                           fcara_b());
        SC_HIP(hipGetLastError());
    }
    if (ncell > 0)
    {
        k_cell_centroid<<<grid_for(ncell, BLK), BLK, 0, s>>>(ndim, 0, ncell, clnds_b(), clfcs_b(), ndcrd_b(),
                                                             fccnd_b(), fcnml_b(), fcara_b(), clcnd_b());
        SC_HIP(hipGetLastError());
    }
    if (nface > 0)
    {
        k_face_flip<<<grid_for(nface, BLK), BLK, 0, s>>>(ndim, nface, fccls_b(), fccnd_b(), fcnml_b(), fcara_b(),
                                                         clcnd_b(), fcnds_b());
        SC_HIP(hipGetLastError());
    }
    if (ncell > 0)
    {
        k_cell_volume<<<grid_for(ncell, BLK), BLK, 0, s>>>(ndim, ncell, clfcs_b(), fccnd_b(), fcnml_b(), fcara_b(),
                                                           clcnd_b(), clvol_b());
        SC_HIP(hipGetLastError());
    }
    sync();
}

void DMesh::build_edge()
{
    AllocScope const scope(stream);
    hipStream_t const s = stream;
    DevBuf<int32_t> coff;
    coff.alloc(nface + 1);
    // Candidate edges per face: fcnds(f, 0) of them.
    if (nface > 0)
    {
        k_col0<<<grid_for(nface, BLK), BLK, 0, s>>>(nface, fcnds_b(), coff.ptr);
        SC_HIP(hipGetLastError());
    }
    int32_t const ncand = static_cast<int32_t>(exclusive_scan<int32_t>(coff.ptr, coff.ptr, nface, s));
    DevBuf<int32_t> ea, eb, ncnt, noff, ncur, lst, flag;
    ea.alloc(ncand);
    eb.alloc(ncand);
    ncnt.alloc(nnode + 1);
    noff.alloc(nnode + 1);
    ncur.alloc(nnode + 1);
    SC_HIP(hipMemsetAsync(ncnt.ptr, 0, ncnt.bytes(), s));
    SC_HIP(hipMemsetAsync(ncur.ptr, 0, ncur.bytes(), s));
    if (nface > 0)
    {
        k_edge_cands<<<grid_for(nface, BLK), BLK, 0, s>>>(nface, fcnds_b(), coff.ptr, ea.ptr, eb.ptr, ncnt.ptr);
        SC_HIP(hipGetLastError());
    }
    (void)exclusive_scan<int32_t>(ncnt.ptr, noff.ptr, nnode, s);
    lst.alloc(ncand);
    flag.alloc(ncand + 1);
    if (ncand > 0)
    {
        k_edge_fill<<<grid_for(ncand, BLK), BLK, 0, s>>>(ncand, ea.ptr, noff.ptr, ncur.ptr, lst.ptr);
        SC_HIP(hipGetLastError());
        k_edge_first<<<grid_for(ncand, BLK), BLK, 0, s>>>(ncand, ea.ptr, eb.ptr, noff.ptr, ncnt.ptr, lst.ptr,
                                                          flag.ptr);
        SC_HIP(hipGetLastError());
    }
    DevBuf<int32_t> pos;
    pos.alloc(ncand + 1);
    int32_t const ne = static_cast<int32_t>(exclusive_scan<int32_t>(flag.ptr, pos.ptr, ncand, s));
    nedge = ne;
    ednds.alloc(static_cast<int64_t>(ne) * 2);
    if (ncand > 0)
    {
        k_edge_write<<<grid_for(ncand, BLK), BLK, 0, s>>>(ncand, ea.ptr, eb.ptr, flag.ptr, pos.ptr, ednds.ptr);
        SC_HIP(hipGetLastError());
    }
    sync();
}

std::vector<int32_t> DMesh::fcjcl_of(std::vector<int32_t> const & faces)
{
    std::vector<int32_t> out(faces.size(), 0);
    if (faces.empty())
    {
        return out;
    }
    DevBuf<int32_t> df, dout;
    df.alloc(static_cast<int64_t>(faces.size()));
    dout.alloc(static_cast<int64_t>(faces.size()));
    SC_HIP(hipMemcpyAsync(df.ptr, faces.data(), df.bytes(), hipMemcpyHostToDevice, stream));
    k_gather_fcjcl<<<grid_for(df.count, BLK), BLK, 0, stream>>>(static_cast<int32_t>(df.count), df.ptr, fccls_b(),
                                                                nface, dout.ptr);
    SC_HIP(hipGetLastError());
    SC_HIP(hipMemcpyAsync(out.data(), dout.ptr, dout.bytes(), hipMemcpyDeviceToHost, stream));
    sync();
    return out;
}

// StaticMesh_boundary.cpp:14-78
void DMesh::add_bc(std::string const & name, std::vector<int32_t> const & faces)
{
    AllocScope const scope(stream);
    if (nbound != 0)
    {
        fail(SC_ERUNTIME, "StaticMesh::add_bc: boundary is already built");
    }
    if (name.empty())
    {
        fail(SC_EVALUE, "StaticMesh::add_bc: empty name");
    }
    for (auto const & b : bcs)
    {
        if (b.name == name)
        {
            fail(SC_EVALUE, "StaticMesh::add_bc: duplicated name \"" + name + "\"");
        }
    }
    std::vector<char> claimed(static_cast<size_t>(nface), 0);
    for (auto const & b : bcs)
    {
        for (int32_t f : b.facn0)
        {
            claimed[static_cast<size_t>(f)] = 1;
        }
    }
    for (int32_t f : faces)
    {
        if (f < 0 || f >= nface)
        {
            fail(SC_EVALUE, "StaticMesh::add_bc: face " + std::to_string(f) + " out of range [0, "
                                + std::to_string(nface) + ")");
        }
    }
    std::vector<int32_t> const jcl = fcjcl_of(faces);
    for (size_t i = 0; i < faces.size(); ++i)
    {
        int32_t const f = faces[i];
        if (jcl[i] >= 0)
        {
            fail(SC_EVALUE, "StaticMesh::add_bc: face " + std::to_string(f) + " is not a boundary face");
        }
        if (claimed[static_cast<size_t>(f)])
        {
            fail(SC_EVALUE, "StaticMesh::add_bc: face " + std::to_string(f) + " is already claimed");
        }
        claimed[static_cast<size_t>(f)] = 1;
    }
    MeshBc b;
    b.name = name;
    b.facn0 = faces;
    b.facn1.assign(faces.size(), -1);
    bcs.push_back(std::move(b));
}

// StaticMesh_boundary.cpp:103-182
void DMesh::build_boundary()
{
    AllocScope const scope(stream);
    hipStream_t const s = stream;
    if (nbound != 0)
    {
        fail(SC_ERUNTIME, "StaticMesh::build_boundary: boundary is already built");
    }
    DevBuf<int32_t> flag, pos;
    flag.alloc(nface + 1);
    pos.alloc(nface + 1);
    if (nface > 0)
    {
        k_bnd_flag<<<grid_for(nface, BLK), BLK, 0, s>>>(nface, fccls_b(), flag.ptr);
        SC_HIP(hipGetLastError());
    }
    int32_t const nb = static_cast<int32_t>(exclusive_scan<int32_t>(flag.ptr, pos.ptr, nface, s));
    nbound = nb;
    bndfcs.alloc(static_cast<int64_t>(nb) * BFREL);
    fill<int32_t>(bndfcs.ptr, bndfcs.count, -1, s);
    std::vector<int32_t> user;
    std::vector<int32_t> rows;
    int32_t ibfc = 0;
    for (size_t ib = 0; ib < bcs.size(); ++ib)
    {
        for (size_t k = 0; k < bcs[ib].facn0.size(); ++k)
        {
            user.push_back(bcs[ib].facn0[k]);
            rows.push_back(bcs[ib].facn0[k]);
            rows.push_back(static_cast<int32_t>(ib));
            rows.push_back(-1);
            bcs[ib].facn1[k] = ibfc++;
        }
    }
    int32_t const nuser = ibfc;
    if (nuser > nb)
    {
        fail(SC_ERUNTIME, "StaticMesh::build_boundary: more BC faces than boundary faces");
    }
    if (nuser > 0)
    {
        SC_HIP(hipMemcpyAsync(bndfcs.ptr, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice, s));
        DevBuf<int32_t> du;
        du.alloc(nuser);
        SC_HIP(hipMemcpyAsync(du.ptr, user.data(), du.bytes(), hipMemcpyHostToDevice, s));
        k_unflag<<<grid_for(nuser, BLK), BLK, 0, s>>>(nuser, du.ptr, flag.ptr);
        SC_HIP(hipGetLastError());
        sync();
    }
    int32_t const nleft = nb - nuser;
    if (nleft > 0)
    {
        int32_t const nrest = static_cast<int32_t>(exclusive_scan<int32_t>(flag.ptr, pos.ptr, nface, s));
        if (nrest != nleft)
        {
            fail(SC_ERUNTIME, "StaticMesh::build_boundary: inconsistent boundary count");
        }
        DevBuf<int32_t> rest;
        rest.alloc(nleft);
        int32_t const ibc = static_cast<int32_t>(bcs.size());
        k_bnd_rest<<<grid_for(nface, BLK), BLK, 0, s>>>(nface, flag.ptr, pos.ptr, nuser, ibc, bndfcs.ptr, rest.ptr);
        SC_HIP(hipGetLastError());
        MeshBc b;
        b.name = "<NONAME>";
        b.facn0.resize(static_cast<size_t>(nleft));
        b.facn1.resize(static_cast<size_t>(nleft));
        SC_HIP(hipMemcpyAsync(b.facn0.data(), rest.ptr, rest.bytes(), hipMemcpyDeviceToHost, s));
        sync();
        for (int32_t i = 0; i < nleft; ++i)
        {
            b.facn1[static_cast<size_t>(i)] = nuser + i;
        }
        bcs.push_back(std::move(b));
    }
    sync();
}

// StaticMesh_boundary.cpp:185-261
void DMesh::build_ghost()
{
    AllocScope const scope(stream);
    hipStream_t const s = stream;
    if (ngstcell != 0)
    {
        fail(SC_ERUNTIME, "StaticMesh::build_ghost: ghost entities already built");
    }
    int32_t const nb = nbound;
    DevBuf<int32_t> gface, gnode;
    gface.alloc(nb + 1);
    gnode.alloc(nb + 1);
    if (nb > 0)
    {
        k_ghost_count<<<grid_for(nb, BLK), BLK, 0, s>>>(nb, bndfcs.ptr, fccls_b(), cltpn_b(), clnds_b(), fcnds_b(),
                                                        gface.ptr, gnode.ptr);
        SC_HIP(hipGetLastError());
    }
    int32_t const gf = static_cast<int32_t>(exclusive_scan<int32_t>(gface.ptr, gface.ptr, nb, s));
    int32_t const gn = static_cast<int32_t>(exclusive_scan<int32_t>(gnode.ptr, gnode.ptr, nb, s));
    int64_t const nd = ndim;
    grow<double>(ndcrd, nnode, gn, nd, 0.0, s);
    grow<double>(fccnd, nface, gf, nd, 0.0, s);
    grow<double>(fcnml, nface, gf, nd, 0.0, s);
    grow<double>(fcara, nface, gf, 1, 0.0, s);
    grow<double>(clcnd, ncell, nb, nd, 0.0, s);
    grow<double>(clvol, ncell, nb, 1, 0.0, s);
    grow<int32_t>(fctpn, nface, gf, 1, 0, s);
    grow<int32_t>(cltpn, ncell, nb, 1, 0, s);
    grow<int32_t>(clgrp, ncell, nb, 1, -1, s);
    grow<int32_t>(fcnds, nface, gf, FCMND + 1, -1, s);
    grow<int32_t>(fccls, nface, gf, FCREL, -1, s);
    grow<int32_t>(clnds, ncell, nb, CLMND + 1, -1, s);
    grow<int32_t>(clfcs, ncell, nb, CLMFC + 1, -1, s);
    ngstnode = gn;
    ngstface = gf;
    ngstcell = nb;
    if (nb > 0)
    {
        k_ghost_fill<<<grid_for(nb, BLK), BLK, 0, s>>>(ndim, nb, bndfcs.ptr, gnode.ptr, gface.ptr, fccls_b(),
                                                       cltpn_b(), clgrp_b(), clnds_b(), clfcs_b(), fctpn_b(),
                                                       fcnds_b(), ndcrd_b(), fccnd_b(), fcnml_b());
        SC_HIP(hipGetLastError());
    }
    if (gf > 0 && (ndim == 2 || ndim == 3))
    {
        k_face_metric<<<grid_for(gf, BLK), BLK, 0, s>>>(ndim, -gf, 0, fcnds_b(), ndcrd_b(), fccnd_b(), fcnml_b(),
                                                        fcara_b());
        SC_HIP(hipGetLastError());
    }
    if (nb > 0)
    {
        k_cell_centroid<<<grid_for(nb, BLK), BLK, 0, s>>>(ndim, -nb, 0, clnds_b(), clfcs_b(), ndcrd_b(), fccnd_b(),
                                                          fcnml_b(), fcara_b(), clcnd_b());
        SC_HIP(hipGetLastError());
        k_ghost_volume<<<grid_for(nb, BLK), BLK, 0, s>>>(ndim, nb, clfcs_b(), fccls_b(), fccnd_b(), fcnml_b(),
                                                         fcara_b(), clcnd_b(), clvol_b());
        SC_HIP(hipGetLastError());
    }
    sync();
}

} /* end namespace scamd */

using scamd::DMesh;
using scamd::guard;

extern "C"
{

int scmesh_create(int ndim, int32_t nnode, int32_t nface, int32_t ncell, void ** out)
{
    return guard([&]() { *out = new DMesh(ndim, nnode, nface, ncell); });
}

void scmesh_destroy(void * h) { delete static_cast<DMesh *>(h); }

// ndim nnode nface ncell nbound ngstnode ngstface ngstcell nedge nbcs
int scmesh_counts(void * h, int64_t * out)
{
    DMesh const & m = *static_cast<DMesh *>(h);
    scamd::AllocScope const alloc_scope_(m.stream);
    out[0] = m.ndim;
    out[1] = m.nnode;
    out[2] = m.nface;
    out[3] = m.ncell;
    out[4] = m.nbound;
    out[5] = m.ngstnode;
    out[6] = m.ngstface;
    out[7] = m.ngstcell;
    out[8] = m.nedge;
    out[9] = static_cast<int64_t>(m.bcs.size());
    return 0;
}

// info: rows, cols (0 for a 1-D table), nghost, is_int
int scmesh_table_info(void * h, int id, int64_t * info)
{
    return guard(
        [&]()
        {
            int64_t rows, cols, ngh;
            bool is_int;
            void * p;
            static_cast<DMesh *>(h)->table(id, rows, cols, ngh, is_int, p);
            info[0] = rows;
            info[1] = cols;
            info[2] = ngh;
            info[3] = is_int ? 1 : 0;
        });
}

int scmesh_table_ptr(void * h, int id, void ** ptr)
{
    return guard(
        [&]()
        {
            int64_t rows, cols, ngh;
            bool is_int;
            static_cast<DMesh *>(h)->table(id, rows, cols, ngh, is_int, *ptr);
        });
}

static int mesh_copy(void * h, int id, void * host, int64_t nbytes, bool down)
{
    return guard(
        [&]()
        {
            DMesh & m = *static_cast<DMesh *>(h);
            scamd::AllocScope const alloc_scope_(m.stream);
            int64_t rows, cols, ngh;
            bool is_int;
            void * p;
            m.table(id, rows, cols, ngh, is_int, p);
            int64_t const want = rows * (cols > 0 ? cols : 1) * (is_int ? 4 : 8);
            if (nbytes != want)
            {
                scamd::fail(scamd::SC_EVALUE, "StaticMesh: table byte size mismatch");
            }
            if (want == 0)
            {
                return;
            }
            if (down)
            {
                SC_HIP(hipMemcpyAsync(host, p, static_cast<size_t>(want), hipMemcpyDeviceToHost, m.stream));
            }
            else
            {
                SC_HIP(hipMemcpyAsync(p, host, static_cast<size_t>(want), hipMemcpyHostToDevice, m.stream));
            }
            m.sync();
        });
}

int scmesh_download(void * h, int id, void * host, int64_t nbytes) { return mesh_copy(h, id, host, nbytes, true); }

int scmesh_upload(void * h, int id, void const * host, int64_t nbytes)
{
    return mesh_copy(h, id, const_cast<void *>(host), nbytes, false);
}

int scmesh_build_interior(void * h, int do_metric, int do_edge)
{
    return guard([&]() { static_cast<DMesh *>(h)->build_interior(do_metric != 0, do_edge != 0); });
}

int scmesh_build_edge(void * h)
{
    return guard([&]() { static_cast<DMesh *>(h)->build_edge(); });
}

int scmesh_build_boundary(void * h)
{
    return guard([&]() { static_cast<DMesh *>(h)->build_boundary(); });
}

int scmesh_build_ghost(void * h)
{
    return guard([&]() { static_cast<DMesh *>(h)->build_ghost(); });
}

int scmesh_add_bc(void * h, char const * name, int32_t const * faces, int64_t n)
{
    return guard(
        [&]()
        {
            std::vector<int32_t> f(faces, faces + n);
            static_cast<DMesh *>(h)->add_bc(name, f);
        });
}

// BC group ibc: name into buf, facn columns 0/1 (may be null); returns the
// face count through *n.
int scmesh_bc(void * h, int ibc, char * name, int nname, int32_t * facn0, int32_t * facn1, int64_t * n)
{
    return guard(
        [&]()
        {
            DMesh const & m = *static_cast<DMesh *>(h);
            scamd::AllocScope const alloc_scope_(m.stream);
            if (ibc < 0 || static_cast<size_t>(ibc) >= m.bcs.size())
            {
                scamd::fail(scamd::SC_EINDEX, "StaticMesh::bc: index out of range");
            }
            scamd::MeshBc const & b = m.bcs[static_cast<size_t>(ibc)];
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
            *n = static_cast<int64_t>(b.facn0.size());
        });
}

// Device bytes a mesh holds (its tables, ghost rows included).
int scmesh_device_bytes(void * h, int64_t * out)
{
    return guard(
        [&]()
        {
            auto & m = *static_cast<DMesh *>(h);
            int64_t b = 0;
            for (auto const * x : {&m.ndcrd, &m.fccnd, &m.fcnml, &m.fcara, &m.clcnd, &m.clvol})
            {
                b += static_cast<int64_t>(x->bytes());
            }
            for (auto const * x : {&m.fctpn, &m.cltpn, &m.clgrp, &m.fcnds, &m.fccls, &m.clnds, &m.clfcs, &m.ednds,
                                   &m.bndfcs})
            {
                b += static_cast<int64_t>(x->bytes());
            }
            *out = b;
        });
}

} /* end extern "C" */
