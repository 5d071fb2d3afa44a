/*
 * Oracle (test infrastructure only): CPU restatement of solvcon's StaticMesh
 * builders.  Each routine cites the reference lines it follows.  Arrays the
 * reference leaves uninitialised (malloc'ed SimpleArray, e.g. the unused
 * columns of clnds) are zero here.
 */

#include "mesh.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <set>
#include <stdexcept>

namespace orc
{

namespace
{

struct FaceDef
{
    int32_t type;
    int nnd;
    int nd[4];
};

struct CellFaces
{
    int nface;
    FaceDef face[6];
};

// Per-cell-type face tables with local (0-based) node positions; the node
// order of each face is the one FaceBuilder::add_* emits
// (StaticMesh_interior.cpp:41-308).
CellFaces const & cell_faces(int32_t tpn)
{
    static CellFaces const none{0, {}};
    static CellFaces const line{2, {{POINT, 1, {0}}, {POINT, 1, {1}}}};
    static CellFaces const quad{4, {{LINE, 2, {0, 1}}, {LINE, 2, {1, 2}}, {LINE, 2, {2, 3}}, {LINE, 2, {3, 0}}}};
    static CellFaces const tri{3, {{LINE, 2, {0, 1}}, {LINE, 2, {1, 2}}, {LINE, 2, {2, 0}}}};
    static CellFaces const hex{6,
                               {{QUADRILATERAL, 4, {0, 3, 2, 1}},
                                {QUADRILATERAL, 4, {1, 2, 6, 5}},
                                {QUADRILATERAL, 4, {4, 5, 6, 7}},
                                {QUADRILATERAL, 4, {0, 4, 7, 3}},
                                {QUADRILATERAL, 4, {0, 1, 5, 4}},
                                {QUADRILATERAL, 4, {2, 3, 7, 6}}}};
    static CellFaces const tet{4,
                               {{TRIANGLE, 3, {0, 2, 1}},
                                {TRIANGLE, 3, {0, 1, 3}},
                                {TRIANGLE, 3, {0, 3, 2}},
                                {TRIANGLE, 3, {1, 2, 3}}}};
    static CellFaces const psm{5,
                               {{TRIANGLE, 3, {0, 1, 2}},
                                {TRIANGLE, 3, {3, 5, 4}},
                                {QUADRILATERAL, 4, {0, 3, 4, 1}},
                                {QUADRILATERAL, 4, {0, 2, 5, 3}},
                                {QUADRILATERAL, 4, {1, 4, 5, 2}}}};
    static CellFaces const pyr{5,
                               {{TRIANGLE, 3, {0, 4, 3}},
                                {TRIANGLE, 3, {1, 4, 0}},
                                {TRIANGLE, 3, {2, 4, 1}},
                                {TRIANGLE, 3, {3, 4, 2}},
                                {QUADRILATERAL, 4, {0, 3, 2, 1}}}};
    switch (tpn)
    {
    case LINE: return line;
    case QUADRILATERAL: return quad;
    case TRIANGLE: return tri;
    case HEXAHEDRON: return hex;
    case TETRAHEDRON: return tet;
    case PRISM: return psm;
    case PYRAMID: return pyr;
    default: return none;
    }
}

} /* end namespace */

int celltype_nface(int32_t tpn) { return cell_faces(tpn).nface; }

Mesh::Mesh(int ndim_in, int32_t nnode_in, int32_t nface_in, int32_t ncell_in)
    : ndim(ndim_in)
    , nnode(nnode_in)
    , nface(nface_in)
    , ncell(ncell_in)
{
    // StaticMesh.hpp:290-311
    ndcrd = make2<double>(nnode, ndim, 0.0);
    fccnd = make2<double>(nface, ndim, 0.0);
    fcnml = make2<double>(nface, ndim, 0.0);
    fcara = make1<double>(nface, 0.0);
    clcnd = make2<double>(ncell, ndim, 0.0);
    clvol = make1<double>(ncell, 0.0);
    fctpn = make1<int32_t>(nface, 0);
    cltpn = make1<int32_t>(ncell, 0);
    clgrp = make1<int32_t>(ncell, 0);
    fcnds = make2<int32_t>(nface, FCMND + 1, 0);
    fccls = make2<int32_t>(nface, FCREL, 0);
    clnds = make2<int32_t>(ncell, CLMND + 1, 0);
    clfcs = make2<int32_t>(ncell, CLMFC + 1, 0);
    ednds = make2<int32_t>(0, 2, 0);
    bndfcs = make2<int32_t>(0, BFREL, 0);
}

// StaticMesh.hpp:363-374
void Mesh::build_interior(bool do_metric, bool do_edge)
{
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

// FaceBuilder (StaticMesh_interior.cpp:16-543) and build_faces_from_cells (:551-564).
void Mesh::build_faces_from_cells()
{
    // count_mface (:340-349)
    int64_t mface = 0;
    for (int32_t icl = 0; icl < ncell; ++icl)
    {
        mface += celltype_nface(cltpn(icl));
    }
    Table<int32_t> bclfcs = make2<int32_t>(ncell, CLMFC + 1, -1);
    Table<int32_t> bfctpn = make1<int32_t>(mface, -1);
    Table<int32_t> bfcnds = make2<int32_t>(mface, FCMND + 1, -1);
    Table<int32_t> bfccls = make2<int32_t>(mface, FCREL, -1);

    // populate (:351-386)
    int32_t ifc = 0;
    for (int32_t icl = 0; icl < ncell; ++icl)
    {
        CellFaces const & cf = cell_faces(cltpn(icl));
        if (cf.nface == 0)
        {
            continue;
        }
        bclfcs(icl, 0) = cf.nface;
        for (int k = 0; k < cf.nface; ++k)
        {
            FaceDef const & fd = cf.face[k];
            bfctpn(ifc) = fd.type;
            bfcnds(ifc, 0) = fd.nnd;
            bclfcs(icl, k + 1) = ifc;
            for (int m = 0; m < fd.nnd; ++m)
            {
                bfcnds(ifc, m + 1) = clnds(icl, fd.nd[m] + 1);
            }
            ++ifc;
        }
    }

    // make_ndfcs (:388-423): faces touching each node, ascending face order.
    std::vector<std::vector<int32_t>> ndfcs(static_cast<size_t>(nnode));
    for (int64_t f = 0; f < mface; ++f)
    {
        for (int m = 1; m <= bfcnds(f, 0); ++m)
        {
            ndfcs[static_cast<size_t>(bfcnds(f, m))].push_back(static_cast<int32_t>(f));
        }
    }

    // make_dedupmap (:428-477)
    auto check_match = [&](int64_t a, int64_t b)
    {
        int cond = 0;
        for (int jnf = 1; jnf <= bfcnds(b, 0); ++jnf)
        {
            for (int inf = 1; inf <= bfcnds(a, 0); ++inf)
            {
                if (bfcnds(a, inf) == bfcnds(b, jnf))
                {
                    ++cond;
                    break;
                }
            }
        }
        return cond;
    };
    std::vector<int64_t> dedup(static_cast<size_t>(mface));
    std::iota(dedup.begin(), dedup.end(), 0);
    for (int64_t f = 0; f < mface; ++f)
    {
        if (dedup[f] != f)
        {
            continue;
        }
        int32_t const ind = bfcnds(f, 1);
        for (int32_t g : ndfcs[static_cast<size_t>(ind)])
        {
            if (g != f && bfctpn(g) == bfctpn(f) && check_match(f, g) == bfcnds(g, 0))
            {
                dedup[g] = f;
            }
        }
    }

    // remap_face (:479-523)
    std::vector<int32_t> map2(static_cast<size_t>(mface));
    int32_t jfc = 0;
    for (int64_t f = 0; f < mface; ++f)
    {
        if (dedup[f] == f)
        {
            for (int m = 0; m <= FCMND; ++m)
            {
                bfcnds(jfc, m) = bfcnds(f, m);
            }
            bfctpn(jfc) = bfctpn(f);
            map2[f] = jfc;
            ++jfc;
        }
        else
        {
            map2[f] = map2[dedup[f]];
        }
    }
    nface = jfc;
    for (int32_t icl = 0; icl < ncell; ++icl)
    {
        for (int ifl = 1; ifl <= bclfcs(icl, 0); ++ifl)
        {
            int32_t const nf = map2[bclfcs(icl, ifl)];
            bclfcs(icl, ifl) = nf;
            if (bfccls(nf, 0) == -1)
            {
                bfccls(nf, 0) = icl;
            }
            else if (bfccls(nf, 1) == -1)
            {
                bfccls(nf, 1) = icl;
            }
        }
    }

    // rebuild_* and the metric remakes (:525-564)
    fctpn = make1<int32_t>(nface, 0);
    fcnds = make2<int32_t>(nface, FCMND + 1, 0);
    fccls = make2<int32_t>(nface, FCREL, 0);
    for (int32_t f = 0; f < nface; ++f)
    {
        fctpn(f) = bfctpn(f);
        for (int m = 0; m <= FCMND; ++m)
        {
            fcnds(f, m) = bfcnds(f, m);
        }
        for (int m = 0; m < FCREL; ++m)
        {
            fccls(f, m) = bfccls(f, m);
        }
    }
    fccnd = make2<double>(nface, ndim, 0.0);
    fcnml = make2<double>(nface, ndim, 0.0);
    fcara = make1<double>(nface, 0.0);
    for (int32_t icl = 0; icl < ncell; ++icl)
    {
        for (int m = 0; m <= CLMFC; ++m)
        {
            clfcs(icl, m) = bclfcs(icl, m);
        }
    }
}

// StaticMesh_interior.cpp:566-610
void Mesh::build_edge()
{
    std::set<std::pair<int32_t, int32_t>> known;
    std::vector<std::pair<int32_t, int32_t>> edges;
    for (int32_t f = 0; f < nface; ++f)
    {
        int32_t const nnd = fcnds(f, 0);
        for (int32_t inf = 1; inf <= nnd; ++inf)
        {
            int32_t const inx = (nnd == inf) ? 1 : inf + 1;
            int32_t a = fcnds(f, inf);
            int32_t b = fcnds(f, inx);
            if (a > b)
            {
                std::swap(a, b);
            }
            if (known.insert({a, b}).second)
            {
                edges.emplace_back(a, b);
            }
        }
    }
    ednds = make2<int32_t>(static_cast<ptrdiff_t>(edges.size()), 2, 0);
    for (size_t i = 0; i < edges.size(); ++i)
    {
        ednds(static_cast<ptrdiff_t>(i), 0) = edges[i].first;
        ednds(static_cast<ptrdiff_t>(i), 1) = edges[i].second;
    }
}

namespace
{

// 3D face centroid: area-weighted sub-triangle centroids about the node
// average (StaticMesh_interior.cpp:650-701, StaticMesh_boundary.cpp:396-446).
void face_centroid_3d(Mesh & m, int32_t f)
{
    std::array<std::array<double, 3>, FCMND + 2> cfd;
    cfd[0][0] = cfd[0][1] = cfd[0][2] = 0.0;
    int32_t const nnd = m.fcnds(f, 0);
    for (int32_t inf = 1; inf <= nnd; ++inf)
    {
        int32_t const ind = m.fcnds(f, inf);
        for (int d = 0; d < 3; ++d)
        {
            cfd[inf][d] = m.ndcrd(ind, d);
            cfd[0][d] += m.ndcrd(ind, d);
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
    m.fccnd(f, 0) = m.fccnd(f, 1) = m.fccnd(f, 2) = 0.0;
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
        double const vob = std::sqrt(dw0 * dw0 + dw1 * dw1 + dw2 * dw2);
        for (int d = 0; d < 3; ++d)
        {
            m.fccnd(f, d) += crd[d] * vob;
        }
        voc += vob;
    }
    for (int d = 0; d < 3; ++d)
    {
        m.fccnd(f, d) /= voc;
    }
}

void face_centroid_2d(Mesh & m, int32_t f)
{
    int32_t ind = m.fcnds(f, 1);
    m.fccnd(f, 0) = m.ndcrd(ind, 0);
    m.fccnd(f, 1) = m.ndcrd(ind, 1);
    ind = m.fcnds(f, 2);
    m.fccnd(f, 0) += m.ndcrd(ind, 0);
    m.fccnd(f, 1) += m.ndcrd(ind, 1);
    m.fccnd(f, 0) /= 2;
    m.fccnd(f, 1) /= 2;
}

// StaticMesh_interior.cpp:705-721, StaticMesh_boundary.cpp:450-466
void face_normal_2d(Mesh & m, int32_t f)
{
    int32_t const i1 = m.fcnds(f, 1);
    int32_t const i2 = m.fcnds(f, 2);
    m.fcnml(f, 0) = m.ndcrd(i2, 1) - m.ndcrd(i1, 1);
    m.fcnml(f, 1) = m.ndcrd(i1, 0) - m.ndcrd(i2, 0);
    m.fcara(f) = std::sqrt(m.fcnml(f, 0) * m.fcnml(f, 0) + m.fcnml(f, 1) * m.fcnml(f, 1));
    m.fcnml(f, 0) /= m.fcara(f);
    m.fcnml(f, 1) /= m.fcara(f);
}

// StaticMesh_interior.cpp:722-768, StaticMesh_boundary.cpp:467-503
void face_normal_3d(Mesh & m, int32_t f)
{
    std::array<std::array<double, 3>, FCMND> rv;
    int32_t const nnd = m.fcnds(f, 0);
    for (int32_t inf = 0; inf < nnd; ++inf)
    {
        int32_t const ind = m.fcnds(f, inf + 1);
        for (int d = 0; d < 3; ++d)
        {
            rv[inf][d] = m.ndcrd(ind, d) - m.fccnd(f, d);
        }
    }
    m.fcnml(f, 0) = rv[nnd - 1][1] * rv[0][2] - rv[nnd - 1][2] * rv[0][1];
    m.fcnml(f, 1) = rv[nnd - 1][2] * rv[0][0] - rv[nnd - 1][0] * rv[0][2];
    m.fcnml(f, 2) = rv[nnd - 1][0] * rv[0][1] - rv[nnd - 1][1] * rv[0][0];
    for (int32_t k = 1; k < nnd; ++k)
    {
        m.fcnml(f, 0) += rv[k - 1][1] * rv[k][2] - rv[k - 1][2] * rv[k][1];
        m.fcnml(f, 1) += rv[k - 1][2] * rv[k][0] - rv[k - 1][0] * rv[k][2];
        m.fcnml(f, 2) += rv[k - 1][0] * rv[k][1] - rv[k - 1][1] * rv[k][0];
    }
    m.fcara(f) = std::sqrt(m.fcnml(f, 0) * m.fcnml(f, 0) + m.fcnml(f, 1) * m.fcnml(f, 1) + m.fcnml(f, 2) * m.fcnml(f, 2));
    m.fcnml(f, 0) /= m.fcara(f);
    m.fcnml(f, 1) /= m.fcara(f);
    m.fcnml(f, 2) /= m.fcara(f);
    m.fcara(f) /= 2.0;
}

// Cell centroid from the pyramids on each face about the node average
// (StaticMesh_interior.cpp:802-836,882-922; StaticMesh_boundary.cpp:506-592).
void cell_centroid(Mesh & m, int32_t icl)
{
    int const nd = m.ndim;
    double crd[3] = {0.0, 0.0, 0.0};
    int32_t const nnd = m.clnds(icl, 0);
    for (int32_t inc = 1; inc <= nnd; ++inc)
    {
        int32_t const ind = m.clnds(icl, inc);
        for (int d = 0; d < nd; ++d)
        {
            crd[d] += m.ndcrd(ind, d);
        }
    }
    for (int d = 0; d < nd; ++d)
    {
        crd[d] /= nnd;
    }
    double voc = 0.0;
    for (int d = 0; d < nd; ++d)
    {
        m.clcnd(icl, d) = 0.0;
    }
    int32_t const nfc = m.clfcs(icl, 0);
    for (int32_t ifl = 1; ifl <= nfc; ++ifl)
    {
        int32_t const f = m.clfcs(icl, ifl);
        double du[3];
        double dot = 0.0;
        for (int d = 0; d < nd; ++d)
        {
            du[d] = crd[d] - m.fccnd(f, d);
        }
        if (nd == 2)
        {
            dot = du[0] * m.fcnml(f, 0) + du[1] * m.fcnml(f, 1);
        }
        else
        {
            dot = du[0] * m.fcnml(f, 0) + du[1] * m.fcnml(f, 1) + du[2] * m.fcnml(f, 2);
        }
        double const vob = std::fabs(dot * m.fcara(f));
        voc += vob;
        for (int d = 0; d < nd; ++d)
        {
            double const dv = m.fccnd(f, d) + du[d] / (nd + 1);
            m.clcnd(icl, d) += dv * vob;
        }
    }
    for (int d = 0; d < nd; ++d)
    {
        m.clcnd(icl, d) /= voc;
    }
}

} /* end namespace */

// StaticMesh_interior.cpp:623-976.  Only the centroid branch of the cell
// centre is restated: use_incenter is false for every StaticMesh
// (StaticMesh.hpp:418) and nothing sets it.
void Mesh::calc_metric()
{
    for (int32_t f = 0; f < nface; ++f)
    {
        if (ndim == 2)
        {
            face_centroid_2d(*this, f);
        }
        else if (ndim == 3)
        {
            face_centroid_3d(*this, f);
        }
    }
    for (int32_t f = 0; f < nface; ++f)
    {
        if (ndim == 2)
        {
            face_normal_2d(*this, f);
        }
        else if (ndim == 3)
        {
            face_normal_3d(*this, f);
        }
    }
    if (ndim == 2 || ndim == 3)
    {
        for (int32_t icl = 0; icl < ncell; ++icl)
        {
            cell_centroid(*this, icl);
        }
    }
    // Volumes, with reorder_face so that every normal points out of fccls[0]
    // (:927-975).
    for (int32_t icl = 0; icl < ncell; ++icl)
    {
        clvol(icl) = 0.0;
        int32_t const nfc = clfcs(icl, 0);
        for (int32_t it = 1; it <= nfc; ++it)
        {
            int32_t const f = clfcs(icl, it);
            double vol = 0.0;
            for (int d = 0; d < ndim; ++d)
            {
                vol += (fccnd(f, d) - clcnd(icl, d)) * fcnml(f, d);
            }
            vol *= fcara(f);
            bool flip = false;
            if (vol < 0.0)
            {
                flip = (fccls(f, 0) == icl);
                vol = -vol;
            }
            else
            {
                flip = (fccls(f, 0) != icl);
            }
            if (flip)
            {
                int32_t const nnd = fcnds(f, 0);
                int32_t rev[FCMND];
                for (int32_t j = 0; j < nnd; ++j)
                {
                    rev[j] = fcnds(f, nnd - j);
                }
                for (int32_t j = 0; j < nnd; ++j)
                {
                    fcnds(f, j + 1) = rev[j];
                }
                for (int d = 0; d < ndim; ++d)
                {
                    fcnml(f, d) = -fcnml(f, d);
                }
            }
            clvol(icl) += vol;
        }
        clvol(icl) /= ndim;
    }
}

// StaticMesh_boundary.cpp:14-78 (name and face validation).
void Mesh::add_bc(std::string const & name, std::vector<int32_t> const & faces)
{
    if (nbound != 0)
    {
        set_error(ERR_RUNTIME, "StaticMesh::add_bc: boundary is already built");
        throw std::runtime_error("add_bc");
    }
    if (name.empty())
    {
        set_error(ERR_VALUE, "StaticMesh::add_bc: empty name");
        throw std::invalid_argument("add_bc");
    }
    for (auto const & b : bcs)
    {
        if (b.name == name)
        {
            set_error(ERR_VALUE, "StaticMesh::add_bc: duplicated name");
            throw std::invalid_argument("add_bc");
        }
    }
    std::vector<bool> claimed(static_cast<size_t>(nface), false);
    for (auto const & b : bcs)
    {
        for (int32_t f : b.facn0)
        {
            claimed[static_cast<size_t>(f)] = true;
        }
    }
    for (int32_t f : faces)
    {
        if (f < 0 || f >= nface)
        {
            set_error(ERR_VALUE, "StaticMesh::add_bc: face out of range");
            throw std::invalid_argument("add_bc");
        }
        if (fccls(f, 1) >= 0)
        {
            set_error(ERR_VALUE, "StaticMesh::add_bc: face is not a boundary face");
            throw std::invalid_argument("add_bc");
        }
        if (claimed[static_cast<size_t>(f)])
        {
            set_error(ERR_VALUE, "StaticMesh::add_bc: face is already claimed");
            throw std::invalid_argument("add_bc");
        }
        claimed[static_cast<size_t>(f)] = true;
    }
    MeshBc b;
    b.name = name;
    b.facn0 = faces;
    b.facn1.assign(faces.size(), -1);
    bcs.push_back(std::move(b));
}

// StaticMesh_boundary.cpp:103-182
void Mesh::build_boundary()
{
    if (nbound != 0)
    {
        set_error(ERR_RUNTIME, "StaticMesh::build_boundary: boundary is already built");
        throw std::runtime_error("build_boundary");
    }
    for (int32_t f = 0; f < nface; ++f)
    {
        if (fccls(f, 1) < 0)
        {
            nbound += 1;
        }
    }
    bndfcs = make2<int32_t>(nbound, BFREL, -1);
    std::vector<int32_t> allfacn;
    allfacn.reserve(static_cast<size_t>(nbound));
    for (int32_t f = 0; f < nface; ++f)
    {
        if (fccls(f, 1) < 0)
        {
            allfacn.push_back(f);
        }
    }
    std::vector<bool> specified(static_cast<size_t>(nbound), false);
    int32_t ibfc = 0;
    int32_t nleft = nbound;
    for (size_t ib = 0; ib < bcs.size(); ++ib)
    {
        MeshBc & b = bcs[ib];
        for (size_t k = 0; k < b.facn0.size(); ++k)
        {
            bndfcs(ibfc, 0) = b.facn0[k];
            bndfcs(ibfc, 1) = static_cast<int32_t>(ib);
            b.facn1[k] = ibfc;
            auto it = std::find(allfacn.begin(), allfacn.end(), b.facn0[k]);
            if (it != allfacn.end())
            {
                specified[static_cast<size_t>(it - allfacn.begin())] = true;
                --nleft;
            }
            ++ibfc;
        }
    }
    if (nleft != 0)
    {
        MeshBc b;
        b.name = "<NONAME>";
        int32_t const ib = static_cast<int32_t>(bcs.size());
        for (size_t s = 0; s < allfacn.size(); ++s)
        {
            if (!specified[s])
            {
                bndfcs(ibfc, 0) = allfacn[s];
                bndfcs(ibfc, 1) = ib;
                b.facn0.push_back(allfacn[s]);
                b.facn1.push_back(ibfc);
                ++ibfc;
            }
        }
        bcs.push_back(std::move(b));
    }
}

// StaticMesh_boundary.cpp:185-261
void Mesh::build_ghost()
{
    int64_t gnode = 0;
    int64_t gface = 0;
    for (int32_t ib = 0; ib < nbound; ++ib)
    {
        int32_t const f = bndfcs(ib, 0);
        int32_t const icl = fccls(f, 0);
        gface += celltype_nface(cltpn(icl)) - 1;
        gnode += clnds(icl, 0) - fcnds(f, 0);
    }
    ngstnode = static_cast<int32_t>(gnode);
    ngstface = static_cast<int32_t>(gface);
    ngstcell = nbound;

    auto grow2d = [](Table<double> & t, int32_t ngst, int32_t nb, int w)
    {
        Table<double> n = make2<double>(ngst + nb, w, 0.0, ngst);
        for (int32_t i = 0; i < nb; ++i)
        {
            for (int j = 0; j < w; ++j)
            {
                n(i, j) = t(i, j);
            }
        }
        t = std::move(n);
    };
    auto grow1d = [](Table<double> & t, int32_t ngst, int32_t nb)
    {
        Table<double> n = make1<double>(ngst + nb, 0.0, ngst);
        for (int32_t i = 0; i < nb; ++i)
        {
            n(i) = t(i);
        }
        t = std::move(n);
    };
    auto grow2i = [](Table<int32_t> & t, int32_t ngst, int32_t nb, int w, int32_t init)
    {
        Table<int32_t> n = make2<int32_t>(ngst + nb, w, init, ngst);
        for (int32_t i = 0; i < nb; ++i)
        {
            for (int j = 0; j < w; ++j)
            {
                n(i, j) = t(i, j);
            }
        }
        t = std::move(n);
    };
    auto grow1i = [](Table<int32_t> & t, int32_t ngst, int32_t nb, int32_t init)
    {
        Table<int32_t> n = make1<int32_t>(ngst + nb, init, ngst);
        for (int32_t i = 0; i < nb; ++i)
        {
            n(i) = t(i);
        }
        t = std::move(n);
    };
    grow2d(ndcrd, ngstnode, nnode, ndim);
    grow2d(fccnd, ngstface, nface, ndim);
    grow2d(fcnml, ngstface, nface, ndim);
    grow1d(fcara, ngstface, nface);
    grow2d(clcnd, ngstcell, ncell, ndim);
    grow1d(clvol, ngstcell, ncell);
    grow1i(fctpn, ngstface, nface, 0);
    grow1i(cltpn, ngstcell, ncell, 0);
    grow1i(clgrp, ngstcell, ncell, -1);
    grow2i(fcnds, ngstface, nface, FCMND + 1, -1);
    grow2i(fccls, ngstface, nface, FCREL, -1);
    grow2i(clnds, ngstcell, ncell, CLMND + 1, -1);
    grow2i(clfcs, ngstcell, ncell, CLMFC + 1, -1);
    fill_ghost();
}

// StaticMesh_boundary.cpp:281-627
void Mesh::fill_ghost()
{
    std::vector<int32_t> gstndmap(static_cast<size_t>(nnode), nnode);
    int32_t ignd = -1;
    int32_t igfc = -1;
    for (int32_t igcl = -1; igcl >= -ngstcell; --igcl)
    {
        int32_t const ibfc = bndfcs(-igcl - 1, 0);
        int32_t const icl = fccls(ibfc, 0);
        cltpn(igcl) = cltpn(icl);
        clgrp(igcl) = clgrp(icl);
        for (int inl = 0; inl <= CLMND; ++inl)
        {
            clnds(igcl, inl) = clnds(icl, inl);
        }
        for (int32_t inl = 1; inl <= clnds(icl, 0); ++inl)
        {
            int32_t const ind = clnds(icl, inl);
            bool found = false;
            for (int32_t inf = 1; inf <= fcnds(ibfc, 0); ++inf)
            {
                if (ind == fcnds(ibfc, inf))
                {
                    found = true;
                    break;
                }
            }
            if (!found)
            {
                gstndmap[static_cast<size_t>(ind)] = ignd;
                clnds(igcl, inl) = ignd;
                double dist = 0.0;
                for (int d = 0; d < ndim; ++d)
                {
                    dist += (fccnd(ibfc, d) - ndcrd(ind, d)) * fcnml(ibfc, d);
                }
                for (int d = 0; d < ndim; ++d)
                {
                    ndcrd(ignd, d) = ndcrd(ind, d) + 2 * dist * fcnml(ibfc, d);
                }
                ignd -= 1;
            }
        }
        fccls(ibfc, 1) = igcl;
        for (int ifl = 0; ifl <= CLMFC; ++ifl)
        {
            clfcs(igcl, ifl) = clfcs(icl, ifl);
        }
        for (int32_t ifl = 1; ifl <= clfcs(icl, 0); ++ifl)
        {
            int32_t const f = clfcs(icl, ifl);
            if (f == ibfc)
            {
                continue;
            }
            fctpn(igfc) = fctpn(f);
            fccls(igfc, 0) = igcl;
            clfcs(igcl, ifl) = igfc;
            for (int inf = 0; inf <= FCMND; ++inf)
            {
                fcnds(igfc, inf) = fcnds(f, inf);
            }
            for (int32_t inf = 1; inf <= fcnds(igfc, 0); ++inf)
            {
                int32_t const ind = fcnds(igfc, inf);
                if (gstndmap[static_cast<size_t>(ind)] != nnode)
                {
                    fcnds(igfc, inf) = gstndmap[static_cast<size_t>(ind)];
                }
            }
            igfc -= 1;
        }
        for (int32_t inl = 1; inl <= clnds(icl, 0); ++inl)
        {
            gstndmap[static_cast<size_t>(clnds(icl, inl))] = nnode;
        }
    }

    for (int32_t f = -1; f >= -ngstface; --f)
    {
        if (ndim == 2)
        {
            face_centroid_2d(*this, f);
        }
        else if (ndim == 3)
        {
            face_centroid_3d(*this, f);
        }
    }
    for (int32_t f = -1; f >= -ngstface; --f)
    {
        if (ndim == 2)
        {
            face_normal_2d(*this, f);
        }
        else if (ndim == 3)
        {
            face_normal_3d(*this, f);
        }
    }
    if (ndim == 2 || ndim == 3)
    {
        for (int32_t icl = -1; icl >= -ngstcell; --icl)
        {
            cell_centroid(*this, icl);
        }
    }
    for (int32_t icl = -1; icl >= -ngstcell; --icl)
    {
        clvol(icl) = 0.0;
        for (int32_t it = 1; it <= clfcs(icl, 0); ++it)
        {
            int32_t const f = clfcs(icl, it);
            double vol = 0.0;
            for (int d = 0; d < ndim; ++d)
            {
                vol += (fccnd(f, d) - clcnd(icl, d)) * fcnml(f, d);
            }
            vol *= fcara(f);
            if (vol < 0.0)
            {
                if (fccls(f, 0) == icl)
                {
                    for (int d = 0; d < ndim; ++d)
                    {
                        fcnml(f, d) = -fcnml(f, d);
                    }
                }
                vol = -vol;
            }
            clvol(icl) += vol;
        }
        clvol(icl) /= ndim;
    }
}

} /* end namespace orc */
