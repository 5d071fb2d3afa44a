#pragma once
/*
 * Device-resident unstructured mesh: the counterpart of solvcon's StaticMesh
 * (cpp/solvcon/mesh/StaticMesh.hpp:264-456).  Tables keep the reference
 * layout (row-major, ghost rows stored first) so they can be handed to the
 * drop-in API unchanged; every builder runs on the GPU.
 */

#include "common.hpp"

#include <string>
#include <vector>

namespace scamd
{

constexpr int FCMND = 4; // StaticMesh.hpp:141-145
constexpr int CLMND = 8;
constexpr int CLMFC = 6;
constexpr int FCREL = 4;
constexpr int BFREL = 3;

enum CellTypeId : int32_t
{
    CT_NONE = 0,
    CT_POINT = 1,
    CT_LINE = 2,
    CT_QUAD = 3,
    CT_TRI = 4,
    CT_HEX = 5,
    CT_TET = 6,
    CT_PRISM = 7,
    CT_PYRAMID = 8,
};

// Table ids of the C ABI (scmesh_download/upload).
enum MeshArray : int
{
    MA_NDCRD = 0,
    MA_FCCND,
    MA_FCNML,
    MA_FCARA,
    MA_CLCND,
    MA_CLVOL,
    MA_FCTPN,
    MA_CLTPN,
    MA_CLGRP,
    MA_FCNDS,
    MA_FCCLS,
    MA_CLNDS,
    MA_CLFCS,
    MA_EDNDS,
    MA_BNDFCS,
    MA_COUNT,
};

struct MeshBc
{
    std::string name;
    std::vector<int32_t> facn0;
    std::vector<int32_t> facn1;
};

struct DMesh
{
    int ndim = 0;
    int32_t nnode = 0, nface = 0, ncell = 0;
    int32_t nbound = 0, ngstnode = 0, ngstface = 0, ngstcell = 0;
    int32_t nedge = 0;
    StreamHolder sh;                // declared before the buffers: freed last
    hipStream_t stream = sh.s;

    DevBuf<double> ndcrd, fccnd, fcnml, fcara, clcnd, clvol;
    DevBuf<int32_t> fctpn, cltpn, clgrp, fcnds, fccls, clnds, clfcs, ednds, bndfcs;
    std::vector<MeshBc> bcs;

    DMesh(int nd, int32_t nn, int32_t nf, int32_t nc);
    ~DMesh();

    // Pointers to row 0 of the body (ghost rows at negative indices).
    double * ndcrd_b() const { return ndcrd.ptr + static_cast<int64_t>(ngstnode) * ndim; }
    double * fccnd_b() const { return fccnd.ptr + static_cast<int64_t>(ngstface) * ndim; }
    double * fcnml_b() const { return fcnml.ptr + static_cast<int64_t>(ngstface) * ndim; }
    double * fcara_b() const { return fcara.ptr + ngstface; }
    double * clcnd_b() const { return clcnd.ptr + static_cast<int64_t>(ngstcell) * ndim; }
    double * clvol_b() const { return clvol.ptr + ngstcell; }
    int32_t * fctpn_b() const { return fctpn.ptr + ngstface; }
    int32_t * cltpn_b() const { return cltpn.ptr + ngstcell; }
    int32_t * clgrp_b() const { return clgrp.ptr + ngstcell; }
    int32_t * fcnds_b() const { return fcnds.ptr + static_cast<int64_t>(ngstface) * (FCMND + 1); }
    int32_t * fccls_b() const { return fccls.ptr + static_cast<int64_t>(ngstface) * FCREL; }
    int32_t * clnds_b() const { return clnds.ptr + static_cast<int64_t>(ngstcell) * (CLMND + 1); }
    int32_t * clfcs_b() const { return clfcs.ptr + static_cast<int64_t>(ngstcell) * (CLMFC + 1); }

    void sync() { SC_HIP(hipStreamSynchronize(stream)); }

    // shape of a table: rows (ghost + body), columns, ghost rows, is-int
    void table(int id, int64_t & rows, int64_t & cols, int64_t & ngh, bool & is_int, void *& ptr) const;

    void build_interior(bool do_metric, bool do_edge);
    void build_faces_from_cells();
    void calc_metric();
    void build_edge();
    void add_bc(std::string const & name, std::vector<int32_t> const & faces);
    void build_boundary();
    void build_ghost();
    // fccls(ifc, 1) for the listed faces, in order.
    std::vector<int32_t> fcjcl_of(std::vector<int32_t> const & faces);
};

} /* end namespace scamd */
