#pragma once
/*
 * Oracle (test infrastructure only): CPU restatement of solvcon's StaticMesh
 * builders (cpp/solvcon/mesh/StaticMesh.hpp, StaticMesh_interior.cpp,
 * StaticMesh_boundary.cpp).
 */

#include "common.hpp"

#include <string>
#include <utility>
#include <vector>

namespace orc
{

// Cell type ids, StaticMesh.hpp:35-45.
enum CellTypeId : int32_t
{
    NONCELLTYPE = 0,
    POINT = 1,
    LINE = 2,
    QUADRILATERAL = 3,
    TRIANGLE = 4,
    HEXAHEDRON = 5,
    TETRAHEDRON = 6,
    PRISM = 7,
    PYRAMID = 8,
};

constexpr int FCMND = 4; // StaticMesh.hpp:141-145
constexpr int CLMND = 8;
constexpr int CLMFC = 6;
constexpr int FCREL = 4;
constexpr int BFREL = 3;

// Number of faces of a cell type: CellType::nface() (StaticMesh.hpp:74,116-124).
int celltype_nface(int32_t tpn);

struct MeshBc
{
    std::string name;
    std::vector<int32_t> facn0; // face index in block
    std::vector<int32_t> facn1; // face index in bndfcs
};

struct Mesh
{
    int ndim = 0;
    int32_t nnode = 0;
    int32_t nface = 0;
    int32_t ncell = 0;
    int32_t nbound = 0;
    int32_t ngstnode = 0;
    int32_t ngstface = 0;
    int32_t ngstcell = 0;

    Table<double> ndcrd, fccnd, fcnml, fcara, clcnd, clvol;
    Table<int32_t> fctpn, cltpn, clgrp, fcnds, fccls, clnds, clfcs, ednds, bndfcs;
    std::vector<MeshBc> bcs;

    Mesh(int ndim_in, int32_t nnode_in, int32_t nface_in, int32_t ncell_in);

    int32_t fcrcl(int32_t ifc, int32_t icl) const { return fccls(ifc, 0) + fccls(ifc, 1) - icl; }

    void build_interior(bool do_metric, bool do_edge);
    void build_faces_from_cells();
    void calc_metric();
    void build_edge();
    void add_bc(std::string const & name, std::vector<int32_t> const & faces);
    void build_boundary();
    void build_ghost();

private:
    void fill_ghost();
};

} /* end namespace orc */
