#pragma once
/*
 * Oracle (test infrastructure only): CPU restatement of solvcon's EulerCore
 * (cpp/solvcon/multidim/euler.hpp, euler_ce.cpp, euler_march.cpp,
 * euler_solution.cpp, euler_gradient.cpp, euler_bc.cpp, GradientElement.*).
 */

#include "mesh.hpp"

#include <array>
#include <memory>
#include <vector>

namespace orc
{

enum BcKind : int32_t
{
    BC_NONREFL = 0,
    BC_SLIPWALL = 1,
    BC_INLET = 2,
};

struct EulerBoundary
{
    int32_t kind = BC_NONREFL;
    std::vector<int32_t> faces;
    std::vector<double> value;
};

// GradientElementType (GradientElement.hpp:30-45, GradientElement.cpp:16-79).
struct GeType
{
    int32_t clnfc = -1;
    int32_t nfge = -1;
    double nfge_inverse = 0.0;
    std::array<std::array<int32_t, 3>, 8> faces{};
};
GeType const & getype(int32_t tpn);

struct GradElem
{
    GeType const * ge = nullptr;
    int32_t icl = 0;
    int ndim = 0;
    int32_t clnfc = 0;
    std::array<int32_t, CLMFC> rcls{};
    std::array<std::array<double, 3>, CLMFC> idis{};
    std::array<std::array<double, 3>, CLMFC> jdis{};

    GradElem(Mesh const & m, Table<double> const & cecnd, int32_t icl, double tau);
    std::array<std::array<double, 3>, 3> displacement_matrix(int32_t ifge) const;
    std::array<double, 3> solve_gradient(int32_t ifge, std::array<double, 3> const & udf) const;
};

double ge_determinant(std::array<std::array<double, 3>, 3> const & a, int ndim);
std::array<std::array<double, 3>, 3> ge_adjugate(std::array<std::array<double, 3>, 3> const & a, int ndim);

struct Euler
{
    std::shared_ptr<Mesh> mesh;
    double dt = 0.0;
    int ndim = 0;
    int32_t ncell = 0;
    int32_t ngstcell = 0;
    int32_t neq = 0;
    double sigma0 = 3.0;
    double taumin = 0.0;
    double tauscale = 1.0;
    int nthreads = 1;

    Table<double> cevol, cecnd, sfcnd, sfnml;
    Table<double> so0c, so0n, so0t, so1c, so1n, stm, cflo, cflc, gamma;
    std::vector<EulerBoundary> boundaries;

    Euler(std::shared_ptr<Mesh> m, double time_increment);

    void prepare_ce();
    void init_solution(double gamma, double rho, std::array<double, 3> const & v, double p);
    void update();
    void calc_solt();
    void calc_soln();
    void calc_cfl();
    void calc_dsoln();
    void bc_soln();
    void bc_dsoln();
    void march_substep();
    void march(int64_t steps);
    void add_bc(int32_t kind, std::vector<int32_t> const & faces, std::vector<double> const & value);
    std::array<std::array<double, 3>, 3> normal_matrix(int32_t ifc) const;
};

} /* end namespace orc */
