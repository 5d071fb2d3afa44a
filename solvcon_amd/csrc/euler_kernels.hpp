#pragma once
/*
 * Device arithmetic of the multidimensional Euler CESE march: the flux
 * Jacobian, the per-cell CE flux integral (soln), the CFL and energy
 * clamp (cfl), the gradient element with the W-1/2 weighting and the W-3/4
 * limiter (dsoln), and the ghost-cell boundary conditions.  Every routine
 * evaluates the reference formula in the reference order; the build uses
 * -ffp-contract=off, so results match the CPU march bit for bit.
 *
 * Cells are processed per cell type: the face count, the sub-face count of
 * each face and the gradient-element table are compile-time constants of
 * the type (Traits<T>), so all loops unroll into straight-line code.
 */

#include "mesh.hpp"

#include <type_traits>

namespace scamd
{

constexpr double TINYJ = 1.e-60; // euler_march.cpp:30,89; euler_solution.cpp:108
constexpr double ALMOST_ZERO = 1.e-200; // euler_gradient.cpp:28

enum BcKind : int32_t
{
    BC_NONE = -1,
    BC_NONREFL = 0,
    BC_SLIPWALL = 1,
    BC_INLET = 2,
};

// Per-cell-type constants: faces, sub-faces per face (the face's node
// count) and the fundamental gradient elements (GradientElement.cpp:16-79,
// 1-based face positions).
template <int T>
struct Traits;

template <>
struct Traits<CT_TRI>
{
    static constexpr int ND = 2, NF = 3, S = 6, NFGE = 3;
    __host__ __device__ static constexpr int nsub(int) { return 2; }
    __host__ __device__ static constexpr int soff(int f) { return 2 * f; }
    __host__ __device__ static constexpr int fge(int e, int v)
    {
        return e == 0 ? (v == 0 ? 1 : 2) : e == 1 ? (v == 0 ? 2 : 3) : (v == 0 ? 3 : 1);
    }
};

template <>
struct Traits<CT_QUAD>
{
    static constexpr int ND = 2, NF = 4, S = 8, NFGE = 4;
    __host__ __device__ static constexpr int nsub(int) { return 2; }
    __host__ __device__ static constexpr int soff(int f) { return 2 * f; }
    __host__ __device__ static constexpr int fge(int e, int v) { return v == 0 ? e + 1 : (e + 1) % 4 + 1; }
};

template <>
struct Traits<CT_TET>
{
    static constexpr int ND = 3, NF = 4, S = 12, NFGE = 4;
    __host__ __device__ static constexpr int nsub(int) { return 3; }
    __host__ __device__ static constexpr int soff(int f) { return 3 * f; }
    __host__ __device__ static constexpr int fge(int e, int v)
    {
        constexpr int t[4][3] = {{3, 1, 2}, {2, 1, 4}, {4, 1, 3}, {2, 4, 3}};
        return t[e][v];
    }
};

template <>
struct Traits<CT_HEX>
{
    static constexpr int ND = 3, NF = 6, S = 24, NFGE = 8;
    __host__ __device__ static constexpr int nsub(int) { return 4; }
    __host__ __device__ static constexpr int soff(int f) { return 4 * f; }
    __host__ __device__ static constexpr int fge(int e, int v)
    {
        constexpr int t[8][3] = {{2, 3, 5}, {6, 3, 2}, {4, 3, 6}, {5, 3, 4},
                                 {5, 1, 2}, {2, 1, 6}, {6, 1, 4}, {4, 1, 5}};
        return t[e][v];
    }
};

template <>
struct Traits<CT_PRISM>
{
    static constexpr int ND = 3, NF = 5, S = 18, NFGE = 6;
    __host__ __device__ static constexpr int nsub(int f) { return f < 2 ? 3 : 4; }
    __host__ __device__ static constexpr int soff(int f) { return f < 2 ? 3 * f : 6 + 4 * (f - 2); }
    __host__ __device__ static constexpr int fge(int e, int v)
    {
        constexpr int t[6][3] = {{5, 2, 4}, {3, 2, 5}, {4, 2, 3}, {4, 1, 5}, {5, 1, 3}, {3, 1, 4}};
        return t[e][v];
    }
};

template <>
struct Traits<CT_PYRAMID>
{
    static constexpr int ND = 3, NF = 5, S = 16, NFGE = 6;
    __host__ __device__ static constexpr int nsub(int f) { return f < 4 ? 3 : 4; }
    __host__ __device__ static constexpr int soff(int f) { return 3 * f; }
    __host__ __device__ static constexpr int fge(int e, int v)
    {
        constexpr int t[6][3] = {{1, 5, 2}, {2, 5, 3}, {3, 5, 4}, {4, 5, 1}, {1, 3, 4}, {3, 1, 2}};
        return t[e][v];
    }
};

// Field layout of a type group's geometry table geo[field][cell]:
// CE volumes (CCE then BCEs), BCE centroids, sub-face centroids, sub-face
// normals, and the CFL distance.
template <int T>
struct Layout
{
    using TR = Traits<T>;
    static constexpr int FV = 0;
    static constexpr int FB = TR::NF + 1;
    static constexpr int FC = FB + TR::NF * TR::ND;
    static constexpr int FN = FC + TR::S * TR::ND;
    static constexpr int FD = FN + TR::S * TR::ND;
    static constexpr int NFIELD = FD + 1;
};

// Compile-time loop: f(std::integral_constant<int, 0>) ... f(<N - 1>).
template <int I, int N, class F>
__device__ inline void static_for(F && f)
{
    if constexpr (I < N)
    {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// The fast form of the march kernels (FAST = true) takes two shortcuts
// that leave every value unchanged where they apply, and reports where
// they may not, so the exact form runs instead (the caller's redo pass):
//
// Structural zeros.  The reference evaluates every Jacobian product,
// including the 24 of 75 entries in 3D (8 of 32 in 2D) that are the
// constant 0 for every state (jac_zero).  For a finite x, 0 * x is +-0,
// and adding it leaves a sum unchanged but for the sign of a zero; for an
// infinite or NaN x it is NaN.  The fast form skips these products.  The
// energy row has no structural zero, so when a value that one multiplies
// is not finite the energy row's sum is not finite either (a product with
// a non-finite value is infinite or NaN, and so is every sum it enters),
// and neither is the energy the cell writes (or its so0t).
//
// Shared denominators.  The device's IEEE division (v_div_scale,
// v_rcp_f64, two Newton steps on the reciprocal, the product, one FMA
// residual, v_div_fmas, v_div_fixup) scales its operands and fixes the
// result up only near the ends of the exponent range and at zeros,
// infinities and NaNs: it scales when d or 1 / d or a / d is subnormal,
// when |a| < 2^-969, or when |a / d| >= 2^767.  With 2^-500 <= |d| < 2^92
// and a = 0 or 2^-922 <= |a| < 2^261 it does neither, and its steps are:
// refine r from v_rcp_f64(d), q = a * r, rem = a - d * q (FMA), q + rem * r
// (FMA).  div_recip() refines r once per denominator and div_fast() takes
// the three steps per quotient.  A quotient that is not finite (from an
// operand that is not, or from d = 0) reaches a value the cell writes
// through the energy row, the gradients or the state itself.
//
// FastChk keeps the checks in three vector registers: the greatest biased
// binary exponent seen (frexp, which gives 0 for a zero and for an
// infinity or NaN; numerators biased by 921, denominators by 499 and
// doubled, so each range maps onto [0, 1182]), and a sum that turns NaN
// when a value the cell writes is not finite.  ok() holds when both pass:
// then every quotient equals the division and no structural zero met a
// non-finite value.
__host__ __device__ constexpr bool jac_zero(int nd, int k, int m, int d)
{
    // Row 0 is non-zero only at m = d + 1 (where it is 1) and the energy
    // row nowhere; a momentum row k is non-zero in the density column, in
    // its own momentum column, at d = k - 1, and in another momentum
    // column m at d = m - 1.
    return k == 0             ? m != d + 1
           : k == nd + 1      ? false
           : m == 0 || m == k ? false
           : m == nd + 1      ? d != k - 1
                              : d != k - 1 && d != m - 1;
}

struct FastChk
{
    unsigned top = 0;
    double nan = 0.0;

    __device__ void num(double v)
    {
        unsigned const u = static_cast<unsigned>(__builtin_amdgcn_frexp_exp(v) + 921);
        top = u > top ? u : top;
    }
    __device__ void den(double v)
    {
        unsigned const u = static_cast<unsigned>(2 * (__builtin_amdgcn_frexp_exp(v) + 499));
        top = u > top ? u : top;
    }
    __device__ void finite(double v) { nan = __builtin_fma(v, 0.0, nan); }
    __device__ bool ok() const { return top <= 1182 && nan == 0.0; }
};

__device__ inline double div_recip(double d)
{
    double y = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-d, y, 1.0);
    return __builtin_fma(y, e, y);
}

// a / d from r = div_recip(d); the caller has checked d.
__device__ inline double div_fast(double a, double d, double r, FastChk & c)
{
    c.num(a);
    double q = a * r;
    double const rem = __builtin_fma(-d, q, a);
    return __builtin_fma(rem, r, q);
}

// The flux Jacobian of one state (EulerJacobian<ND>::update,
// euler_march.cpp:26-158), evaluated one equation row at a time: set()
// forms the shared scalars, row<K>() the K-th rows of the flux (f[d]) and
// of the Jacobian (j[m][d]) with the reference's expressions.  A row is
// only live while it is used, which keeps the register footprint of the
// flux integral small.  With FAST, the quotients by the density and its
// square share their reciprocals (div_fast), checked in *chk.
template <int ND, bool FAST = false>
struct JacS;

template <bool FAST>
struct JacS<2, FAST>
{
    double ga, ga1, ga3, ga1h, u1, u2, u3, u4, rho2, v1, v1o2, v2, v2o2, ke2, g1ke2, vs, gretot, getot, pr;
    double rr2; // div_recip(rho2) (FAST)
    FastChk * chk = nullptr;

    __device__ inline void set(double gamma, double const * sol, FastChk * c = nullptr)
    {
        ga = gamma;
        ga1 = ga - 1;
        ga3 = ga - 3;
        ga1h = ga1 / 2;
        u1 = sol[0] + TINYJ;
        u2 = sol[1];
        u3 = sol[2];
        u4 = sol[3];
        rho2 = u1 * u1;
        double r1 = 0.0;
        if constexpr (FAST)
        {
            chk = c;
            chk->den(u1);
            chk->den(rho2);
            r1 = div_recip(u1);
            rr2 = div_recip(rho2);
        }
        auto q1 = [&](double a) { return FAST ? div_fast(a, u1, r1, *chk) : a / u1; };
        v1 = q1(u2);
        v1o2 = v1 * v1;
        v2 = q1(u3);
        v2o2 = v2 * v2;
        ke2 = q1(u2 * u2 + u3 * u3);
        g1ke2 = ga1 * ke2;
        vs = q1(ke2);
        gretot = ga * u4;
        getot = q1(gretot);
        pr = ga1 * u4 - ga1h * ke2;
    }

    template <int K>
    __device__ inline void row(double (&j)[4][2], double (&f)[2]) const
    {
        if constexpr (K == 0)
        {
            f[0] = u2; f[1] = u3;
            j[0][0] = 0; j[0][1] = 0;
            j[1][0] = 1; j[1][1] = 0;
            j[2][0] = 0; j[2][1] = 1;
            j[3][0] = 0; j[3][1] = 0;
        }
        else if constexpr (K == 1)
        {
            f[0] = pr + u2 * v1; f[1] = u2 * v2;
            j[0][0] = -v1o2 + ga1h * vs; j[0][1] = -v1 * v2;
            j[1][0] = -ga3 * v1; j[1][1] = v2;
            j[2][0] = -ga1 * v2; j[2][1] = v1;
            j[3][0] = ga1; j[3][1] = 0;
        }
        else if constexpr (K == 2)
        {
            f[0] = u3 * v1; f[1] = pr + u3 * v2;
            j[0][0] = -v2 * v1; j[0][1] = -v2o2 + ga1h * vs;
            j[1][0] = v2; j[1][1] = -ga1 * v1;
            j[2][0] = v1; j[2][1] = -ga3 * v2;
            j[3][0] = 0; j[3][1] = ga1;
        }
        else
        {
            f[0] = (pr + u4) * v1; f[1] = (pr + u4) * v2;
            auto q2 = [&](double a) { return FAST ? div_fast(a, rho2, rr2, *chk) : a / rho2; };
            j[0][0] = q2((-gretot + g1ke2) * u2); j[0][1] = q2((-gretot + g1ke2) * u3);
            j[1][0] = getot - ga1h * (vs + 2 * v1o2); j[1][1] = -ga1 * v1 * v2;
            j[2][0] = -ga1 * v2 * v1; j[2][1] = getot - ga1h * (vs + 2 * v2o2);
            j[3][0] = ga * v1; j[3][1] = ga * v2;
        }
    }
};

template <bool FAST>
struct JacS<3, FAST>
{
    double ga, ga1, ga3, ga1h, u1, u2, u3, u4, u5, rho2, v1, v1o2, v2, v2o2, v3, v3o2, ke2, g1ke2, vs, gretot, getot,
        pr;
    double rr2; // div_recip(rho2) (FAST)
    FastChk * chk = nullptr;

    __device__ inline void set(double gamma, double const * sol, FastChk * c = nullptr)
    {
        ga = gamma;
        ga1 = ga - 1;
        ga3 = ga - 3;
        ga1h = ga1 / 2;
        u1 = sol[0] + TINYJ;
        u2 = sol[1];
        u3 = sol[2];
        u4 = sol[3];
        u5 = sol[4];
        rho2 = u1 * u1;
        double r1 = 0.0;
        if constexpr (FAST)
        {
            chk = c;
            chk->den(u1);
            chk->den(rho2);
            r1 = div_recip(u1);
            rr2 = div_recip(rho2);
        }
        auto q1 = [&](double a) { return FAST ? div_fast(a, u1, r1, *chk) : a / u1; };
        v1 = q1(u2);
        v1o2 = v1 * v1;
        v2 = q1(u3);
        v2o2 = v2 * v2;
        v3 = q1(u4);
        v3o2 = v3 * v3;
        ke2 = q1(u2 * u2 + u3 * u3 + u4 * u4);
        g1ke2 = ga1 * ke2;
        vs = q1(ke2);
        gretot = ga * u5;
        getot = q1(gretot);
        pr = ga1 * u5 - ga1h * ke2;
    }

    template <int K>
    __device__ inline void row(double (&j)[5][3], double (&f)[3]) const
    {
        if constexpr (K == 0)
        {
            f[0] = u2; f[1] = u3; f[2] = u4;
            j[0][0] = 0; j[0][1] = 0; j[0][2] = 0;
            j[1][0] = 1; j[1][1] = 0; j[1][2] = 0;
            j[2][0] = 0; j[2][1] = 1; j[2][2] = 0;
            j[3][0] = 0; j[3][1] = 0; j[3][2] = 1;
            j[4][0] = 0; j[4][1] = 0; j[4][2] = 0;
        }
        else if constexpr (K == 1)
        {
            f[0] = pr + u2 * v1; f[1] = u2 * v2; f[2] = u2 * v3;
            j[0][0] = -v1o2 + ga1h * vs; j[0][1] = -v1 * v2; j[0][2] = -v1 * v3;
            j[1][0] = -ga3 * v1; j[1][1] = v2; j[1][2] = v3;
            j[2][0] = -ga1 * v2; j[2][1] = v1; j[2][2] = 0;
            j[3][0] = -ga1 * v3; j[3][1] = 0; j[3][2] = v1;
            j[4][0] = ga1; j[4][1] = 0; j[4][2] = 0;
        }
        else if constexpr (K == 2)
        {
            f[0] = u3 * v1; f[1] = pr + u3 * v2; f[2] = u3 * v3;
            j[0][0] = -v2 * v1; j[0][1] = -v2o2 + ga1h * vs; j[0][2] = -v2 * v3;
            j[1][0] = v2; j[1][1] = -ga1 * v1; j[1][2] = 0;
            j[2][0] = v1; j[2][1] = -ga3 * v2; j[2][2] = v3;
            j[3][0] = 0; j[3][1] = -ga1 * v3; j[3][2] = v2;
            j[4][0] = 0; j[4][1] = ga1; j[4][2] = 0;
        }
        else if constexpr (K == 3)
        {
            f[0] = u4 * v1; f[1] = u4 * v2; f[2] = pr + u4 * v3;
            j[0][0] = -v3 * v1; j[0][1] = -v3 * v2; j[0][2] = -v3o2 + ga1h * vs;
            j[1][0] = v3; j[1][1] = 0; j[1][2] = -ga1 * v1;
            j[2][0] = 0; j[2][1] = v3; j[2][2] = -ga1 * v2;
            j[3][0] = v1; j[3][1] = v2; j[3][2] = -ga3 * v3;
            j[4][0] = 0; j[4][1] = 0; j[4][2] = ga1;
        }
        else
        {
            f[0] = (pr + u5) * v1; f[1] = (pr + u5) * v2; f[2] = (pr + u5) * v3;
            auto q2 = [&](double a) { return FAST ? div_fast(a, rho2, rr2, *chk) : a / rho2; };
            j[0][0] = q2((-gretot + g1ke2) * u2);
            j[0][1] = q2((-gretot + g1ke2) * u3);
            j[0][2] = q2((-gretot + g1ke2) * u4);
            j[1][0] = getot - ga1h * (vs + 2 * v1o2); j[1][1] = -ga1 * v1 * v2; j[1][2] = -ga1 * v1 * v3;
            j[2][0] = -ga1 * v2 * v1; j[2][1] = getot - ga1h * (vs + 2 * v2o2); j[2][2] = -ga1 * v2 * v3;
            j[3][0] = -ga1 * v3 * v1; j[3][1] = -ga1 * v3 * v2; j[3][2] = getot - ga1h * (vs + 2 * v3o2);
            j[4][0] = ga * v1; j[4][1] = ga * v2; j[4][2] = ga * v3;
        }
    }
};

// so0t of one row: -(J . so1c) summed over the dimension then the equation
// (calc_solt_impl, euler_march.cpp:170-190).  With FAST the structural
// zeros are skipped; the caller checks the energy row.
template <int ND, bool FAST>
__device__ inline void solt_of(JacS<ND, FAST> const & jac, double const (&s1)[ND + 2][ND], double (&st)[ND + 2])
{
    constexpr int NEQ = ND + 2;
    static_for<0, NEQ>([&](auto kc) {
        constexpr int ieq = decltype(kc)::value;
        double jr[NEQ][ND];
        double fr[ND];
        jac.template row<ieq>(jr, fr);
        double val = 0.0;
#pragma unroll
        for (int idm = 0; idm < ND; ++idm)
        {
#pragma unroll
            for (int jeq = 0; jeq < NEQ; ++jeq)
            {
                if (!FAST || !jac_zero(ND, ieq, jeq, idm))
                {
                    val += jr[jeq][idm] * s1[jeq][idm];
                }
            }
        }
        st[ieq] = -val;
    });
}

// Row-major ND x ND helpers of the boundary conditions (euler_bc.cpp:36-81).
template <int ND>
__device__ inline void bc_matvec(double const (&m)[ND][ND], double const (&v)[ND], double (&r)[ND])
{
#pragma unroll
    for (int i = 0; i < ND; ++i)
    {
        r[i] = 0.0;
#pragma unroll
        for (int k = 0; k < ND; ++k)
        {
            r[i] += m[i][k] * v[k];
        }
    }
}

template <int ND>
__device__ inline void bc_matmat(double const (&a)[ND][ND], double const (&b)[ND][ND], double (&r)[ND][ND])
{
#pragma unroll
    for (int i = 0; i < ND; ++i)
    {
#pragma unroll
        for (int k = 0; k < ND; ++k)
        {
            r[i][k] = 0.0;
#pragma unroll
            for (int m = 0; m < ND; ++m)
            {
                r[i][k] += a[i][m] * b[m][k];
            }
        }
    }
}

// Order-0 ghost state (bc_soln_impl, euler_bc.cpp:147-215).
template <int ND>
__device__ inline void bc_order0(int32_t kind, double const (&mat)[ND][ND], double const * inl, double const (&in)[ND + 2],
                                 double (&out)[ND + 2])
{
    constexpr int NEQ = ND + 2;
    if (kind == BC_NONREFL)
    {
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
            out[k] = in[k];
        }
    }
    else if (kind == BC_SLIPWALL)
    {
        double inv[ND][ND];
#pragma unroll
        for (int a = 0; a < ND; ++a)
        {
#pragma unroll
            for (int b = 0; b < ND; ++b)
            {
                inv[a][b] = mat[b][a];
            }
        }
        out[0] = in[0];
        out[NEQ - 1] = in[NEQ - 1];
        double mom[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            mom[d] = in[1 + d];
        }
        double rot[ND];
        bc_matvec<ND>(mat, mom, rot);
        rot[0] = -rot[0];
        double o[ND];
        bc_matvec<ND>(inv, rot, o);
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            out[1 + d] = o[d];
        }
    }
    else
    {
        double const rho = inl[0];
        double const p = inl[1 + ND];
        double const ga = inl[2 + ND];
        double vsq = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            vsq += inl[1 + d] * inl[1 + d];
        }
        out[0] = rho;
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            out[1 + d] = rho * inl[1 + d];
        }
        out[NEQ - 1] = p / (ga - 1.0) + 0.5 * rho * vsq;
    }
}

// Order-1 ghost gradient (bc_dsoln_impl, euler_bc.cpp:221-320).
template <int ND>
__device__ inline void bc_order1(int32_t kind, double const (&mat)[ND][ND], double const (&in)[ND + 2][ND],
                                 double (&out)[ND + 2][ND])
{
    constexpr int NEQ = ND + 2;
    if (kind == BC_INLET)
    {
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                out[k][d] = 0.0;
            }
        }
        return;
    }
    double inv[ND][ND];
#pragma unroll
    for (int a = 0; a < ND; ++a)
    {
#pragma unroll
        for (int b = 0; b < ND; ++b)
        {
            inv[a][b] = mat[b][a];
        }
    }
    if (kind == BC_NONREFL)
    {
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
            double rot[ND];
            bc_matvec<ND>(mat, in[k], rot);
            rot[0] = 0.0;
            bc_matvec<ND>(inv, rot, out[k]);
        }
        return;
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
    {
        int const k = s == 0 ? 0 : NEQ - 1;
        double rot[ND];
        bc_matvec<ND>(mat, in[k], rot);
        rot[0] = -rot[0];
        bc_matvec<ND>(inv, rot, out[k]);
    }
    double mg[ND][ND];
#pragma unroll
    for (int a = 0; a < ND; ++a)
    {
#pragma unroll
        for (int b = 0; b < ND; ++b)
        {
            mg[a][b] = in[1 + a][b];
        }
    }
    double t1[ND][ND];
    double uv[ND][ND];
    bc_matmat<ND>(mat, mg, t1);
    bc_matmat<ND>(t1, inv, uv);
#pragma unroll
    for (int it = 1; it < ND; ++it)
    {
        uv[0][it] = -uv[0][it];
        uv[it][0] = -uv[it][0];
    }
    double t2[ND][ND];
    double o[ND][ND];
    bc_matmat<ND>(inv, uv, t2);
    bc_matmat<ND>(t2, mat, o);
#pragma unroll
    for (int a = 0; a < ND; ++a)
    {
#pragma unroll
        for (int b = 0; b < ND; ++b)
        {
            out[1 + a][b] = o[a][b];
        }
    }
}

// determinant / adjugate of GradientElement (GradientElement.cpp:383-417).
template <int ND>
__device__ inline double ge_det(double const (&a)[ND][ND])
{
    if constexpr (ND == 2)
    {
        return a[0][0] * a[1][1] - a[0][1] * a[1][0];
    }
    else
    {
        return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
             + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    }
}

template <int ND>
__device__ inline void ge_adj(double const (&a)[ND][ND], double (&r)[ND][ND])
{
    if constexpr (ND == 2)
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
}

} /* end namespace scamd */
