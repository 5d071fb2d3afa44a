/*
 * One-dimensional Euler CESE march on the GPU: the device counterpart of
 * solvcon's Euler1DCore (cpp/solvcon/onedim/Euler1DCore.hpp:39-349 and
 * Euler1DCore.cpp:22-205).
 *
 * State is FP64 in HBM: coord[n], cfl[n], so0[n][3], so1[n][3], gamma[n].
 * Every kernel evaluates the reference arithmetic in the reference order
 * (built with -ffp-contract=off), so results match the CPU march bit for bit.
 *
 * March paths:
 *  - by default the tiled march (k_march_tiled): segments of the tube with
 *    halos in LDS, up to 64 steps per launch and no synchronisation between
 *    blocks, so a short tube spreads over many CUs instead of one;
 *  - with SC_1D_TILED=0, for n <= LDS_MAX_N one persistent workgroup keeps
 *    so0/so1 in LDS and runs every half step with workgroup barriers in
 *    between, and for larger n one fused kernel runs per half step (so0,
 *    cfl and so1 together; each thread owns one new solution point) plus a
 *    one-thread boundary kernel between the two half steps.
 */

#include "common.hpp"

#include <cmath>
#include <cstdlib>
#include <memory>
#include <cstring>
#include <deque>
#include <mutex>
#include <vector>
#include <utility>

namespace scamd
{
namespace
{

constexpr double TINY1D = 1.e-100; // Euler1DCore.hpp:47,182
constexpr double RGAS = 8.31446261815324; // Euler1DCore.hpp:48
constexpr int LDS_MAX_N = 3400; // 48 B per point of so0+so1 in 160 KB of LDS
constexpr unsigned BLOCK = 256;
constexpr unsigned PBLOCK = 1024;

// One solution element: the state of Euler1DKernel (Euler1DCore.hpp:180-282).
struct Elem
{
    double x, xneg, xpos, xctr;
    double u[3], ux[3], f[3], ut[3], ft[3], up[3];
};

// set_value + derive (Euler1DCore.hpp:194-244).
__device__ inline void derive(Elem & e, double ga, double hdt, double x, double xneg, double xpos,
                              double const u[3], double const ux[3])
{
    constexpr double t = TINY1D;
    e.x = x;
    e.xneg = xneg;
    e.xpos = xpos;
    e.xctr = (xpos + xneg) * 0.5;
    double jac[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        e.u[k] = u[k];
        e.ux[k] = ux[k];
    }
    double const u0 = u[0];
    double const u1 = u[1];
    double const u2 = u[2];
    jac[0][0] = 0.0;
    jac[0][1] = 1.0;
    jac[0][2] = 0.0;
    jac[1][0] = (ga - 3.0) / 2.0 * u1 * u1 / (u0 * u0 + t);
    jac[1][1] = -(ga - 3.0) * u1 / (u0 + t);
    jac[1][2] = ga - 1.0;
    jac[2][0] = (ga - 1.0) * u1 * u1 * u1 / (u0 * u0 * u0 + t) - ga * u1 * u2 / (u0 * u0 + t);
    jac[2][1] = ga * u2 / (u0 + t) - 3.0 / 2.0 * (ga - 1.0) * u1 * u1 / (u0 * u0 + t);
    jac[2][2] = ga * u1 / (u0 + t);

    e.f[0] = u1;
    e.f[1] = (ga - 1.0) * u2 + (3.0 - ga) / 2.0 * u1 * u1 / (u0 + t);
    e.f[2] = ga * u1 * u2 / (u0 + t) - (ga - 1.0) / 2.0 * u1 * u1 * u1 / (u0 * u0 + t);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        e.ut[i] = -jac[i][0] * ux[0] - jac[i][1] * ux[1] - jac[i][2] * ux[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        e.ft[i] = jac[i][0] * e.ut[0] + jac[i][1] * e.ut[1] + jac[i][2] * e.ut[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        e.up[i] = u[i] + (x - e.xctr) * ux[i] + hdt * e.ut[i];
    }
}

// calc_flux_ll (Euler1DCore.hpp:246-255) of the left element plus
// calc_flux_lr (:257-266) of the right one, divided by the CE width
// (Euler1DCore.cpp:119-128).
__device__ inline void new_so0(Elem const & xn, Elem const & xp, double hdt, double qdt, double dx, double out[3])
{
    double const dll = xn.xpos - xn.x;
    double const mll = 0.5 * (xn.x + xn.xpos) - xn.xctr;
    double const cll = xn.x - xn.xctr;
    double const dlr = xp.x - xp.xneg;
    double const mlr = 0.5 * (xp.x + xp.xneg) - xp.xctr;
    double const clr = xp.x - xp.xctr;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        double const fll = dll * (xn.u[i] + mll * xn.ux[i]) + hdt * (xn.f[i] - (cll * xn.ut[i]) + (qdt * xn.ft[i]));
        double const flr = dlr * (xp.u[i] + mlr * xp.ux[i]) - hdt * (xp.f[i] - (clr * xp.ut[i]) + (qdt * xp.ft[i]));
        out[i] = (fll + flr) / dx;
    }
}

// march_half_so1_alpha gradient (Euler1DCore.hpp:310-318).
template <int ALPHA>
__device__ inline double alpha_pow(double v)
{
    if constexpr (ALPHA == 1)
    {
        return v;
    }
    else
    {
        return v * v;
    }
}

template <int ALPHA>
__device__ inline void new_so1(Elem const & xn, Elem const & xp, double const utp[3], double out[3])
{
#pragma unroll
    for (int iv = 0; iv < 3; ++iv)
    {
        double const duxn = (utp[iv] - xn.up[iv]) / (xn.xpos - xn.x);
        double const duxp = (xp.up[iv] - utp[iv]) / (xp.x - xp.xneg);
        double const fan = alpha_pow<ALPHA>(fabs(duxn));
        double const fap = alpha_pow<ALPHA>(fabs(duxp));
        out[iv] = (fap * duxn + fan * duxp) / (fap + fan + TINY1D);
    }
}

// update_cfl body (Euler1DCore.cpp:76-93).
__device__ inline double cfl_at(double ga, double hdt, double s0, double s1, double s2, double xm, double x, double xp)
{
    double wspd = s1;
    wspd *= wspd;
    double const ke = wspd / (2.0 * s0);
    double pr = (ga - 1.0) * (s2 - ke);
    pr = (pr + fabs(pr)) / 2.0;
    wspd = sqrt(ga * pr / s0) + sqrt(wspd) / s0;
    double const dxpos = xp - x;
    double const dxneg = x - xm;
    return hdt * wspd / (dxpos < dxneg ? dxpos : dxneg);
}

struct Geom1D
{
    int64_t start;
    int64_t nout;
};

// Loop bounds of every half-step pass (Euler1DCore.cpp:73-76,99-100).
__host__ __device__ inline Geom1D plane(int64_t n, bool odd)
{
    int64_t const start = 2 - (odd ? 1 : 0);
    int64_t const stop = n - 2 - (odd ? 0 : 1);
    int64_t const nout = stop > start ? (stop - start + 1) / 2 : 0;
    return Geom1D{start, nout};
}

__device__ inline void load_elem(Elem & e, int64_t ic, double ga, double hdt, double const * coord, double const * so0,
                                 double const * so1)
{
    double u[3] = {so0[ic * 3 + 0], so0[ic * 3 + 1], so0[ic * 3 + 2]};
    double ux[3] = {so1[ic * 3 + 0], so1[ic * 3 + 1], so1[ic * 3 + 2]};
    derive(e, ga, hdt, coord[ic], coord[ic - 1], coord[ic + 1], u, ux);
}

__global__ void k_cfl(int64_t start, int64_t nout, double hdt, double const * coord, double const * so0,
                      double const * gamma, double * cfl)
{
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nout)
    {
        return;
    }
    int64_t const it = start + 2 * k;
    cfl[it] = cfl_at(gamma[it], hdt, so0[it * 3], so0[it * 3 + 1], so0[it * 3 + 2], coord[it - 1], coord[it],
                     coord[it + 1]);
}

__global__ void k_so0(int64_t start, int64_t nout, double hdt, double qdt, double const * coord, double * so0,
                      double const * so1, double const * gamma)
{
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nout)
    {
        return;
    }
    int64_t const ic = start + 2 * k;
    Elem xn;
    Elem xp;
    load_elem(xn, ic, gamma[ic], hdt, coord, so0, so1);
    load_elem(xp, ic + 2, gamma[ic + 2], hdt, coord, so0, so1);
    double out[3];
    new_so0(xn, xp, hdt, qdt, coord[ic + 2] - coord[ic], out);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        so0[(ic + 1) * 3 + i] = out[i];
    }
}

template <int ALPHA>
__global__ void k_so1(int64_t start, int64_t nout, double hdt, double const * coord, double const * so0, double * so1,
                      double const * gamma)
{
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nout)
    {
        return;
    }
    int64_t const ic = start + 2 * k;
    Elem xn;
    Elem xp;
    load_elem(xn, ic, gamma[ic], hdt, coord, so0, so1);
    load_elem(xp, ic + 2, gamma[ic + 2], hdt, coord, so0, so1);
    double const utp[3] = {so0[(ic + 1) * 3], so0[(ic + 1) * 3 + 1], so0[(ic + 1) * 3 + 2]};
    double out[3];
    new_so1<ALPHA>(xn, xp, utp, out);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        so1[(ic + 1) * 3 + i] = out[i];
    }
}

// treat_boundary_so0/so1 (Euler1DCore.cpp:132-172).
__global__ void k_bnd(int64_t n, double * a, double * b)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
    {
        return;
    }
    for (int i = 0; i < 3; ++i)
    {
        a[3 + i] = a[6 + i];
        a[(n - 2) * 3 + i] = a[(n - 3) * 3 + i];
        if (b != nullptr)
        {
            b[3 + i] = b[6 + i];
            b[(n - 2) * 3 + i] = b[(n - 3) * 3 + i];
        }
    }
}

// One fused half step (march_half1/2_alpha, Euler1DCore.hpp:322-337): the
// new so0 point, the CFL of the input point and the new so1 point.
template <int ALPHA>
__global__ void k_half(int64_t start, int64_t nout, double hdt, double qdt, double const * coord, double * so0,
                       double * so1, double const * gamma, double * cfl)
{
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nout)
    {
        return;
    }
    int64_t const ic = start + 2 * k;
    Elem xn;
    Elem xp;
    double const gn = gamma[ic];
    load_elem(xn, ic, gn, hdt, coord, so0, so1);
    load_elem(xp, ic + 2, gamma[ic + 2], hdt, coord, so0, so1);
    double s0[3];
    new_so0(xn, xp, hdt, qdt, coord[ic + 2] - coord[ic], s0);
    cfl[ic] = cfl_at(gn, hdt, xn.u[0], xn.u[1], xn.u[2], xn.xneg, xn.x, xn.xpos);
    double s1[3];
    new_so1<ALPHA>(xn, xp, s0, s1);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        so0[(ic + 1) * 3 + i] = s0[i];
        so1[(ic + 1) * 3 + i] = s1[i];
    }
}

// Persistent march for n <= LDS_MAX_N: one workgroup, so0 and so1 held in
// LDS across all steps (march_alpha, Euler1DCore.hpp:339-349).
template <int ALPHA>
__global__ void __launch_bounds__(PBLOCK) k_march_lds(int64_t n, int64_t steps, double hdt, double qdt,
                                                      double const * coord, double * gso0, double * gso1,
                                                      double const * gamma, double * cfl)
{
    __shared__ double s0[LDS_MAX_N * 3];
    __shared__ double s1[LDS_MAX_N * 3];
    int const tid = threadIdx.x;
    for (int64_t i = tid; i < n * 3; i += blockDim.x)
    {
        s0[i] = gso0[i];
        s1[i] = gso1[i];
    }
    __syncthreads();
    Geom1D const pe = plane(n, false);
    Geom1D const po = plane(n, true);
    for (int64_t it = 0; it < steps; ++it)
    {
        for (int half = 0; half < 2; ++half)
        {
            Geom1D const p = half == 0 ? pe : po;
            for (int64_t k = tid; k < p.nout; k += blockDim.x)
            {
                int64_t const ic = p.start + 2 * k;
                Elem xn;
                Elem xp;
                double const gn = gamma[ic];
                {
                    double u[3] = {s0[ic * 3], s0[ic * 3 + 1], s0[ic * 3 + 2]};
                    double ux[3] = {s1[ic * 3], s1[ic * 3 + 1], s1[ic * 3 + 2]};
                    derive(xn, gn, hdt, coord[ic], coord[ic - 1], coord[ic + 1], u, ux);
                }
                {
                    int64_t const jc = ic + 2;
                    double u[3] = {s0[jc * 3], s0[jc * 3 + 1], s0[jc * 3 + 2]};
                    double ux[3] = {s1[jc * 3], s1[jc * 3 + 1], s1[jc * 3 + 2]};
                    derive(xp, gamma[jc], hdt, coord[jc], coord[jc - 1], coord[jc + 1], u, ux);
                }
                double n0[3];
                new_so0(xn, xp, hdt, qdt, coord[ic + 2] - coord[ic], n0);
                if (it + 1 == steps)
                {
                    cfl[ic] = cfl_at(gn, hdt, xn.u[0], xn.u[1], xn.u[2], xn.xneg, xn.x, xn.xpos);
                }
                double n1[3];
                new_so1<ALPHA>(xn, xp, n0, n1);
#pragma unroll
                for (int i = 0; i < 3; ++i)
                {
                    s0[(ic + 1) * 3 + i] = n0[i];
                    s1[(ic + 1) * 3 + i] = n1[i];
                }
            }
            __syncthreads();
            if (half == 0)
            {
                if (tid < 3)
                {
                    s0[3 + tid] = s0[6 + tid];
                    s1[3 + tid] = s1[6 + tid];
                }
                else if (tid >= 32 && tid < 35)
                {
                    int const i = tid - 32;
                    s0[(n - 2) * 3 + i] = s0[(n - 3) * 3 + i];
                    s1[(n - 2) * 3 + i] = s1[(n - 3) * 3 + i];
                }
                __syncthreads();
            }
        }
    }
    for (int64_t i = tid; i < n * 3; i += blockDim.x)
    {
        gso0[i] = s0[i];
        gso1[i] = s1[i];
    }
}

// Temporal blocking of the march (march_alpha, Euler1DCore.hpp:339-349):
// block b owns rows [b seg, (b + 1) seg) of [0, n), copies the rows within
// `halo` of them from (in0, in1) into LDS and runs `steps` whole steps there
// without synchronising with other blocks, then writes its own rows to
// (out0, out1).  A row is updated only when both its neighbours lie in the
// block's window; the stale values this leaves next to an inner window edge
// spread by one row per half step, so with halo >= 2 steps + 2 the owned
// rows are exact.  The boundary rows 1 and n - 2 are copied after the first
// half of every step by the block whose window holds them.  With `last`,
// the CFL of the last step is written for the owned rows.  Input and output
// are different buffers: other blocks read their halos from the input.
constexpr int TBLOCK = 512;
constexpr int TROWS = 1024; // most rows of one window (48 KB of LDS)

template <int ALPHA>
__device__ inline void tiled_segment(int64_t blk, double * s0, double * s1, int64_t n, int64_t seg, int64_t halo,
                                     int64_t steps, int last, double hdt, double qdt, double const * coord,
                                     double const * in0, double const * in1, double * out0, double * out1,
                                     double const * gamma, double * cfl)
{
    int64_t const own0 = blk * seg;
    int64_t const own1 = own0 + seg < n ? own0 + seg : n;
    int64_t const lo = own0 - halo > 0 ? own0 - halo : 0;
    int64_t const hi = own1 + halo < n ? own1 + halo : n;
    int const tid = threadIdx.x;
    for (int64_t q = tid; q < (hi - lo) * 3; q += blockDim.x)
    {
        s0[q] = in0[lo * 3 + q];
        s1[q] = in1[lo * 3 + q];
    }
    __syncthreads();
    for (int64_t it = 0; it < steps; ++it)
    {
        for (int half = 0; half < 2; ++half)
        {
            Geom1D const p = plane(n, half == 1);
            // Left elements ic of this plane whose rows ic .. ic + 2 lie in
            // the window.
            int64_t k0 = lo > p.start ? (lo - p.start + 1) / 2 : 0;
            int64_t k1 = hi - 3 >= p.start ? (hi - 3 - p.start) / 2 + 1 : 0;
            k1 = k1 < p.nout ? k1 : p.nout;
            for (int64_t k = k0 + tid; k < k1; k += blockDim.x)
            {
                int64_t const ic = p.start + 2 * k;
                int64_t const l = ic - lo;
                Elem xn;
                Elem xp;
                double const gn = gamma[ic];
                {
                    double u[3] = {s0[l * 3], s0[l * 3 + 1], s0[l * 3 + 2]};
                    double ux[3] = {s1[l * 3], s1[l * 3 + 1], s1[l * 3 + 2]};
                    derive(xn, gn, hdt, coord[ic], coord[ic - 1], coord[ic + 1], u, ux);
                }
                {
                    int64_t const m = l + 2;
                    double u[3] = {s0[m * 3], s0[m * 3 + 1], s0[m * 3 + 2]};
                    double ux[3] = {s1[m * 3], s1[m * 3 + 1], s1[m * 3 + 2]};
                    derive(xp, gamma[ic + 2], hdt, coord[ic + 2], coord[ic + 1], coord[ic + 3], u, ux);
                }
                double n0[3];
                new_so0(xn, xp, hdt, qdt, coord[ic + 2] - coord[ic], n0);
                if (last != 0 && it + 1 == steps && ic >= own0 && ic < own1)
                {
                    cfl[ic] = cfl_at(gn, hdt, xn.u[0], xn.u[1], xn.u[2], xn.xneg, xn.x, xn.xpos);
                }
                double n1[3];
                new_so1<ALPHA>(xn, xp, n0, n1);
#pragma unroll
                for (int i = 0; i < 3; ++i)
                {
                    s0[(l + 1) * 3 + i] = n0[i];
                    s1[(l + 1) * 3 + i] = n1[i];
                }
            }
            __syncthreads();
            if (half == 0)
            {
                if (tid < 3 && lo <= 1 && hi > 2)
                {
                    s0[(1 - lo) * 3 + tid] = s0[(2 - lo) * 3 + tid];
                    s1[(1 - lo) * 3 + tid] = s1[(2 - lo) * 3 + tid];
                }
                else if (tid >= 64 && tid < 67 && lo <= n - 3 && hi > n - 2)
                {
                    int const i = tid - 64;
                    s0[(n - 2 - lo) * 3 + i] = s0[(n - 3 - lo) * 3 + i];
                    s1[(n - 2 - lo) * 3 + i] = s1[(n - 3 - lo) * 3 + i];
                }
                __syncthreads();
            }
        }
    }
    for (int64_t q = tid; q < (own1 - own0) * 3; q += blockDim.x)
    {
        out0[own0 * 3 + q] = s0[(own0 - lo) * 3 + q];
        out1[own0 * 3 + q] = s1[(own0 - lo) * 3 + q];
    }
}


template <int ALPHA>
__global__ void __launch_bounds__(TBLOCK) k_march_tiled(int64_t n, int64_t seg, int64_t halo, int64_t steps, int last,
                                                        double hdt, double qdt, double const * coord,
                                                        double const * in0, double const * in1, double * out0,
                                                        double * out1, double const * gamma, double * cfl)
{
    __shared__ double s0[TROWS * 3];
    __shared__ double s1[TROWS * 3];
    tiled_segment<ALPHA>(blockIdx.x, s0, s1, n, seg, halo, steps, last, hdt, qdt, coord, in0, in1, out0, out1, gamma,
                         cfl);
}

// Several tubes in one launch (the request loop's batched 1D march): each
// tube's blocks are a contiguous range of the grid, and each tube runs its
// own step count.  The table travels by value with the launch.
struct TiledTube
{
    int64_t n;
    int64_t seg;
    int64_t steps;
    int64_t blk0;
    int last;
    double hdt;
    double qdt;
    double const * coord;
    double const * in0;
    double const * in1;
    double * out0;
    double * out1;
    double const * gamma;
    double * cfl;
};
constexpr int MAX_TUBES = 16;
struct TubeSet
{
    int nt;
    TiledTube t[MAX_TUBES];
};

template <int ALPHA>
__global__ void __launch_bounds__(TBLOCK) k_march_tiled_multi(TubeSet const ts)
{
    __shared__ double s0[TROWS * 3];
    __shared__ double s1[TROWS * 3];
    int k = 0;
    for (int j = 1; j < ts.nt; ++j)
    {
        if (static_cast<int64_t>(blockIdx.x) >= ts.t[j].blk0)
        {
            k = j;
        }
    }
    TiledTube const & tb = ts.t[k];
    tiled_segment<ALPHA>(static_cast<int64_t>(blockIdx.x) - tb.blk0, s0, s1, tb.n, tb.seg, 2 * tb.steps + 2,
                         tb.steps, tb.last, tb.hdt, tb.qdt, tb.coord, tb.in0, tb.in1, tb.out0, tb.out1, tb.gamma,
                         tb.cfl);
}

// Derived fields (Euler1DCore.hpp:106-168, Euler1DCore.cpp:37-68,174-205).
__global__ void k_derived(int which, int64_t n, double const * so0, double const * gamma, double * out)
{
    int64_t const it = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (it >= n)
    {
        return;
    }
    double const u0 = so0[it * 3];
    double const u1 = so0[it * 3 + 1];
    double const u2 = so0[it * 3 + 2];
    double const ga = gamma[it];
    double pr = u1;
    pr *= pr;
    pr /= 2.0 * u0 + TINY1D;
    pr = u2 - pr;
    pr *= ga - 1.0;
    double v = 0.0;
    switch (which)
    {
    case 0: v = u0; break;
    case 1: v = u1 / (u0 + TINY1D); break;
    case 2: v = pr; break;
    case 3:
        v = (ga - 1.0) / RGAS;
        v *= u2 / u0 - 0.5 * (u1 * u1) / (u0 * u0);
        break;
    case 4:
        v = u2 / u0;
        v -= 0.5 * ((u1 * u1) / (u0 * u0));
        break;
    default: v = pr / pow(u0, ga); break;
    }
    out[it] = v;
}

} /* end namespace */

// The tiled march: segments of 64-512 rows (about n / 128), up to 64 steps
// per launch with a halo of 2 steps + 2 rows.  The state moves between
// (so0, so1) and the other pair (b0, b1) from launch to launch and ends in
// (so0, so1); both pairs hold n * 3 values.  Enqueued on `stream`.
void tiled_march(int64_t n, double hdt, double qdt, int alpha, int64_t steps, double const * coord,
                 double const * gamma, double * cfl, double * so0, double * so1, double * b0, double * b1,
                 hipStream_t stream)
{
    int64_t seg = (n + 127) / 128;
    seg = seg < 64 ? 64 : seg > 512 ? 512 : seg;
    int64_t const kmax = (TROWS - seg - 4) / 4; // seg + 2 (2 k + 2) <= TROWS
    int64_t const kstep = kmax < 64 ? kmax : 64;
    unsigned const nb = static_cast<unsigned>((n + seg - 1) / seg);
    double * a0 = so0;
    double * a1 = so1;
    for (int64_t done = 0; done < steps;)
    {
        int64_t const k = steps - done < kstep ? steps - done : kstep;
        int const last = done + k == steps ? 1 : 0;
        if (alpha == 1)
        {
            k_march_tiled<1><<<nb, TBLOCK, 0, stream>>>(n, seg, 2 * k + 2, k, last, hdt, qdt, coord, a0, a1, b0,
                                                        b1, gamma, cfl);
        }
        else
        {
            k_march_tiled<2><<<nb, TBLOCK, 0, stream>>>(n, seg, 2 * k + 2, k, last, hdt, qdt, coord, a0, a1, b0,
                                                        b1, gamma, cfl);
        }
        SC_HIP(hipGetLastError());
        std::swap(a0, b0);
        std::swap(a1, b1);
        done += k;
    }
    if (a0 != so0)
    {
        SC_HIP(hipMemcpyAsync(so0, a0, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, stream));
        SC_HIP(hipMemcpyAsync(so1, a1, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, stream));
    }
}

struct Core1D
{
    int64_t n = 0;
    double dt = 0.0;
    StreamHolder sh; // declared before the buffers: freed last
    hipStream_t stream = sh.s;
    DevBuf<double> coord, cfl, so0, so1, gamma;
    DevBuf<double> tso0, tso1; // the other buffers of the tiled march

    Core1D(int64_t ncoord, double time_increment)
        : n(ncoord)
        , dt(time_increment)
    {
        if (char const * env = std::getenv("SC_1D_TILED"))
        {
            tiled = std::atoi(env) != 0 ? 1 : 0;
        }
        AllocScope const scope(stream);
        coord.alloc(n);
        cfl.alloc(n);
        so0.alloc(n * 3);
        so1.alloc(n * 3);
        gamma.alloc(n);
        // Euler1DCore.cpp:30-34: zero-filled tables, gamma = 1.4.
        SC_HIP(hipMemsetAsync(coord.ptr, 0, coord.bytes(), stream));
        SC_HIP(hipMemsetAsync(cfl.ptr, 0, cfl.bytes(), stream));
        SC_HIP(hipMemsetAsync(so0.ptr, 0, so0.bytes(), stream));
        SC_HIP(hipMemsetAsync(so1.ptr, 0, so1.bytes(), stream));
        fill_gamma(1.4);
    }

    ~Core1D() = default;

    void fill_gamma(double g);

    DevBuf<double> & buf(int which)
    {
        switch (which)
        {
        case 0: return coord;
        case 1: return cfl;
        case 2: return so0;
        case 3: return so1;
        case 4: return gamma;
        default: fail(SC_EVALUE, "Euler1DCore: unknown array id");
        }
    }

    double hdt() const { return dt / 2.0; }
    double qdt() const { return hdt() / 2.0; }

    void sync() { SC_HIP(hipStreamSynchronize(stream)); }

    void update_cfl(bool odd)
    {
        Geom1D const p = plane(n, odd);
        if (p.nout > 0)
        {
            k_cfl<<<grid_for(p.nout, BLOCK), BLOCK, 0, stream>>>(p.start, p.nout, hdt(), coord.ptr, so0.ptr,
                                                                 gamma.ptr, cfl.ptr);
            SC_HIP(hipGetLastError());
        }
    }

    void half_so0(bool odd)
    {
        Geom1D const p = plane(n, odd);
        if (p.nout > 0)
        {
            k_so0<<<grid_for(p.nout, BLOCK), BLOCK, 0, stream>>>(p.start, p.nout, hdt(), qdt(), coord.ptr, so0.ptr,
                                                                 so1.ptr, gamma.ptr);
            SC_HIP(hipGetLastError());
        }
    }

    void half_so1(int alpha, bool odd)
    {
        Geom1D const p = plane(n, odd);
        if (p.nout <= 0)
        {
            return;
        }
        if (alpha == 1)
        {
            k_so1<1><<<grid_for(p.nout, BLOCK), BLOCK, 0, stream>>>(p.start, p.nout, hdt(), coord.ptr, so0.ptr,
                                                                    so1.ptr, gamma.ptr);
        }
        else
        {
            k_so1<2><<<grid_for(p.nout, BLOCK), BLOCK, 0, stream>>>(p.start, p.nout, hdt(), coord.ptr, so0.ptr,
                                                                    so1.ptr, gamma.ptr);
        }
        SC_HIP(hipGetLastError());
    }

    void boundary(bool do0, bool do1)
    {
        k_bnd<<<1, 64, 0, stream>>>(n, do0 ? so0.ptr : so1.ptr, (do0 && do1) ? so1.ptr : nullptr);
        SC_HIP(hipGetLastError());
    }

    void half(int alpha, bool odd)
    {
        Geom1D const p = plane(n, odd);
        if (p.nout <= 0)
        {
            return;
        }
        if (alpha == 1)
        {
            k_half<1><<<grid_for(p.nout, BLOCK), BLOCK, 0, stream>>>(p.start, p.nout, hdt(), qdt(), coord.ptr,
                                                                     so0.ptr, so1.ptr, gamma.ptr, cfl.ptr);
        }
        else
        {
            k_half<2><<<grid_for(p.nout, BLOCK), BLOCK, 0, stream>>>(p.start, p.nout, hdt(), qdt(), coord.ptr,
                                                                     so0.ptr, so1.ptr, gamma.ptr, cfl.ptr);
        }
        SC_HIP(hipGetLastError());
    }

    // The tiled march (tiled_march below), the other buffer pair being
    // (tso0, tso1), allocated on first use.
    void march_tiled(int alpha, int64_t steps)
    {
        if (tso0.ptr == nullptr)
        {
            AllocScope const scope(stream);
            tso0.alloc(n * 3);
            tso1.alloc(n * 3);
        }
        tiled_march(n, hdt(), qdt(), alpha, steps, coord.ptr, gamma.ptr, cfl.ptr, so0.ptr, so1.ptr, tso0.ptr,
                    tso1.ptr, stream);
    }

    int tiled = 1; // march with march_tiled (else the single-workgroup or per-half-step kernels)

    void march(int alpha, int64_t steps)
    {
        if (steps <= 0)
        {
            return;
        }
        if (tiled != 0)
        {
            march_tiled(alpha, steps);
            return;
        }
        if (n <= LDS_MAX_N)
        {
            // The persistent kernel writes the CFL of the last step only: the
            // reference overwrites cfl every half step, so only the last
            // step's values survive either way.
            if (alpha == 1)
            {
                k_march_lds<1><<<1, PBLOCK, 0, stream>>>(n, steps, hdt(), qdt(), coord.ptr, so0.ptr, so1.ptr,
                                                         gamma.ptr, cfl.ptr);
            }
            else
            {
                k_march_lds<2><<<1, PBLOCK, 0, stream>>>(n, steps, hdt(), qdt(), coord.ptr, so0.ptr, so1.ptr,
                                                         gamma.ptr, cfl.ptr);
            }
            SC_HIP(hipGetLastError());
            return;
        }
        for (int64_t it = 0; it < steps; ++it)
        {
            half(alpha, false);
            boundary(true, true);
            half(alpha, true);
        }
    }

    void derived(int which, double * dout)
    {
        k_derived<<<grid_for(n, BLOCK), BLOCK, 0, stream>>>(which, n, so0.ptr, gamma.ptr, dout);
        SC_HIP(hipGetLastError());
    }
};

namespace
{
__global__ void k_fill(int64_t n, double * a, double v)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
    {
        a[i] = v;
    }
}
} /* end namespace */

void Core1D::fill_gamma(double g)
{
    k_fill<<<grid_for(n, BLOCK), BLOCK, 0, stream>>>(n, gamma.ptr, g);
    SC_HIP(hipGetLastError());
}

// The march that allocates nothing (sc1d_march_noalloc): the other buffer
// pair of the tiled march lives in this file-scope device array, one copy
// per GPU, for tubes of at most NOALLOC_MAX points.  Host code never uses
// the array's name as an address: the address of each device's copy comes
// from hipGetSymbolAddress, once per device.
constexpr int64_t NOALLOC_MAX = 131071;
__device__ __align__(256) double scratch[2 * 3 * (NOALLOC_MAX + 1)];

constexpr int NOALLOC_DEVICES = 64;
std::mutex noalloc_mu;
double * noalloc_ptr[NOALLOC_DEVICES] = {};

// The current device's copy of the scratch array.
double * noalloc_scratch()
{
    int dev = 0;
    SC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_DEVICES)
    {
        fail(SC_ERUNTIME, "sc1d_march_noalloc: device index out of range");
    }
    std::lock_guard<std::mutex> const lock(noalloc_mu);
    if (noalloc_ptr[dev] == nullptr)
    {
        void * ptr = nullptr;
        SC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        noalloc_ptr[dev] = static_cast<double *>(ptr);
    }
    return noalloc_ptr[dev];
}

// The request loop's 1D engine: admitted tubes (each a Core1D whose
// tables it takes over) march together, one launch per round for all of
// them (MAX_TUBES per launch), each tube advancing up to its own tiled
// step count per round.  A tube that has run its steps has its so0, so1
// and cfl drained to pinned host memory on the engine's stream.
struct DServe1D
{
    StreamHolder sh; // declared first: destroyed last
    hipStream_t stream = sh.s;

    struct Slot
    {
        bool used = false;
        bool marching = false;
        Core1D * core = nullptr;
        int alpha = 2;
        int64_t steps_left = 0;
        bool in_other = false; // state in (tso0, tso1)
        double * host = nullptr;
        hipEvent_t done = nullptr;
    };
    std::vector<Slot> slots;
    std::deque<hipEvent_t> chunks;

    explicit DServe1D(int64_t nslot)
    {
        if (nslot < 1 || nslot > (1 << 20))
        {
            fail(SC_EVALUE, "1D serve engine: between 1 and 1048576 slots");
        }
        slots.resize(static_cast<size_t>(nslot));
    }

    ~DServe1D()
    {
        (void)hipStreamSynchronize(stream);
        for (auto ev : chunks)
        {
            (void)hipEventDestroy(ev);
        }
        for (auto & sl : slots)
        {
            release(sl);
        }
    }

    static void release(Slot & sl)
    {
        if (sl.done != nullptr)
        {
            (void)hipEventDestroy(sl.done);
        }
        if (sl.host != nullptr)
        {
            (void)hipHostFree(sl.host);
        }
        delete sl.core;
        sl = Slot{};
    }

    // Take over a tube; returns the slot or -1 when every slot is in use.
    int64_t admit(Core1D * c, int alpha, int64_t steps)
    {
        if (alpha != 1 && alpha != 2)
        {
            fail(SC_EVALUE, "alpha must be 1 or 2");
        }
        if (steps < 1)
        {
            fail(SC_EVALUE, "1D serve engine: a solve needs at least one step");
        }
        int64_t slot = -1;
        for (size_t k = 0; k < slots.size(); ++k)
        {
            if (!slots[k].used)
            {
                slot = static_cast<int64_t>(k);
                break;
            }
        }
        if (slot < 0)
        {
            return -1;
        }
        if (c->tso0.ptr == nullptr)
        {
            AllocScope const scope(c->stream);
            c->tso0.alloc(c->n * 3);
            c->tso1.alloc(c->n * 3);
        }
        Slot sl;
        SC_HIP(hipHostMalloc(reinterpret_cast<void **>(&sl.host), sizeof(double) * c->n * 7));
        hipEvent_t ev = nullptr;
        try
        {
            SC_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
            SC_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            SC_HIP(hipEventRecord(ev, c->stream));
            SC_HIP(hipStreamWaitEvent(stream, ev, 0));
            (void)hipEventDestroy(ev);
        }
        catch (...)
        {
            if (ev != nullptr)
            {
                (void)hipEventDestroy(ev);
            }
            release(sl);
            throw;
        }
        sl.used = true;
        sl.marching = true;
        sl.core = c;
        sl.alpha = alpha;
        sl.steps_left = steps;
        slots[slot] = sl;
        return slot;
    }

    void drain(Slot & sl)
    {
        Core1D & c = *sl.core;
        double const * a0 = sl.in_other ? c.tso0.ptr : c.so0.ptr;
        double const * a1 = sl.in_other ? c.tso1.ptr : c.so1.ptr;
        SC_HIP(hipMemcpyAsync(sl.host, a0, sizeof(double) * 3 * c.n, hipMemcpyDeviceToHost, stream));
        SC_HIP(hipMemcpyAsync(sl.host + 3 * c.n, a1, sizeof(double) * 3 * c.n, hipMemcpyDeviceToHost, stream));
        SC_HIP(hipMemcpyAsync(sl.host + 6 * c.n, c.cfl.ptr, sizeof(double) * c.n, hipMemcpyDeviceToHost, stream));
        SC_HIP(hipEventRecord(sl.done, stream));
    }

    // One round: every marching tube advances min(its tiled step count,
    // the steps it has left).
    void step()
    {
        for (int alpha = 1; alpha <= 2; ++alpha)
        {
            TubeSet ts{};
            std::vector<int64_t> members;
            int64_t blocks = 0;
            auto flush = [&]()
            {
                if (ts.nt == 0)
                {
                    return;
                }
                unsigned const nb = static_cast<unsigned>(blocks);
                if (alpha == 1)
                {
                    k_march_tiled_multi<1><<<nb, TBLOCK, 0, stream>>>(ts);
                }
                else
                {
                    k_march_tiled_multi<2><<<nb, TBLOCK, 0, stream>>>(ts);
                }
                SC_HIP(hipGetLastError());
                for (int64_t k : members)
                {
                    Slot & sl = slots[k];
                    sl.in_other = !sl.in_other;
                    if (sl.steps_left == 0)
                    {
                        sl.marching = false;
                        drain(sl);
                    }
                }
                ts = TubeSet{};
                members.clear();
                blocks = 0;
            };
            for (size_t k = 0; k < slots.size(); ++k)
            {
                Slot & sl = slots[k];
                if (!sl.marching || sl.alpha != alpha)
                {
                    continue;
                }
                Core1D & c = *sl.core;
                int64_t seg = (c.n + 127) / 128;
                seg = seg < 64 ? 64 : seg > 512 ? 512 : seg;
                int64_t const kmax = (TROWS - seg - 4) / 4;
                int64_t const kstep = kmax < 64 ? kmax : 64;
                int64_t const kk = sl.steps_left < kstep ? sl.steps_left : kstep;
                TiledTube & t = ts.t[ts.nt++];
                t.n = c.n;
                t.seg = seg;
                t.steps = kk;
                t.blk0 = blocks;
                t.last = kk == sl.steps_left ? 1 : 0;
                t.hdt = c.hdt();
                t.qdt = c.qdt();
                t.coord = c.coord.ptr;
                t.in0 = sl.in_other ? c.tso0.ptr : c.so0.ptr;
                t.in1 = sl.in_other ? c.tso1.ptr : c.so1.ptr;
                t.out0 = sl.in_other ? c.so0.ptr : c.tso0.ptr;
                t.out1 = sl.in_other ? c.so1.ptr : c.tso1.ptr;
                t.gamma = c.gamma.ptr;
                t.cfl = c.cfl.ptr;
                blocks += (c.n + seg - 1) / seg;
                sl.steps_left -= kk;
                members.push_back(static_cast<int64_t>(k));
                if (ts.nt == MAX_TUBES)
                {
                    flush();
                }
            }
            flush();
        }
        hipEvent_t ev = nullptr;
        SC_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        SC_HIP(hipEventRecord(ev, stream));
        chunks.push_back(ev);
    }

    int64_t pending()
    {
        while (!chunks.empty())
        {
            hipError_t const e = hipEventQuery(chunks.front());
            if (e == hipErrorNotReady)
            {
                break;
            }
            SC_HIP(e);
            (void)hipEventDestroy(chunks.front());
            chunks.pop_front();
        }
        return static_cast<int64_t>(chunks.size());
    }

    Slot & at(int64_t slot)
    {
        if (slot < 0 || slot >= static_cast<int64_t>(slots.size()))
        {
            fail(SC_EINDEX, "1D serve engine: slot out of range");
        }
        Slot & sl = slots[slot];
        if (!sl.used)
        {
            fail(SC_EVALUE, "1D serve engine: slot " + std::to_string(slot) + " holds no solve");
        }
        return sl;
    }

    int query(int64_t slot)
    {
        Slot & sl = at(slot);
        if (sl.marching)
        {
            return 0;
        }
        hipError_t const e = hipEventQuery(sl.done);
        if (e == hipErrorNotReady)
        {
            return 0;
        }
        SC_HIP(e);
        return 1;
    }

    // so0 [n*3], so1 [n*3], cfl [n] of a finished tube (any may be null);
    // frees the slot and the tube.
    void result(int64_t slot, double * so0, double * so1, double * cfl)
    {
        Slot & sl = at(slot);
        if (sl.marching)
        {
            fail(SC_ERUNTIME, "1D serve engine: slot " + std::to_string(slot) + " is still marching");
        }
        SC_HIP(hipEventSynchronize(sl.done));
        int64_t const n = sl.core->n;
        if (so0 != nullptr)
        {
            std::memcpy(so0, sl.host, sizeof(double) * 3 * n);
        }
        if (so1 != nullptr)
        {
            std::memcpy(so1, sl.host + 3 * n, sizeof(double) * 3 * n);
        }
        if (cfl != nullptr)
        {
            std::memcpy(cfl, sl.host + 6 * n, sizeof(double) * n);
        }
        release(sl);
    }
};

} /* end namespace scamd */

using scamd::Core1D;
using scamd::DServe1D;
using scamd::guard;

extern "C"
{

int sc1d_create(int64_t ncoord, double time_increment, void ** out)
{
    return guard(
        [&]()
        {
            // Euler1DCore.cpp:25-28
            if (ncoord % 2 == 0)
            {
                scamd::fail(scamd::SC_EVALUE, "ncoord cannot be even");
            }
            if (ncoord < 5)
            {
                scamd::fail(scamd::SC_EVALUE, "ncoord must be at least 5");
            }
            auto * h = new Core1D(ncoord, time_increment);
            h->sync();
            *out = h;
        });
}

int sc1d_clone(void * h, void ** out)
{
    return guard(
        [&]()
        {
            auto & a = *static_cast<Core1D *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            auto * b = new Core1D(a.n, a.dt);
            for (int w = 0; w < 5; ++w)
            {
                SC_HIP(hipMemcpyAsync(b->buf(w).ptr, a.buf(w).ptr, a.buf(w).bytes(), hipMemcpyDeviceToDevice,
                                      a.stream));
            }
            a.sync();
            b->sync();
            *out = b;
        });
}

void sc1d_destroy(void * h) { delete static_cast<Core1D *>(h); }

int sc1d_info(void * h, int64_t * ncoord, double * time_increment)
{
    auto & a = *static_cast<Core1D *>(h);
    scamd::AllocScope const alloc_scope_(a.stream);
    *ncoord = a.n;
    *time_increment = a.dt;
    return 0;
}

int sc1d_device_ptr(void * h, int which, double ** ptr, int64_t * count)
{
    return guard(
        [&]()
        {
            auto & b = static_cast<Core1D *>(h)->buf(which);
            *ptr = b.ptr;
            *count = b.count;
        });
}

int sc1d_upload(void * h, int which, double const * host, int64_t count)
{
    return guard(
        [&]()
        {
            auto & a = *static_cast<Core1D *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            auto & b = a.buf(which);
            if (count != b.count)
            {
                scamd::fail(scamd::SC_EVALUE, "Euler1DCore: array size mismatch");
            }
            SC_HIP(hipMemcpyAsync(b.ptr, host, b.bytes(), hipMemcpyHostToDevice, a.stream));
            a.sync();
        });
}

int sc1d_download(void * h, int which, double * host, int64_t count)
{
    return guard(
        [&]()
        {
            auto & a = *static_cast<Core1D *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            auto & b = a.buf(which);
            if (count != b.count)
            {
                scamd::fail(scamd::SC_EVALUE, "Euler1DCore: array size mismatch");
            }
            SC_HIP(hipMemcpyAsync(host, b.ptr, b.bytes(), hipMemcpyDeviceToHost, a.stream));
            a.sync();
        });
}

int sc1d_update_cfl(void * h, int odd_plane)
{
    return guard([&]() { auto & a = *static_cast<Core1D *>(h); a.update_cfl(odd_plane != 0); a.sync(); });
}

int sc1d_march_half_so0(void * h, int odd_plane)
{
    return guard([&]() { auto & a = *static_cast<Core1D *>(h); a.half_so0(odd_plane != 0); a.sync(); });
}

int sc1d_march_half_so1(void * h, int alpha, int odd_plane)
{
    return guard(
        [&]()
        {
            if (alpha != 1 && alpha != 2)
            {
                scamd::fail(scamd::SC_EVALUE, "alpha must be 1 or 2");
            }
            auto & a = *static_cast<Core1D *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            a.half_so1(alpha, odd_plane != 0);
            a.sync();
        });
}

int sc1d_treat_boundary_so0(void * h)
{
    return guard([&]() { auto & a = *static_cast<Core1D *>(h); a.boundary(true, false); a.sync(); });
}

int sc1d_treat_boundary_so1(void * h)
{
    return guard([&]() { auto & a = *static_cast<Core1D *>(h); a.boundary(false, true); a.sync(); });
}

int sc1d_setup_march(void * h)
{
    return guard([&]() { auto & a = *static_cast<Core1D *>(h); a.update_cfl(false); a.sync(); });
}

int sc1d_march_half(void * h, int alpha, int second)
{
    return guard(
        [&]()
        {
            if (alpha != 1 && alpha != 2)
            {
                scamd::fail(scamd::SC_EVALUE, "alpha must be 1 or 2");
            }
            auto & a = *static_cast<Core1D *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            a.half(alpha, second != 0);
            a.sync();
        });
}

int sc1d_march(void * h, int alpha, int64_t steps)
{
    return guard(
        [&]()
        {
            if (alpha != 1 && alpha != 2)
            {
                scamd::fail(scamd::SC_EVALUE, "alpha must be 1 or 2");
            }
            auto & a = *static_cast<Core1D *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            a.march(alpha, steps);
            a.sync();
        });
}

int sc1d_march_async(void * h, int alpha, int64_t steps)
{
    return guard(
        [&]()
        {
            if (alpha != 1 && alpha != 2)
            {
                scamd::fail(scamd::SC_EVALUE, "alpha must be 1 or 2");
            }
            static_cast<Core1D *>(h)->march(alpha, steps);
        });
}

int sc1d_sync(void * h)
{
    return guard([&]() { static_cast<Core1D *>(h)->sync(); });
}

int sc1d_query(void * h, int * done)
{
    return guard(
        [&]()
        {
            hipError_t const e = hipStreamQuery(static_cast<Core1D *>(h)->stream);
            if (e == hipErrorNotReady)
            {
                *done = 0;
                return;
            }
            SC_HIP(e);
            *done = 1;
        });
}

int sc1d_derived(void * h, int which, double * host_out)
{
    return guard(
        [&]()
        {
            auto & a = *static_cast<Core1D *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            if (which < 0 || which > 5)
            {
                scamd::fail(scamd::SC_EVALUE, "unknown derived field");
            }
            scamd::DevBuf<double> tmp;
            tmp.alloc(a.n);
            a.derived(which, tmp.ptr);
            SC_HIP(hipMemcpyAsync(host_out, tmp.ptr, tmp.bytes(), hipMemcpyDeviceToHost, a.stream));
            a.sync();
        });
}

int sc1d_march_noalloc(double const * coord, double const * gamma, double * so0, double * so1, double * cfl,
                       int64_t ncoord, double time_increment, int alpha, int64_t steps, void * stream)
{
    return guard(
        [&]()
        {
            if (ncoord % 2 == 0)
            {
                scamd::fail(scamd::SC_EVALUE, "ncoord cannot be even");
            }
            if (ncoord < 5)
            {
                scamd::fail(scamd::SC_EVALUE, "ncoord must be at least 5");
            }
            if (ncoord > scamd::NOALLOC_MAX)
            {
                scamd::fail(scamd::SC_EVALUE, "sc1d_march_noalloc: ncoord " + std::to_string(ncoord) +
                                                  " exceeds the scratch maximum of " +
                                                  std::to_string(scamd::NOALLOC_MAX) + " points");
            }
            if (alpha != 1 && alpha != 2)
            {
                scamd::fail(scamd::SC_EVALUE, "alpha must be 1 or 2");
            }
            if (steps < 0)
            {
                scamd::fail(scamd::SC_EVALUE, "steps must be >= 0");
            }
            if (coord == nullptr || gamma == nullptr || so0 == nullptr || so1 == nullptr || cfl == nullptr)
            {
                scamd::fail(scamd::SC_EVALUE, "sc1d_march_noalloc: null array");
            }
            double * const b0 = scamd::noalloc_scratch();
            double * const b1 = b0 + 3 * (scamd::NOALLOC_MAX + 1);
            double const hdt = time_increment / 2.0;
            scamd::tiled_march(ncoord, hdt, hdt / 2.0, alpha, steps, coord, gamma, cfl, so0, so1, b0, b1,
                               static_cast<hipStream_t>(stream));
        });
}

int scserve1d_create(int64_t max_slots, void ** out)
{
    return guard([&]() { *out = new DServe1D(max_slots); });
}

void scserve1d_destroy(void * a) { delete static_cast<DServe1D *>(a); }

int scserve1d_admit(void * a, void * core, int alpha, int64_t steps, int64_t * slot)
{
    return guard(
        [&]() { *slot = static_cast<DServe1D *>(a)->admit(static_cast<Core1D *>(core), alpha, steps); });
}

int scserve1d_step(void * a)
{
    return guard([&]() { static_cast<DServe1D *>(a)->step(); });
}

int scserve1d_pending(void * a, int64_t * chunks)
{
    return guard([&]() { *chunks = static_cast<DServe1D *>(a)->pending(); });
}

int scserve1d_query(void * a, int64_t slot, int * done)
{
    return guard([&]() { *done = static_cast<DServe1D *>(a)->query(slot); });
}

int scserve1d_result(void * a, int64_t slot, double * so0, double * so1, double * cfl)
{
    return guard([&]() { static_cast<DServe1D *>(a)->result(slot, so0, so1, cfl); });
}

} /* end extern "C" */
