/*
 * One-dimensional space-time CESE scalar engine on the GPU: the device
 * counterpart of solvcon's spacetime SolverBase (cpp/solvcon/spacetime/
 * core.hpp:706-1036, core.cpp:14-116) with the element kernels of
 * kernel/linear_scalar.hpp:45-87, kernel/inviscid_burgers.hpp:51-109 and the
 * default hooks of Kernel::reset (core.hpp:520-536).
 *
 * Layout as in the reference: coordinates xc[nx] (nx = 2*ncelm + 5, two
 * mirrored ghost points each side), so0/so1[nx][nvar] row-major, cfl[nx].
 * Solution element j of a plane sits at xindex 2 + 2j + odd, conservation
 * element j at 3 + 2j + odd; only variable 0 is marched.
 *
 * One half step is one kernel: each thread owns one conservation element,
 * forms the new so0 at its centre, the CFL number of the solution element
 * there and the new so1, and the threads of the first and last elements
 * also write the periodic ghost points (treat_boundary_so0/so1,
 * core.hpp:982-1004).  Up to ST_LDS_MAX points the whole march runs in one
 * persistent workgroup with coordinates and state in LDS.
 */

#include "common.hpp"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace scamd
{
namespace
{

constexpr unsigned ST_BLOCK = 256;
constexpr unsigned ST_PBLOCK = 1024;
constexpr int64_t ST_LDS_MAX = 6400; // xc, so0, so1 at 8 B each in 150 KB of LDS

enum StKind : int
{
    K_DEFAULT = 0,
    K_LINEAR = 1,
    K_BURGERS = 2,
};

// Coordinates and variable-0 state as seen by one element evaluation; the
// accessors take an xindex.  The same code runs on global memory and on LDS.
struct StView
{
    double const * xc;
    double const * u;
    double const * ux;
    int64_t stride;
    double hdt;
    double qdt;

    __device__ double x(int64_t y) const { return xc[y]; }
    __device__ double xneg(int64_t y) const { return xc[y - 1]; }
    __device__ double xpos(int64_t y) const { return xc[y + 1]; }
    __device__ double xctr(int64_t y) const { return (xneg(y) + xpos(y)) / 2; }
    __device__ double dxneg(int64_t y) const { return x(y) - xneg(y); }
    __device__ double dxpos(int64_t y) const { return xpos(y) - x(y); }
    __device__ double so0(int64_t y) const { return u[y * stride]; }
    __device__ double so1(int64_t y) const { return ux[y * stride]; }
};

template <int K>
__device__ inline double se_xn(StView const & v, int64_t y)
{
    if constexpr (K == K_DEFAULT)
    {
        return 0.0;
    }
    else
    {
        double const displacement = 0.5 * (v.x(y) + v.xneg(y)) - v.xctr(y);
        return v.dxneg(y) * (v.so0(y) + displacement * v.so1(y));
    }
}

template <int K>
__device__ inline double se_xp(StView const & v, int64_t y)
{
    if constexpr (K == K_DEFAULT)
    {
        return 0.0;
    }
    else
    {
        double const displacement = 0.5 * (v.x(y) + v.xpos(y)) - v.xctr(y);
        return v.dxpos(y) * (v.so0(y) + displacement * v.so1(y));
    }
}

// tn (sign +1) and tp (sign -1): the flux through the temporal faces.
template <int K, int SIGN>
__device__ inline double se_t(StView const & v, int64_t y)
{
    if constexpr (K == K_DEFAULT)
    {
        return 0.0;
    }
    else if constexpr (K == K_LINEAR)
    {
        double const displacement = v.x(y) - v.xctr(y);
        double ret = v.so0(y);
        ret += displacement * v.so1(y);
        if constexpr (SIGN > 0)
        {
            ret += v.qdt * v.so1(y);
        }
        else
        {
            ret -= v.qdt * v.so1(y);
        }
        return v.hdt * ret;
    }
    else
    {
        double const displacement = v.x(y) - v.xctr(y);
        double const u = v.so0(y);
        double const u_2 = u * u;
        double ret = 0.5 * u_2;
        ret += displacement * u * v.so1(y);
        if constexpr (SIGN > 0)
        {
            ret += v.qdt * u_2 * v.so1(y);
        }
        else
        {
            ret -= v.qdt * u_2 * v.so1(y);
        }
        return v.hdt * ret;
    }
}

template <int K>
__device__ inline double se_so0p(StView const & v, int64_t y)
{
    if constexpr (K == K_DEFAULT)
    {
        return v.so0(y);
    }
    else
    {
        double ret = v.so0(y);
        ret += (v.x(y) - v.xctr(y)) * v.so1(y);
        ret -= v.hdt * v.so1(y);
        return ret;
    }
}

// update_cfl of the solution element at y with so0 value u0.
template <int K>
__device__ inline double se_cfl(StView const & v, int64_t y, double u0)
{
    if constexpr (K == K_DEFAULT)
    {
        return 0.0;
    }
    else
    {
        double const dn = v.dxneg(y);
        double const dp = v.dxpos(y);
        double const hdx = dp < dn ? dp : dn; // std::min(dxneg, dxpos)
        if constexpr (K == K_LINEAR)
        {
            return v.hdt / hdx;
        }
        else
        {
            return fabs(u0) * v.hdt / hdx;
        }
    }
}

// Celm::calc_so0 (core.hpp:645-653) of the CE centred at X.
template <int K>
__device__ inline double ce_so0(StView const & v, int64_t X)
{
    double const flux_ll = se_xp<K>(v, X - 1) + se_t<K, -1>(v, X - 1);
    double const flux_ur = se_xn<K>(v, X + 1) - se_t<K, -1>(v, X + 1);
    return (flux_ll + flux_ur) / (v.xpos(X) - v.xneg(X));
}

template <int A>
__device__ inline double alpha_pow(double a)
{
    if constexpr (A == 0)
    {
        return 1.0;
    }
    else if constexpr (A == 1)
    {
        return a;
    }
    else
    {
        return a * a;
    }
}

// Celm::calc_so1_alpha (core.hpp:655-672) with the new so0 value utp at X.
template <int K, int A>
__device__ inline double ce_so1(StView const & v, int64_t X, double utp)
{
    double const upn = se_so0p<K>(v, X - 1);
    double const upp = se_so0p<K>(v, X + 1);
    double const duxn = (utp - upn) / v.dxpos(X - 1);
    double const duxp = (upp - utp) / v.dxneg(X + 1);
    double const fan = alpha_pow<A>(fabs(duxn));
    double const fap = alpha_pow<A>(fabs(duxp));
    return (fap * duxn + fan * duxp) / (fap + fan + DBL_MIN);
}

struct Plane
{
    int64_t first; // xindex of the first conservation element
    int64_t count;
};

__host__ __device__ inline Plane ce_plane(int64_t ncelm, bool odd)
{
    // ic in [odd ? -1 : 0, ncelm), X = 3 + 2 ic + odd (core.hpp:950-957).
    return odd ? Plane{2, ncelm + 1} : Plane{3, ncelm};
}

__host__ __device__ inline Plane se_plane(int64_t ncelm, bool odd)
{
    // ic in [odd ? -1 : 0, nselm), y = 2 + 2 ic + odd (core.hpp:960-968).
    return odd ? Plane{1, ncelm + 2} : Plane{2, ncelm + 1};
}

struct StArgs
{
    int64_t ncelm;
    int64_t nvar;
    double hdt;
    double qdt;
    double const * xc;
    double * so0;
    double * so1;
    double * cfl;
};

__device__ inline StView view_of(StArgs const & a)
{
    return StView{a.xc, a.so0, a.so1, a.nvar, a.hdt, a.qdt};
}

template <int K>
__global__ void k_st_so0(StArgs a, Plane p)
{
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < p.count)
    {
        int64_t const X = p.first + 2 * k;
        a.so0[X * a.nvar] = ce_so0<K>(view_of(a), X);
    }
}

template <int K, int A>
__global__ void k_st_so1(StArgs a, Plane p)
{
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < p.count)
    {
        int64_t const X = p.first + 2 * k;
        StView const v = view_of(a);
        a.so1[X * a.nvar] = ce_so1<K, A>(v, X, v.so0(X));
    }
}

template <int K>
__global__ void k_st_cfl(StArgs a, Plane p)
{
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < p.count)
    {
        int64_t const y = p.first + 2 * k;
        StView const v = view_of(a);
        a.cfl[y] = se_cfl<K>(v, y, v.so0(y));
    }
}

// treat_boundary_so0/so1 (core.hpp:982-1004): periodic odd-plane ghosts.
__global__ void k_st_bnd(int64_t ncelm, int64_t nvar, double * arr)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        double const left_in = arr[3 * nvar];
        double const right_in = arr[(2 * ncelm + 1) * nvar];
        arr[1 * nvar] = right_in;
        arr[(2 * ncelm + 3) * nvar] = left_in;
    }
}

// One element of a fused half step.  First half (odd == false):
// march_half_so0(false), treat_boundary_so0, update_cfl(true),
// march_half_so1(false), treat_boundary_so1 (core.hpp:1006-1014); the
// solution element at X belongs to the odd plane, and the threads of the
// first and last CEs also write the ghost points 2 ncelm + 3 and 1.  Second
// half: march_half_so0(true), update_cfl(false), march_half_so1(true)
// (core.hpp:1016-1024).
template <int K, int A>
__device__ inline void st_half_elem(StView const & v, int64_t ncelm, bool odd, int64_t k, double * so0, double * so1,
                                    int64_t stride, double * cfl)
{
    int64_t const X = (odd ? 2 : 3) + 2 * k;
    double const s0 = ce_so0<K>(v, X);
    double const c = se_cfl<K>(v, X, s0);
    double const s1 = ce_so1<K, A>(v, X, s0);
    so0[X * stride] = s0;
    so1[X * stride] = s1;
    if (cfl != nullptr)
    {
        cfl[X] = c;
    }
    if (!odd)
    {
        if (k == ncelm - 1)
        {
            so0[1 * stride] = s0;
            so1[1 * stride] = s1;
            if (cfl != nullptr)
            {
                cfl[1] = se_cfl<K>(v, 1, s0);
            }
        }
        if (k == 0)
        {
            int64_t const g = 2 * ncelm + 3;
            so0[g * stride] = s0;
            so1[g * stride] = s1;
            if (cfl != nullptr)
            {
                cfl[g] = se_cfl<K>(v, g, s0);
            }
        }
    }
}

template <int K, int A>
__global__ void k_st_half(StArgs a, int odd)
{
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const count = odd ? a.ncelm + 1 : a.ncelm;
    if (k < count)
    {
        st_half_elem<K, A>(view_of(a), a.ncelm, odd != 0, k, a.so0, a.so1, a.nvar, a.cfl);
    }
}

// The whole march_alpha<A>(steps) in one workgroup for nx <= ST_LDS_MAX:
// coordinates and variable 0 of so0/so1 live in LDS; the CFL numbers of the
// last step are written (every step overwrites them in the reference, so
// only the last step's survive).
template <int K, int A>
__global__ void __launch_bounds__(ST_PBLOCK) k_st_march_lds(StArgs a, int64_t steps)
{
    __shared__ double sx[ST_LDS_MAX];
    __shared__ double s0[ST_LDS_MAX];
    __shared__ double s1[ST_LDS_MAX];
    int64_t const nx = 2 * a.ncelm + 5;
    for (int64_t i = threadIdx.x; i < nx; i += blockDim.x)
    {
        sx[i] = a.xc[i];
        s0[i] = a.so0[i * a.nvar];
        s1[i] = a.so1[i * a.nvar];
    }
    __syncthreads();
    StView const v{sx, s0, s1, 1, a.hdt, a.qdt};
    for (int64_t it = 0; it < steps; ++it)
    {
        double * cfl = it + 1 == steps ? a.cfl : nullptr;
        for (int h = 0; h < 2; ++h)
        {
            bool const odd = h == 1;
            int64_t const count = odd ? a.ncelm + 1 : a.ncelm;
            for (int64_t k = threadIdx.x; k < count; k += blockDim.x)
            {
                st_half_elem<K, A>(v, a.ncelm, odd, k, s0, s1, 1, cfl);
            }
            __syncthreads();
        }
    }
    for (int64_t i = threadIdx.x; i < nx; i += blockDim.x)
    {
        a.so0[i * a.nvar] = s0[i];
        a.so1[i * a.nvar] = s1[i];
    }
}

// One element value (what: 0 xn, 1 xp, 2 tn, 3 tp, 4 so0p, 5 update_cfl,
// 6 calc_so0, 7 calc_so1_alpha) of variable iv at xindex y; store != 0 with
// what == 5 also writes cfl[y].
template <int K>
__global__ void k_st_eval(StArgs a, int what, int alpha, int64_t y, int64_t iv, int store, double * out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
    {
        return;
    }
    StView const v{a.xc, a.so0 + iv, a.so1 + iv, a.nvar, a.hdt, a.qdt};
    double r = 0.0;
    switch (what)
    {
    case 0: r = se_xn<K>(v, y); break;
    case 1: r = se_xp<K>(v, y); break;
    case 2: r = se_t<K, 1>(v, y); break;
    case 3: r = se_t<K, -1>(v, y); break;
    case 4: r = se_so0p<K>(v, y); break;
    case 5:
        // update_cfl reads variable 0 (inviscid_burgers.hpp:104-108).
        r = se_cfl<K>(StView{a.xc, a.so0, a.so1, a.nvar, a.hdt, a.qdt}, y, a.so0[y * a.nvar]);
        if (store != 0)
        {
            a.cfl[y] = r;
        }
        break;
    case 6: r = ce_so0<K>(v, y); break;
    default:
        if (alpha == 0)
        {
            r = ce_so1<K, 0>(v, y, v.so0(y));
        }
        else if (alpha == 1)
        {
            r = ce_so1<K, 1>(v, y, v.so0(y));
        }
        else
        {
            r = ce_so1<K, 2>(v, y, v.so0(y));
        }
        break;
    }
    out[0] = r;
}

// get_so0p (core.hpp:812-824) of one plane.
template <int K>
__global__ void k_st_so0p(StArgs a, Plane p, int64_t iv, double * out)
{
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < p.count)
    {
        StView const v{a.xc, a.so0 + iv, a.so1 + iv, a.nvar, a.hdt, a.qdt};
        out[k] = se_so0p<K>(v, p.first + 2 * k);
    }
}

// Shortest decimal that reads back to v, the way the reference's messages
// print numbers (std::format "{}"): the shorter of the fixed and the
// scientific forms, fixed on a tie.
std::string fmt_num(double v)
{
    char fx[400];
    char sc[64];
    fx[0] = '\0';
    for (int k = 0; k <= 17; ++k)
    {
        std::snprintf(fx, sizeof(fx), "%.*f", k, v);
        if (std::strtod(fx, nullptr) == v)
        {
            break;
        }
    }
    for (int p = 0; p <= 16; ++p)
    {
        std::snprintf(sc, sizeof(sc), "%.*e", p, v);
        if (std::strtod(sc, nullptr) == v)
        {
            break;
        }
    }
    std::string const f(fx);
    std::string const e(sc);
    return (std::strtod(fx, nullptr) == v && f.size() <= e.size()) ? f : e;
}

} /* end namespace */

// Grid::init_from_array (core.cpp:48-98).
void st_grid_from(double const * xloc, int64_t n, double * xc)
{
    if (n < 2)
    {
        fail(SC_EVALUE, "Grid::init_from_array(xloc) invalid arguments: xloc.size()=" + std::to_string(n)
                            + " smaller than 2");
    }
    for (int64_t it = 0; it + 1 < n; ++it)
    {
        if (xloc[it] >= xloc[it + 1])
        {
            fail(SC_EVALUE, "Grid::init_from_array(xloc) invalid arguments: xloc[" + std::to_string(it)
                                + "]=" + fmt_num(xloc[it]) + " >= xloc[" + std::to_string(it + 1)
                                + "]=" + fmt_num(xloc[it + 1]));
        }
    }
    constexpr int64_t B = 2;
    int64_t const ncelm = n - 1;
    int64_t const nx = ncelm * 2 + (1 + B * 2);
    for (int64_t it = 0; it < n; ++it)
    {
        xc[it * 2 + B] = xloc[it];
    }
    for (int64_t it = 0; it < ncelm; ++it)
    {
        int64_t const ref = it * 2 + B + 1;
        xc[ref] = (xc[ref - 1] + xc[ref + 1]) / 2;
    }
    for (int64_t it = 1; it <= B; ++it)
    {
        xc[B - it] = xc[B] + xc[B] - xc[B + it];
        int64_t const ref = nx - B - 1;
        xc[ref + it] = xc[ref] + xc[ref] - xc[ref - it];
    }
}

// Grid::Grid(xmin, xmax, ncelm) (core.cpp:14-45).
void st_grid(double xmin, double xmax, int64_t ncelm, double * xc)
{
    std::string const head = "Grid::Grid(xmin=" + fmt_num(xmin) + ", xmax=" + fmt_num(xmax)
                             + ", ncelm=" + std::to_string(ncelm) + ")";
    if (ncelm < 1)
    {
        fail(SC_EVALUE, head + " invalid argument: ncelm smaller than 1");
    }
    if (xmin >= xmax)
    {
        fail(SC_EVALUE, head + " invalid arguments: xmin >= xmax");
    }
    double const xspace = (xmax - xmin) / static_cast<double>(ncelm);
    std::vector<double> xloc(static_cast<size_t>(ncelm + 1));
    xloc[0] = xmin;
    for (int64_t it = 1; it < ncelm; ++it)
    {
        xloc[it] = xloc[it - 1] + xspace;
    }
    xloc[ncelm] = xmax;
    st_grid_from(xloc.data(), ncelm + 1, xc);
}

struct StCore
{
    int kind = K_LINEAR;
    int64_t nx = 0;
    int64_t ncelm = 0;
    int64_t nvar = 1;
    double dt = 0.0;
    StreamHolder sh; // declared before the buffers: freed last
    hipStream_t stream = sh.s;
    DevBuf<double> xc, so0, so1, cfl;

    StCore(int k, int64_t n, int64_t nv, double time_increment)
        : kind(k)
        , nx(n)
        , ncelm((n - 5) / 2)
        , nvar(nv)
        , dt(time_increment)
    {
        AllocScope const scope(stream);
        xc.alloc(nx);
        so0.alloc(nx * nvar);
        so1.alloc(nx * nvar);
        cfl.alloc(nx);
        SC_HIP(hipMemsetAsync(xc.ptr, 0, xc.bytes(), stream));
        SC_HIP(hipMemsetAsync(so0.ptr, 0, so0.bytes(), stream));
        SC_HIP(hipMemsetAsync(so1.ptr, 0, so1.bytes(), stream));
        SC_HIP(hipMemsetAsync(cfl.ptr, 0, cfl.bytes(), stream));
    }

    ~StCore() = default;

    void sync() { SC_HIP(hipStreamSynchronize(stream)); }

    DevBuf<double> & buf(int which)
    {
        switch (which)
        {
        case 0: return xc;
        case 1: return so0;
        case 2: return so1;
        case 3: return cfl;
        default: fail(SC_EVALUE, "spacetime solver: unknown array id");
        }
    }

    // Field::set_time_increment (core.cpp:109-114).
    StArgs args() const { return StArgs{ncelm, nvar, 0.5 * dt, 0.25 * dt, xc.ptr, so0.ptr, so1.ptr, cfl.ptr}; }

    template <int K>
    void so0_pass(bool odd)
    {
        Plane const p = ce_plane(ncelm, odd);
        k_st_so0<K><<<grid_for(p.count, ST_BLOCK), ST_BLOCK, 0, stream>>>(args(), p);
        SC_HIP(hipGetLastError());
    }

    template <int K, int A>
    void so1_pass(bool odd)
    {
        Plane const p = ce_plane(ncelm, odd);
        k_st_so1<K, A><<<grid_for(p.count, ST_BLOCK), ST_BLOCK, 0, stream>>>(args(), p);
        SC_HIP(hipGetLastError());
    }

    template <int K>
    void cfl_pass(bool odd)
    {
        Plane const p = se_plane(ncelm, odd);
        k_st_cfl<K><<<grid_for(p.count, ST_BLOCK), ST_BLOCK, 0, stream>>>(args(), p);
        SC_HIP(hipGetLastError());
    }

    void bnd(double * arr)
    {
        k_st_bnd<<<1, 64, 0, stream>>>(ncelm, nvar, arr);
        SC_HIP(hipGetLastError());
    }

    template <int K, int A>
    void half(bool odd)
    {
        int64_t const count = odd ? ncelm + 1 : ncelm;
        k_st_half<K, A><<<grid_for(count, ST_BLOCK), ST_BLOCK, 0, stream>>>(args(), odd ? 1 : 0);
        SC_HIP(hipGetLastError());
    }

    template <int K, int A>
    void march_k(int64_t steps)
    {
        if (steps <= 0)
        {
            return;
        }
        if (nx <= ST_LDS_MAX)
        {
            k_st_march_lds<K, A><<<1, ST_PBLOCK, 0, stream>>>(args(), steps);
            SC_HIP(hipGetLastError());
            return;
        }
        for (int64_t it = 0; it < steps; ++it)
        {
            half<K, A>(false);
            half<K, A>(true);
        }
    }

    template <int K>
    void stage_k(int stage, int alpha, bool odd)
    {
        switch (stage)
        {
        case 0: so0_pass<K>(odd); break;
        case 1:
            if (alpha == 0)
            {
                so1_pass<K, 0>(odd);
            }
            else if (alpha == 1)
            {
                so1_pass<K, 1>(odd);
            }
            else
            {
                so1_pass<K, 2>(odd);
            }
            break;
        case 2: cfl_pass<K>(odd); break;
        case 3: bnd(so0.ptr); break;
        case 4: bnd(so1.ptr); break;
        case 5:
        case 6:
            if (alpha == 0)
            {
                half<K, 0>(stage == 6);
            }
            else if (alpha == 1)
            {
                half<K, 1>(stage == 6);
            }
            else
            {
                half<K, 2>(stage == 6);
            }
            break;
        default: fail(SC_EVALUE, "spacetime solver: unknown stage");
        }
    }

    void stage(int st, int alpha, bool odd)
    {
        switch (kind)
        {
        case K_DEFAULT: stage_k<K_DEFAULT>(st, alpha, odd); break;
        case K_LINEAR: stage_k<K_LINEAR>(st, alpha, odd); break;
        default: stage_k<K_BURGERS>(st, alpha, odd); break;
        }
    }

    template <int K>
    void march_a(int alpha, int64_t steps)
    {
        if (alpha == 0)
        {
            march_k<K, 0>(steps);
        }
        else if (alpha == 1)
        {
            march_k<K, 1>(steps);
        }
        else
        {
            march_k<K, 2>(steps);
        }
    }

    void march(int alpha, int64_t steps)
    {
        switch (kind)
        {
        case K_DEFAULT: march_a<K_DEFAULT>(alpha, steps); break;
        case K_LINEAR: march_a<K_LINEAR>(alpha, steps); break;
        default: march_a<K_BURGERS>(alpha, steps); break;
        }
    }

    double eval(int what, int alpha, int64_t y, int64_t iv, int store)
    {
        DevBuf<double> out;
        out.alloc(1);
        switch (kind)
        {
        case K_DEFAULT:
            k_st_eval<K_DEFAULT><<<1, 64, 0, stream>>>(args(), what, alpha, y, iv, store, out.ptr);
            break;
        case K_LINEAR:
            k_st_eval<K_LINEAR><<<1, 64, 0, stream>>>(args(), what, alpha, y, iv, store, out.ptr);
            break;
        default: k_st_eval<K_BURGERS><<<1, 64, 0, stream>>>(args(), what, alpha, y, iv, store, out.ptr); break;
        }
        SC_HIP(hipGetLastError());
        double r = 0.0;
        SC_HIP(hipMemcpyAsync(&r, out.ptr, sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        return r;
    }

    void so0p(int64_t iv, bool odd, double * host, int64_t n)
    {
        // get_so0p walks the nselm - odd solution elements from index 0.
        Plane const r{odd ? 3 : 2, ncelm + 1 - (odd ? 1 : 0)};
        if (n != r.count)
        {
            fail(SC_EVALUE, "get_so0p: output size mismatch");
        }
        DevBuf<double> out;
        out.alloc(r.count);
        switch (kind)
        {
        case K_DEFAULT:
            k_st_so0p<K_DEFAULT><<<grid_for(r.count, ST_BLOCK), ST_BLOCK, 0, stream>>>(args(), r, iv, out.ptr);
            break;
        case K_LINEAR:
            k_st_so0p<K_LINEAR><<<grid_for(r.count, ST_BLOCK), ST_BLOCK, 0, stream>>>(args(), r, iv, out.ptr);
            break;
        default:
            k_st_so0p<K_BURGERS><<<grid_for(r.count, ST_BLOCK), ST_BLOCK, 0, stream>>>(args(), r, iv, out.ptr);
            break;
        }
        SC_HIP(hipGetLastError());
        SC_HIP(hipMemcpyAsync(host, out.ptr, out.bytes(), hipMemcpyDeviceToHost, stream));
        sync();
    }
};

} /* end namespace scamd */

using scamd::guard;
using scamd::StCore;

extern "C"
{

int scst_grid(double xmin, double xmax, int64_t ncelm, double * xc, int64_t nx)
{
    return guard(
        [&]()
        {
            if (ncelm >= 1 && nx != 2 * ncelm + 5)
            {
                scamd::fail(scamd::SC_EVALUE, "Grid: coordinate buffer size mismatch");
            }
            scamd::st_grid(xmin, xmax, ncelm, xc);
        });
}

int scst_grid_from(double const * xloc, int64_t n, double * xc, int64_t nx)
{
    return guard(
        [&]()
        {
            if (n >= 2 && nx != 2 * (n - 1) + 5)
            {
                scamd::fail(scamd::SC_EVALUE, "Grid: coordinate buffer size mismatch");
            }
            scamd::st_grid_from(xloc, n, xc);
        });
}

int scst_create(int kind, double const * xc, int64_t nx, int64_t nvar, double time_increment, void ** out)
{
    return guard(
        [&]()
        {
            if (kind < 0 || kind > 2)
            {
                scamd::fail(scamd::SC_EVALUE, "spacetime solver: unknown kernel kind");
            }
            if (nx < 7 || (nx - 5) % 2 != 0)
            {
                scamd::fail(scamd::SC_EVALUE, "spacetime solver: coordinate count must be 2*ncelm+5 with ncelm >= 1");
            }
            if (nvar < 1)
            {
                scamd::fail(scamd::SC_EVALUE, "spacetime solver: nvar must be at least 1");
            }
            auto * h = new StCore(kind, nx, nvar, time_increment);
            SC_HIP(hipMemcpyAsync(h->xc.ptr, xc, h->xc.bytes(), hipMemcpyHostToDevice, h->stream));
            h->sync();
            *out = h;
        });
}

int scst_clone(void * h, void ** out)
{
    return guard(
        [&]()
        {
            auto & a = *static_cast<StCore *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            auto * b = new StCore(a.kind, a.nx, a.nvar, a.dt);
            for (int w = 0; w < 4; ++w)
            {
                SC_HIP(hipMemcpyAsync(b->buf(w).ptr, a.buf(w).ptr, a.buf(w).bytes(), hipMemcpyDeviceToDevice,
                                      a.stream));
            }
            a.sync();
            b->sync();
            *out = b;
        });
}

void scst_destroy(void * h) { delete static_cast<StCore *>(h); }

int scst_set_time_increment(void * h, double time_increment)
{
    static_cast<StCore *>(h)->dt = time_increment;
    return 0;
}

int scst_upload(void * h, int which, double const * host, int64_t count)
{
    return guard(
        [&]()
        {
            auto & a = *static_cast<StCore *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            auto & b = a.buf(which);
            if (count != b.count)
            {
                scamd::fail(scamd::SC_EVALUE, "spacetime solver: array size mismatch");
            }
            SC_HIP(hipMemcpyAsync(b.ptr, host, b.bytes(), hipMemcpyHostToDevice, a.stream));
            a.sync();
        });
}

int scst_download(void * h, int which, double * host, int64_t count)
{
    return guard(
        [&]()
        {
            auto & a = *static_cast<StCore *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            auto & b = a.buf(which);
            if (count != b.count)
            {
                scamd::fail(scamd::SC_EVALUE, "spacetime solver: array size mismatch");
            }
            SC_HIP(hipMemcpyAsync(host, b.ptr, b.bytes(), hipMemcpyDeviceToHost, a.stream));
            a.sync();
        });
}

int scst_stage(void * h, int stage, int alpha, int odd_plane)
{
    return guard(
        [&]()
        {
            if (alpha < 0 || alpha > 2)
            {
                scamd::fail(scamd::SC_EVALUE, "alpha must be 0, 1 or 2");
            }
            auto & a = *static_cast<StCore *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            a.stage(stage, alpha, odd_plane != 0);
            a.sync();
        });
}

int scst_march(void * h, int alpha, int64_t steps)
{
    return guard(
        [&]()
        {
            if (alpha < 0 || alpha > 2)
            {
                scamd::fail(scamd::SC_EVALUE, "alpha must be 0, 1 or 2");
            }
            auto & a = *static_cast<StCore *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            a.march(alpha, steps);
            a.sync();
        });
}

int scst_march_async(void * h, int alpha, int64_t steps)
{
    return guard(
        [&]()
        {
            if (alpha < 0 || alpha > 2)
            {
                scamd::fail(scamd::SC_EVALUE, "alpha must be 0, 1 or 2");
            }
            static_cast<StCore *>(h)->march(alpha, steps);
        });
}

int scst_sync(void * h)
{
    return guard([&]() { static_cast<StCore *>(h)->sync(); });
}

int scst_eval(void * h, int what, int alpha, int64_t xindex, int64_t iv, int store, double * out)
{
    return guard(
        [&]()
        {
            auto & a = *static_cast<StCore *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            if (what < 0 || what > 7)
            {
                scamd::fail(scamd::SC_EVALUE, "spacetime solver: unknown element value");
            }
            if (iv < 0 || iv >= a.nvar)
            {
                scamd::fail(scamd::SC_EINDEX, "out of nvar range");
            }
            int64_t const lo = what >= 6 ? 2 : 1;
            if (xindex < lo || xindex >= a.nx - lo)
            {
                scamd::fail(scamd::SC_EINDEX, "spacetime solver: element index out of range");
            }
            *out = a.eval(what, alpha, xindex, iv, store);
        });
}

int scst_so0p(void * h, int64_t iv, int odd_plane, double * host_out, int64_t n)
{
    return guard(
        [&]()
        {
            auto & a = *static_cast<StCore *>(h);
            scamd::AllocScope const alloc_scope_(a.stream);
            if (iv < 0 || iv >= a.nvar)
            {
                scamd::fail(scamd::SC_EINDEX, "get_so0p(): out of nvar range");
            }
            a.so0p(iv, odd_plane != 0, host_out, n);
        });
}

} /* end extern "C" */
