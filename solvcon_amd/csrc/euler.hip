/*
 * Multidimensional Euler CESE march on the GPU: the device counterpart of
 * solvcon's EulerCore (cpp/solvcon/multidim/euler.hpp:56-211 and
 * euler_ce.cpp, euler_march.cpp, euler_solution.cpp, euler_gradient.cpp,
 * euler_bc.cpp).
 *
 * Layout in HBM.  Body cells are grouped by cell type and, inside a group,
 * ordered along a Hilbert curve of their centroids; ghost rows keep the
 * reference order and come first.  A row is therefore 0..G-1 for ghosts and
 * G + (internal position) for body cells.  Solution tables are row-major
 * records as in the reference, table[row * width + component] (so0 and so0t
 * with neq components, so1 with neq * ndim, the solution points with
 * ndim), so gathering a neighbour's state touches one or two cache lines
 * instead of one line per component.  Each type group
 * stores its CE geometry as geo[field * n + cell] (CE volumes, BCE
 * centroids, sub-face centroids and normals, CFL distance) and its
 * neighbour rows as nbr[face * n + cell].
 *
 * One substep of march() is one fused kernel per type group: the CE flux
 * integral (calc_soln), the ghost update of every boundary face the cell
 * owns (bc_soln from the pre-CFL state, bc_dsoln from so1c), the CFL and
 * energy clamp (calc_cfl) and the gradient element with its weighting and
 * limiter (calc_dsoln).  The neighbours' so0t is recomputed from the
 * Jacobian that the flux integral builds anyway, so so0t is written only
 * once at the end of march() (identical to the reference's last
 * calc_solt).  The fine-grained stage methods run the same device code one
 * stage at a time, reading so0t from memory the way the reference does.
 *
 * The API tables (cevol, cecnd, so0c, ...) keep the reference layout and
 * numbering; they are produced and consumed by gather/scatter kernels.
 */

#include "euler_kernels.hpp"
#include "mesh.hpp"
#include "scan.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <vector>

// Minimum waves per SIMD requested for the per-cell march kernel.
#ifndef SC_CELL_WAVES
#define SC_CELL_WAVES 1
#endif

namespace scamd
{

void sort_pairs_u64(uint64_t * kin, uint64_t * kout, int32_t * vin, int32_t * vout, int64_t n, hipStream_t s);

namespace
{

constexpr unsigned BLK = 256;

enum EulerArray : int
{
    EA_CEVOL = 0,
    EA_CECND,
    EA_SFCND,
    EA_SFNML,
    EA_SO0C,
    EA_SO0N,
    EA_SO0T,
    EA_SO1C,
    EA_SO1N,
    EA_STM,
    EA_CFLO,
    EA_CFLC,
    EA_GAMMA,
    EA_COUNT,
};

enum Mode : int
{
    M_SOLN = 1,
    M_BC = 2,
    M_CFL = 4,
    M_DSOLN = 8,
    M_STMEM = 16,
    M_BATCH = 32,
    M_W3 = 64,  // compiled for three wavefronts per SIMD (at most 168 registers)
    M_W2 = 128, // two wavefronts per SIMD (at most 256 registers)
    M_W4 = 256, // four wavefronts per SIMD (at most 128 registers)
    M_WAVES = M_W2 | M_W3 | M_W4,
    M_STAGE = 512, // stage the block's own rows in LDS for neighbour reads
    M_GDMA = 1024, // prefetch the sub-face geometry into LDS (LDS DMA)
    M_STNEXT = 4096, // also write so0t of the level being written (chained so0t)
    M_FAST = 8192, // skip the Jacobian's structural zeros, share reciprocals; report where inexact
    M_FUSED = M_SOLN | M_BC | M_CFL | M_DSOLN,
};

// Minimum wavefronts per SIMD a k_cell variant is compiled for; the
// register budget per lane is 512 / waves.
constexpr int waves_of(int mode)
{
    return (mode & M_W4) != 0 ? 4 : (mode & M_W3) != 0 ? 3 : (mode & M_W2) != 0 ? 2 : SC_CELL_WAVES;
}

// Kernel roles: 0 the fused kernel, 1 soln + BC + CFL, 2 dsoln.  Each
// role is compiled in a few variants: a minimum wave count per SIMD (which
// caps registers), whether the block's own rows are staged in LDS, whether
// the fused kernel reads so0t from memory and whether the sub-face geometry
// is prefetched into LDS.  A variant code is waves + 10 * staged + 100 *
// so0t read + 1000 * geometry prefetch.
constexpr int NVAR = 5;

constexpr int mode_waves(int waves)
{
    return waves == 4 ? M_W4 : waves == 3 ? M_W3 : waves == 2 ? M_W2 : 0;
}

constexpr int role_variant(int nd, int role, int v)
{
    // 0 marks an unused slot.
    if (nd == 2)
    {
        return role == 0 ? (v == 0 ? 1 : v == 1 ? 3 : v == 2 ? 11 : v == 3 ? 1001 : 1003) : (v == 0 ? 1 : 0);
    }
    if (role == 0)
    {
        return v == 0 ? 1 : v == 1 ? 2 : v == 2 ? 102 : v == 3 ? 112 : 1102;
    }
    return role == 1 ? (v == 0 ? 1 : v == 1 ? 11 : 0) : (v == 0 ? 1 : v == 1 ? 2 : v == 2 ? 12 : 0);
}

// Mode bits of a variant code: waves + 10 when LDS-staged + 100 when the
// fused kernel reads the neighbours' so0t from memory (calc_solt first)
// instead of recomputing it + 1000 when the sub-face geometry is prefetched
// into LDS.
constexpr int variant_mode(int code)
{
    return mode_waves(code % 10) | ((code / 10) % 10 != 0 ? M_STAGE : 0) | ((code / 100) % 10 != 0 ? M_STMEM : 0) |
           ((code / 1000) % 10 != 0 ? M_GDMA : 0);
}

// The batched march runs the fused kernel that recomputes so0t, at three
// wavefronts per SIMD for triangles, two for 3D cells and one for quads
// (the fastest such variants of the plain march that do not spill, or that
// win despite spilling in 3D).
constexpr int batch_waves(int type)
{
    return type == CT_TRI ? M_W3 : type == CT_QUAD ? 0 : M_W2;
}

// Fields of the sub-face geometry prefetch: the sub-face centroids and
// normals in 2D (2 S d fields), the centroids alone in 3D (S d).  LDS bytes
// per block (64 cells per wave, four waves), and whether the blocks that the
// mode's wave count keeps resident on a CU fit in its 160 KB.
constexpr int gdma_fields(int type)
{
    return type == CT_TRI ? 2 * 6 * 2 : type == CT_QUAD ? 2 * 8 * 2 : type == CT_TET ? 12 * 3 : type == CT_HEX ? 24 * 3
           : type == CT_PRISM ? 18 * 3 : type == CT_PYRAMID ? 16 * 3 : 0;
}

constexpr int gdma_bytes(int type)
{
    return 4 * gdma_fields(type) * 64 * 8;
}

constexpr bool gdma_fits(int type, int mode)
{
    int const w = (mode & M_W4) != 0 ? 4 : (mode & M_W3) != 0 ? 3 : (mode & M_W2) != 0 ? 2 : 1;
    return gdma_bytes(type) > 0 && w * gdma_bytes(type) <= 160 * 1024;
}

constexpr bool variant_built(int nd, int mode)
{
    bool const soln = (mode & M_SOLN) != 0;
    bool const dsoln = (mode & M_DSOLN) != 0;
    int const role = soln && dsoln ? 0 : soln ? 1 : 2;
    if (role != 0 && (mode & M_GDMA) != 0)
    {
        return false;
    }
    int const vmask = M_WAVES | M_STAGE | (role == 0 ? M_STMEM | M_GDMA : 0);
    if ((mode & vmask) == 0)
    {
        return true;
    }
    if ((mode & M_BATCH) != 0)
    {
        return (mode & (vmask & ~M_GDMA)) == (nd == 2 ? M_W3 : M_W2) && (nd == 2 || (mode & M_GDMA) == 0);
    }
    for (int v = 0; v < NVAR; ++v)
    {
        int const c = role_variant(nd, role, v);
        if (c != 0 && (variant_mode(c) & vmask) == (mode & vmask))
        {
            return true;
        }
    }
    return false;
}

struct KP
{
    int64_t R;
    int64_t G;
    double hdt;
    double qdt;
    double sigma0;
    double taumin;
    double tauscale;
    double const * s0c;
    double const * s1c;
    double * s0n;
    double * s1n;
    double * st;
    double * stn; // so0t of the level being written (chained so0t)
    double const * gam;
    double * cflo;
    double * cflc;
    double const * cen;
    int32_t const * gkind;
    int32_t const * ginl;
    double const * gmat;
    double const * inlet;
    // Batched solves: the instance rank of every body cell (internal order)
    // and the half/quarter time increments of every rank.
    int32_t const * inst;
    double const * bhdt;
    double const * bqdt;
    // The fast/exact pair of every plain launch: the fast kernel stores seq
    // in *redo when one of its cells may differ from the exact arithmetic,
    // and the exact kernel launched after it with gate set returns at once
    // unless *redo holds that seq (then it rewrites the whole group).
    unsigned long long * redo = nullptr;
    unsigned long long seq = 0;
    int gate = 0;
};

struct GV
{
    int64_t off;
    int64_t n;   // cells of the group: the stride of its SoA tables
    int32_t const * nbr;
    double const * geo;
    int64_t cnt; // cells to process, a prefix of the group
};

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

// --------------------------------------------------------------------------
// The per-cell march kernel.
// --------------------------------------------------------------------------

// Select between two values that have both been loaded (no branch).
__device__ inline double pick(bool first, double a, double b) { return first ? a : b; }

template <int T, int MODE>
__global__ void __launch_bounds__(BLK, waves_of(MODE)) k_cell(KP const p, GV const g)
{
    using TR = Traits<T>;
    using LY = Layout<T>;
    constexpr int ND = TR::ND;
    constexpr int NF = TR::NF;
    constexpr int NEQ = ND + 2;
    constexpr int NFGE = TR::NFGE;

    constexpr bool FAST = (MODE & M_FAST) != 0;
    static_assert(!FAST || (MODE & M_BATCH) == 0, "the batched kernels are exact");
    // redo[0]: the sequence number of the last fast launch that reported a
    // cell; redo[1]: exact passes run; redo[2]: set by an exact pass, so the
    // rest of the march runs the exact form alone (the fast kernels return
    // at once) instead of both kernels of every pair.
    if constexpr (FAST)
    {
        if (p.redo[2] != 0)
        {
            return;
        }
    }
    else
    {
        if (p.gate != 0)
        {
            if (*p.redo != p.seq && p.redo[2] == 0)
            {
                return;
            }
            if (blockIdx.x == 0 && threadIdx.x == 0)
            {
                atomicAdd(p.redo + 1, 1ull); // exact passes run
                p.redo[2] = 1;
            }
        }
    }
    FastChk chk; // FAST: whether every shortcut of this cell equals the exact arithmetic

    // Wave priority.  At one to three waves per SIMD a wave that waits to
    // issue loads behind a wave busy with FP64 arithmetic leaves the memory
    // pipeline idle.  A wave therefore runs at a raised priority while it
    // issues the loads that start its work (the block's staged rows, the
    // neighbour lists and, in 3D, the first face's neighbour row) and the
    // loads of the gradient (in 3D the whole section through the
    // displacement matrices, in 2D its neighbour gathers), and at the base
    // priority for the arithmetic in between.  The placement is the fastest
    // of those measured per dimension.  Only the issue order changes, no
    // value.
    //   prio_load(): raise; prio_math(): lower.
    auto prio_load = [] { __builtin_amdgcn_s_setprio(2); };
    auto prio_math = [] { __builtin_amdgcn_s_setprio(0); };
    prio_load();

    int64_t const i0 = static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x;
    int64_t const i = i0 + threadIdx.x;
    // Batched launches may cover cells of no marching instance: the serving
    // arena's free blocks, finished instances and the tail of an instance's
    // last block all carry an instance whose half time increment is
    // negative.  An instance owns whole blocks, so a block whose first cell
    // is dead returns at once, before any load.
    if constexpr ((MODE & M_BATCH) != 0)
    {
        if (p.bhdt[p.inst[g.off + i0]] < 0.0)
        {
            return;
        }
    }

    // The block's own rows of the current level (so0c, so1c, so0t, gamma,
    // solution points) are staged in LDS with coalesced loads; a neighbour
    // that is one of them is then read from LDS instead of being gathered
    // from global memory.  In SFC order most neighbours are.
    constexpr bool STAGE = (MODE & M_STAGE) != 0 && (MODE & (M_SOLN | M_DSOLN)) != 0;
    constexpr bool STG_ST = ((MODE & M_SOLN) != 0 && (MODE & M_STMEM) != 0) || (MODE & (M_SOLN | M_DSOLN)) == M_DSOLN;
    constexpr bool STG_GM = (MODE & M_SOLN) != 0;
    constexpr int O_S1 = BLK * NEQ;
    constexpr int O_ST = O_S1 + BLK * NEQ * ND;
    constexpr int O_GM = O_ST + (STG_ST ? BLK * NEQ : 0);
    constexpr int O_CN = O_GM + (STG_GM ? BLK : 0);
    constexpr int LW = STAGE ? O_CN + BLK * ND : 1;
    __shared__ double lds[LW];
    // Sub-face geometry prefetch: each wave copies the sub-face centroids
    // (and in 2D the normals) of its 64 cells into its own LDS region with
    // LDS-DMA loads of 16 B per lane (lanes 0-31 one field, lanes 32-63 the
    // next, two cells each), so the geometry arrives in one batch while the
    // neighbour rows are gathered.  Field G0 + j of lane l's cell is then
    // gw[j * 64 + l].  The host launches it only when the field stride is
    // even and the table 16-byte aligned, so every load is aligned; a lane
    // reads at most one cell past the group's count, which is still inside
    // the next field.
    constexpr bool GDMA = (MODE & M_GDMA) != 0 && (MODE & M_SOLN) != 0;
    constexpr int G0 = LY::FC;
    constexpr int GN = gdma_fields(T);
    static_assert(!GDMA || (GN == (ND == 2 ? 2 : 1) * TR::S * ND && GN % 2 == 0), "prefetched fields");
    constexpr int GW = GDMA ? GN * 64 : 1;
    __shared__ double gsh[GDMA ? (BLK / 64) * GW : 1];
    int const lane = threadIdx.x & 63;
    double * const gw = gsh + (GDMA ? (threadIdx.x >> 6) * GW : 0);
    if constexpr (GDMA)
    {
        int64_t const c = i0 + (threadIdx.x & ~63u) + 2 * (lane & 31);
        bool live = c < g.cnt;
        if constexpr ((MODE & M_BATCH) != 0)
        {
            // No loads for the dead tail of an instance's last block, so a
            // wave of tail cells issues none and may return early.
            live = live && p.bhdt[p.inst[g.off + c]] >= 0.0;
        }
        if (live)
        {
#pragma unroll
            for (int q = 0; q < GN / 2; ++q)
            {
                int64_t const fld = G0 + 2 * q + (lane >> 5);
                __builtin_amdgcn_global_load_lds((void const __attribute__((address_space(1))) *)(g.geo + fld * g.n + c),
                                                 (void __attribute__((address_space(3))) *)(gw + q * 128), 16, 0, 0);
            }
        }
    }
    int64_t const rbase = p.G + g.off + i0;
    int const nrow = static_cast<int>(g.cnt - i0 < BLK ? g.cnt - i0 : BLK);
    if constexpr (STAGE)
    {
        for (int q = threadIdx.x; q < nrow * NEQ; q += BLK)
        {
            lds[q] = p.s0c[rbase * NEQ + q];
        }
        for (int q = threadIdx.x; q < nrow * NEQ * ND; q += BLK)
        {
            lds[O_S1 + q] = p.s1c[rbase * (NEQ * ND) + q];
        }
        if constexpr (STG_ST)
        {
            for (int q = threadIdx.x; q < nrow * NEQ; q += BLK)
            {
                lds[O_ST + q] = p.st[rbase * NEQ + q];
            }
        }
        if constexpr (STG_GM)
        {
            for (int q = threadIdx.x; q < nrow; q += BLK)
            {
                lds[O_GM + q] = p.gam[rbase + q];
            }
        }
        for (int q = threadIdx.x; q < nrow * ND; q += BLK)
        {
            lds[O_CN + q] = p.cen[rbase * ND + q];
        }
        __syncthreads();
    }
    if (i >= g.cnt)
    {
        return;
    }
    if constexpr ((MODE & M_BATCH) != 0)
    {
        if (p.bhdt[p.inst[g.off + i]] < 0.0)
        {
            return;
        }
    }
    int64_t const n = g.n;
    int64_t const row = p.G + g.off + i;
    double const * geo = g.geo + i;
    // One sub-face geometry field of this cell, from the prefetch or HBM.
    // The CE geometry, the neighbour lists and the rows this kernel writes
    // stream through once per substep: their loads and stores are
    // non-temporal, which leaves the L2 to the solution rows that the
    // neighbours read again.
    auto gsub = [&](int fld) -> double
    {
        if constexpr (GDMA)
        {
            if (fld < G0 + GN)
            {
                return gw[(fld - G0) * 64 + lane];
            }
        }
        return __builtin_nontemporal_load(geo + fld * n);
    };
    double hdt = p.hdt;
    double qdt = p.qdt;
    if constexpr ((MODE & M_BATCH) != 0)
    {
        int32_t const r = p.inst[g.off + i];
        hdt = p.bhdt[r];
        qdt = p.bqdt[r];
    }

    // a / d with a shared reciprocal r = recip(d) in the fast form, the
    // division in the exact form.
    auto recip = [&](double d) -> double
    {
        if constexpr (FAST)
        {
            chk.den(d);
            return div_recip(d);
        }
        return 0.0;
    };
    auto quot = [&](double a, double d, double r) -> double
    {
        if constexpr (FAST)
        {
            return div_fast(a, d, r, chk);
        }
        return a / d;
    };

    int64_t nb[NF];
    int lr[NF]; // the neighbour's row in the LDS copy, or -1
#pragma unroll
    for (int f = 0; f < NF; ++f)
    {
        nb[f] = __builtin_nontemporal_load(g.nbr + f * n + i);
        lr[f] = (STAGE && nb[f] >= rbase && nb[f] < rbase + nrow) ? static_cast<int>(nb[f] - rbase) : -1;
    }
    if constexpr (ND == 2)
    {
        prio_math();
    }
    // Neighbour solution points: CE centroid (body) or mesh centroid (ghost).
    // A neighbour value is read from both places with no branch: from LDS
    // at its staged row (row 0 when not staged) and from global memory at
    // its row, where every staged lane reads the block's first row instead;
    // the lanes of one load then touch only the cache lines of the
    // neighbours that are not staged.
    //
    // In 3D the solution points and, when so0t is read from memory, the
    // neighbours' so0t are loaded again where they are used instead of
    // being held in registers from the flux loop to the gradient: the same
    // values (the tables are not written by this kernel), a shorter live
    // range, fewer spills.
    constexpr bool RELOAD_CEN = ND == 3;
    constexpr bool RELOAD_ST = ND == 3 && ((MODE & M_STMEM) != 0 || (MODE & M_SOLN) == 0);
    auto cen_at = [&](int f, int d) -> double
    {
        bool const inb = lr[f] >= 0;
        double const * cg = p.cen + (inb ? rbase : nb[f]) * ND;
        double const * cl = lds + O_CN + (inb ? lr[f] : 0) * ND;
        return pick(inb, cl[d], cg[d]);
    };
    auto st_at = [&](int f, int k) -> double
    {
        bool const inb = lr[f] >= 0;
        int const l = inb ? lr[f] : 0;
        int64_t const jg = inb ? rbase : nb[f];
        return pick(inb, lds[O_ST + l * NEQ + k], p.st[jg * NEQ + k]);
    };
    double jce[RELOAD_CEN ? 1 : NF][ND];
    if constexpr (!RELOAD_CEN)
    {
#pragma unroll
        for (int f = 0; f < NF; ++f)
        {
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                jce[f][d] = cen_at(f, d);
            }
        }
    }
    double st[RELOAD_ST ? 1 : NF][NEQ];
    double u0[NEQ];

    if constexpr ((MODE & M_SOLN) != 0)
    {
        if constexpr (GDMA)
        {
            // The compiler does not track LDS-DMA loads: wait for them here,
            // and let no LDS read move above the wait.
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        // calc_soln_impl (euler_march.cpp:219-302).
        double acc[NEQ];
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
            acc[k] = 0.0;
        }
        static_for<0, NF>([&](auto fc) {
            constexpr int f = decltype(fc)::value;
            int64_t const j = nb[f];
            double js[NEQ];
            double j1[NEQ][ND];
            bool const inb = lr[f] >= 0;
            int const l = inb ? lr[f] : 0;
            int64_t const jg = inb ? rbase : j;
#pragma unroll
            for (int k = 0; k < NEQ; ++k)
            {
                js[k] = pick(inb, lds[l * NEQ + k], p.s0c[jg * NEQ + k]);
#pragma unroll
                for (int d = 0; d < ND; ++d)
                {
                    j1[k][d] = pick(inb, lds[O_S1 + l * (NEQ * ND) + k * ND + d], p.s1c[jg * (NEQ * ND) + k * ND + d]);
                }
            }
            // This face's neighbour so0t (from memory or recomputed) and
            // solution point.
            double stf[NEQ];
            if constexpr ((MODE & M_STMEM) != 0)
            {
#pragma unroll
                for (int k = 0; k < NEQ; ++k)
                {
                    stf[k] = pick(inb, lds[O_ST + l * NEQ + k], p.st[jg * NEQ + k]);
                }
            }
            double jcef[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                if constexpr (RELOAD_CEN)
                {
                    jcef[d] = cen_at(f, d);
                }
                else
                {
                    jcef[d] = jce[f][d];
                }
            }
            double const jgam = pick(inb, lds[O_GM + l], p.gam[jg]);
            double const bvol = __builtin_nontemporal_load(geo + (LY::FV + 1 + f) * n);
#pragma unroll
            for (int k = 0; k < NEQ; ++k)
            {
                double fusp = js[k];
#pragma unroll
                for (int d = 0; d < ND; ++d)
                {
                    fusp += (geo[(LY::FB + f * ND + d) * n] - jcef[d]) * j1[k][d];
                }
                acc[k] += fusp * bvol;
            }
            if constexpr (ND == 3 && f == 0)
            {
                prio_math();
            }
            JacS<ND, FAST> jac;
            jac.set(jgam, js, &chk);
            if constexpr ((MODE & M_STMEM) == 0)
            {
                solt_of<ND>(jac, j1, stf);
            }
            if constexpr (!RELOAD_ST)
            {
#pragma unroll
                for (int k = 0; k < NEQ; ++k)
                {
                    st[f][k] = stf[k];
                }
            }
            constexpr int NS = TR::nsub(f);
            // The solution at every sub-face centroid of this face, then
            // the flux through each sub-face one Jacobian row (equation)
            // at a time, so only one row is live; acc[k] still receives
            // its sub-face terms in sub-face order.
            double usfc[NS][NEQ];
#pragma unroll
            for (int s = 0; s < NS; ++s)
            {
                int const sf = TR::soff(f) + s;
#pragma unroll
                for (int k = 0; k < NEQ; ++k)
                {
                    usfc[s][k] = qdt * stf[k];
#pragma unroll
                    for (int d = 0; d < ND; ++d)
                    {
                        usfc[s][k] += (gsub(LY::FC + sf * ND + d) - jcef[d]) * j1[k][d];
                    }
                }
            }
            static_for<0, NEQ>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                double jr[NEQ][ND];
                double fr[ND];
                jac.template row<k>(jr, fr);
#pragma unroll
                for (int s = 0; s < NS; ++s)
                {
                    int const sf = TR::soff(f) + s;
                    double dot = 0.0;
#pragma unroll
                    for (int d2 = 0; d2 < ND; ++d2)
                    {
                        double dfcn = fr[d2];
#pragma unroll
                        for (int m = 0; m < NEQ; ++m)
                        {
                            if (!FAST || !jac_zero(ND, k, m, d2))
                            {
                                dfcn += jr[m][d2] * usfc[s][m];
                            }
                        }
                        dot += dfcn * gsub(LY::FN + sf * ND + d2);
                    }
                    acc[k] -= hdt * dot;
                }
            });
        });
        double const cvol = __builtin_nontemporal_load(geo + LY::FV * n);
        double const rcv = recip(cvol);
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
            u0[k] = quot(acc[k], cvol, rcv);
        }
    }

    if constexpr ((MODE & M_BC) != 0)
    {
        // bc_soln (from the pre-CFL state) and bc_dsoln (from so1c) of the
        // boundary faces this cell owns (euler_bc.cpp:143-320).
#pragma unroll
        for (int f = 0; f < NF; ++f)
        {
            int64_t const j = nb[f];
            if (j < p.G)
            {
                int32_t const kind = p.gkind[j];
                if (kind >= 0)
                {
                    double mat[ND][ND];
#pragma unroll
                    for (int a = 0; a < ND; ++a)
                    {
#pragma unroll
                        for (int b = 0; b < ND; ++b)
                        {
                            mat[a][b] = p.gmat[(a * ND + b) * p.G + j];
                        }
                    }
                    double o0[NEQ];
                    bc_order0<ND>(kind, mat, p.inlet + static_cast<int64_t>(p.ginl[j]) * (ND + 3), u0, o0);
#pragma unroll
                    for (int k = 0; k < NEQ; ++k)
                    {
                        p.s0n[(j) * NEQ + k] = o0[k];
                    }
                    double in1[NEQ][ND];
#pragma unroll
                    for (int k = 0; k < NEQ; ++k)
                    {
#pragma unroll
                        for (int d = 0; d < ND; ++d)
                        {
                            in1[k][d] = p.s1c[(row) * (NEQ * ND) + k * ND + d];
                        }
                    }
                    double o1[NEQ][ND];
                    bc_order1<ND>(kind, mat, in1, o1);
#pragma unroll
                    for (int k = 0; k < NEQ; ++k)
                    {
#pragma unroll
                        for (int d = 0; d < ND; ++d)
                        {
                            p.s1n[(j) * (NEQ * ND) + k * ND + d] = o1[k][d];
                        }
                    }
                    if constexpr ((MODE & M_STNEXT) != 0)
                    {
                        // so0t of the ghost row for the level being
                        // written (calc_solt_impl, euler_march.cpp:161-191).
                        JacS<ND, FAST> gj;
                        gj.set(p.gam[j], o0, &chk);
                        double gst[NEQ];
                        solt_of<ND>(gj, o1, gst);
                        if constexpr (FAST)
                        {
                            chk.finite(gst[NEQ - 1]);
                        }
#pragma unroll
                        for (int k = 0; k < NEQ; ++k)
                        {
                            p.stn[(j) * NEQ + k] = gst[k];
                        }
                    }
                }
            }
        }
    }

    double cflc_v = 0.0;
    if constexpr ((MODE & M_CFL) != 0)
    {
        // calc_cfl (euler_solution.cpp:114-159).
        if constexpr ((MODE & M_SOLN) == 0)
        {
#pragma unroll
            for (int k = 0; k < NEQ; ++k)
            {
                u0[k] = p.s0n[(row) * NEQ + k];
            }
        }
        double const dist = __builtin_nontemporal_load(geo + LY::FD * n);
        double const ga = p.gam[row];
        double const ga1 = ga - 1.0;
        double const density = u0[0];
        double momsq = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            double const mom = u0[1 + d];
            momsq += mom * mom;
        }
        double const ke = momsq / (2.0 * density);
        double const energy = u0[NEQ - 1];
        double const pr = ga1 * (energy - ke);
        double const pra = (pr + fabs(pr)) / 2.0;
        double const wspd = sqrt(ga * pra / density) + sqrt(momsq) / density;
        double const co = hdt * wspd / dist;
        __builtin_nontemporal_store(co, p.cflo + row);
        cflc_v = (co - 1.0) * pra / (pra + TINYJ) + 1.0;
        __builtin_nontemporal_store(cflc_v, p.cflc + row);
        u0[NEQ - 1] = pra / ga1 + ke + TINYJ;
    }
    if constexpr ((MODE & (M_SOLN | M_CFL)) != 0)
    {
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
            __builtin_nontemporal_store(u0[k], p.s0n + (row) * NEQ + k);
            if constexpr (FAST && (MODE & M_SOLN) != 0)
            {
                chk.finite(u0[k]);
            }
        }
    }

    if constexpr ((MODE & M_DSOLN) != 0)
    {
        if constexpr (ND == 3)
        {
            prio_load();
        }
        // calc_dsoln (euler_gradient.cpp:23-130) with GradientElement
        // (GradientElement.cpp:137-213).
        if constexpr ((MODE & M_CFL) == 0)
        {
            cflc_v = p.cflc[row];
#pragma unroll
            for (int k = 0; k < NEQ; ++k)
            {
                u0[k] = p.s0n[(row) * NEQ + k];
            }
        }
        if constexpr ((MODE & M_SOLN) == 0 && !RELOAD_ST)
        {
#pragma unroll
            for (int f = 0; f < NF; ++f)
            {
#pragma unroll
                for (int k = 0; k < NEQ; ++k)
                {
                    st[f][k] = st_at(f, k);
                }
            }
        }
        double const acfl = fabs(cflc_v);
        double const sgm0 = p.sigma0 / acfl;
        double const tau = p.taumin + acfl * p.tauscale;
        double ice[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            ice[d] = p.cen[(row) * ND + d];
        }
        double gp[NF][ND];
#pragma unroll
        for (int f = 0; f < NF; ++f)
        {
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                double const mid = geo[(LY::FB + f * ND + d) * n];
                double jc;
                if constexpr (RELOAD_CEN)
                {
                    jc = cen_at(f, d);
                }
                else
                {
                    jc = jce[f][d];
                }
                gp[f][d] = mid + tau * (jc - mid);
            }
        }
        double avg[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            avg[d] = 0.0;
        }
#pragma unroll
        for (int f = 0; f < NF; ++f)
        {
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                avg[d] += gp[f][d];
            }
        }
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            avg[d] /= NF;
        }
        // calc_gge_centroid (GradientElement.cpp:81-133).
        double cnd[ND];
        double voc = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            cnd[d] = 0.0;
        }
#pragma unroll
        for (int e = 0; e < NFGE; ++e)
        {
            double sub[ND];
            double dst[ND][ND];
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                sub[d] = avg[d];
            }
#pragma unroll
            for (int v = 0; v < ND; ++v)
            {
                int const fl = TR::fge(e, v) - 1;
#pragma unroll
                for (int d = 0; d < ND; ++d)
                {
                    sub[d] += gp[fl][d];
                    dst[v][d] = gp[fl][d] - avg[d];
                }
            }
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                sub[d] /= static_cast<double>(ND + 1);
            }
            double const vob = ge_det<ND>(dst);
            voc += vob;
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                cnd[d] += sub[d] * vob;
            }
        }
        double const rvoc = recip(voc);
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            cnd[d] = quot(cnd[d], voc, rvoc);
        }
        double idis[NF][ND];
        double udf[NF][NEQ];
        if constexpr (ND == 2)
        {
            prio_load();
        }
#pragma unroll
        for (int f = 0; f < NF; ++f)
        {
            double jdis[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                idis[f][d] = gp[f][d] - cnd[d];
                double jc;
                if constexpr (RELOAD_CEN)
                {
                    jc = cen_at(f, d);
                }
                else
                {
                    jc = jce[f][d];
                }
                jdis[d] = (gp[f][d] - jc) + ice[d] - cnd[d];
            }
            int64_t const j = nb[f];
            bool const inb = lr[f] >= 0;
            int const l = inb ? lr[f] : 0;
            int64_t const jg = inb ? rbase : j;
#pragma unroll
            for (int k = 0; k < NEQ; ++k)
            {
                double stv;
                if constexpr (RELOAD_ST)
                {
                    stv = st_at(f, k);
                }
                else
                {
                    stv = st[f][k];
                }
                double val = pick(inb, lds[l * NEQ + k], p.s0c[jg * NEQ + k]) + hdt * stv - u0[k];
#pragma unroll
                for (int d = 0; d < ND; ++d)
                {
                    val += jdis[d] * pick(inb, lds[O_S1 + l * (NEQ * ND) + k * ND + d], p.s1c[jg * (NEQ * ND) + k * ND + d]);
                }
                udf[f][k] = val;
            }
        }
        if constexpr (ND == 2)
        {
            prio_math();
        }
        // Displacement-matrix adjugates and determinants of every FGE
        // depend only on geometry.  In 2D all adjugates stay in registers
        // across the equations; in 3D (8 x 9 values for a hexahedron) each
        // is formed again per equation, which gives the same values and
        // keeps the register footprint down.
        constexpr bool ADJ_ONCE = NFGE * ND * ND <= 16;
        double adj[ADJ_ONCE ? NFGE : 1][ND][ND];
        double det[NFGE];
        double rdet[NFGE]; // recip(det[e])
#pragma unroll
        for (int e = 0; e < NFGE; ++e)
        {
            double dst[ND][ND];
#pragma unroll
            for (int v = 0; v < ND; ++v)
            {
#pragma unroll
                for (int d = 0; d < ND; ++d)
                {
                    dst[v][d] = idis[TR::fge(e, v) - 1][d];
                }
            }
            if constexpr (ADJ_ONCE)
            {
                ge_adj<ND>(dst, adj[e]);
            }
            det[e] = ge_det<ND>(dst);
            rdet[e] = recip(det[e]);
        }
        if constexpr (ND == 3)
        {
            prio_math();
        }
        constexpr double ofg1 = 1.0 / NFGE;
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
            double gk[NFGE][ND];
            double wk[NFGE];
            double wacc = 0.0;
#pragma unroll
            for (int e = 0; e < NFGE; ++e)
            {
                double ae[ND][ND];
                if constexpr (ADJ_ONCE)
                {
#pragma unroll
                    for (int a = 0; a < ND; ++a)
                    {
#pragma unroll
                        for (int b = 0; b < ND; ++b)
                        {
                            ae[a][b] = adj[e][a][b];
                        }
                    }
                }
                else
                {
                    double dst[ND][ND];
#pragma unroll
                    for (int v = 0; v < ND; ++v)
                    {
#pragma unroll
                        for (int d = 0; d < ND; ++d)
                        {
                            dst[v][d] = idis[TR::fge(e, v) - 1][d];
                        }
                    }
                    ge_adj<ND>(dst, ae);
                }
                double r[ND];
#pragma unroll
                for (int a = 0; a < ND; ++a)
                {
                    r[a] = 0.0;
#pragma unroll
                    for (int b = 0; b < ND; ++b)
                    {
                        r[a] += ae[a][b] * udf[TR::fge(e, b) - 1][k];
                    }
                }
                double sq = 0.0;
#pragma unroll
                for (int d = 0; d < ND; ++d)
                {
                    gk[e][d] = quot(r[d], det[e], rdet[e]);
                    sq += gk[e][d] * gk[e][d];
                }
                double const wgt = 1.0 / sqrt(sq + ALMOST_ZERO);
                wacc += wgt;
                wk[e] = wgt;
            }
            double wmax = 0.0;
            double wmin = 0.0;
            // The weights are 1 / sqrt(sq + 1e-200) in [7e-155, 1e100] (or
            // 0, or NaN) and their sum is no smaller than any: the quotients
            // lie in [2^-850, 1] and take no scaling, and one that is not
            // finite reaches the gradients the cell writes.
            double rwa = 0.0;
            if constexpr (FAST)
            {
                rwa = div_recip(wacc);
            }
#pragma unroll
            for (int e = 0; e < NFGE; ++e)
            {
                double w;
                if constexpr (FAST)
                {
                    w = wk[e] * rwa;
                    double const rem = __builtin_fma(-wacc, w, wk[e]);
                    w = __builtin_fma(rem, rwa, w);
                }
                else
                {
                    w = wk[e] / wacc;
                }
                w -= ofg1;
                wk[e] = w;
                wmax = fmax(wmax, w);
                wmin = fmin(wmin, w);
            }
            double const sm = fmin((1.0 - ofg1) / (wmax + ALMOST_ZERO), -ofg1 / (wmin - ALMOST_ZERO));
            double const smax = fmin(sm, sgm0);
            double out[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                out[d] = 0.0;
            }
#pragma unroll
            for (int e = 0; e < NFGE; ++e)
            {
                double const w = ofg1 + smax * wk[e];
#pragma unroll
                for (int d = 0; d < ND; ++d)
                {
                    out[d] += w * gk[e][d];
                }
            }
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                p.s1n[(row) * (NEQ * ND) + k * ND + d] = out[d];
                if constexpr (FAST)
                {
                    chk.finite(out[d]);
                }
            }
        }
        if constexpr ((MODE & M_STNEXT) != 0)
        {
            // so0t of this row for the level being written (calc_solt_impl,
            // euler_march.cpp:161-191), from the values just stored.
            double s1w[NEQ][ND];
#pragma unroll
            for (int k = 0; k < NEQ; ++k)
            {
#pragma unroll
                for (int d = 0; d < ND; ++d)
                {
                    s1w[k][d] = p.s1n[(row) * (NEQ * ND) + k * ND + d];
                }
            }
            JacS<ND, FAST> rj;
            rj.set(p.gam[row], u0, &chk);
            double rst[NEQ];
            solt_of<ND>(rj, s1w, rst);
            if constexpr (FAST)
            {
                chk.finite(rst[NEQ - 1]);
            }
#pragma unroll
            for (int k = 0; k < NEQ; ++k)
            {
                __builtin_nontemporal_store(rst[k], p.stn + (row) * NEQ + k);
            }
        }
    }
    if constexpr (FAST)
    {
        if (!chk.ok())
        {
            *p.redo = p.seq;
        }
    }
}

// calc_solt over every row, ghosts included (euler_march.cpp:161-191).
template <int ND>
__global__ void k_solt(KP const p)
{
    constexpr int NEQ = ND + 2;
    int64_t const r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= p.R)
    {
        return;
    }
    double s0[NEQ];
    double s1[NEQ][ND];
#pragma unroll
    for (int k = 0; k < NEQ; ++k)
    {
        s0[k] = p.s0c[(r) * NEQ + k];
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            s1[k][d] = p.s1c[(r) * (NEQ * ND) + k * ND + d];
        }
    }
    JacS<ND> jac;
    jac.set(p.gam[r], s0);
    double st[NEQ];
    solt_of<ND>(jac, s1, st);
#pragma unroll
    for (int k = 0; k < NEQ; ++k)
    {
        p.st[(r) * NEQ + k] = st[k];
    }
}

// Ghost updates for the stand-alone bc_soln / bc_dsoln (euler_bc.cpp:143-320).
template <int ND, bool ORDER1>
__global__ void k_bc(KP const p, int32_t const * gsrc)
{
    constexpr int NEQ = ND + 2;
    int64_t const j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= p.G)
    {
        return;
    }
    int32_t const kind = p.gkind[j];
    if (kind < 0)
    {
        return;
    }
    int64_t const src = gsrc[j];
    double mat[ND][ND];
#pragma unroll
    for (int a = 0; a < ND; ++a)
    {
#pragma unroll
        for (int b = 0; b < ND; ++b)
        {
            mat[a][b] = p.gmat[(a * ND + b) * p.G + j];
        }
    }
    if constexpr (!ORDER1)
    {
        double in[NEQ];
        double out[NEQ];
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
            in[k] = p.s0n[(src) * NEQ + k];
        }
        bc_order0<ND>(kind, mat, p.inlet + static_cast<int64_t>(p.ginl[j]) * (ND + 3), in, out);
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
            p.s0n[(j) * NEQ + k] = out[k];
        }
    }
    else
    {
        double in[NEQ][ND];
        double out[NEQ][ND];
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                in[k][d] = p.s1c[(src) * (NEQ * ND) + k * ND + d];
            }
        }
        bc_order1<ND>(kind, mat, in, out);
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
#pragma unroll
            for (int d = 0; d < ND; ++d)
            {
                p.s1n[(j) * (NEQ * ND) + k * ND + d] = out[k][d];
            }
        }
    }
}

// --------------------------------------------------------------------------
// CE geometry (prepare_ce_2d / prepare_ce_3d, euler_ce.cpp:78-347).
// --------------------------------------------------------------------------

struct MeshView
{
    int32_t const * clfcs;
    int32_t const * fccls;
    int32_t const * fcnds;
    double const * ndcrd;
    double const * clcnd;
    double const * fccnd;
    int64_t G;
    int32_t const * int_of_ref;
    int32_t const * ref_of_int;
};

template <int T>
__global__ void k_prepare(MeshView const m, int64_t off, int64_t n, int32_t * nbr, double * geo, double * cen, int64_t R)
{
    using TR = Traits<T>;
    using LY = Layout<T>;
    constexpr int ND = TR::ND;
    constexpr int NF = TR::NF;
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
    {
        return;
    }
    int32_t const icl = m.ref_of_int[off + i];
    int32_t const * cf = m.clfcs + static_cast<int64_t>(icl) * (CLMFC + 1);
    double * gi = geo + i;
    double volc = 0.0;
    double cc[ND];
    double crdi[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d)
    {
        cc[d] = 0.0;
        crdi[d] = m.clcnd[static_cast<int64_t>(icl) * ND + d];
    }
#pragma unroll
    for (int f = 0; f < NF; ++f)
    {
        int32_t const ifc = cf[f + 1];
        int32_t const * fc = m.fccls + static_cast<int64_t>(ifc) * FCREL;
        int32_t const jcl = fc[0] + fc[1] - icl;
        nbr[f * n + i] = jcl < 0 ? static_cast<int32_t>(m.G + jcl) : static_cast<int32_t>(m.G + m.int_of_ref[jcl]);
        double crde[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            crde[d] = m.clcnd[static_cast<int64_t>(jcl) * ND + d];
        }
        int32_t const * fn = m.fcnds + static_cast<int64_t>(ifc) * (FCMND + 1);
        if constexpr (ND == 2)
        {
            double const crd10 = m.ndcrd[static_cast<int64_t>(fn[1]) * 2];
            double const crd11 = m.ndcrd[static_cast<int64_t>(fn[1]) * 2 + 1];
            double const crd20 = m.ndcrd[static_cast<int64_t>(fn[2]) * 2];
            double const crd21 = m.ndcrd[static_cast<int64_t>(fn[2]) * 2 + 1];
            double const cndi0 = (crd10 + crd20 + crdi[0]) / 3.0;
            double const cndi1 = (crd11 + crd21 + crdi[1]) / 3.0;
            double const voli = fabs((crd10 - crdi[0]) * (crd21 - crdi[1]) - (crd11 - crdi[1]) * (crd20 - crdi[0])) / 2.0;
            double const cnde0 = (crd10 + crd20 + crde[0]) / 3.0;
            double const cnde1 = (crd11 + crd21 + crde[1]) / 3.0;
            double const vole = fabs((crd10 - crde[0]) * (crd21 - crde[1]) - (crd11 - crde[1]) * (crd20 - crde[0])) / 2.0;
            double const volb = voli + vole;
            gi[(LY::FV + 1 + f) * n] = volb;
            gi[(LY::FB + f * 2) * n] = (cndi0 * voli + cnde0 * vole) / volb;
            gi[(LY::FB + f * 2 + 1) * n] = (cndi1 * voli + cnde1 * vole) / volb;
            volc += volb;
            cc[0] += (cndi0 * voli + cnde0 * vole);
            cc[1] += (cndi1 * voli + cnde1 * vole);
            double const cross = (crd10 - crde[0]) * (crd21 - crde[1]) - (crd11 - crde[1]) * (crd20 - crde[0]);
            double const sgn = (cross > 0.0) ? 1.0 : -1.0;
            int const s0 = TR::soff(f);
            gi[(LY::FC + s0 * 2) * n] = (crde[0] + crd10) / 2.0;
            gi[(LY::FC + s0 * 2 + 1) * n] = (crde[1] + crd11) / 2.0;
            double const s0dx = crde[0] - crd10;
            double const s0dy = crde[1] - crd11;
            gi[(LY::FN + s0 * 2) * n] = -s0dy * sgn;
            gi[(LY::FN + s0 * 2 + 1) * n] = s0dx * sgn;
            gi[(LY::FC + (s0 + 1) * 2) * n] = (crd20 + crde[0]) / 2.0;
            gi[(LY::FC + (s0 + 1) * 2 + 1) * n] = (crd21 + crde[1]) / 2.0;
            double const s1dx = crd20 - crde[0];
            double const s1dy = crd21 - crde[1];
            gi[(LY::FN + (s0 + 1) * 2) * n] = -s1dy * sgn;
            gi[(LY::FN + (s0 + 1) * 2 + 1) * n] = s1dx * sgn;
        }
        else
        {
            constexpr int MAXN = 4;
            int const nnd = TR::nsub(f);
            double pfc[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                pfc[d] = m.fccnd[static_cast<int64_t>(ifc) * 3 + d];
            }
            double crd[MAXN + 1][3];
#pragma unroll
            for (int k = 0; k < MAXN; ++k)
            {
                if (k < nnd)
                {
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                    {
                        crd[k][d] = m.ndcrd[static_cast<int64_t>(fn[k + 1]) * 3 + d];
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                crd[nnd][d] = crd[0][d];
            }
            double volb = 0.0;
            double bc[3] = {0.0, 0.0, 0.0};
            double const voe = (fc[0] == icl) ? 0.5 : -0.5;
#pragma unroll
            for (int k = 0; k < MAXN; ++k)
            {
                if (k >= nnd)
                {
                    break;
                }
                double const u0 = crd[k][0] - pfc[0];
                double const u1 = crd[k][1] - pfc[1];
                double const u2 = crd[k][2] - pfc[2];
                double const v0 = crd[k + 1][0] - pfc[0];
                double const v1 = crd[k + 1][1] - pfc[1];
                double const v2 = crd[k + 1][2] - pfc[2];
                double const w0 = u1 * v2 - u2 * v1;
                double const w1 = u2 * v0 - u0 * v2;
                double const w2 = u0 * v1 - u1 * v0;
                double const voli = fabs(w0 * (crdi[0] - pfc[0]) + w1 * (crdi[1] - pfc[1]) + w2 * (crdi[2] - pfc[2])) / 6.0;
                double const vole = fabs(w0 * (crde[0] - pfc[0]) + w1 * (crde[1] - pfc[1]) + w2 * (crde[2] - pfc[2])) / 6.0;
                volb += voli + vole;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    double const cndi = (crd[k][d] + crd[k + 1][d] + pfc[d] + crdi[d]) / 4.0;
                    double const cnde = (crd[k][d] + crd[k + 1][d] + pfc[d] + crde[d]) / 4.0;
                    bc[d] += cndi * voli + cnde * vole;
                }
                int const si = TR::soff(f) + k;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    gi[(LY::FC + si * 3 + d) * n] = (crde[d] + crd[k][d] + crd[k + 1][d]) / 3.0;
                }
                double const su0 = crd[k][0] - crde[0];
                double const su1 = crd[k][1] - crde[1];
                double const su2 = crd[k][2] - crde[2];
                double const sv0 = crd[k + 1][0] - crde[0];
                double const sv1 = crd[k + 1][1] - crde[1];
                double const sv2 = crd[k + 1][2] - crde[2];
                gi[(LY::FN + si * 3) * n] = (su1 * sv2 - su2 * sv1) * voe;
                gi[(LY::FN + si * 3 + 1) * n] = (su2 * sv0 - su0 * sv2) * voe;
                gi[(LY::FN + si * 3 + 2) * n] = (su0 * sv1 - su1 * sv0) * voe;
            }
            gi[(LY::FV + 1 + f) * n] = volb;
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                gi[(LY::FB + f * 3 + d) * n] = bc[d] / volb;
            }
            volc += volb;
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                cc[d] += bc[d];
            }
        }
    }
    gi[LY::FV * n] = volc;
    int64_t const row = m.G + off + i;
#pragma unroll
    for (int d = 0; d < ND; ++d)
    {
        cen[(row) * ND + d] = cc[d] / volc;
    }
}

// Minimum CCE-to-BCE centroid distance of calc_cfl (euler_solution.cpp:118-130);
// it depends on geometry only, so it is evaluated once per geometry change.
template <int T>
__global__ void k_dist(int64_t G, int64_t off, int64_t n, double * geo, double const * cen, int64_t R)
{
    using TR = Traits<T>;
    using LY = Layout<T>;
    constexpr int ND = TR::ND;
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
    {
        return;
    }
    int64_t const row = G + off + i;
    double dist = DBL_MAX;
#pragma unroll
    for (int f = 0; f < TR::NF; ++f)
    {
        double d2 = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            double const diff = geo[(LY::FB + f * ND + d) * n + i] - cen[(row) * ND + d];
            d2 += diff * diff;
        }
        double const s = sqrt(d2);
        dist = (s < dist) ? s : dist;
    }
    geo[LY::FD * n + i] = dist;
}

// Ghost rows of the solution-point table: the mesh centroid of each ghost.
__global__ void k_cen_ghost(int ndim, int64_t G, double const * clcnd_full, double * cen, int64_t R)
{
    int64_t const r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < G)
    {
        for (int d = 0; d < ndim; ++d)
        {
            cen[r * ndim + d] = clcnd_full[r * ndim + d];
        }
    }
}

// --------------------------------------------------------------------------
// Layout conversion between the API tables (reference order, row-major)
// and the internal tables.
// --------------------------------------------------------------------------

__device__ inline int64_t irow_of(int64_t r, int64_t G, int32_t const * int_of_ref)
{
    return r < G ? r : G + int_of_ref[r - G];
}

// Row-major [R][w] <-> structure-of-arrays [w][R].
template <bool DOWN>
__global__ void k_state(int64_t R, int64_t G, int w, int32_t const * int_of_ref, double * api, double * soa)
{
    int64_t const r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= R)
    {
        return;
    }
    int64_t const ir = irow_of(r, G, int_of_ref);
    for (int c = 0; c < w; ++c)
    {
        if (DOWN)
        {
            api[r * w + c] = soa[ir * w + c];
        }
        else
        {
            soa[ir * w + c] = api[r * w + c];
        }
    }
}

// CE geometry of one type group <-> cevol/cecnd/sfcnd/sfnml.
template <int T, bool DOWN>
__global__ void k_geom(int which, int64_t G, int64_t off, int64_t n, int32_t const * ref_of_int, double * api,
                       double * geo, double * cen, int64_t R)
{
    using TR = Traits<T>;
    using LY = Layout<T>;
    constexpr int ND = TR::ND;
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
    {
        return;
    }
    int64_t const r = ref_of_int[off + i];
    int64_t const row = G + off + i;
    if (which == EA_CEVOL)
    {
        double * a = api + (G + r) * (CLMFC + 1);
        for (int c = 0; c <= TR::NF; ++c)
        {
            if (DOWN)
            {
                a[c] = geo[(LY::FV + c) * n + i];
            }
            else
            {
                geo[(LY::FV + c) * n + i] = a[c];
            }
        }
    }
    else if (which == EA_CECND)
    {
        double * a = api + (G + r) * (CLMFC + 1) * ND;
        for (int d = 0; d < ND; ++d)
        {
            if (DOWN)
            {
                a[d] = cen[(row) * ND + d];
            }
            else
            {
                cen[(row) * ND + d] = a[d];
            }
        }
        for (int c = 0; c < TR::NF * ND; ++c)
        {
            if (DOWN)
            {
                a[ND + c] = geo[(LY::FB + c) * n + i];
            }
            else
            {
                geo[(LY::FB + c) * n + i] = a[ND + c];
            }
        }
    }
    else
    {
        int const base = which == EA_SFCND ? LY::FC : LY::FN;
        double * a = api + r * (CLMFC * FCMND) * ND;
        for (int f = 0; f < TR::NF; ++f)
        {
            for (int s = 0; s < TR::nsub(f); ++s)
            {
                for (int d = 0; d < ND; ++d)
                {
                    double & x = a[(f * FCMND + s) * ND + d];
                    double & y = geo[(base + (TR::soff(f) + s) * ND + d) * n + i];
                    if (DOWN)
                    {
                        x = y;
                    }
                    else
                    {
                        y = x;
                    }
                }
            }
        }
    }
}

// init_solution (euler_solution.cpp:79-98): gamma on every row, the uniform
// conserved state on the body rows of so0n.
__global__ void k_init(int64_t R, int64_t G, int neq, int ndim, double ga, double rho, double v0, double v1, double v2,
                       double energy, double * gam, double * s0n)
{
    int64_t const r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= R)
    {
        return;
    }
    gam[r] = ga;
    if (r >= G)
    {
        double const v[3] = {v0, v1, v2};
        s0n[r * neq] = rho;
        for (int d = 0; d < ndim; ++d)
        {
            s0n[r * neq + 1 + d] = rho * v[d];
        }
        s0n[r * neq + neq - 1] = energy;
    }
}

// Space-filling-curve code of the centroid for the cell ordering; the cell
// type sits in the top four bits so that the sort also groups the types.
__device__ inline uint64_t spread2(uint64_t x)
{
    x &= 0x3fffffffULL;
    x = (x | (x << 16)) & 0x0000ffff0000ffffULL;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffULL;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

__device__ inline uint64_t spread3(uint64_t x)
{
    x &= 0xfffffULL;
    x = (x | (x << 32)) & 0x001f00000000ffffULL;
    x = (x | (x << 16)) & 0x001f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

// Hilbert index of a point of an n-dimensional grid of 2^bits cells per
// axis, from its integer coordinates: Skilling's transform ("Programming the
// Hilbert curve", AIP Conf. Proc. 707, 2004) turns the coordinates into the
// transposed index in place, whose bits interleaved from the top (x[0]
// first) are the index.  Cells consecutive along a Hilbert curve are
// always face neighbours, so a block of consecutive cells is a compact
// region and fewer of its cells' neighbours lie outside it than along a
// Morton curve, which jumps.
template <int N>
__device__ inline uint64_t hilbert_code(uint64_t (&x)[3], int bits)
{
    uint64_t const m = 1ULL << (bits - 1);
    for (uint64_t q = m; q > 1; q >>= 1)
    {
        uint64_t const p = q - 1;
        for (int i = 0; i < N; ++i)
        {
            if (x[i] & q)
            {
                x[0] ^= p;
            }
            else
            {
                uint64_t const t = (x[0] ^ x[i]) & p;
                x[0] ^= t;
                x[i] ^= t;
            }
        }
    }
    for (int i = 1; i < N; ++i)
    {
        x[i] ^= x[i - 1];
    }
    uint64_t t = 0;
    for (uint64_t q = m; q > 1; q >>= 1)
    {
        if (x[N - 1] & q)
        {
            t ^= q - 1;
        }
    }
    for (int i = 0; i < N; ++i)
    {
        x[i] ^= t;
    }
    return N == 2 ? (spread2(x[0]) << 1) | spread2(x[1])
                  : (spread3(x[0]) << 2) | (spread3(x[1]) << 1) | spread3(x[2]);
}

// Sort keys of the body cells: the cell type in the top 4 bits, then (for
// batched solves) the instance rank in the next 20, then the Hilbert index
// of the centroid.
__global__ void k_keys(int ndim, int32_t ncell, int32_t const * cltpn, double const * clcnd, double const * lo,
                       double const * hi, int use_sfc, int32_t const * rank, uint64_t * keys, int32_t * vals)
{
    int32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncell)
    {
        return;
    }
    uint64_t code = 0;
    if (use_sfc)
    {
        int const bits = rank != nullptr ? (ndim == 2 ? 20 : 13) : (ndim == 2 ? 30 : 20);
        double const scale = static_cast<double>((1ULL << bits) - 1);
        uint64_t q[3] = {0, 0, 0};
        for (int d = 0; d < ndim; ++d)
        {
            double const w = hi[d] - lo[d];
            double t = w > 0 ? (clcnd[static_cast<int64_t>(i) * ndim + d] - lo[d]) / w : 0.0;
            t = t < 0 ? 0 : (t > 1 ? 1 : t);
            q[d] = static_cast<uint64_t>(t * scale);
        }
        code = ndim == 2 ? hilbert_code<2>(q, bits) : hilbert_code<3>(q, bits);
    }
    if (rank != nullptr)
    {
        code = (static_cast<uint64_t>(rank[i] & 0xfffff) << 40) | (code & 0xffffffffffULL);
    }
    keys[i] = (static_cast<uint64_t>(cltpn[i] & 0xf) << 60) | (code & 0x0fffffffffffffffULL);
    vals[i] = i;
}

// Bounding box of the cell centroids in two passes: every workgroup of the
// first pass reduces a grid-strided share into part[block][6] (lo then hi);
// one workgroup of the second pass (nin = number of parts, clcnd null)
// reduces the parts into out[6].
__global__ void k_bbox(int ndim, int64_t n, double const * clcnd, double const * part_in, double * out)
{
    __shared__ double lo[BLK][3];
    __shared__ double hi[BLK][3];
    double l[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double h[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        for (int d = 0; d < 3; ++d)
        {
            if (clcnd != nullptr && d < ndim)
            {
                double const x = clcnd[i * ndim + d];
                l[d] = fmin(l[d], x);
                h[d] = fmax(h[d], x);
            }
            else if (clcnd == nullptr)
            {
                l[d] = fmin(l[d], part_in[i * 6 + d]);
                h[d] = fmax(h[d], part_in[i * 6 + 3 + d]);
            }
        }
    }
    for (int d = 0; d < 3; ++d)
    {
        lo[threadIdx.x][d] = l[d];
        hi[threadIdx.x][d] = h[d];
    }
    __syncthreads();
    for (unsigned s = blockDim.x / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
        {
            for (int d = 0; d < 3; ++d)
            {
                lo[threadIdx.x][d] = fmin(lo[threadIdx.x][d], lo[threadIdx.x + s][d]);
                hi[threadIdx.x][d] = fmax(hi[threadIdx.x][d], hi[threadIdx.x + s][d]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        for (int d = 0; d < 3; ++d)
        {
            out[blockIdx.x * 6 + d] = lo[0][d];
            out[blockIdx.x * 6 + 3 + d] = hi[0][d];
        }
    }
}

// Histogram of the cell types 0..8 (other values count as 0): counts
// gather in LDS, one global atomic per workgroup and type.
__global__ void k_type_count(int32_t ncell, int32_t const * cltpn, int32_t * cnt)
{
    __shared__ int32_t loc[9];
    if (threadIdx.x < 9)
    {
        loc[threadIdx.x] = 0;
    }
    __syncthreads();
    int32_t const stride = gridDim.x * blockDim.x;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ncell; i += stride)
    {
        int32_t const t = cltpn[i];
        atomicAdd(&loc[(t >= 0 && t <= 8) ? t : 0], 1);
    }
    __syncthreads();
    if (threadIdx.x < 9 && loc[threadIdx.x] != 0)
    {
        atomicAdd(&cnt[threadIdx.x], loc[threadIdx.x]);
    }
}

__global__ void k_invert(int32_t n, int32_t const * ref_of_int, int32_t * int_of_ref)
{
    int32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        int_of_ref[ref_of_int[i]] = i;
    }
}

// Orthonormal frame with the outward unit normal as its first row
// (normal_matrix<2>/<3>, euler_bc.cpp:87-128).
__device__ void normal_frame(int ndim, double const * n, double (&m)[3][3])
{
    if (ndim == 2)
    {
        m[0][0] = n[0];
        m[0][1] = n[1];
        m[1][0] = -n[1];
        m[1][1] = n[0];
        return;
    }
    double ax[3] = {0.0, 0.0, 0.0};
    double const a0 = fabs(n[0]);
    double const a1 = fabs(n[1]);
    double const a2 = fabs(n[2]);
    if (a0 <= a1 && a0 <= a2)
    {
        ax[0] = 1.0;
    }
    else if (a1 <= a2)
    {
        ax[1] = 1.0;
    }
    else
    {
        ax[2] = 1.0;
    }
    double t1[3] = {n[1] * ax[2] - n[2] * ax[1], n[2] * ax[0] - n[0] * ax[2], n[0] * ax[1] - n[1] * ax[0]};
    double const len = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    t1[0] = t1[0] / len;
    t1[1] = t1[1] / len;
    t1[2] = t1[2] / len;
    double const t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
    for (int d = 0; d < 3; ++d)
    {
        m[0][d] = n[d];
        m[1][d] = t1[d];
        m[2][d] = t2[d];
    }
}

// Boundary tables per ghost row: the interior row and the face frame.
__global__ void k_ghost_meta(int ndim, int64_t G, int32_t const * bndfcs, int32_t const * fccls_b,
                             double const * fcnml_b, int32_t const * int_of_ref, int32_t * gsrc, double * gmat)
{
    int64_t const j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= G)
    {
        return;
    }
    // Ghost row j is ghost cell -(G - j), which mirrors bndfcs row G-j-1.
    int64_t const k = G - 1 - j;
    int32_t const ifc = bndfcs[k * BFREL];
    int32_t const icl = fccls_b[static_cast<int64_t>(ifc) * FCREL];
    gsrc[j] = static_cast<int32_t>(G + int_of_ref[icl]);
    double n[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ndim; ++d)
    {
        n[d] = fcnml_b[static_cast<int64_t>(ifc) * ndim + d];
    }
    double m[3][3];
    normal_frame(ndim, n, m);
    for (int a = 0; a < ndim; ++a)
    {
        for (int b = 0; b < ndim; ++b)
        {
            gmat[(a * ndim + b) * G + j] = m[a][b];
        }
    }
}

__global__ void k_face_frame(int ndim, int32_t ifc, double const * fcnml_b, double * out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
    {
        return;
    }
    double n[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ndim; ++d)
    {
        n[d] = fcnml_b[static_cast<int64_t>(ifc) * ndim + d];
    }
    double m[3][3];
    normal_frame(ndim, n, m);
    for (int a = 0; a < ndim; ++a)
    {
        for (int b = 0; b < ndim; ++b)
        {
            out[a * ndim + b] = m[a][b];
        }
    }
}

__global__ void k_set_bc(int32_t n, int32_t const * faces, int32_t const * fccls_b, int64_t G, int32_t kind,
                         int32_t inl, int32_t * gkind, int32_t * ginl)
{
    int32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        int32_t const gcl = fccls_b[static_cast<int64_t>(faces[i]) * FCREL + 1];
        int64_t const j = G + gcl;
        gkind[j] = kind;
        ginl[j] = inl;
    }
}

// Sums over the body cells for diagnostics: density * volume and the CFL
// range (EulerField.calc_overall_mass / calc_cfl_range).
__global__ void k_reduce(int64_t R, int64_t G, int neq, double const * s0n, double const * clvol_b, int32_t const * ref_of_int,
                         double const * cflo, double * part)
{
    __shared__ double sm[BLK];
    __shared__ double mn[BLK];
    __shared__ double mx[BLK];
    double a = 0.0;
    double lo = DBL_MAX;
    double hi = -DBL_MAX;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < R - G;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    {
        a += s0n[(G + i) * neq] * clvol_b[ref_of_int[i]];
        lo = fmin(lo, cflo[G + i]);
        hi = fmax(hi, cflo[G + i]);
    }
    sm[threadIdx.x] = a;
    mn[threadIdx.x] = lo;
    mx[threadIdx.x] = hi;
    __syncthreads();
    for (unsigned s = blockDim.x / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
        {
            sm[threadIdx.x] += sm[threadIdx.x + s];
            mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + s]);
            mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        part[blockIdx.x * 3] = sm[0];
        part[blockIdx.x * 3 + 1] = mn[0];
        part[blockIdx.x * 3 + 2] = mx[0];
    }
}

// Instance rank of every body cell in internal order, and the histogram
// of (cell type, rank) that gives each type group its per-rank prefixes.
__global__ void k_batch_inst(int32_t n, int32_t const * ref_of_int, int32_t const * rank_ref,
                             int32_t const * cltpn, int32_t ninst, int32_t * inst, int32_t * hist)
{
    int32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        int32_t const c = ref_of_int[i];
        int32_t const r = rank_ref[c];
        inst[i] = r;
        atomicAdd(&hist[static_cast<int64_t>(cltpn[c] & 0xf) * ninst + r], 1);
    }
}

// Largest calc_cfl number of every instance (of the whole mesh when inst is
// null) at time increment 1 (hdt 0.5) from the so0n state (euler_solution.cpp:114-159), as the bit pattern of a
// non-negative double so that an integer maximum orders it.
template <int T>
__global__ void k_batch_cflmax(int64_t R, int64_t G, int64_t off, int64_t n, double const * geo,
                               double const * s0n, double const * gam, int32_t const * inst,
                               unsigned long long * cmax)
{
    using LY = Layout<T>;
    constexpr int ND = Traits<T>::ND;
    constexpr int NEQ = ND + 2;
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u);
    if (i0 >= n)
    {
        return; // the whole wavefront is past the end
    }
    bool const valid = i < n;
    unsigned long long bits = 0;
    int32_t slot = 0;
    if (valid)
    {
        int64_t const row = G + off + i;
        double u0[NEQ];
#pragma unroll
        for (int k = 0; k < NEQ; ++k)
        {
            u0[k] = s0n[(row) * NEQ + k];
        }
        double const dist = geo[LY::FD * n + i];
        double const ga = gam[row];
        double const density = u0[0];
        double momsq = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d)
        {
            momsq += u0[1 + d] * u0[1 + d];
        }
        double const ke = momsq / (2.0 * density);
        double const pr = (ga - 1.0) * (u0[NEQ - 1] - ke);
        double const pra = (pr + fabs(pr)) / 2.0;
        double const wspd = sqrt(ga * pra / density) + sqrt(momsq) / density;
        double const v = 0.5 * wspd / dist;
        bits = v >= 0.0 ? static_cast<unsigned long long>(__double_as_longlong(v)) : 0ull;
        slot = inst != nullptr ? inst[off + i] : 0;
    }
    // One atomic per wavefront when all its cells belong to one instance
    // (always for a single mesh): a maximum over the lanes first.
    int32_t const s0 = __shfl(slot, 0);
    if (__all(!valid || slot == s0))
    {
        for (int m = 32; m > 0; m >>= 1)
        {
            unsigned long long const o = __shfl_xor(bits, m);
            bits = o > bits ? o : bits;
        }
        if ((threadIdx.x & 63u) == 0)
        {
            atomicMax(&cmax[s0], bits);
        }
    }
    else if (valid)
    {
        atomicMax(&cmax[slot], bits);
    }
}

} /* end namespace */

struct EulerBoundary
{
    int32_t kind;
    std::vector<int32_t> faces;
    std::vector<double> value;
};

struct DEuler
{
    DMesh * mesh = nullptr;
    int ndim = 0;
    int neq = 0;
    int64_t G = 0;
    int64_t N = 0;
    int64_t R = 0;
    double dt = 0.0;
    double sigma0 = 3.0; // euler.hpp:183-188
    double taumin = 0.0;
    double tauscale = 1.0;
    bool use_sfc = true;
    // Kernel form of a type group: the fused kernel in variant vf, or the
    // split pair (soln + BC + CFL, then dsoln) in variants v1 and v2 (codes
    // of role_variant).
    struct Form
    {
        int split = 0;
        int vf = 1;
        int v1 = 1;
        int v2 = 1;
    };
    int tuned = 0;            // 0 until the first substep has timed the forms
    std::vector<Form> gform;  // per group
    StreamHolder sh;            // declared before the buffers: freed last
    hipStream_t stream = sh.s;

    // With use_graph, whole steps (two substeps, the half-step pair) are
    // enqueued once under stream capture and replayed as a hipGraph.  The
    // kernels take every parameter by value, so a graph is replayed only
    // while its key (the launch parameters, the forms and the active ranks)
    // is unchanged.  Off by default: on this stack a replayed step costs
    // more than launching its kernels directly.
    int use_graph = 0;
    hipGraphExec_t pgraph[2] = {nullptr, nullptr}; // plain march, by buffer parity
    std::vector<unsigned char> pgraph_key[2];
    std::vector<hipGraphExec_t> bgraphs; // batched march, one per active rank range

    struct Group
    {
        int32_t type = 0;
        int64_t off = 0;
        int64_t n = 0;
        DevBuf<int32_t> nbr;
        DevBuf<double> geo;
    };
    std::vector<Group> groups;
    DevBuf<int32_t> int_of_ref, ref_of_int;
    // st[b] holds so0t of the level in s0[b], s1[b].
    DevBuf<double> cen, s0[2], s1[2], st[2], stm, gam, cflo, cflc;
    int cur = 0;
    // The fast/exact launch pairs' flag, the count of exact passes that ran
    // and the exact-only switch of the current march (k_cell), and the
    // sequence number of the last fast launch.
    DevBuf<unsigned long long> redo;
    unsigned long long redo_seq = 0;

    std::vector<EulerBoundary> bcs;
    DevBuf<int32_t> gkind, ginl, gsrc;
    DevBuf<double> gmat, inlet;

    // Batched solves of independent instances in one disjoint-union mesh:
    // ranks order the instances by descending step count, cells are sorted
    // by (type, rank, Hilbert index), so the cells still marching at any
    // step are a prefix of every type group (bcum[group][rank]).
    int64_t ninst = 0;
    std::vector<int64_t> bsteps;       // by rank, non-increasing
    std::vector<int32_t> brank_of;     // instance id -> rank
    std::vector<std::vector<int64_t>> bcum;
    DevBuf<int32_t> brank_ref, binst;
    DevBuf<double> bhdt, bqdt;
    // Completion times of a timed batched march: an event is recorded on
    // the stream when the ranks [first, second) have run their last step,
    // and bdone_ms[rank] receives its time from the march's start.
    bool btime = false;
    std::vector<std::pair<int64_t, int64_t>> bdone_rng;
    std::vector<hipEvent_t> bdone_ev;
    std::vector<double> bdone_ms;

    DEuler(DMesh * m, double time_increment, bool sfc, int64_t nbatch = 0, int32_t const * inst_of_cell = nullptr,
           int64_t const * steps = nullptr);
    ~DEuler()
    {
        (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (auto ev : bdone_ev) // left by a timed batched march that failed
        {
            (void)hipEventDestroy(ev);
        }
    }

    void sync() { SC_HIP(hipStreamSynchronize(stream)); }

    // Destroys the graph executables; their launches must have completed.
    void drop_graphs()
    {
        for (int c = 0; c < 2; ++c)
        {
            if (pgraph[c] != nullptr)
            {
                (void)hipGraphExecDestroy(pgraph[c]);
                pgraph[c] = nullptr;
            }
            pgraph_key[c].clear();
        }
        for (auto x : bgraphs)
        {
            (void)hipGraphExecDestroy(x);
        }
        bgraphs.clear();
    }

    // Captures what enqueue() puts on the stream into a graph executable.
    // The buffer parity is restored if the capture fails.
    template <class F>
    hipGraphExec_t capture(F && enqueue)
    {
        int const c0 = cur;
        hipGraph_t g = nullptr;
        SC_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try
        {
            enqueue();
        }
        catch (...)
        {
            (void)hipStreamEndCapture(stream, &g);
            if (g != nullptr)
            {
                (void)hipGraphDestroy(g);
            }
            cur = c0;
            throw;
        }
        SC_HIP(hipStreamEndCapture(stream, &g));
        hipGraphExec_t ex = nullptr;
        hipError_t const rc = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SC_HIP(rc);
        return ex;
    }

    KP params(bool swap_after = false) const
    {
        (void)swap_after;
        KP p;
        p.R = R;
        p.G = G;
        p.hdt = dt * 0.5;
        p.qdt = dt * 0.25;
        p.sigma0 = sigma0;
        p.taumin = taumin;
        p.tauscale = tauscale;
        p.s0c = s0[cur].ptr;
        p.s1c = s1[cur].ptr;
        p.s0n = s0[1 - cur].ptr;
        p.s1n = s1[1 - cur].ptr;
        p.st = st[cur].ptr;
        p.stn = st[1 - cur].ptr;
        p.gam = gam.ptr;
        p.cflo = cflo.ptr;
        p.cflc = cflc.ptr;
        p.cen = cen.ptr;
        p.gkind = gkind.ptr;
        p.ginl = ginl.ptr;
        p.gmat = gmat.ptr;
        p.inlet = inlet.ptr;
        p.inst = binst.ptr;
        p.bhdt = bhdt.ptr;
        p.bqdt = bqdt.ptr;
        p.redo = redo.ptr;
        p.seq = 0;
        p.gate = 0;
        return p;
    }

    // Every plain launch is a pair: the fast form over the group, then the
    // exact form, which returns at once unless a cell of the fast launch
    // reported a shortcut that may not equal the exact arithmetic (a value
    // that is not finite or lies near the ends of the exponent range); it
    // then rewrites the whole group from the same inputs.  Each pair has
    // its own sequence number, so a flag left by an earlier pair does not
    // count; one replayed from a graph reuses its numbers, and a stale
    // match only costs an exact pass.
    template <int T, int MODE>
    void launch_cell(KP const & p, GV const & g, unsigned nb)
    {
        if constexpr (variant_built(Traits<T>::ND, MODE) && ((MODE & M_GDMA) == 0 || gdma_fits(T, MODE)))
        {
            KP q = p;
            q.seq = ++redo_seq;
            q.gate = 0;
            k_cell<T, MODE | M_FAST><<<nb, BLK, 0, stream>>>(q, g);
            q.gate = 1;
            k_cell<T, MODE><<<nb, BLK, 0, stream>>>(q, g);
        }
        else
        {
            fail(SC_ERUNTIME, "EulerCore: kernel variant not built for this cell type");
        }
    }

    // One kernel of mode MODE over the first cnt cells of a type group; a
    // negative cnt launches one block that processes no cell, which loads
    // the kernel's code so a later timed launch does not pay for it.
    template <int MODE>
    void run_group(KP const & p, Group const & gr, int64_t cnt)
    {
        if (cnt == 0)
        {
            return;
        }
        GV const g{gr.off, gr.n, gr.nbr.ptr, gr.geo.ptr, cnt < 0 ? 0 : cnt};
        unsigned const nb = cnt < 0 ? 1u : grid_for(cnt, BLK);
        switch (gr.type)
        {
        case CT_TRI: launch_cell<CT_TRI, MODE>(p, g, nb); break;
        case CT_QUAD: launch_cell<CT_QUAD, MODE>(p, g, nb); break;
        case CT_TET: launch_cell<CT_TET, MODE>(p, g, nb); break;
        case CT_HEX: launch_cell<CT_HEX, MODE>(p, g, nb); break;
        case CT_PRISM: launch_cell<CT_PRISM, MODE>(p, g, nb); break;
        case CT_PYRAMID: launch_cell<CT_PYRAMID, MODE>(p, g, nb); break;
        default: fail(SC_ERUNTIME, "EulerCore: unsupported cell type in group");
        }
        SC_HIP(hipGetLastError());
    }

    // The geometry prefetch issues 16-byte loads of two cells each: it needs
    // an even field stride, a 16-byte aligned table and room in LDS for the
    // blocks the variant keeps resident.
    static bool has_gdma(int code) { return (code / 1000) % 10 != 0; }
    static bool gdma_ok(Group const & gr, int code)
    {
        return gr.n % 2 == 0 && reinterpret_cast<uintptr_t>(gr.geo.ptr) % 16 == 0 &&
               gdma_fits(gr.type, variant_mode(code));
    }

    // run_group over a whole group with the variant chosen at run time.
    template <int MODE>
    void run_group_v(KP const & p, Group const & gr, int code, bool warm = false)
    {
        int64_t const cnt = warm ? -1 : gr.n;
        if (has_gdma(code) && !gdma_ok(gr, code))
        {
            code -= 1000; // the same variant without the geometry prefetch
        }
        switch (code)
        {
        case 2: run_group<MODE | variant_mode(2)>(p, gr, cnt); break;
        case 3: run_group<MODE | variant_mode(3)>(p, gr, cnt); break;
        case 11: run_group<MODE | variant_mode(11)>(p, gr, cnt); break;
        case 12: run_group<MODE | variant_mode(12)>(p, gr, cnt); break;
        case 102: run_group<MODE | variant_mode(102)>(p, gr, cnt); break;
        case 112: run_group<MODE | variant_mode(112)>(p, gr, cnt); break;
        case 1001: run_group<MODE | variant_mode(1001)>(p, gr, cnt); break;
        case 1003: run_group<MODE | variant_mode(1003)>(p, gr, cnt); break;
        case 1102: run_group<MODE | variant_mode(1102)>(p, gr, cnt); break;
        default: run_group<MODE>(p, gr, cnt); break;
        }
    }

    template <int MODE>
    void run_cells()
    {
        KP const p = params();
        for (auto const & gr : groups)
        {
            run_group<MODE>(p, gr, gr.n);
        }
    }

    // One fused substep of the batched march over the cells of the first
    // nact ranks of every type group.
    void run_cells_batch(int64_t ra, int64_t rb)
    {
        KP const p = params();
        for (size_t gi = 0; gi < groups.size(); ++gi)
        {
            auto & gr = groups[gi];
            int64_t const lo = bcum[gi][ra];
            int64_t const cnt = bcum[gi][rb] - lo;
            if (cnt <= 0)
            {
                continue;
            }
            GV const g{gr.off + lo, gr.n, gr.nbr.ptr + lo, gr.geo.ptr + lo, cnt};
            unsigned const nb = grid_for(cnt, BLK);
            constexpr int MB = M_FUSED | M_BATCH;
            constexpr int MT = MB | batch_waves(CT_TRI);
            // Triangles prefetch their sub-face geometry when the range
            // starts at an even cell (always, unless rank chunks split it).
            bool const tri_gdma = lo % 2 == 0 && gdma_ok(gr, 1003);
            switch (gr.type)
            {
            case CT_TRI:
                if (tri_gdma)
                {
                    k_cell<CT_TRI, MT | M_GDMA><<<nb, BLK, 0, stream>>>(p, g);
                }
                else
                {
                    k_cell<CT_TRI, MT><<<nb, BLK, 0, stream>>>(p, g);
                }
                break;
            case CT_QUAD: k_cell<CT_QUAD, MB | batch_waves(CT_QUAD)><<<nb, BLK, 0, stream>>>(p, g); break;
            case CT_TET: k_cell<CT_TET, MB | batch_waves(CT_TET)><<<nb, BLK, 0, stream>>>(p, g); break;
            case CT_HEX: k_cell<CT_HEX, MB | batch_waves(CT_HEX)><<<nb, BLK, 0, stream>>>(p, g); break;
            case CT_PRISM: k_cell<CT_PRISM, MB | batch_waves(CT_PRISM)><<<nb, BLK, 0, stream>>>(p, g); break;
            case CT_PYRAMID: k_cell<CT_PYRAMID, MB | batch_waves(CT_PYRAMID)><<<nb, BLK, 0, stream>>>(p, g); break;
            default: fail(SC_ERUNTIME, "EulerCore: unsupported cell type in group");
            }
            SC_HIP(hipGetLastError());
        }
    }

    // Every instance marches its own step count: two fused substeps per
    // step over the ranks still marching, then calc_solt once.  A finished
    // instance keeps its state in place: the buffer swap of the others
    // returns to the same parity after each whole step.
    //
    // The ranks are marched in chunks of consecutive ranks holding about
    // chunk_cells cells (0: one chunk of every rank).  A chunk runs all its
    // steps before the next starts, so its working set can stay in the
    // last-level cache from step to step.  Every chunk swaps the buffers an
    // even number of times, so each instance still ends in the same buffer,
    // and each instance's arithmetic is unchanged.
    int64_t chunk_cells = 0;

    int64_t rank_cells(int64_t ra, int64_t rb) const
    {
        int64_t c = 0;
        for (auto const & cum : bcum)
        {
            c += cum[rb] - cum[ra];
        }
        return c;
    }

    // With use_graph, the step's launches are captured once per range of
    // active ranks and the graph is replayed for every step of that range:
    // a step swaps the buffers twice, so each replay starts from the same
    // buffer parity as the capture.
    void march_ranks(int64_t ra, int64_t rb)
    {
        int64_t const maxs = bsteps[ra];
        int64_t nact = rb;
        hipGraphExec_t g = nullptr;
        int64_t gact = -1;
        auto done = [&](int64_t lo, int64_t hi)
        {
            if (btime && lo < hi)
            {
                hipEvent_t ev = nullptr;
                SC_HIP(hipEventCreate(&ev));
                bdone_ev.push_back(ev);
                bdone_rng.emplace_back(lo, hi);
                SC_HIP(hipEventRecord(ev, stream));
            }
        };
        for (int64_t step = 0; step < maxs; ++step)
        {
            int64_t const was = nact;
            while (nact > ra && bsteps[nact - 1] <= step)
            {
                --nact;
            }
            done(nact, was);
            if (use_graph != 0)
            {
                if (g == nullptr || gact != nact)
                {
                    g = capture(
                        [&]()
                        {
                            for (int sub = 0; sub < 2; ++sub)
                            {
                                update();
                                run_cells_batch(ra, nact);
                            }
                        });
                    bgraphs.push_back(g);
                    gact = nact;
                }
                SC_HIP(hipGraphLaunch(g, stream));
                continue;
            }
            for (int sub = 0; sub < 2; ++sub)
            {
                update();
                run_cells_batch(ra, nact);
            }
        }
        done(ra, nact);
    }

    void march_batch()
    {
        int64_t ra = 0;
        while (ra < ninst)
        {
            int64_t rb = ninst;
            if (chunk_cells > 0)
            {
                rb = ra + 1;
                while (rb < ninst && rank_cells(ra, rb + 1) <= chunk_cells)
                {
                    ++rb;
                }
            }
            march_ranks(ra, rb);
            ra = rb;
        }
        if (!bgraphs.empty())
        {
            sync();
            for (auto x : bgraphs)
            {
                (void)hipGraphExecDestroy(x);
            }
            bgraphs.clear();
        }
        if (ninst > 0 && bsteps[0] > 0)
        {
            calc_solt();
        }
    }

    void batch_set_dt(double const * dt_inst);
    void batch_cfl_dt(double cfl, double * dt_inst);
    double cfl_max();

    void calc_solt()
    {
        KP const p = params();
        if (R == 0)
        {
            return;
        }
        if (ndim == 2)
        {
            k_solt<2><<<grid_for(R, BLK), BLK, 0, stream>>>(p);
        }
        else
        {
            k_solt<3><<<grid_for(R, BLK), BLK, 0, stream>>>(p);
        }
        SC_HIP(hipGetLastError());
    }

    void bc_soln()
    {
        if (G == 0)
        {
            return;
        }
        KP const p = params();
        if (ndim == 2)
        {
            k_bc<2, false><<<grid_for(G, BLK), BLK, 0, stream>>>(p, gsrc.ptr);
        }
        else
        {
            k_bc<3, false><<<grid_for(G, BLK), BLK, 0, stream>>>(p, gsrc.ptr);
        }
        SC_HIP(hipGetLastError());
    }

    void bc_dsoln()
    {
        if (G == 0)
        {
            return;
        }
        KP const p = params();
        if (ndim == 2)
        {
            k_bc<2, true><<<grid_for(G, BLK), BLK, 0, stream>>>(p, gsrc.ptr);
        }
        else
        {
            k_bc<3, true><<<grid_for(G, BLK), BLK, 0, stream>>>(p, gsrc.ptr);
        }
        SC_HIP(hipGetLastError());
    }

    void update() { cur = 1 - cur; }

    // One substep.  Every type group runs either the fused kernel or the
    // split pair that reads so0t from memory: calc_solt over every row
    // (once per substep when any group is split), then soln + BC + CFL,
    // then dsoln.  Both forms compute the same values; the split pair
    // needs fewer registers per kernel.
    static constexpr int M_SPLIT1 = M_SOLN | M_BC | M_CFL | M_STMEM;
    static constexpr int M_SPLIT2 = M_DSOLN | M_STMEM;

    // Chained so0t.  When a kept form reads so0t from memory, the kernels
    // that write the next level also write its so0t (body rows where so1n
    // is written, ghost rows where the boundary conditions write them) into
    // st[1 - cur], so the next substep finds it in st[cur] without a
    // calc_solt pass.  Ghost rows of boundary faces without a condition are
    // never written during a march: solt_both() computes so0t of both levels
    // when a chain starts, and those rows stay valid.  chain_ready is false
    // until then, and every march starts a new chain.
    bool chain_ready = false;

    bool chain_forms() const
    {
        for (auto const & f : gform)
        {
            if (needs_solt(f))
            {
                return true;
            }
        }
        return false;
    }

    void solt_both()
    {
        calc_solt();
        if (R == 0)
        {
            return;
        }
        KP q = params();
        q.s0c = s0[1 - cur].ptr;
        q.s1c = s1[1 - cur].ptr;
        q.st = st[1 - cur].ptr;
        if (ndim == 2)
        {
            k_solt<2><<<grid_for(R, BLK), BLK, 0, stream>>>(q);
        }
        else
        {
            k_solt<3><<<grid_for(R, BLK), BLK, 0, stream>>>(q);
        }
        SC_HIP(hipGetLastError());
    }

    // Starts the chain eagerly (outside any graph capture) when it is due.
    void start_chain()
    {
        if (tuned != 0 && !chain_ready && chain_forms())
        {
            solt_both();
            chain_ready = true;
        }
    }

    void substep_fused()
    {
        if (tuned == 0)
        {
            tune_substep();
            chain_ready = false;
            return;
        }
        bool const chain = chain_forms();
        if (chain && !chain_ready)
        {
            solt_both();
            chain_ready = true;
        }
        update();
        KP const p = params();
        for (size_t gi = 0; gi < groups.size(); ++gi)
        {
            run_form(p, groups[gi], gform[gi], chain);
        }
        chain_ready = chain;
    }

    // Whether a form reads so0t from memory, so calc_solt must run first.
    static bool reads_solt(int code) { return (code / 100) % 10 != 0; }
    static bool needs_solt(Form const & f) { return f.split != 0 || reads_solt(f.vf); }

    void run_role(KP const & p, Group const & gr, int role, int code, bool warm = false, bool chain = false)
    {
        if (role == 0)
        {
            if (chain)
            {
                run_group_v<M_FUSED | M_STNEXT>(p, gr, code, warm);
            }
            else
            {
                run_group_v<M_FUSED>(p, gr, code, warm);
            }
        }
        else if (role == 1)
        {
            if (chain)
            {
                run_group_v<M_SPLIT1 | M_STNEXT>(p, gr, code, warm);
            }
            else
            {
                run_group_v<M_SPLIT1>(p, gr, code, warm);
            }
        }
        else
        {
            if (chain)
            {
                run_group_v<M_SPLIT2 | M_STNEXT>(p, gr, code, warm);
            }
            else
            {
                run_group_v<M_SPLIT2>(p, gr, code, warm);
            }
        }
    }

    void run_form(KP const & p, Group const & gr, Form const & f, bool chain = false)
    {
        if (f.split != 0)
        {
            run_role(p, gr, 1, f.v1, false, chain);
            run_role(p, gr, 2, f.v2, false, chain);
        }
        else
        {
            run_role(p, gr, 0, f.vf, false, chain);
        }
    }

    // The autotuner: the first substep of a march runs every compiled
    // variant of every kernel role of every group (later variants rewrite
    // the same outputs with the same values), times each with HIP events
    // over a few runs back to back, and keeps per group the fastest fused
    // variant or the fastest split1 + split2 pair, counting calc_solt once
    // if any group is split.
    void tune_substep()
    {
        update();
        KP const p = params();
        size_t const ng = groups.size();
        size_t const nslot = ng * 3 * NVAR;
        // Load every variant's code first (one empty block each).
        calc_solt();
        for (size_t gi = 0; gi < ng; ++gi)
        {
            for (int role = 0; role < 3; ++role)
            {
                for (int v = 0; v < NVAR; ++v)
                {
                    int const c = role_variant(ndim, role, v);
                    if (c != 0 && (!has_gdma(c) || gdma_ok(groups[gi], c)))
                    {
                        run_role(p, groups[gi], role, c, true, true);
                    }
                }
            }
        }
        // Every variant runs TUNE_RUNS times back to back and keeps the
        // fastest of all but the first run: timed right after another
        // variant, a kernel meets that variant's cache contents, and
        // variants a few percent apart would swap places from run to run.
        constexpr int TUNE_RUNS = 4;
        // Event pairs by (slot, run): slot 0 times calc_solt, slot 1 + sl
        // the variant of timing slot sl with the chained so0t write (as
        // every kernel of a substep whose forms read so0t runs), slot
        // 1 + nslot + sl a fused variant that recomputes so0t without it (as
        // it runs when no form reads so0t).  Everything is enqueued first
        // and the stream synchronised once.
        std::vector<hipEvent_t> ev(2 * TUNE_RUNS * (1 + 2 * nslot), nullptr);
        auto evp = [&](size_t slot, int r, int b) -> hipEvent_t & { return ev[(slot * TUNE_RUNS + r) * 2 + b]; };
        auto rec = [&](size_t slot, int r, int b)
        {
            hipEvent_t & e = evp(slot, r, b);
            if (e == nullptr)
            {
                SC_HIP(hipEventCreate(&e));
            }
            SC_HIP(hipEventRecord(e, stream));
        };
        for (int r = 0; r < TUNE_RUNS; ++r)
        {
            rec(0, r, 0);
            calc_solt();
            rec(0, r, 1);
        }
        for (size_t gi = 0; gi < ng; ++gi)
        {
            for (int role = 0; role < 3; ++role)
            {
                for (int v = 0; v < NVAR; ++v)
                {
                    int const c = role_variant(ndim, role, v);
                    if (c == 0 || (has_gdma(c) && !gdma_ok(groups[gi], c)))
                    {
                        continue;
                    }
                    size_t const sl = (gi * 3 + role) * NVAR + v;
                    for (int chain = 0; chain < 2; ++chain)
                    {
                        if (chain == 0 && (role != 0 || reads_solt(c)))
                        {
                            continue;
                        }
                        size_t const slot = chain != 0 ? 1 + sl : 1 + nslot + sl;
                        for (int r = 0; r < TUNE_RUNS; ++r)
                        {
                            rec(slot, r, 0);
                            run_role(p, groups[gi], role, c, false, chain != 0);
                            rec(slot, r, 1);
                        }
                    }
                }
            }
        }
        SC_HIP(hipStreamSynchronize(stream));
        auto best = [&](size_t slot)
        {
            float b = -1.0f;
            for (int r = 1; r < TUNE_RUNS; ++r)
            {
                float t = 0.0f;
                SC_HIP(hipEventElapsedTime(&t, evp(slot, r, 0), evp(slot, r, 1)));
                b = b < 0.0f ? t : std::min(b, t);
            }
            return b;
        };
        float const solt = best(0);
        std::vector<float> tm(nslot, -1.0f); // chained
        std::vector<float> tp(nslot, -1.0f); // plain fused that recomputes so0t
        for (size_t sl = 0; sl < nslot; ++sl)
        {
            if (evp(1 + sl, 0, 0) != nullptr)
            {
                tm[sl] = best(1 + sl);
            }
            if (evp(1 + nslot + sl, 0, 0) != nullptr)
            {
                tp[sl] = best(1 + nslot + sl);
            }
        }
        for (auto & x : ev)
        {
            if (x != nullptr)
            {
                (void)hipEventDestroy(x);
            }
        }
        if (std::getenv("SC_EULER_TUNE_LOG") != nullptr)
        {
            std::fprintf(stderr, "EulerCore tuning (ms; chained so0t write, plain in brackets): calc_solt %.4f\n", solt);
            for (size_t gi = 0; gi < ng; ++gi)
            {
                for (int role = 0; role < 3; ++role)
                {
                    std::fprintf(stderr, "  type %d cells %lld %s", groups[gi].type,
                                 static_cast<long long>(groups[gi].n),
                                 role == 0 ? "fused " : role == 1 ? "split1" : "split2");
                    for (int v = 0; v < NVAR; ++v)
                    {
                        size_t const sl = (gi * 3 + role) * NVAR + v;
                        int const c = role_variant(ndim, role, v);
                        if (tm[sl] >= 0.0f)
                        {
                            std::fprintf(stderr, "  w%d%s%s%s %.4f", c % 10, (c / 10) % 10 != 0 ? "+lds" : "",
                                         (c / 100) % 10 != 0 ? "+so0t" : "", (c / 1000) % 10 != 0 ? "+gdma" : "", tm[sl]);
                            if (tp[sl] >= 0.0f)
                            {
                                std::fprintf(stderr, " [%.4f]", tp[sl]);
                            }
                        }
                    }
                    std::fprintf(stderr, "\n");
                }
            }
        }
        // Two configurations.  Plain: every group runs its fastest fused
        // variant that recomputes so0t, without the chained write.  Chained:
        // every group takes its fastest form of all, each timed with the
        // chained write (so0t is then chained from the kernels that write
        // each level, see substep_fused, and no calc_solt runs per substep).
        std::vector<Form> fast(ng), pick(ng);
        double no_solt = 0.0;
        double with_solt = 0.0;
        for (size_t gi = 0; gi < ng; ++gi)
        {
            double tf = -1.0;
            double tc = -1.0;
            double best1 = -1.0;
            double best2 = -1.0;
            int bc1 = 1;
            int bc2 = 1;
            for (int role = 0; role < 3; ++role)
            {
                for (int v = 0; v < NVAR; ++v)
                {
                    size_t const sl = (gi * 3 + role) * NVAR + v;
                    int const c = role_variant(ndim, role, v);
                    if (role == 0 && tp[sl] >= 0.0f && (tf < 0.0 || tp[sl] < tf))
                    {
                        tf = tp[sl];
                        fast[gi].vf = c;
                    }
                    if (tm[sl] < 0.0f)
                    {
                        continue;
                    }
                    if (role == 0 && (tc < 0.0 || tm[sl] < tc))
                    {
                        tc = tm[sl];
                        pick[gi] = Form{};
                        pick[gi].vf = c;
                    }
                    else if (role == 1 && (best1 < 0.0 || tm[sl] < best1))
                    {
                        best1 = tm[sl];
                        bc1 = c;
                    }
                    else if (role == 2 && (best2 < 0.0 || tm[sl] < best2))
                    {
                        best2 = tm[sl];
                        bc2 = c;
                    }
                }
            }
            if (best1 >= 0.0 && best2 >= 0.0 && (tc < 0.0 || best1 + best2 < tc))
            {
                tc = best1 + best2;
                pick[gi] = Form{1, 1, bc1, bc2};
            }
            no_solt += tf;
            with_solt += tc;
        }
        bool const chained = with_solt < no_solt && std::any_of(pick.begin(), pick.end(), [](Form const & f)
                                                                { return needs_solt(f); });
        for (size_t gi = 0; gi < ng; ++gi)
        {
            gform[gi] = chained ? pick[gi] : fast[gi];
        }
        tuned = 1;
    }

    // Form of the group of one cell type: the fused variant code vf, or
    // 10000 + 100 v1 + v2 for the split pair; 0 when the mesh has no such
    // cells, -1 before the first substep has tuned.
    int form_code(int32_t type) const
    {
        if (tuned == 0)
        {
            return -1;
        }
        for (size_t gi = 0; gi < groups.size(); ++gi)
        {
            if (groups[gi].type == type)
            {
                Form const & f = gform[gi];
                if (f.split != 0)
                {
                    return 10000 + 100 * f.v1 + f.v2;
                }
                return has_gdma(f.vf) && !gdma_ok(groups[gi], f.vf) ? f.vf - 1000 : f.vf;
            }
        }
        return 0;
    }

    bool has_variant(int role, int code) const
    {
        for (int v = 0; v < NVAR; ++v)
        {
            if (role_variant(ndim, role, v) == code)
            {
                return true;
            }
        }
        return false;
    }

    // What the launches of a step depend on: the kernel parameters at its
    // start (buffer parity included) and every group's form.
    std::vector<unsigned char> step_key() const
    {
        KP const p = params();
        std::vector<unsigned char> k(sizeof(KP));
        std::memcpy(k.data(), &p, sizeof(KP));
        for (size_t gi = 0; gi < groups.size(); ++gi)
        {
            Group const & gr = groups[gi];
            GV const g{gr.off, gr.n, gr.nbr.ptr, gr.geo.ptr, gr.n};
            unsigned char const * a = reinterpret_cast<unsigned char const *>(&g);
            unsigned char const * f = reinterpret_cast<unsigned char const *>(&gform[gi]);
            k.insert(k.end(), a, a + sizeof(GV));
            k.insert(k.end(), f, f + sizeof(Form));
        }
        return k;
    }

    // march (euler_march.cpp:350-359): after the tuning substep, whole
    // steps are replayed from a captured graph when use_graph is set.
    void march(int64_t steps)
    {
        int64_t sub = 2 * steps;
        chain_ready = false;
        // Each march starts with the fast kernels of its pairs enabled.
        fill<unsigned long long>(redo.ptr + 2, 1, 0ull, stream);
        if (sub > 0 && tuned == 0)
        {
            substep_fused();
            --sub;
        }
        if (sub > 0)
        {
            start_chain();
        }
        if (use_graph != 0 && sub >= 2)
        {
            std::vector<unsigned char> const key = step_key();
            hipGraphExec_t & g = pgraph[cur];
            if (g == nullptr || key != pgraph_key[cur])
            {
                if (g != nullptr)
                {
                    sync(); // launches of the old graph may still be queued
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
                g = capture(
                    [&]()
                    {
                        substep_fused();
                        substep_fused();
                    });
                pgraph_key[cur] = key;
            }
            for (; sub >= 2; sub -= 2)
            {
                SC_HIP(hipGraphLaunch(g, stream));
            }
        }
        for (; sub > 0; --sub)
        {
            substep_fused();
        }
        if (steps > 0)
        {
            calc_solt();
        }
    }

    void march_substep()
    {
        chain_ready = false;
        substep_fused();
        calc_solt();
    }

    void prepare_ce();
    void init_solution(double ga, double rho, double const * v, double p);
    void add_bc(int32_t kind, std::vector<int32_t> const & faces, std::vector<double> const & value);
    void refresh_bc();
    void table(int id, int64_t & rows, int64_t & w, int64_t & ngh, int & rank);
    void copy(int id, double * host, int64_t count, bool down);
};

DEuler::DEuler(DMesh * m, double time_increment, bool sfc, int64_t nbatch, int32_t const * inst_of_cell,
               int64_t const * steps)
    : mesh(m)
    , dt(time_increment)
    , use_sfc(sfc)
    , ninst(nbatch)
{
    if (m->ndim != 2 && m->ndim != 3)
    {
        fail(SC_EVALUE, "EulerCore: ndim must be 2 or 3");
    }
    AllocScope const scope(stream);
    ndim = m->ndim;
    neq = ndim + 2;
    G = m->ngstcell;
    N = m->ncell;
    R = G + N;
    hipStream_t const s = stream;
    m->sync();

    // Cell types present.
    DevBuf<int32_t> tcnt;
    tcnt.alloc(9);
    SC_HIP(hipMemsetAsync(tcnt.ptr, 0, tcnt.bytes(), s));
    if (N > 0)
    {
        k_type_count<<<std::min<int64_t>(grid_for(N, BLK), 2048), BLK, 0, s>>>(static_cast<int32_t>(N), m->cltpn_b(), tcnt.ptr);
        SC_HIP(hipGetLastError());
    }
    int32_t cnt[9];
    SC_HIP(hipMemcpyAsync(cnt, tcnt.ptr, sizeof(cnt), hipMemcpyDeviceToHost, s));
    sync();
    for (int t = 0; t < 9; ++t)
    {
        bool const ok = ndim == 2 ? (t == CT_TRI || t == CT_QUAD)
                                  : (t == CT_TET || t == CT_HEX || t == CT_PRISM || t == CT_PYRAMID);
        if (cnt[t] > 0 && !ok)
        {
            fail(SC_EVALUE, "EulerCore: cell type " + std::to_string(t) + " is not supported in "
                                + std::to_string(ndim) + "D");
        }
    }

    // Batched instances: rank by descending step count (stable in the
    // instance id), rank of every cell in the reference order.
    if (ninst > 0)
    {
        if (ninst > (1 << 20))
        {
            fail(SC_EVALUE, "EulerCore batch: at most 1048576 instances");
        }
        std::vector<int32_t> order(static_cast<size_t>(ninst));
        for (int64_t k = 0; k < ninst; ++k)
        {
            if (steps[k] < 1)
            {
                fail(SC_EVALUE, "EulerCore batch: every instance needs at least one step");
            }
            order[k] = static_cast<int32_t>(k);
        }
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return steps[a] > steps[b]; });
        brank_of.assign(static_cast<size_t>(ninst), 0);
        bsteps.assign(static_cast<size_t>(ninst), 0);
        for (int64_t r = 0; r < ninst; ++r)
        {
            brank_of[order[r]] = static_cast<int32_t>(r);
            bsteps[r] = steps[order[r]];
        }
        std::vector<int32_t> rank_ref(static_cast<size_t>(N));
        for (int64_t c = 0; c < N; ++c)
        {
            int32_t const k = inst_of_cell[c];
            if (k < 0 || k >= ninst)
            {
                fail(SC_EVALUE, "EulerCore batch: instance id of cell " + std::to_string(c) + " out of range");
            }
            rank_ref[c] = brank_of[k];
        }
        brank_ref.alloc(N);
        SC_HIP(hipMemcpyAsync(brank_ref.ptr, rank_ref.data(), brank_ref.bytes(), hipMemcpyHostToDevice, s));
        sync();
    }

    // Cell order: by type, then along the Hilbert curve of the centroids.
    int_of_ref.alloc(N);
    ref_of_int.alloc(N);
    if (N > 0)
    {
        int const nparts = static_cast<int>(std::min<int64_t>(grid_for(N, BLK), 1024));
        DevBuf<double> box, parts;
        box.alloc(6);
        parts.alloc(6 * nparts);
        k_bbox<<<nparts, BLK, 0, s>>>(ndim, N, m->clcnd_b(), nullptr, parts.ptr);
        SC_HIP(hipGetLastError());
        k_bbox<<<1, BLK, 0, s>>>(ndim, nparts, nullptr, parts.ptr, box.ptr);
        SC_HIP(hipGetLastError());
        DevBuf<uint64_t> kin, kout;
        DevBuf<int32_t> vin;
        kin.alloc(N);
        kout.alloc(N);
        vin.alloc(N);
        k_keys<<<grid_for(N, BLK), BLK, 0, s>>>(ndim, static_cast<int32_t>(N), m->cltpn_b(), m->clcnd_b(), box.ptr,
                                                box.ptr + 3, use_sfc ? 1 : 0, brank_ref.ptr, kin.ptr, vin.ptr);
        SC_HIP(hipGetLastError());
        sort_pairs_u64(kin.ptr, kout.ptr, vin.ptr, ref_of_int.ptr, N, s);
        k_invert<<<grid_for(N, BLK), BLK, 0, s>>>(static_cast<int32_t>(N), ref_of_int.ptr, int_of_ref.ptr);
        SC_HIP(hipGetLastError());
    }
    int64_t off = 0;
    for (int t = 0; t < 9; ++t)
    {
        if (cnt[t] == 0)
        {
            continue;
        }
        Group gr;
        gr.type = t;
        gr.off = off;
        gr.n = cnt[t];
        off += cnt[t];
        groups.push_back(std::move(gr));
    }
    if (ninst > 0)
    {
        binst.alloc(N);
        DevBuf<int32_t> hist;
        hist.alloc(9 * ninst);
        SC_HIP(hipMemsetAsync(hist.ptr, 0, hist.bytes(), s));
        if (N > 0)
        {
            k_batch_inst<<<grid_for(N, BLK), BLK, 0, s>>>(static_cast<int32_t>(N), ref_of_int.ptr, brank_ref.ptr,
                                                           m->cltpn_b(), static_cast<int32_t>(ninst), binst.ptr,
                                                           hist.ptr);
            SC_HIP(hipGetLastError());
        }
        std::vector<int32_t> h(static_cast<size_t>(9 * ninst));
        SC_HIP(hipMemcpyAsync(h.data(), hist.ptr, hist.bytes(), hipMemcpyDeviceToHost, s));
        sync();
        for (auto & gr : groups)
        {
            std::vector<int64_t> cum(static_cast<size_t>(ninst + 1), 0);
            for (int64_t r = 0; r < ninst; ++r)
            {
                cum[r + 1] = cum[r] + h[static_cast<size_t>(gr.type * ninst + r)];
            }
            bcum.push_back(std::move(cum));
        }
        bhdt.alloc(ninst);
        bqdt.alloc(ninst);
        std::vector<double> d(static_cast<size_t>(ninst), time_increment);
        batch_set_dt(d.data());
    }
    // SC_EULER_GRAPH=1 makes replaying captured graphs of whole steps the
    // default instead of launching their kernels one by one.
    if (char const * env = std::getenv("SC_EULER_GRAPH"))
    {
        use_graph = std::atoi(env) != 0 ? 1 : 0;
    }
    // The form of every group is timed by the first substep unless
    // SC_EULER_FORM fixes it: a comma-separated list of entries, each a
    // form for every group or "T:form" for the group of cell type T.  A
    // form is fV (the fused kernel in variant V) or sA/B (the split pair in
    // variants A and B), where a variant is the minimum wavefronts per SIMD
    // followed by 's' when LDS-staged and 't' when the fused kernel reads
    // so0t from memory (f1, f3, f1s, f2t, s1s/2).  A group whose type lacks
    // the variant runs f1.
    gform.assign(groups.size(), Form{});
    if (char const * env = std::getenv("SC_EULER_FORM"))
    {
        auto parse_var = [](std::string const & t, int & code)
        {
            if (t.empty() || t[0] < '1' || t[0] > '4')
            {
                return false;
            }
            code = t[0] - '0';
            size_t k = 1;
            if (k < t.size() && t[k] == 's')
            {
                code += 10;
                ++k;
            }
            if (k < t.size() && t[k] == 't')
            {
                code += 100;
                ++k;
            }
            if (k < t.size() && t[k] == 'g')
            {
                code += 1000;
                ++k;
            }
            return k == t.size();
        };
        auto parse_form = [&](std::string const & c, Form & want)
        {
            if (c.size() >= 2 && c[0] == 'f')
            {
                return parse_var(c.substr(1), want.vf);
            }
            size_t const sl = c.find('/');
            if (c.size() >= 4 && c[0] == 's' && sl != std::string::npos)
            {
                want.split = 1;
                return parse_var(c.substr(1, sl - 1), want.v1) && parse_var(c.substr(sl + 1), want.v2);
            }
            return false;
        };
        std::string const all(env);
        bool any = false;
        size_t pos = 0;
        while (pos <= all.size())
        {
            size_t const end = std::min(all.find(',', pos), all.size());
            std::string item = all.substr(pos, end - pos);
            pos = end + 1;
            int type = -1;
            size_t const colon = item.find(':');
            if (colon != std::string::npos)
            {
                type = std::atoi(item.substr(0, colon).c_str());
                item = item.substr(colon + 1);
            }
            Form want;
            if (!parse_form(item, want))
            {
                continue;
            }
            bool const avail = want.split != 0 ? has_variant(1, want.v1) && has_variant(2, want.v2)
                                               : has_variant(0, want.vf);
            for (size_t gi = 0; gi < groups.size(); ++gi)
            {
                if (type < 0 || groups[gi].type == type)
                {
                    gform[gi] = avail ? want : Form{};
                }
            }
            any = true;
        }
        if (any)
        {
            tuned = 1;
        }
    }
    if (ninst > 0)
    {
        tuned = 1; // the batched march runs the fused kernel
    }
    for (auto & gr : groups)
    {
        int nf = 0;
        int nfield = 0;
        switch (gr.type)
        {
        case CT_TRI: nf = Traits<CT_TRI>::NF; nfield = Layout<CT_TRI>::NFIELD; break;
        case CT_QUAD: nf = Traits<CT_QUAD>::NF; nfield = Layout<CT_QUAD>::NFIELD; break;
        case CT_TET: nf = Traits<CT_TET>::NF; nfield = Layout<CT_TET>::NFIELD; break;
        case CT_HEX: nf = Traits<CT_HEX>::NF; nfield = Layout<CT_HEX>::NFIELD; break;
        case CT_PRISM: nf = Traits<CT_PRISM>::NF; nfield = Layout<CT_PRISM>::NFIELD; break;
        case CT_PYRAMID: nf = Traits<CT_PYRAMID>::NF; nfield = Layout<CT_PYRAMID>::NFIELD; break;
        default: break;
        }
        gr.nbr.alloc(static_cast<int64_t>(nf) * gr.n);
        gr.geo.alloc(static_cast<int64_t>(nfield) * gr.n);
    }

    // Solution tables (euler_solution.cpp:19-41): zero-filled.
    cen.alloc(static_cast<int64_t>(ndim) * R);
    for (int b = 0; b < 2; ++b)
    {
        s0[b].alloc(static_cast<int64_t>(neq) * R);
        s1[b].alloc(static_cast<int64_t>(neq) * ndim * R);
    }
    st[0].alloc(static_cast<int64_t>(neq) * R);
    st[1].alloc(static_cast<int64_t>(neq) * R);
    stm.alloc(static_cast<int64_t>(neq) * R);
    gam.alloc(R);
    cflo.alloc(R);
    cflc.alloc(R);
    for (auto * b : {&s0[0], &s0[1], &s1[0], &s1[1], &st[0], &st[1], &stm, &gam, &cflo, &cflc, &cen})
    {
        if (b->count)
        {
            SC_HIP(hipMemsetAsync(b->ptr, 0, b->bytes(), s));
        }
    }
    redo.alloc(3);
    SC_HIP(hipMemsetAsync(redo.ptr, 0, redo.bytes(), s));
    // Boundary tables.
    gkind.alloc(G);
    ginl.alloc(G);
    gsrc.alloc(G);
    gmat.alloc(static_cast<int64_t>(ndim) * ndim * G);
    fill<int32_t>(gkind.ptr, G, BC_NONE, s);
    fill<int32_t>(ginl.ptr, G, 0, s);
    if (G > 0)
    {
        if (m->nbound < G)
        {
            fail(SC_ERUNTIME, "EulerCore: mesh has fewer boundary faces than ghost cells");
        }
        k_ghost_meta<<<grid_for(G, BLK), BLK, 0, s>>>(ndim, G, m->bndfcs.ptr, m->fccls_b(), m->fcnml_b(),
                                                      int_of_ref.ptr, gsrc.ptr, gmat.ptr);
        SC_HIP(hipGetLastError());
    }
    inlet.alloc(1);
    fill<double>(inlet.ptr, 1, 0.0, s);
    prepare_ce();
}

// Time increments of the batched instances, given per instance id; the
// kernels read them per rank (hdt = dt/2, qdt = dt/4 as for one solve).
void DEuler::batch_set_dt(double const * dt_inst)
{
    AllocScope const scope(stream);
    std::vector<double> h(static_cast<size_t>(ninst));
    std::vector<double> q(static_cast<size_t>(ninst));
    for (int64_t k = 0; k < ninst; ++k)
    {
        h[brank_of[k]] = dt_inst[k] * 0.5;
        q[brank_of[k]] = dt_inst[k] * 0.25;
    }
    SC_HIP(hipMemcpyAsync(bhdt.ptr, h.data(), bhdt.bytes(), hipMemcpyHostToDevice, stream));
    SC_HIP(hipMemcpyAsync(bqdt.ptr, q.data(), bqdt.bytes(), hipMemcpyHostToDevice, stream));
    sync();
}

// dt of every instance such that its largest initial calc_cfl number is cfl
// (from the current so0n), then set.
void DEuler::batch_cfl_dt(double cfl, double * dt_inst)
{
    AllocScope const scope(stream);
    DevBuf<unsigned long long> cm;
    cm.alloc(ninst);
    SC_HIP(hipMemsetAsync(cm.ptr, 0, cm.bytes(), stream));
    double const * s0n = s0[1 - cur].ptr;
    for (auto & gr : groups)
    {
        unsigned const nb = grid_for(gr.n, BLK);
        switch (gr.type)
        {
#define SC_CM(TT)                                                                                              \
    case TT:                                                                                                   \
        k_batch_cflmax<TT><<<nb, BLK, 0, stream>>>(R, G, gr.off, gr.n, gr.geo.ptr, s0n, gam.ptr, binst.ptr, cm.ptr); \
        break;
            SC_CM(CT_TRI)
            SC_CM(CT_QUAD)
            SC_CM(CT_TET)
            SC_CM(CT_HEX)
            SC_CM(CT_PRISM)
            SC_CM(CT_PYRAMID)
#undef SC_CM
        default: break;
        }
        SC_HIP(hipGetLastError());
    }
    std::vector<unsigned long long> h(static_cast<size_t>(ninst));
    SC_HIP(hipMemcpyAsync(h.data(), cm.ptr, cm.bytes(), hipMemcpyDeviceToHost, stream));
    sync();
    for (int64_t k = 0; k < ninst; ++k)
    {
        double v;
        std::memcpy(&v, &h[brank_of[k]], sizeof(v));
        if (!(v > 0.0))
        {
            fail(SC_EVALUE, "EulerCore batch: instance " + std::to_string(k) + " has no positive wave speed");
        }
        dt_inst[k] = cfl / v;
    }
    batch_set_dt(dt_inst);
}

// Largest calc_cfl number over the body cells at dt = 1, from so0n.
double DEuler::cfl_max()
{
    AllocScope const scope(stream);
    DevBuf<unsigned long long> cm;
    cm.alloc(1);
    SC_HIP(hipMemsetAsync(cm.ptr, 0, cm.bytes(), stream));
    double const * s0n = s0[1 - cur].ptr;
    for (auto & gr : groups)
    {
        unsigned const nb = grid_for(gr.n, BLK);
        switch (gr.type)
        {
#define SC_CM(TT)                                                                                              \
    case TT:                                                                                                   \
        k_batch_cflmax<TT><<<nb, BLK, 0, stream>>>(R, G, gr.off, gr.n, gr.geo.ptr, s0n, gam.ptr, nullptr, cm.ptr); \
        break;
            SC_CM(CT_TRI)
            SC_CM(CT_QUAD)
            SC_CM(CT_TET)
            SC_CM(CT_HEX)
            SC_CM(CT_PRISM)
            SC_CM(CT_PYRAMID)
#undef SC_CM
        default: break;
        }
        SC_HIP(hipGetLastError());
    }
    unsigned long long h = 0;
    SC_HIP(hipMemcpyAsync(&h, cm.ptr, sizeof(h), hipMemcpyDeviceToHost, stream));
    sync();
    double v;
    std::memcpy(&v, &h, sizeof(v));
    return v;
}

void DEuler::prepare_ce()
{
    AllocScope const scope(stream);
    hipStream_t const s = stream;
    mesh->sync();
    MeshView mv;
    mv.clfcs = mesh->clfcs_b();
    mv.fccls = mesh->fccls_b();
    mv.fcnds = mesh->fcnds_b();
    mv.ndcrd = mesh->ndcrd_b();
    mv.clcnd = mesh->clcnd_b();
    mv.fccnd = mesh->fccnd_b();
    mv.G = G;
    mv.int_of_ref = int_of_ref.ptr;
    mv.ref_of_int = ref_of_int.ptr;
    if (G > 0)
    {
        k_cen_ghost<<<grid_for(G, BLK), BLK, 0, s>>>(ndim, G, mesh->clcnd.ptr, cen.ptr, R);
        SC_HIP(hipGetLastError());
    }
    for (auto & gr : groups)
    {
        unsigned const nb = grid_for(gr.n, BLK);
        switch (gr.type)
        {
#define SC_PREP(TT)                                                                                   \
    case TT:                                                                                          \
        k_prepare<TT><<<nb, BLK, 0, s>>>(mv, gr.off, gr.n, gr.nbr.ptr, gr.geo.ptr, cen.ptr, R);      \
        SC_HIP(hipGetLastError());                                                                    \
        k_dist<TT><<<nb, BLK, 0, s>>>(G, gr.off, gr.n, gr.geo.ptr, cen.ptr, R);                      \
        break;
            SC_PREP(CT_TRI)
            SC_PREP(CT_QUAD)
            SC_PREP(CT_TET)
            SC_PREP(CT_HEX)
            SC_PREP(CT_PRISM)
            SC_PREP(CT_PYRAMID)
#undef SC_PREP
        default: break;
        }
        SC_HIP(hipGetLastError());
    }
    sync();
}

void DEuler::init_solution(double ga, double rho, double const * v, double p)
{
    AllocScope const scope(stream);
    // euler_solution.cpp:60-98
    if (ga <= 1.0)
    {
        fail(SC_EVALUE, "EulerCore::init_solution: gamma " + std::to_string(ga) + " must be > 1");
    }
    if (rho <= 0.0)
    {
        fail(SC_EVALUE, "EulerCore::init_solution: rho " + std::to_string(rho) + " must be > 0");
    }
    if (p < 0.0)
    {
        fail(SC_EVALUE, "EulerCore::init_solution: pressure " + std::to_string(p) + " must be >= 0");
    }
    double vsq = 0.0;
    for (int d = 0; d < ndim; ++d)
    {
        vsq += v[d] * v[d];
    }
    double const energy = p / (ga - 1.0) + 0.5 * rho * vsq;
    if (R > 0)
    {
        k_init<<<grid_for(R, BLK), BLK, 0, stream>>>(R, G, neq, ndim, ga, rho, v[0], v[1], ndim == 3 ? v[2] : 0.0,
                                                     energy, gam.ptr, s0[1 - cur].ptr);
        SC_HIP(hipGetLastError());
    }
    sync();
}

// euler_bc.cpp:324-391
void DEuler::add_bc(int32_t kind, std::vector<int32_t> const & faces, std::vector<double> const & value)
{
    AllocScope const scope(stream);
    int32_t const nface = mesh->nface;
    for (int32_t f : faces)
    {
        if (f < 0 || f >= nface)
        {
            fail(SC_EVALUE, "EulerCore::add_bc: face " + std::to_string(f) + " out of range [0, " + std::to_string(nface)
                                + ")");
        }
    }
    std::vector<int32_t> const jcl = mesh->fcjcl_of(faces);
    for (size_t i = 0; i < faces.size(); ++i)
    {
        if (jcl[i] >= 0)
        {
            fail(SC_EVALUE, "EulerCore::add_bc: face " + std::to_string(faces[i]) + " is not a boundary face");
        }
        if (-jcl[i] > G)
        {
            fail(SC_EVALUE, "EulerCore::add_bc: face " + std::to_string(faces[i]) + " has no ghost cell");
        }
    }
    if (kind == BC_INLET)
    {
        size_t const expect = static_cast<size_t>(ndim) + 3;
        if (value.size() != expect)
        {
            fail(SC_EVALUE, "EulerCore::add_bc: inlet value size " + std::to_string(value.size())
                                + " must equal ndim+3 = " + std::to_string(expect));
        }
        double const rho = value[0];
        double const pr = value[1 + ndim];
        double const ga = value[2 + ndim];
        if (ga <= 1.0)
        {
            fail(SC_EVALUE, "EulerCore::add_bc: inlet gamma must be > 1");
        }
        if (rho <= 0.0)
        {
            fail(SC_EVALUE, "EulerCore::add_bc: inlet rho must be > 0");
        }
        if (pr < 0.0)
        {
            fail(SC_EVALUE, "EulerCore::add_bc: inlet pressure must be >= 0");
        }
    }
    else if (kind != BC_NONREFL && kind != BC_SLIPWALL)
    {
        fail(SC_EVALUE, "EulerCore::add_bc: unknown boundary kind");
    }
    EulerBoundary b;
    b.kind = kind;
    b.faces = faces;
    b.value = value;
    bcs.push_back(std::move(b));
    refresh_bc();
}

// The ghost of a face registered more than once takes the state of its last
// registration, exactly as the reference's in-order passes leave it.
void DEuler::refresh_bc()
{
    AllocScope const scope(stream);
    hipStream_t const s = stream;
    fill<int32_t>(gkind.ptr, G, BC_NONE, s);
    fill<int32_t>(ginl.ptr, G, 0, s);
    std::vector<double> vals;
    int32_t ninl = 0;
    for (auto const & b : bcs)
    {
        if (b.kind == BC_INLET)
        {
            vals.insert(vals.end(), b.value.begin(), b.value.end());
            ++ninl;
        }
    }
    inlet.alloc(std::max<int64_t>(1, static_cast<int64_t>(vals.size())));
    if (!vals.empty())
    {
        SC_HIP(hipMemcpyAsync(inlet.ptr, vals.data(), sizeof(double) * vals.size(), hipMemcpyHostToDevice, s));
    }
    else
    {
        fill<double>(inlet.ptr, 1, 0.0, s);
    }
    int32_t iinl = 0;
    for (auto const & b : bcs)
    {
        if (b.faces.empty())
        {
            if (b.kind == BC_INLET)
            {
                ++iinl;
            }
            continue;
        }
        DevBuf<int32_t> df;
        df.alloc(static_cast<int64_t>(b.faces.size()));
        SC_HIP(hipMemcpyAsync(df.ptr, b.faces.data(), df.bytes(), hipMemcpyHostToDevice, s));
        int32_t const inl = b.kind == BC_INLET ? iinl : 0;
        k_set_bc<<<grid_for(df.count, BLK), BLK, 0, s>>>(static_cast<int32_t>(df.count), df.ptr, mesh->fccls_b(), G,
                                                         b.kind, inl, gkind.ptr, ginl.ptr);
        SC_HIP(hipGetLastError());
        sync();
        if (b.kind == BC_INLET)
        {
            ++iinl;
        }
    }
    (void)ninl;
    sync();
}

void DEuler::table(int id, int64_t & rows, int64_t & w, int64_t & ngh, int & rank)
{
    ngh = G;
    switch (id)
    {
    case EA_CEVOL: rows = R; w = CLMFC + 1; rank = 2; break;
    case EA_CECND: rows = R; w = (CLMFC + 1) * ndim; rank = 2; break;
    case EA_SFCND:
    case EA_SFNML: rows = N; w = CLMFC * FCMND * ndim; rank = 3; ngh = 0; break;
    case EA_SO0C:
    case EA_SO0N:
    case EA_SO0T:
    case EA_STM: rows = R; w = neq; rank = 2; break;
    case EA_SO1C:
    case EA_SO1N: rows = R; w = neq * ndim; rank = 3; break;
    case EA_CFLO:
    case EA_CFLC:
    case EA_GAMMA: rows = R; w = 1; rank = 1; break;
    default: fail(SC_EVALUE, "EulerCore: unknown table id");
    }
}

void DEuler::copy(int id, double * host, int64_t count, bool down)
{
    AllocScope const scope(stream);
    hipStream_t const s = stream;
    int64_t rows, w, ngh;
    int rank;
    table(id, rows, w, ngh, rank);
    if (count != rows * w)
    {
        fail(SC_EVALUE, "EulerCore: table size mismatch");
    }
    if (count == 0)
    {
        return;
    }
    DevBuf<double> api;
    api.alloc(count);
    if (down)
    {
        SC_HIP(hipMemsetAsync(api.ptr, 0, api.bytes(), s));
    }
    else
    {
        SC_HIP(hipMemcpyAsync(api.ptr, host, api.bytes(), hipMemcpyHostToDevice, s));
    }
    if (id == EA_CEVOL || id == EA_CECND || id == EA_SFCND || id == EA_SFNML)
    {
        for (auto & gr : groups)
        {
            unsigned const nb = grid_for(gr.n, BLK);
            switch (gr.type)
            {
#define SC_GEOM(TT)                                                                                                 \
    case TT:                                                                                                        \
        if (down)                                                                                                   \
        {                                                                                                           \
            k_geom<TT, true><<<nb, BLK, 0, s>>>(id, G, gr.off, gr.n, ref_of_int.ptr, api.ptr, gr.geo.ptr, cen.ptr, R); \
        }                                                                                                           \
        else                                                                                                        \
        {                                                                                                           \
            k_geom<TT, false><<<nb, BLK, 0, s>>>(id, G, gr.off, gr.n, ref_of_int.ptr, api.ptr, gr.geo.ptr, cen.ptr, R); \
            if (id == EA_CECND)                                                                                     \
            {                                                                                                       \
                k_dist<TT><<<nb, BLK, 0, s>>>(G, gr.off, gr.n, gr.geo.ptr, cen.ptr, R);                             \
            }                                                                                                       \
        }                                                                                                           \
        break;
                SC_GEOM(CT_TRI)
                SC_GEOM(CT_QUAD)
                SC_GEOM(CT_TET)
                SC_GEOM(CT_HEX)
                SC_GEOM(CT_PRISM)
                SC_GEOM(CT_PYRAMID)
#undef SC_GEOM
            default: break;
            }
            SC_HIP(hipGetLastError());
        }
    }
    else
    {
        double * soa = nullptr;
        switch (id)
        {
        case EA_SO0C: soa = s0[cur].ptr; break;
        case EA_SO0N: soa = s0[1 - cur].ptr; break;
        case EA_SO0T: soa = st[cur].ptr; break;
        case EA_STM: soa = stm.ptr; break;
        case EA_SO1C: soa = s1[cur].ptr; break;
        case EA_SO1N: soa = s1[1 - cur].ptr; break;
        case EA_CFLO: soa = cflo.ptr; break;
        case EA_CFLC: soa = cflc.ptr; break;
        case EA_GAMMA: soa = gam.ptr; break;
        default: break;
        }
        if (down)
        {
            k_state<true><<<grid_for(R, BLK), BLK, 0, s>>>(R, G, static_cast<int>(w), int_of_ref.ptr, api.ptr, soa);
        }
        else
        {
            k_state<false><<<grid_for(R, BLK), BLK, 0, s>>>(R, G, static_cast<int>(w), int_of_ref.ptr, api.ptr, soa);
        }
        SC_HIP(hipGetLastError());
    }
    if (down)
    {
        SC_HIP(hipMemcpyAsync(host, api.ptr, api.bytes(), hipMemcpyDeviceToHost, s));
    }
    sync();
}

// --------------------------------------------------------------------------
// The serving arena: continuous batching of independent solves.
//
// One arena per dimension holds, at fixed capacity, the tables the batched
// fused kernel reads: per cell type a group of SoA geometry and neighbour
// lists, and the row tables (ghost rows first, then each type group).  A
// request is set up as a plain EulerCore of its own mesh (device mesh
// build, CE geometry, state, boundary conditions, dt) and admitted by
// copying its tables into free whole-block ranges of the arena, with its
// neighbour rows renumbered.  Every substep is one fused launch per type
// group over the arena's used prefix; blocks of free ranges and of
// finished solves return at once (their instance's half time increment is
// negative).  A solve that has run its steps is marked finished and its
// state drained to pinned host memory on the arena's stream; its ranges
// are reused after its result is taken.  Each solve's arithmetic is that
// of the plain march: its cells read only its own rows.
// --------------------------------------------------------------------------

namespace
{

// First-fit allocator of [0, cap) in whole ranges; merges on release.
struct RangeAlloc
{
    int64_t cap = 0;
    std::map<int64_t, int64_t> free_; // start -> end

    void init(int64_t c)
    {
        cap = c;
        free_.clear();
        if (c > 0)
        {
            free_[0] = c;
        }
    }
    int64_t take(int64_t n)
    {
        if (n <= 0)
        {
            return 0;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it)
        {
            if (it->second - it->first >= n)
            {
                int64_t const a = it->first;
                int64_t const e = it->second;
                free_.erase(it);
                if (a + n < e)
                {
                    free_[a + n] = e;
                }
                return a;
            }
        }
        return -1;
    }
    void give(int64_t a, int64_t n)
    {
        if (n <= 0)
        {
            return;
        }
        int64_t e = a + n;
        auto it = free_.lower_bound(a);
        if (it != free_.end() && it->first == e)
        {
            e = it->second;
            it = free_.erase(it);
        }
        if (it != free_.begin())
        {
            auto pv = std::prev(it);
            if (pv->second == a)
            {
                a = pv->first;
                free_.erase(pv);
            }
        }
        free_[a] = e;
    }
    // End of the used prefix.
    int64_t hwm() const
    {
        if (free_.empty())
        {
            return cap;
        }
        auto last = std::prev(free_.end());
        return last->second == cap ? last->first : cap;
    }
    int64_t free_total() const
    {
        int64_t t = 0;
        for (auto const & kv : free_)
        {
            t += kv.second - kv.first;
        }
        return t;
    }
};

// Renumber an instance's neighbour rows into the arena: ghost row r to
// ghost0 + r, body row G + b of instance group k (its cells
// [goff[k], goff[k] + gn[k]) of the instance's body order) to
// abase[k] + b - goff[k].
struct RowMap
{
    int64_t G;
    int64_t ghost0;
    int ng;
    int64_t goff[4];
    int64_t gn[4];
    int64_t abase[4];
};

__device__ inline int64_t map_row(RowMap const & m, int64_t r)
{
    if (r < m.G)
    {
        return m.ghost0 + r;
    }
    int64_t const b = r - m.G;
    int64_t out = -1;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        if (k < m.ng && b >= m.goff[k] && b < m.goff[k] + m.gn[k])
        {
            out = m.abase[k] + (b - m.goff[k]);
        }
    }
    return out;
}

__global__ void k_arena_nbr(int nf, int64_t n, int32_t const * src, int64_t cap, int32_t * dst, RowMap m)
{
    int64_t const t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nf * n)
    {
        return;
    }
    int64_t const f = t / n;
    int64_t const i = t - f * n;
    dst[f * cap + i] = static_cast<int32_t>(map_row(m, src[f * n + i]));
}

__global__ void k_arena_ginl(int64_t n, int32_t const * src, int32_t add, int32_t * dst)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
    {
        dst[i] = src[i] + add;
    }
}

__global__ void k_arena_set_dt(double * bhdt, double * bqdt, int64_t slot, double h, double q)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        bhdt[slot] = h;
        bqdt[slot] = q;
    }
}

} /* end namespace */

struct DArena
{
    int ndim = 0;
    int neq = 0;
    int64_t Gcap = 0;
    int64_t Ncap = 0;
    int64_t Rcap = 0;
    int64_t max_slots = 0;
    int64_t inl_cap = 0;
    double sigma0 = 3.0;
    double taumin = 0.0;
    double tauscale = 1.0;
    StreamHolder sh; // declared before the buffers: freed last
    hipStream_t stream = sh.s;

    struct AGroup
    {
        int32_t type = 0;
        int64_t off = 0;
        int64_t cap = 0;
        int nf = 0;
        int nfield = 0;
        DevBuf<int32_t> nbr;
        DevBuf<double> geo;
        RangeAlloc ra;
    };
    std::vector<AGroup> groups;
    DevBuf<double> cen, s0[2], s1[2], st[2], gam, cflo, cflc, gmat, inlet;
    DevBuf<int32_t> gkind, ginl, binst;
    DevBuf<double> bhdt, bqdt; // per slot; entry max_slots is the dead instance
    RangeAlloc ghosts, inlets;
    int cur = 0;

    struct Slot
    {
        bool used = false;
        bool marching = false;
        bool done_ok = false;
        int64_t steps_left = 0;
        int64_t ncell = 0;
        int64_t G = 0;
        int64_t ghost0 = 0;
        int64_t inl0 = 0;
        int64_t ninl = 0;
        std::vector<int> gidx;        // arena group of each instance group
        std::vector<int64_t> start;   // arena cell start per instance group
        std::vector<int64_t> count;   // cells per instance group
        std::vector<int64_t> ioff;    // instance body offset per instance group
        std::vector<int32_t> ref_of_int;
        double * host = nullptr;      // pinned: so0 then so1 body rows, internal order
        hipEvent_t done = nullptr;
    };
    std::vector<Slot> slots;
    std::vector<int64_t> live;     // marching slots, in admission order
    std::deque<hipEvent_t> chunks; // events after enqueued step chunks

    DArena(int nd, int64_t const * type_cap, int64_t ghost_cap, int64_t nslot)
    {
        if (nd != 2 && nd != 3)
        {
            fail(SC_EVALUE, "serve arena: ndim must be 2 or 3");
        }
        if (nslot < 1 || nslot > (1 << 20))
        {
            fail(SC_EVALUE, "serve arena: between 1 and 1048576 slots");
        }
        if (ghost_cap < 0)
        {
            fail(SC_EVALUE, "serve arena: negative ghost capacity");
        }
        AllocScope const scope(stream);
        ndim = nd;
        neq = nd + 2;
        max_slots = nslot;
        inl_cap = 64 * nslot;
        int const types2[] = {CT_QUAD, CT_TRI};
        int const types3[] = {CT_HEX, CT_TET, CT_PRISM, CT_PYRAMID};
        int const * tl = nd == 2 ? types2 : types3;
        int const nt = nd == 2 ? 2 : 4;
        int64_t off = 0;
        for (int k = 0; k < nt; ++k)
        {
            int const t = tl[k];
            int64_t c = type_cap[t];
            if (c < 0)
            {
                fail(SC_EVALUE, "serve arena: negative cell capacity");
            }
            if (c == 0)
            {
                continue;
            }
            c = (c + BLK - 1) / BLK * BLK;
            AGroup g;
            g.type = t;
            g.off = off;
            g.cap = c;
            switch (t)
            {
            case CT_TRI: g.nf = Traits<CT_TRI>::NF; g.nfield = Layout<CT_TRI>::NFIELD; break;
            case CT_QUAD: g.nf = Traits<CT_QUAD>::NF; g.nfield = Layout<CT_QUAD>::NFIELD; break;
            case CT_TET: g.nf = Traits<CT_TET>::NF; g.nfield = Layout<CT_TET>::NFIELD; break;
            case CT_HEX: g.nf = Traits<CT_HEX>::NF; g.nfield = Layout<CT_HEX>::NFIELD; break;
            case CT_PRISM: g.nf = Traits<CT_PRISM>::NF; g.nfield = Layout<CT_PRISM>::NFIELD; break;
            case CT_PYRAMID: g.nf = Traits<CT_PYRAMID>::NF; g.nfield = Layout<CT_PYRAMID>::NFIELD; break;
            default: break;
            }
            g.nbr.alloc(static_cast<int64_t>(g.nf) * c);
            g.geo.alloc(static_cast<int64_t>(g.nfield) * c);
            // Finite fill: dead blocks never read it, but the geometry
            // prefetch of a live block's tail may load it.
            SC_HIP(hipMemsetAsync(g.geo.ptr, 0, g.geo.bytes(), stream));
            fill<int32_t>(g.nbr.ptr, static_cast<int64_t>(g.nf) * c, 0, stream);
            g.ra.init(c);
            off += c;
            groups.push_back(std::move(g));
        }
        Ncap = off;
        Gcap = ghost_cap;
        Rcap = Gcap + Ncap;
        cen.alloc(static_cast<int64_t>(ndim) * Rcap);
        for (int b = 0; b < 2; ++b)
        {
            s0[b].alloc(static_cast<int64_t>(neq) * Rcap);
            s1[b].alloc(static_cast<int64_t>(neq) * ndim * Rcap);
            st[b].alloc(static_cast<int64_t>(neq) * Rcap);
        }
        gam.alloc(Rcap);
        cflo.alloc(Rcap);
        cflc.alloc(Rcap);
        for (auto * b : {&s0[0], &s0[1], &s1[0], &s1[1], &st[0], &st[1], &gam, &cflo, &cflc, &cen})
        {
            if (b->count)
            {
                SC_HIP(hipMemsetAsync(b->ptr, 0, b->bytes(), stream));
            }
        }
        gkind.alloc(Gcap);
        ginl.alloc(Gcap);
        gmat.alloc(static_cast<int64_t>(ndim) * ndim * Gcap);
        fill<int32_t>(gkind.ptr, Gcap, BC_NONE, stream);
        fill<int32_t>(ginl.ptr, Gcap, 0, stream);
        inlet.alloc(inl_cap * (ndim + 3));
        binst.alloc(Ncap);
        fill<int32_t>(binst.ptr, Ncap, static_cast<int32_t>(max_slots), stream);
        bhdt.alloc(max_slots + 1);
        bqdt.alloc(max_slots + 1);
        fill<double>(bhdt.ptr, max_slots + 1, -1.0, stream);
        fill<double>(bqdt.ptr, max_slots + 1, -1.0, stream);
        ghosts.init(Gcap);
        inlets.init(inl_cap);
        slots.resize(static_cast<size_t>(max_slots));
        sync();
    }

    ~DArena()
    {
        (void)hipStreamSynchronize(stream);
        for (auto ev : chunks)
        {
            (void)hipEventDestroy(ev);
        }
        for (auto & sl : slots)
        {
            release_host(sl);
        }
    }

    void sync() { SC_HIP(hipStreamSynchronize(stream)); }

    static void release_host(Slot & sl)
    {
        if (sl.done != nullptr)
        {
            (void)hipEventDestroy(sl.done);
            sl.done = nullptr;
        }
        if (sl.host != nullptr)
        {
            (void)hipHostFree(sl.host);
            sl.host = nullptr;
        }
    }

    int group_of(int32_t type) const
    {
        for (size_t k = 0; k < groups.size(); ++k)
        {
            if (groups[k].type == type)
            {
                return static_cast<int>(k);
            }
        }
        return -1;
    }

    KP params() const
    {
        KP p;
        p.R = Rcap;
        p.G = Gcap;
        p.hdt = 0.0;
        p.qdt = 0.0;
        p.sigma0 = sigma0;
        p.taumin = taumin;
        p.tauscale = tauscale;
        p.s0c = s0[cur].ptr;
        p.s1c = s1[cur].ptr;
        p.s0n = s0[1 - cur].ptr;
        p.s1n = s1[1 - cur].ptr;
        p.st = st[cur].ptr;
        p.stn = st[1 - cur].ptr;
        p.gam = gam.ptr;
        p.cflo = cflo.ptr;
        p.cflc = cflc.ptr;
        p.cen = cen.ptr;
        p.gkind = gkind.ptr;
        p.ginl = ginl.ptr;
        p.gmat = gmat.ptr;
        p.inlet = inlet.ptr;
        p.inst = binst.ptr;
        p.bhdt = bhdt.ptr;
        p.bqdt = bqdt.ptr;
        return p;
    }

    // Admit a set-up solver for `steps` whole steps; returns the slot, or
    // -1 when its ranges do not fit now.  The solver's tables are copied on
    // the arena's stream after the solver's stream work; the solver's
    // stream then waits for the copies, so it may be destroyed at once.
    int64_t admit(DEuler & e, int64_t steps)
    {
        if (e.ninst != 0)
        {
            fail(SC_EVALUE, "serve arena: a batched solver cannot be admitted");
        }
        if (e.ndim != ndim)
        {
            fail(SC_EVALUE, "serve arena: the solver's ndim differs from the arena's");
        }
        if (steps < 1)
        {
            fail(SC_EVALUE, "serve arena: a solve needs at least one step");
        }
        if (e.sigma0 != sigma0 || e.taumin != taumin || e.tauscale != tauscale)
        {
            fail(SC_EVALUE, "serve arena: the solver's sigma0/taumin/tauscale differ from the arena's");
        }
        if (e.groups.size() > 4)
        {
            fail(SC_ERUNTIME, "serve arena: too many cell-type groups");
        }
        int64_t slot = -1;
        for (int64_t k = 0; k < max_slots; ++k)
        {
            if (!slots[k].used)
            {
                slot = k;
                break;
            }
        }
        if (slot < 0)
        {
            return -1;
        }
        int64_t ninl = 0;
        for (auto const & b : e.bcs)
        {
            if (b.kind == BC_INLET)
            {
                ++ninl;
            }
        }
        // Take the ranges, or give back what was taken.
        std::vector<int> gidx;
        std::vector<int64_t> start;
        for (auto const & g : e.groups)
        {
            int const a = group_of(g.type);
            if (a < 0)
            {
                fail(SC_EVALUE, "serve arena: no capacity for cell type " + std::to_string(g.type));
            }
            gidx.push_back(a);
        }
        bool ok = true;
        for (size_t k = 0; k < e.groups.size() && ok; ++k)
        {
            int64_t const nb = (e.groups[k].n + BLK - 1) / BLK * BLK;
            int64_t const a = groups[gidx[k]].ra.take(nb);
            if (a < 0)
            {
                ok = false;
            }
            else
            {
                start.push_back(a);
            }
        }
        int64_t const g0 = ok ? ghosts.take(e.G) : -1;
        ok = ok && g0 >= 0;
        int64_t const i0 = ok ? inlets.take(ninl) : -1;
        ok = ok && i0 >= 0;
        if (!ok)
        {
            for (size_t k = 0; k < start.size(); ++k)
            {
                groups[gidx[k]].ra.give(start[k], (e.groups[k].n + BLK - 1) / BLK * BLK);
            }
            if (g0 >= 0)
            {
                ghosts.give(g0, e.G);
            }
            return -1;
        }
        AllocScope const scope(stream);
        Slot & sl = slots[slot];
        sl = Slot{};
        sl.used = true;
        sl.marching = true;
        sl.steps_left = steps;
        sl.ncell = e.N;
        sl.G = e.G;
        sl.ghost0 = g0;
        sl.inl0 = i0;
        sl.ninl = ninl;
        sl.gidx = gidx;
        sl.start = start;
        for (auto const & g : e.groups)
        {
            sl.count.push_back(g.n);
            sl.ioff.push_back(g.off);
        }
        try
        {
            copy_in(e, sl, slot);
        }
        catch (...)
        {
            free_slot(slot);
            throw;
        }
        live.push_back(slot);
        return slot;
    }

    void copy_in(DEuler & e, Slot & sl, int64_t slot)
    {
        hipStream_t const s = stream;
        int const nd = ndim;
        int const q = neq;
        // The instance's reference order of its body cells.
        sl.ref_of_int.resize(static_cast<size_t>(e.N));
        if (e.N > 0)
        {
            SC_HIP(hipMemcpyAsync(sl.ref_of_int.data(), e.ref_of_int.ptr, e.ref_of_int.bytes(), hipMemcpyDeviceToHost,
                                  e.stream));
        }
        e.sync();
        hipEvent_t ev = nullptr;
        SC_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        RowMap m{};
        m.G = e.G;
        m.ghost0 = sl.ghost0;
        m.ng = static_cast<int>(e.groups.size());
        for (size_t k = 0; k < e.groups.size(); ++k)
        {
            m.goff[k] = e.groups[k].off;
            m.gn[k] = e.groups[k].n;
            m.abase[k] = Gcap + groups[sl.gidx[k]].off + sl.start[k];
        }
        // The instance's so0n / so1n (s*[1 - cur]) land in the arena's
        // so0n / so1n, and its other level likewise.
        int const bmap[2] = {e.cur == cur ? 0 : 1, e.cur == cur ? 1 : 0};
        auto rows = [&](double const * src, double * dst, int64_t w, int64_t r_src, int64_t r_dst, int64_t n)
        {
            if (n > 0)
            {
                SC_HIP(hipMemcpyAsync(dst + r_dst * w, src + r_src * w, sizeof(double) * w * n,
                                      hipMemcpyDeviceToDevice, s));
            }
        };
        auto all_rows = [&](int64_t r_src, int64_t r_dst, int64_t n)
        {
            rows(e.cen.ptr, cen.ptr, nd, r_src, r_dst, n);
            for (int b = 0; b < 2; ++b)
            {
                rows(e.s0[b].ptr, s0[bmap[b]].ptr, q, r_src, r_dst, n);
                rows(e.s1[b].ptr, s1[bmap[b]].ptr, q * nd, r_src, r_dst, n);
            }
            rows(e.gam.ptr, gam.ptr, 1, r_src, r_dst, n);
            rows(e.cflo.ptr, cflo.ptr, 1, r_src, r_dst, n);
            rows(e.cflc.ptr, cflc.ptr, 1, r_src, r_dst, n);
        };
        try
        {
            for (size_t k = 0; k < e.groups.size(); ++k)
            {
                auto const & eg = e.groups[k];
                auto & ag = groups[sl.gidx[k]];
                int64_t const a = sl.start[k];
                SC_HIP(hipMemcpy2DAsync(ag.geo.ptr + a, sizeof(double) * ag.cap, eg.geo.ptr, sizeof(double) * eg.n,
                                        sizeof(double) * eg.n, ag.nfield, hipMemcpyDeviceToDevice, s));
                k_arena_nbr<<<grid_for(ag.nf * eg.n, BLK), BLK, 0, s>>>(ag.nf, eg.n, eg.nbr.ptr, ag.cap,
                                                                        ag.nbr.ptr + a, m);
                SC_HIP(hipGetLastError());
                all_rows(e.G + eg.off, Gcap + ag.off + a, eg.n);
                int64_t const nb = (eg.n + BLK - 1) / BLK * BLK;
                fill<int32_t>(binst.ptr + ag.off + a, eg.n, static_cast<int32_t>(slot), s);
                fill<int32_t>(binst.ptr + ag.off + a + eg.n, nb - eg.n, static_cast<int32_t>(max_slots), s);
            }
            if (e.G > 0)
            {
                all_rows(0, sl.ghost0, e.G);
                SC_HIP(hipMemcpyAsync(gkind.ptr + sl.ghost0, e.gkind.ptr, sizeof(int32_t) * e.G,
                                      hipMemcpyDeviceToDevice, s));
                k_arena_ginl<<<grid_for(e.G, BLK), BLK, 0, s>>>(e.G, e.ginl.ptr, static_cast<int32_t>(sl.inl0),
                                                                ginl.ptr + sl.ghost0);
                SC_HIP(hipGetLastError());
                SC_HIP(hipMemcpy2DAsync(gmat.ptr + sl.ghost0, sizeof(double) * Gcap, e.gmat.ptr, sizeof(double) * e.G,
                                        sizeof(double) * e.G, nd * nd, hipMemcpyDeviceToDevice, s));
            }
            if (sl.ninl > 0)
            {
                SC_HIP(hipMemcpyAsync(inlet.ptr + sl.inl0 * (nd + 3), e.inlet.ptr, sizeof(double) * sl.ninl * (nd + 3),
                                      hipMemcpyDeviceToDevice, s));
            }
            if (nd == 3)
            {
                // The 3D kernels read so0t: compute it for both levels of
                // the admitted rows; the kernels that write a level then
                // write its so0t (rows no kernel writes keep theirs).
                std::vector<std::pair<int64_t, int64_t>> rr;
                rr.emplace_back(sl.ghost0, e.G);
                for (size_t k = 0; k < e.groups.size(); ++k)
                {
                    rr.emplace_back(Gcap + groups[sl.gidx[k]].off + sl.start[k], e.groups[k].n);
                }
                for (int b = 0; b < 2; ++b)
                {
                    for (auto const & r : rr)
                    {
                        if (r.second <= 0)
                        {
                            continue;
                        }
                        KP t = params();
                        t.s0c = s0[b].ptr + r.first * q;
                        t.s1c = s1[b].ptr + r.first * q * nd;
                        t.gam = gam.ptr + r.first;
                        t.st = st[b].ptr + r.first * q;
                        t.R = r.second;
                        k_solt<3><<<grid_for(r.second, BLK), BLK, 0, s>>>(t);
                        SC_HIP(hipGetLastError());
                    }
                }
            }
            k_arena_set_dt<<<1, 64, 0, s>>>(bhdt.ptr, bqdt.ptr, slot, e.dt * 0.5, e.dt * 0.25);
            SC_HIP(hipGetLastError());
            SC_HIP(hipHostMalloc(reinterpret_cast<void **>(&sl.host), sizeof(double) * e.N * q * (1 + nd)));
            SC_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
            // The solver's buffers are freed on its own stream: order that
            // after the copies.
            SC_HIP(hipEventRecord(ev, s));
            SC_HIP(hipStreamWaitEvent(e.stream, ev, 0));
        }
        catch (...)
        {
            (void)hipEventDestroy(ev);
            throw;
        }
        (void)hipEventDestroy(ev);
    }

    // Return a slot's ranges; its cells become dead first.
    void free_slot(int64_t slot)
    {
        Slot & sl = slots[slot];
        if (!sl.used)
        {
            return;
        }
        AllocScope const scope(stream);
        k_arena_set_dt<<<1, 64, 0, stream>>>(bhdt.ptr, bqdt.ptr, slot, -1.0, -1.0);
        SC_HIP(hipGetLastError());
        for (size_t k = 0; k < sl.start.size(); ++k)
        {
            auto & ag = groups[sl.gidx[k]];
            int64_t const nb = (sl.count[k] + BLK - 1) / BLK * BLK;
            fill<int32_t>(binst.ptr + ag.off + sl.start[k], nb, static_cast<int32_t>(max_slots), stream);
            ag.ra.give(sl.start[k], nb);
        }
        ghosts.give(sl.ghost0, sl.G);
        inlets.give(sl.inl0, sl.ninl);
        release_host(sl);
        sl = Slot{};
    }

    // Each group's launch covers the blocks from the first to the last
    // marching solve's range in it; blocks of free ranges and finished
    // solves within that span return at once.
    void launch_groups()
    {
        KP const p = params();
        for (size_t gi = 0; gi < groups.size(); ++gi)
        {
            auto & gr = groups[gi];
            int64_t lo = gr.cap;
            int64_t hi = 0;
            for (int64_t slot : live)
            {
                Slot const & sl = slots[slot];
                for (size_t k = 0; k < sl.gidx.size(); ++k)
                {
                    if (sl.gidx[k] == static_cast<int>(gi))
                    {
                        lo = std::min(lo, sl.start[k]);
                        hi = std::max(hi, sl.start[k] + sl.count[k]);
                    }
                }
            }
            if (hi <= lo)
            {
                continue;
            }
            GV const g{gr.off + lo, gr.cap, gr.nbr.ptr + lo, gr.geo.ptr + lo, hi - lo};
            unsigned const nb = grid_for(hi - lo, BLK);
            constexpr int MB = M_FUSED | M_BATCH;
            constexpr int MT = MB | batch_waves(CT_TRI);
            // 3D: the fused kernel that reads so0t, stages the block's rows
            // in LDS and writes the next level's so0t (chained so0t, the
            // form the plain march keeps for tetrahedra and prisms).
            constexpr int M3 = MB | M_W2 | M_STMEM | M_STAGE | M_STNEXT;
            switch (gr.type)
            {
            case CT_TRI: k_cell<CT_TRI, MT | M_GDMA><<<nb, BLK, 0, stream>>>(p, g); break;
            case CT_QUAD: k_cell<CT_QUAD, MB | batch_waves(CT_QUAD)><<<nb, BLK, 0, stream>>>(p, g); break;
            case CT_TET: k_cell<CT_TET, M3><<<nb, BLK, 0, stream>>>(p, g); break;
            case CT_HEX: k_cell<CT_HEX, M3><<<nb, BLK, 0, stream>>>(p, g); break;
            case CT_PRISM: k_cell<CT_PRISM, M3><<<nb, BLK, 0, stream>>>(p, g); break;
            case CT_PYRAMID: k_cell<CT_PYRAMID, M3><<<nb, BLK, 0, stream>>>(p, g); break;
            default: fail(SC_ERUNTIME, "serve arena: unsupported cell type in group");
            }
            SC_HIP(hipGetLastError());
        }
    }

    // Enqueue k whole steps of every marching solve.  A solve whose steps
    // end within them is marked finished after its last step, and its
    // so0n and so1n body rows are drained to its pinned buffer.
    void step(int64_t k)
    {
        AllocScope const scope(stream);
        for (int64_t it = 0; it < k; ++it)
        {
            if (live.empty())
            {
                break;
            }
            for (int sub = 0; sub < 2; ++sub)
            {
                cur = 1 - cur;
                launch_groups();
            }
            std::vector<int64_t> still;
            for (int64_t slot : live)
            {
                Slot & sl = slots[slot];
                if (--sl.steps_left > 0)
                {
                    still.push_back(slot);
                    continue;
                }
                sl.marching = false;
                k_arena_set_dt<<<1, 64, 0, stream>>>(bhdt.ptr, bqdt.ptr, slot, -1.0, -1.0);
                SC_HIP(hipGetLastError());
                int64_t const q = neq;
                for (size_t g = 0; g < sl.start.size(); ++g)
                {
                    auto const & ag = groups[sl.gidx[g]];
                    int64_t const r = Gcap + ag.off + sl.start[g];
                    SC_HIP(hipMemcpyAsync(sl.host + (sl.ioff[g]) * q, s0[1 - cur].ptr + r * q,
                                          sizeof(double) * sl.count[g] * q, hipMemcpyDeviceToHost, stream));
                    SC_HIP(hipMemcpyAsync(sl.host + sl.ncell * q + sl.ioff[g] * q * ndim, s1[1 - cur].ptr + r * q * ndim,
                                          sizeof(double) * sl.count[g] * q * ndim, hipMemcpyDeviceToHost, stream));
                }
                SC_HIP(hipEventRecord(sl.done, stream));
            }
            live.swap(still);
        }
        hipEvent_t ev = nullptr;
        SC_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        SC_HIP(hipEventRecord(ev, stream));
        chunks.push_back(ev);
    }

    // Enqueued step chunks not yet completed.
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

    // 1 when the slot's result has been drained, 0 while it marches.
    int query(int64_t slot)
    {
        Slot & sl = slots[slot];
        if (!sl.used)
        {
            fail(SC_EVALUE, "serve arena: slot " + std::to_string(slot) + " holds no solve");
        }
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

    // The drained so0n [ncell][neq] and so1n [ncell][neq][ndim] in the
    // reference cell order (either pointer may be null); frees the slot.
    void result(int64_t slot, double * so0_out, double * so1_out)
    {
        Slot & sl = slots[slot];
        if (!sl.used)
        {
            fail(SC_EVALUE, "serve arena: slot " + std::to_string(slot) + " holds no solve");
        }
        if (sl.marching)
        {
            fail(SC_ERUNTIME, "serve arena: slot " + std::to_string(slot) + " is still marching");
        }
        SC_HIP(hipEventSynchronize(sl.done));
        int64_t const q = neq;
        int64_t const w1 = q * ndim;
        double const * h0 = sl.host;
        double const * h1 = sl.host + sl.ncell * q;
        for (int64_t c = 0; c < sl.ncell; ++c)
        {
            int64_t const r = sl.ref_of_int[c];
            if (so0_out != nullptr)
            {
                std::memcpy(so0_out + r * q, h0 + c * q, sizeof(double) * q);
            }
            if (so1_out != nullptr)
            {
                std::memcpy(so1_out + r * w1, h1 + c * w1, sizeof(double) * w1);
            }
        }
        free_slot(slot);
    }

    int64_t device_bytes() const
    {
        int64_t b = 0;
        for (auto const & g : groups)
        {
            b += static_cast<int64_t>(g.nbr.bytes() + g.geo.bytes());
        }
        for (auto const * x : {&cen, &s0[0], &s0[1], &s1[0], &s1[1], &st[0], &st[1], &gam, &cflo, &cflc, &gmat, &inlet,
                               &bhdt, &bqdt})
        {
            b += static_cast<int64_t>(x->bytes());
        }
        for (auto const * x : {&gkind, &ginl, &binst})
        {
            b += static_cast<int64_t>(x->bytes());
        }
        return b;
    }
};

} /* end namespace scamd */

using scamd::DArena;
using scamd::DEuler;
using scamd::DMesh;
using scamd::guard;

extern "C"
{

int sceuler_create(void * mesh, double time_increment, int use_sfc, void ** out)
{
    return guard([&]() { *out = new DEuler(static_cast<DMesh *>(mesh), time_increment, use_sfc != 0); });
}

// Batched EulerCore over a disjoint-union mesh of ninst instances:
// inst_of_cell[ncell] gives the instance of every body cell (reference
// order), steps[ninst] the step count of every instance.
int sceuler_create_batch(void * mesh, int64_t ninst, int32_t const * inst_of_cell, int64_t const * steps, int use_sfc,
                         void ** out)
{
    return guard(
        [&]()
        {
            if (ninst < 1)
            {
                scamd::fail(scamd::SC_EVALUE, "EulerCore batch: ninst must be at least 1");
            }
            *out = new DEuler(static_cast<DMesh *>(mesh), 1.0, use_sfc != 0, ninst, inst_of_cell, steps);
        });
}

int sceuler_batch_set_dt(void * h, double const * dt)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            if (e.ninst == 0)
            {
                scamd::fail(scamd::SC_ERUNTIME, "EulerCore: not a batched solver");
            }
            e.batch_set_dt(dt);
        });
}

int sceuler_batch_cfl_dt(void * h, double cfl, double * dt_out)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            if (e.ninst == 0)
            {
                scamd::fail(scamd::SC_ERUNTIME, "EulerCore: not a batched solver");
            }
            e.batch_cfl_dt(cfl, dt_out);
        });
}

// dt that puts the largest calc_cfl number of the current so0n at cfl (no
// reference counterpart: the reference takes dt from its caller).
int sceuler_cfl_dt(void * h, double cfl, double * dt)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            double const v = e.cfl_max();
            if (!(v > 0.0))
            {
                scamd::fail(scamd::SC_EVALUE, "EulerCore: no positive wave speed");
            }
            *dt = cfl / v;
        });
}

int sceuler_march_batch(void * h, double * ms_total)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            if (e.ninst == 0)
            {
                scamd::fail(scamd::SC_ERUNTIME, "EulerCore: not a batched solver");
            }
            hipEvent_t a = nullptr;
            hipEvent_t b = nullptr;
            auto drop_done = [&]()
            {
                for (auto ev : e.bdone_ev)
                {
                    (void)hipEventDestroy(ev);
                }
                e.bdone_ev.clear();
                e.bdone_rng.clear();
            };
            drop_done();
            e.bdone_ms.assign(static_cast<size_t>(e.ninst), -1.0);
            e.btime = ms_total != nullptr;
            if (ms_total != nullptr)
            {
                SC_HIP(hipEventCreate(&a));
                SC_HIP(hipEventCreate(&b));
                SC_HIP(hipEventRecord(a, e.stream));
            }
            e.march_batch();
            e.btime = false;
            if (ms_total != nullptr)
            {
                SC_HIP(hipEventRecord(b, e.stream));
            }
            e.sync();
            if (ms_total != nullptr)
            {
                float ms = 0.0f;
                SC_HIP(hipEventElapsedTime(&ms, a, b));
                *ms_total = ms;
                for (size_t i = 0; i < e.bdone_ev.size(); ++i)
                {
                    float t = 0.0f;
                    SC_HIP(hipEventElapsedTime(&t, a, e.bdone_ev[i]));
                    for (int64_t r = e.bdone_rng[i].first; r < e.bdone_rng[i].second; ++r)
                    {
                        e.bdone_ms[static_cast<size_t>(r)] = t;
                    }
                }
                drop_done();
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
        });
}

int sceuler_batch_done_ms(void * h, double * out, int64_t n)
{
    return guard(
        [&]()
        {
            DEuler const & e = *static_cast<DEuler const *>(h);
            if (e.ninst == 0)
            {
                scamd::fail(scamd::SC_ERUNTIME, "EulerCore: not a batched solver");
            }
            if (n != e.ninst)
            {
                scamd::fail(scamd::SC_EVALUE, "EulerCore batch: completion times need one slot per instance");
            }
            if (e.bdone_ms.size() != static_cast<size_t>(e.ninst) || e.bdone_ms[0] < 0.0)
            {
                scamd::fail(scamd::SC_ERUNTIME, "EulerCore batch: no timed march_batch yet");
            }
            for (int64_t k = 0; k < n; ++k)
            {
                out[k] = e.bdone_ms[static_cast<size_t>(e.brank_of[static_cast<size_t>(k)])];
            }
        });
}

void sceuler_destroy(void * h) { delete static_cast<DEuler *>(h); }

// ndim ncell ngstcell neq
int sceuler_counts(void * h, int64_t * out)
{
    DEuler const & e = *static_cast<DEuler *>(h);
    scamd::AllocScope const alloc_scope_(e.stream);
    out[0] = e.ndim;
    out[1] = e.N;
    out[2] = e.G;
    out[3] = e.neq;
    return 0;
}

// which: 0 sigma0, 1 taumin, 2 tauscale, 3 time_increment, 4 1 once the
// kernel forms are tuned (else -1), 5 the batch chunk size in cells, 6 1 when
// whole steps replay a captured hipGraph (else 0), 100 + t the form code of
// cell type t
double sceuler_get_param(void * h, int which)
{
    DEuler const & e = *static_cast<DEuler *>(h);
    scamd::AllocScope const alloc_scope_(e.stream);
    switch (which)
    {
    case 0: return e.sigma0;
    case 1: return e.taumin;
    case 2: return e.tauscale;
    case 4: return e.tuned != 0 ? 1.0 : -1.0;
    case 5: return static_cast<double>(e.chunk_cells);
    case 6: return static_cast<double>(e.use_graph);
    case 7:
    {
        unsigned long long n = 0;
        if (hipMemcpyAsync(&n, e.redo.ptr + 1, sizeof(n), hipMemcpyDeviceToHost, e.stream) != hipSuccess ||
            hipStreamSynchronize(e.stream) != hipSuccess)
        {
            return -1.0;
        }
        return static_cast<double>(n);
    }
    default:
        if (which >= 100 && which < 116)
        {
            return static_cast<double>(e.form_code(which - 100));
        }
        return e.dt;
    }
}

int sceuler_set_param(void * h, int which, double v)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            switch (which)
            {
            case 0: e.sigma0 = v; break;
            case 1: e.taumin = v; break;
            case 2: e.tauscale = v; break;
            case 3:
                if (e.ninst > 0)
                {
                    scamd::fail(scamd::SC_ERUNTIME, "EulerCore: a batched solver sets dt per instance");
                }
                e.dt = v;
                break;
            case 5:
                if (v < 0.0)
                {
                    scamd::fail(scamd::SC_EVALUE, "EulerCore: the batch chunk size must be >= 0 cells");
                }
                e.chunk_cells = static_cast<int64_t>(v);
                break;
            case 6: e.use_graph = v != 0.0 ? 1 : 0; break;
            default: scamd::fail(scamd::SC_EVALUE, "EulerCore: read-only or unknown parameter");
            }
        });
}

// info: rows, width (elements per row), nghost, rank
int sceuler_table_info(void * h, int id, int64_t * info)
{
    return guard(
        [&]()
        {
            int64_t rows, w, ngh;
            int rank;
            static_cast<DEuler *>(h)->table(id, rows, w, ngh, rank);
            info[0] = rows;
            info[1] = w;
            info[2] = ngh;
            info[3] = rank;
        });
}

int sceuler_download(void * h, int id, double * host, int64_t count)
{
    return guard([&]() { static_cast<DEuler *>(h)->copy(id, host, count, true); });
}

int sceuler_upload(void * h, int id, double const * host, int64_t count)
{
    return guard([&]() { static_cast<DEuler *>(h)->copy(id, const_cast<double *>(host), count, false); });
}

// stage: 0 prepare_ce, 1 update, 2 calc_solt, 3 calc_soln, 4 calc_cfl,
// 5 calc_dsoln, 6 bc_soln, 7 bc_dsoln, 8 march_substep, 9 clear_bc
int sceuler_stage(void * h, int stage)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            if (e.ninst > 0 && (stage == 3 || stage == 4 || stage == 5 || stage == 8))
            {
                scamd::fail(scamd::SC_ERUNTIME, "EulerCore: a batched solver marches with march_batch");
            }
            switch (stage)
            {
            case 0: e.prepare_ce(); break;
            case 1: e.update(); break;
            case 2: e.calc_solt(); break;
            case 3: e.run_cells<scamd::M_SOLN | scamd::M_STMEM>(); break;
            case 4: e.run_cells<scamd::M_CFL>(); break;
            case 5: e.run_cells<scamd::M_DSOLN | scamd::M_STMEM>(); break;
            case 6: e.bc_soln(); break;
            case 7: e.bc_dsoln(); break;
            case 8: e.march_substep(); break;
            case 9:
                e.bcs.clear();
                e.refresh_bc();
                break;
            default: scamd::fail(scamd::SC_EVALUE, "EulerCore: unknown stage");
            }
            e.sync();
        });
}

int sceuler_march(void * h, int64_t steps)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            if (e.ninst > 0)
            {
                scamd::fail(scamd::SC_ERUNTIME, "EulerCore: a batched solver marches with march_batch");
            }
            e.march(steps);
            e.sync();
        });
}

// Enqueue a march without waiting for it.
int sceuler_march_async(void * h, int64_t steps)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            if (e.ninst > 0)
            {
                scamd::fail(scamd::SC_ERUNTIME, "EulerCore: a batched solver marches with march_batch");
            }
            e.march(steps);
        });
}

int sceuler_sync(void * h)
{
    return guard([&]() { static_cast<DEuler *>(h)->sync(); });
}

// *done = 1 when all work enqueued on the solver's stream has finished,
// 0 while some is running (no reference counterpart; for request loops).
int sceuler_query(void * h, int * done)
{
    return guard(
        [&]()
        {
            hipError_t const e = hipStreamQuery(static_cast<DEuler *>(h)->stream);
            if (e == hipErrorNotReady)
            {
                *done = 0;
                return;
            }
            SC_HIP(e);
            *done = 1;
        });
}

// March `steps` steps and return the average duration, in milliseconds, of
// one substep's fused kernels (all type groups) and the duration of the
// whole march, from HIP events recorded on the solver's stream around every
// launch; the host waits only once, at the end.
int sceuler_march_timed(void * h, int64_t steps, double * ms_substep, double * ms_total)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            if (e.ninst > 0)
            {
                scamd::fail(scamd::SC_ERUNTIME, "EulerCore: a batched solver marches with march_batch");
            }
            int64_t const nsub = 2 * steps;
            std::vector<hipEvent_t> ev(static_cast<size_t>(2 * nsub + 2));
            for (auto & x : ev)
            {
                SC_HIP(hipEventCreate(&x));
            }
            SC_HIP(hipEventRecord(ev[2 * nsub], e.stream));
            int64_t counted = 0;
            e.chain_ready = false;
            e.start_chain();
            for (int64_t it = 0; it < nsub; ++it)
            {
                if (e.tuned == 0)
                {
                    e.substep_fused(); // the tuning substep, not counted
                    e.start_chain();
                    SC_HIP(hipEventRecord(ev[2 * it], e.stream));
                    SC_HIP(hipEventRecord(ev[2 * it + 1], e.stream));
                    continue;
                }
                SC_HIP(hipEventRecord(ev[2 * it], e.stream));
                e.substep_fused();
                SC_HIP(hipEventRecord(ev[2 * it + 1], e.stream));
                ++counted;
            }
            if (steps > 0)
            {
                e.calc_solt();
            }
            SC_HIP(hipEventRecord(ev[2 * nsub + 1], e.stream));
            SC_HIP(hipEventSynchronize(ev[2 * nsub + 1]));
            double sub = 0.0;
            for (int64_t it = 0; it < nsub; ++it)
            {
                float ms = 0.0f;
                SC_HIP(hipEventElapsedTime(&ms, ev[2 * it], ev[2 * it + 1]));
                sub += ms;
            }
            float tot = 0.0f;
            SC_HIP(hipEventElapsedTime(&tot, ev[2 * nsub], ev[2 * nsub + 1]));
            *ms_substep = counted > 0 ? sub / static_cast<double>(counted) : 0.0;
            *ms_total = tot;
            for (auto & x : ev)
            {
                (void)hipEventDestroy(x);
            }
        });
}

int sceuler_init_solution(void * h, double gamma, double rho, double const * v, int nv, double p)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            if (nv != e.ndim)
            {
                scamd::fail(scamd::SC_EVALUE, "EulerCore::init_solution: velocity size " + std::to_string(nv)
                                                  + " must equal ndim " + std::to_string(e.ndim));
            }
            double vel[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < nv; ++d)
            {
                vel[d] = v[d];
            }
            e.init_solution(gamma, rho, vel, p);
        });
}

int sceuler_add_bc(void * h, int kind, int32_t const * faces, int64_t n, double const * value, int64_t nv)
{
    return guard(
        [&]()
        {
            std::vector<int32_t> f(faces, faces + n);
            std::vector<double> v(value, value + nv);
            static_cast<DEuler *>(h)->add_bc(kind, f, v);
        });
}

// get_normal_matrix (euler_bc.cpp:393-425) of face ifc into out[ndim*ndim].
int sceuler_normal_matrix(void * h, int32_t ifc, double * out)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            if (ifc < 0 || ifc >= e.mesh->nface)
            {
                scamd::fail(scamd::SC_EINDEX, "EulerCore: face " + std::to_string(ifc) + " out of range");
            }
            scamd::DevBuf<double> mat;
            mat.alloc(static_cast<int64_t>(e.ndim) * e.ndim);
            scamd::k_face_frame<<<1, 64, 0, e.stream>>>(e.ndim, ifc, e.mesh->fcnml_b(), mat.ptr);
            SC_HIP(hipGetLastError());
            SC_HIP(hipMemcpyAsync(out, mat.ptr, mat.bytes(), hipMemcpyDeviceToHost, e.stream));
            e.sync();
        });
}

// Diagnostics: out = {sum(rho * clvol) over body cells, min cflo, max cflo}.
int sceuler_reduce(void * h, double * out)
{
    return guard(
        [&]()
        {
            DEuler & e = *static_cast<DEuler *>(h);
            scamd::AllocScope const alloc_scope_(e.stream);
            constexpr int NB = 256;
            scamd::DevBuf<double> part;
            part.alloc(NB * 3);
            scamd::k_reduce<<<NB, scamd::BLK, 0, e.stream>>>(e.R, e.G, e.neq, e.s0[1 - e.cur].ptr, e.mesh->clvol_b(),
                                                             e.ref_of_int.ptr, e.cflo.ptr, part.ptr);
            SC_HIP(hipGetLastError());
            std::vector<double> hp(NB * 3);
            SC_HIP(hipMemcpyAsync(hp.data(), part.ptr, part.bytes(), hipMemcpyDeviceToHost, e.stream));
            e.sync();
            double s = 0.0;
            double lo = DBL_MAX;
            double hi = -DBL_MAX;
            for (int b = 0; b < NB; ++b)
            {
                s += hp[b * 3];
                lo = std::min(lo, hp[b * 3 + 1]);
                hi = std::max(hi, hp[b * 3 + 2]);
            }
            out[0] = s;
            out[1] = lo;
            out[2] = hi;
        });
}

// Device bytes a solver holds (its own tables; the mesh's are counted by
// scmesh_device_bytes).
int sceuler_device_bytes(void * h, int64_t * out)
{
    return guard(
        [&]()
        {
            auto & e = *static_cast<DEuler *>(h);
            int64_t b = 0;
            for (auto const & g : e.groups)
            {
                b += static_cast<int64_t>(g.nbr.bytes() + g.geo.bytes());
            }
            for (auto const * x : {&e.cen, &e.s0[0], &e.s0[1], &e.s1[0], &e.s1[1], &e.st[0], &e.st[1], &e.stm, &e.gam,
                                   &e.cflo, &e.cflc, &e.gmat, &e.inlet, &e.bhdt, &e.bqdt})
            {
                b += static_cast<int64_t>(x->bytes());
            }
            for (auto const * x : {&e.int_of_ref, &e.ref_of_int, &e.gkind, &e.ginl, &e.gsrc, &e.brank_ref, &e.binst})
            {
                b += static_cast<int64_t>(x->bytes());
            }
            *out = b;
        });
}

int scserve_create(int ndim, int64_t const * type_cap, int64_t ghost_cap, int64_t max_slots, void ** out)
{
    return guard([&]() { *out = new DArena(ndim, type_cap, ghost_cap, max_slots); });
}

void scserve_destroy(void * a) { delete static_cast<DArena *>(a); }

int scserve_admit(void * a, void * euler, int64_t steps, int64_t * slot)
{
    return guard([&]() { *slot = static_cast<DArena *>(a)->admit(*static_cast<DEuler *>(euler), steps); });
}

int scserve_step(void * a, int64_t steps)
{
    return guard(
        [&]()
        {
            if (steps < 0)
            {
                scamd::fail(scamd::SC_EVALUE, "serve arena: steps must be >= 0");
            }
            static_cast<DArena *>(a)->step(steps);
        });
}

int scserve_pending(void * a, int64_t * chunks)
{
    return guard([&]() { *chunks = static_cast<DArena *>(a)->pending(); });
}

int scserve_query(void * a, int64_t slot, int * done)
{
    return guard(
        [&]()
        {
            auto & ar = *static_cast<DArena *>(a);
            if (slot < 0 || slot >= ar.max_slots)
            {
                scamd::fail(scamd::SC_EINDEX, "serve arena: slot out of range");
            }
            *done = ar.query(slot);
        });
}

int scserve_result(void * a, int64_t slot, double * so0, double * so1)
{
    return guard(
        [&]()
        {
            auto & ar = *static_cast<DArena *>(a);
            if (slot < 0 || slot >= ar.max_slots)
            {
                scamd::fail(scamd::SC_EINDEX, "serve arena: slot out of range");
            }
            ar.result(slot, so0, so1);
        });
}

int scserve_sync(void * a)
{
    return guard([&]() { static_cast<DArena *>(a)->sync(); });
}

// out[0] used slots, [1] marching slots, [2] device bytes, [3] cell
// capacity, [4] ghost capacity, [5] free cells, [6] free ghost rows,
// [7] slot capacity.
int scserve_info(void * a, int64_t * out)
{
    return guard(
        [&]()
        {
            auto & ar = *static_cast<DArena *>(a);
            int64_t used = 0;
            int64_t marching = 0;
            for (auto const & sl : ar.slots)
            {
                used += sl.used ? 1 : 0;
                marching += sl.marching ? 1 : 0;
            }
            int64_t fc = 0;
            for (auto const & g : ar.groups)
            {
                fc += g.ra.free_total();
            }
            out[0] = used;
            out[1] = marching;
            out[2] = ar.device_bytes();
            out[3] = ar.Ncap;
            out[4] = ar.Gcap;
            out[5] = fc;
            out[6] = ar.ghosts.free_total();
            out[7] = ar.max_slots;
        });
}

} /* end extern "C" */
