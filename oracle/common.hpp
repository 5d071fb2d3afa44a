#pragma once
/*
 * Oracle (test infrastructure only): a CPU restatement of solvcon's CESE
 * Euler march, used by tests/ and the CPU-baseline leg of bench.py as the
 * checker.  Nothing in the product library links or calls this code.
 *
 * This header holds the ghost-aware row-major table used by the restatement.
 * It follows the indexing of the reference SimpleArray: a table of shape
 * [nghost + nbody, ...] stores its ghost rows first, and row index -k
 * addresses ghost row k (cpp/solvcon/buffer/SimpleArray.hpp:2241-2271,
 * 2354-2362).
 */

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace orc
{

template <typename T>
struct Table
{
    std::vector<T> data;
    std::vector<ptrdiff_t> shape; // full shape, shape[0] includes the ghosts
    ptrdiff_t nghost = 0;
    ptrdiff_t s0 = 0; // row stride
    ptrdiff_t s1 = 0; // stride of axis 1 (rank 3 only)

    Table() = default;
    Table(std::vector<ptrdiff_t> shp, T init, ptrdiff_t ngh = 0) { remake(std::move(shp), init, ngh); }

    void remake(std::vector<ptrdiff_t> shp, T init, ptrdiff_t ngh = 0)
    {
        shape = std::move(shp);
        nghost = ngh;
        size_t n = 1;
        for (ptrdiff_t s : shape)
        {
            n *= static_cast<size_t>(s);
        }
        data.assign(n, init);
        s0 = 1;
        for (size_t i = 1; i < shape.size(); ++i)
        {
            s0 *= shape[i];
        }
        s1 = shape.size() > 2 ? shape[2] : 1;
    }

    ptrdiff_t nbody() const { return shape.empty() ? 0 : shape[0] - nghost; }
    size_t size() const { return data.size(); }

    T & operator()(ptrdiff_t i) { return data[static_cast<size_t>((i + nghost) * s0)]; }
    T const & operator()(ptrdiff_t i) const { return data[static_cast<size_t>((i + nghost) * s0)]; }
    T & operator()(ptrdiff_t i, ptrdiff_t j) { return data[static_cast<size_t>((i + nghost) * s0 + j * s1)]; }
    T const & operator()(ptrdiff_t i, ptrdiff_t j) const { return data[static_cast<size_t>((i + nghost) * s0 + j * s1)]; }
    T & operator()(ptrdiff_t i, ptrdiff_t j, ptrdiff_t k) { return data[static_cast<size_t>((i + nghost) * s0 + j * s1 + k)]; }
    T const & operator()(ptrdiff_t i, ptrdiff_t j, ptrdiff_t k) const { return data[static_cast<size_t>((i + nghost) * s0 + j * s1 + k)]; }
};

// Rank-2 tables index (i, j) as i*s0 + j; fix s1 for that case.
template <typename T>
inline Table<T> make2(ptrdiff_t n0, ptrdiff_t n1, T init, ptrdiff_t ngh = 0)
{
    Table<T> t({n0, n1}, init, ngh);
    t.s1 = 1;
    return t;
}

template <typename T>
inline Table<T> make1(ptrdiff_t n0, T init, ptrdiff_t ngh = 0)
{
    Table<T> t({n0}, init, ngh);
    t.s1 = 0;
    return t;
}

template <typename T>
inline Table<T> make3(ptrdiff_t n0, ptrdiff_t n1, ptrdiff_t n2, T init, ptrdiff_t ngh = 0)
{
    Table<T> t({n0, n1, n2}, init, ngh);
    return t;
}

// Error kinds shared by the C entry points; they mirror the exception
// classes the reference raises (ValueError, IndexError, RuntimeError).
enum ErrKind : int
{
    ERR_OK = 0,
    ERR_VALUE = 1,
    ERR_INDEX = 2,
    ERR_RUNTIME = 3,
};

void set_error(int kind, char const * msg);

} /* end namespace orc */
