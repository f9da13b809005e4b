// v2c_search.hpp -- cell search on a non-uniform axis (interpolation.c:1536-1573).
//
// The reference bisects the gridline array.  Its result is the number of
// interior gridlines g[1..n-2] not greater than x; NaN lands in the last cell
// because every comparison with it fails.  The same count is found here from a
// table of equal-width bins and a short walk.  Each bin stores the count at
// its left edge together with the gridlines g[s-1..s+2] around that cell; the
// bin width is at most the smallest spacing, so at most one gridline lies
// inside a bin and the walk takes 0-2 steps.  In the common case one load of
// the bin gives both the cell and every gridline the interpolation stencils
// need, instead of log2(n) dependent loads and a second round of gridline
// loads.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <vector>

namespace v2c {

// Cell s of an axis with its gridlines g[k] = grid[s-1+k], k = 0..3 (entries
// past either end of the axis hold NaN; no stencil reads them),
// inv = 1.0 / (grid[s+1] - grid[s]) and the reciprocals of the two
// pressure-node half-spacings a stencil in this cell divides by:
// rc[d] = 1.0 / ((g[d+2] - g[d]) / 2.0), d = 0 below the cell midline, 1 above
// (each rounded once, from the value the reference's expression rounds to).
// 64 bytes: one cache line.
struct Axis4 { double g[4]; double inv; double rc[2]; int s, pad; };

// a / b rounded to nearest, computed from rb = RN(1/b): y = RN(a rb) is within
// one ulp of a/b, the residual a - b y is exact in one fused multiply-add, and
// RN(y + rb (a - b y)) is the correctly rounded quotient (Markstein's
// correction theorem; it holds for normal a, b with a/b away from overflow and
// underflow, which covers every use here: cell-local coordinates and stencil
// weights, |a/b| <= 2 or a == 0).  The result equals the IEEE division a / b
// bit for bit, at the cost of one multiply and two FMAs instead of a
// division sequence.
__host__ __device__ inline double div_rcp(double a, double b, double rb)
{
    const double y = a * rb;
    const double e = fma(-b, y, a);
    return fma(e, rb, y);
}

__host__ __device__ inline double pick(const Axis4 &A, int m)
{
    const int k = m - A.s + 1;
    return k <= 0 ? A.g[0] : k == 1 ? A.g[1] : k == 2 ? A.g[2] : A.g[3];
}

__host__ __device__ inline Axis4 axis_cell(const double *__restrict__ grid, int n, int s)
{
    Axis4 a;
    a.s = s;
    a.pad = 0;
    for (int k = 0; k < 4; k++) {
        const int i = s - 1 + k;
        a.g[k] = (i >= 0 && i < n) ? grid[i] : NAN;
    }
    a.inv = 1.0 / (grid[s + 1] - grid[s]);
    for (int d = 0; d < 2; d++) a.rc[d] = 1.0 / ((a.g[d + 2] - a.g[d]) / 2.0);
    return a;
}

// bins[j]: the cell at the bin's left edge.  cells[s]: cell s (0 <= s <= n-2).
__host__ __device__ inline Axis4 axis_find(const double *__restrict__ grid, int n, const Axis4 *__restrict__ bins,
                                           const Axis4 *__restrict__ cells, int nb, double b0, double binv,
                                           double x)
{
    if (x != x) return cells[n - 2];
    const double t = (x - b0) * binv;
    const int j = t > 0 ? (t < (double)(nb - 1) ? (int)t : nb - 1) : 0;
    const Axis4 e = bins[j];
    int s = e.s;
    const bool up = s < n - 2 && !(e.g[2] > x), down = s > 0 && e.g[1] > x;
    if (!up && !down) return e;
    double lo = e.g[1], hi = e.g[2];
    while (s < n - 2 && !(hi > x)) { s++; lo = hi; hi = grid[s + 1]; }
    while (s > 0 && lo > x) { s--; hi = lo; lo = grid[s]; }
    return cells[s];
}

// The reference's bisection, used to fill the tables.
inline int host_bisect(const double *g, int n, double x)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (g[mid] > x) hi = mid; else lo = mid;
    }
    return lo > n - 2 ? n - 2 : lo;
}

// Tables of one strictly increasing axis: bin width = the smallest spacing,
// at most 2^22 bins (wider bins on extreme grids only lengthen the walk).
inline void make_tables(const double *g, int n, std::vector<Axis4> &bins, std::vector<Axis4> &cells,
                        double &b0, double &binv)
{
    double hmin = g[1] - g[0];
    for (int i = 1; i + 1 < n; i++) hmin = std::min(hmin, g[i + 1] - g[i]);
    const double range = g[n - 1] - g[0];
    const long cap = 1L << 22;
    double h = hmin;
    if (range / h + 2 > (double)cap) h = range / (double)(cap - 2);
    const long nb = (long)(range / h) + 2;
    cells.resize((size_t)(n - 1));
    for (int s = 0; s < n - 1; s++) cells[(size_t)s] = axis_cell(g, n, s);
    bins.resize((size_t)nb);
    for (long j = 0; j < nb; j++) bins[(size_t)j] = cells[(size_t)host_bisect(g, n, g[0] + (double)j * h)];
    b0 = g[0];
    binv = 1.0 / h;
}

}  // namespace v2c
