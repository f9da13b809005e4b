// v2c_search.hpp -- cell search on a non-uniform axis (interpolation.c:1536-1573).
//
// The reference bisects the gridline array.  Its result is the number of
// interior gridlines g[1..n-2] not greater than x; NaN lands in the last cell
// because every comparison with it fails.  The same count is found here from a
// table of equal-width bins and a short walk: each bin stores the count at its
// left edge and the two gridlines bounding that cell; the bin width is at most
// the smallest spacing, so at most one gridline lies inside a bin and the walk
// takes 0-2 steps (one dependent load each) instead of log2(n).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <vector>

namespace v2c {

struct AxisBin { double glo, ghi; int s, pad; };

__host__ __device__ inline int axis_find(const double *__restrict__ g, int n, const AxisBin *__restrict__ bins,
                                         int nb, double b0, double binv, double x)
{
    if (x != x) return n - 2;
    const double t = (x - b0) * binv;
    const int j = t > 0 ? (t < (double)(nb - 1) ? (int)t : nb - 1) : 0;
    const AxisBin e = bins[j];
    int s = e.s;
    double lo = e.glo, hi = e.ghi;
    while (s < n - 2 && !(hi > x)) { s++; lo = hi; hi = g[s + 1]; }
    while (s > 0 && lo > x) { s--; hi = lo; lo = g[s]; }
    return s;
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

// Table of one strictly increasing axis: bin width = the smallest spacing,
// at most 2^22 bins (wider bins on extreme grids only lengthen the walk).
inline std::vector<AxisBin> make_bins(const double *g, int n, double &b0, double &binv)
{
    double hmin = g[1] - g[0];
    for (int i = 1; i + 1 < n; i++) hmin = std::min(hmin, g[i + 1] - g[i]);
    const double range = g[n - 1] - g[0];
    const long cap = 1L << 22;
    double h = hmin;
    if (range / h + 2 > (double)cap) h = range / (double)(cap - 2);
    const long nb = (long)(range / h) + 2;
    std::vector<AxisBin> b((size_t)nb);
    for (long j = 0; j < nb; j++) {
        const int s = host_bisect(g, n, g[0] + (double)j * h);
        b[(size_t)j] = AxisBin{g[s], g[s + 1], s, 0};
    }
    b0 = g[0];
    binv = 1.0 / h;
    return b;
}

}  // namespace v2c
