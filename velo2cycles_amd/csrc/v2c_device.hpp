// v2c_device.hpp -- device-side pieces of the per-marker hot path (gfx950).
//
// Every routine restates the reference formula named in its comment with the
// reference's operation order.  The library is compiled with
// -ffp-contract=off, so the arithmetic that decides cell indices and marker
// positions (searches, weights, bilinear sums, position updates) rounds exactly
// like the reference's gcc/x86-64 build; divisions by gridline spacings use the
// cells' precomputed reciprocals through div_rcp, which returns the IEEE
// quotient bit for bit.  Only libm calls (exp, log, pow, sin, ...) may differ,
// by a few ulp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/v2c.h"
#include "v2c_search.hpp"
#include "v2c_vmath.hpp"

namespace v2c {

// Marker SoA field slots (same order as the v2c_markers struct).
enum MF {
    MF_X = 0, MF_Y, MF_K, MF_P, MF_XX, MF_XY, MF_EXX, MF_EXY, MF_E, MF_PHI, MF_VX, MF_VY,
    MF_C, MF_D, MF_W, MF_A, MF_B, MF_L, MF_L0, MF_V0, MF_FS, MF_C0, NMF
};

// Extension fields of the per-marker local update (not reference state): the
// fault marker's own RSF parameters and state, the solver outputs and the
// warm-start guesses.
enum MX {
    MX_V = 0, MX_PHI, MX_A, MX_B, MX_L, MX_V0, MX_MU0,       // RSF inputs / state
    MX_ETA, MX_ETAVP, MX_SY,                                  // outputs
    MX_GETA, MX_GV,                                           // warm-start guesses
    MX_ROOT_ETA, MX_ROOT_V,                                   // raw Newton roots
    NMX
};

struct DevMat {
    double nu[V2C_MAX_TYPES], dh[V2C_MAX_TYPES], dv[V2C_MAX_TYPES], ss[V2C_MAX_TYPES];
    double mm[V2C_MAX_TYPES], gg[V2C_MAX_TYPES], n0[V2C_MAX_TYPES], n1[V2C_MAX_TYPES];
    double ro[V2C_MAX_TYPES], bb[V2C_MAX_TYPES], aa[V2C_MAX_TYPES], cp[V2C_MAX_TYPES];
    double kt[V2C_MAX_TYPES], kf[V2C_MAX_TYPES], kp[V2C_MAX_TYPES], ht[V2C_MAX_TYPES];
    double ssn1[V2C_MAX_TYPES];     // ss^(mm-1), computed once on the host
    // per-type quotients of the viscosity laws, computed once on the host
    // (IEEE division there and here round alike): nu/ss^(mm-1), 1/mm, (mm-1)/mm
    double k1[V2C_MAX_TYPES], rn[V2C_MAX_TYPES], ex1[V2C_MAX_TYPES];
    int im[V2C_MAX_TYPES];
};

// Every pointer in the argument structs points to global memory.  Saying so
// in the type (device pass only; the host sees plain pointers of the same
// size) keeps global loads with a scalar base even when a struct is itself
// read from memory, as the batched ensemble launches do; through a plain
// pointer loaded from memory the compiler would have to emit flat loads.
#ifdef __HIP_DEVICE_COMPILE__
#define V2C_GLOBAL __attribute__((address_space(1)))
#else
#define V2C_GLOBAL
#endif

struct DevGrid {
    V2C_GLOBAL const double *gx, *gy;
    int xnumx, ynumy;
    double xsize, ysize;
    // search tables: bins of width h <= the smallest spacing from the first
    // gridline (bin j = (x - b0) * binv) and per-cell records
    V2C_GLOBAL const Axis4 *bx, *by, *cellx, *celly;
    double bx0, by0, bxinv, byinv;
    int nbx, nby;
    // allinteri's outer half-cell bounds (interpolation.c:1701-1705):
    // (gx[0]+gx[1])/2, (gx[n-2]+gx[n-1])/2 and the same in y
    double xr_lo, xr_hi, yr_lo, yr_hi;
};

// Everything a kernel needs, passed by value.
struct KArgs {
    DevGrid g;
    v2c_params p;
    V2C_GLOBAL const DevMat *mat;
    V2C_GLOBAL double *mk; // NMF fields x cap
    long cap;              // field stride
    V2C_GLOBAL int8_t *t;  // marker types
    long n;                // marker count
    V2C_GLOBAL const uint32_t *order;      // sorted position -> marker id
    V2C_GLOBAL const uint32_t *skey;       // sorted position -> sub-cell key
    V2C_GLOBAL const uint32_t *sub_start;  // sub-cell key -> first sorted position (4*ncell+1)
    V2C_GLOBAL uint32_t *key;              // per marker (physical) sub-cell key
    V2C_GLOBAL double *nod;                // V2C_NF fields x nodenum
    long nodenum;
    long ncell;
    V2C_GLOBAL double *part;               // node records of the scatter passes
    V2C_GLOBAL double *prop;               // per-marker properties of marker2nodes (NPROP x cap)
    double timesum, timestep;
    V2C_GLOBAL int *counters;              // device counters (errors, unsupported types)
    int write_friction;    // write the calc_frictionpara marker outputs
    V2C_GLOBAL uint8_t *fclass;            // per marker: which constant friction set it holds (0: none)
    V2C_GLOBAL double *mx;                 // NMX extension fields x cap
    V2C_GLOBAL const uint8_t *fault;       // per-marker fault flag
    V2C_GLOBAL unsigned long long *lu_stats;  // local-update statistics
    V2C_GLOBAL const uint32_t *fault_list;  // the fault markers' ids (nfault)
    long nfault;
    V2C_GLOBAL const float *mlp, *mlp_norm; // warm-start network: FP32 parameters, normalisation
};

__device__ __forceinline__ double *NOD(const KArgs &a, int f) { return a.nod + (long)f * a.nodenum; }
__device__ __forceinline__ double &MK(const KArgs &a, int f, long i) { return a.mk[(long)f * a.cap + i]; }
__device__ __forceinline__ double &MXF(const KArgs &a, int f, long i) { return a.mx[(long)f * a.cap + i]; }

// ---- small helpers ---------------------------------------------------------
__device__ __forceinline__ double ref_abs(double a) { return a >= 0 ? a : -a; }
__device__ __forceinline__ double ref_minv(double a, double b) { return a <= b ? a : b; }
__device__ __forceinline__ double clampd(double v, double lo, double hi)
{
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

__device__ __forceinline__ Axis4 find_x(const DevGrid &g, double x)
{
    return axis_find(g.gx, g.xnumx, g.bx, g.cellx, g.nbx, g.bx0, g.bxinv, x);
}
__device__ __forceinline__ Axis4 find_y(const DevGrid &g, double y)
{
    return axis_find(g.gy, g.ynumy, g.by, g.celly, g.nby, g.by0, g.byinv, y);
}

// m3: nodal index (nodenum < 2^29, checked at create).
struct Stencil { uint32_t m3; double e, n; };


// Stencils at (x, y) in cell (X.s, Y.s).  Every gridline they use lies in
// X.g / Y.g: after the reference's clamps the stencil's first gridline is
// g[d] with d in {0,1} (three-point stencils) or {0,1,2} (two-point), so a
// couple of selects replace gridline loads.
__device__ __forceinline__ void sel2(const Axis4 &A, int d, double &a, double &b)
{
    a = d == 0 ? A.g[0] : (d == 1 ? A.g[1] : A.g[2]);
    b = d == 0 ? A.g[1] : (d == 1 ? A.g[2] : A.g[3]);
}
__device__ __forceinline__ void sel3(const Axis4 &A, int d, double &a, double &b, double &c)
{
    a = d ? A.g[1] : A.g[0];
    b = d ? A.g[2] : A.g[1];
    c = d ? A.g[3] : A.g[2];
}

// Pressure-node stencil (interpolation.c:1942-1948, 1216-1224, 1723-1730).
__device__ __forceinline__ Stencil centre_stencil(const DevGrid &g, const Axis4 &X, const Axis4 &Y,
                                                  double x, double y)
{
    int m10 = X.s, m20 = Y.s;
    if (x > (X.g[1] + X.g[2]) / 2.0) m10++;
    if (y > (Y.g[1] + Y.g[2]) / 2.0) m20++;
    if (m10 < 1) m10 = 1;
    if (m10 > g.xnumx - 2) m10 = g.xnumx - 2;
    if (m20 < 1) m20 = 1;
    if (m20 > g.ynumy - 2) m20 = g.ynumy - 2;
    double a0, a1, a2, b0, b1, b2;            // grid[m-1], grid[m], grid[m+1]
    // after the clamps m - s is 0 or 1 on both axes (a cell's own midline
    // decides), so the half-spacing's reciprocal is in the cell record
    const int dx = m10 - X.s, dy = m20 - Y.s;
    sel3(X, dx, a0, a1, a2);
    sel3(Y, dy, b0, b1, b2);
    Stencil s;
    s.e = div_rcp(x - (a0 + a1) / 2.0, (a2 - a0) / 2.0, dx ? X.rc[1] : X.rc[0]);
    s.n = div_rcp(y - (b0 + b1) / 2.0, (b2 - b0) / 2.0, dy ? Y.rc[1] : Y.rc[0]);
    s.m3 = (uint32_t)(m10 * g.ynumy + m20);
    return s;
}

// Basic-node stencil without ghost nodes (interpolation.c:1069-1083, 1711-1717).
__device__ __forceinline__ Stencil basic_stencil(const DevGrid &g, const Axis4 &X, const Axis4 &Y,
                                                 double x, double y)
{
    int m10 = X.s, m20 = Y.s;
    if (m10 < 1) m10 = 1;
    if (m10 > g.xnumx - 3) m10 = g.xnumx - 3;
    if (m20 < 1) m20 = 1;
    if (m20 > g.ynumy - 3) m20 = g.ynumy - 3;
    double a0, a1, b0, b1;                    // grid[m], grid[m+1]
    sel2(X, m10 - X.s + 1, a0, a1);
    sel2(Y, m20 - Y.s + 1, b0, b1);
    Stencil s;
    // the cell's own spacing unless a clamp moved the stencil (edge cells)
    s.e = m10 == X.s ? div_rcp(x - a0, a1 - a0, X.inv) : (x - a0) / (a1 - a0);
    s.n = m20 == Y.s ? div_rcp(y - b0, b1 - b0, Y.inv) : (y - b0) / (b1 - b0);
    s.m3 = (uint32_t)(m10 * g.ynumy + m20);
    return s;
}

// vx-node stencil (interpolation.c:1377-1388, 1738-1744).  The x clamp of
// allinteri is a no-op: the searched cell is already in [0, xnumx-2].
__device__ __forceinline__ Stencil vx_stencil(const DevGrid &g, const Axis4 &X, const Axis4 &Y,
                                              double x, double y)
{
    int m20 = Y.s;
    if (y < (Y.g[1] + Y.g[2]) / 2.0) m20 -= 1;
    if (m20 < 0) m20 = 0;
    if (m20 > g.ynumy - 3) m20 = g.ynumy - 3;
    double b0, b1, b2;                        // grid[m], grid[m+1], grid[m+2]
    const int dy = m20 - Y.s + 1;             // 0 or 1 after the clamps
    sel3(Y, dy, b0, b1, b2);
    Stencil s;
    s.e = div_rcp(x - X.g[1], X.g[2] - X.g[1], X.inv);
    s.n = div_rcp(y - (b0 + b1) / 2.0, (b2 - b0) / 2.0, dy ? Y.rc[1] : Y.rc[0]);
    s.m3 = (uint32_t)(X.s * g.ynumy + m20);
    return s;
}

// vy-node stencil (interpolation.c:1434-1445, 1749-1755); the y clamp of
// allinteri is likewise a no-op.
__device__ __forceinline__ Stencil vy_stencil(const DevGrid &g, const Axis4 &X, const Axis4 &Y,
                                              double x, double y)
{
    int m10 = X.s;
    if (x < (X.g[1] + X.g[2]) / 2.0) m10 -= 1;
    if (m10 < 0) m10 = 0;
    if (m10 > g.xnumx - 3) m10 = g.xnumx - 3;
    double a0, a1, a2;
    const int dx = m10 - X.s + 1;             // 0 or 1 after the clamps
    sel3(X, dx, a0, a1, a2);
    Stencil s;
    s.e = div_rcp(x - (a0 + a1) / 2.0, (a2 - a0) / 2.0, dx ? X.rc[1] : X.rc[0]);
    s.n = div_rcp(y - Y.g[1], Y.g[2] - Y.g[1], Y.inv);
    s.m3 = (uint32_t)(m10 * g.ynumy + Y.s);
    return s;
}

// All four stencils of one point (basic interpolation.c:1069-1083, 1711-1717;
// centre :1216-1224, 1723-1730; vx :1377-1388, 1738-1744; vy :1434-1445,
// 1749-1755), each weight with the reference's formula and operands.  The x
// clamp of the vx stencil and the y clamp of the vy stencil are no-ops: the
// searched cell is already in range.
struct Stencil4 { Stencil b, c, vx, vy; };

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ Stencil4 stencils4_full(const DevGrid &g, const Axis4 &X, const Axis4 &Y,
                                                   double x, double y)
{
    const int nx = g.xnumx, ny = g.ynumy;
    const double xmid = (X.g[1] + X.g[2]) / 2.0, ymid = (Y.g[1] + Y.g[2]) / 2.0;
    const int bx = clampi(X.s, 1, nx - 3), by = clampi(Y.s, 1, ny - 3);
    const int cx = clampi(X.s + (x > xmid ? 1 : 0), 1, nx - 2), cy = clampi(Y.s + (y > ymid ? 1 : 0), 1, ny - 2);
    const int vyx = clampi(X.s - (x < xmid ? 1 : 0), 0, nx - 3), vxy = clampi(Y.s - (y < ymid ? 1 : 0), 0, ny - 3);
    double a0, a1, a2, b0, b1, b2;
    Stencil4 o;
    sel2(X, bx - X.s + 1, a0, a1);
    sel2(Y, by - Y.s + 1, b0, b1);
    o.b.e = (x - a0) / (a1 - a0);
    o.b.n = (y - b0) / (b1 - b0);
    o.b.m3 = (uint32_t)(bx * ny + by);
    sel3(X, cx - X.s, a0, a1, a2);
    sel3(Y, cy - Y.s, b0, b1, b2);
    o.c.e = (x - (a0 + a1) / 2.0) / ((a2 - a0) / 2.0);
    o.c.n = (y - (b0 + b1) / 2.0) / ((b2 - b0) / 2.0);
    o.c.m3 = (uint32_t)(cx * ny + cy);
    sel3(Y, vxy - Y.s + 1, b0, b1, b2);
    o.vx.e = (x - X.g[1]) / (X.g[2] - X.g[1]);
    o.vx.n = (y - (b0 + b1) / 2.0) / ((b2 - b0) / 2.0);
    o.vx.m3 = (uint32_t)(X.s * ny + vxy);
    sel3(X, vyx - X.s + 1, a0, a1, a2);
    o.vy.e = (x - (a0 + a1) / 2.0) / ((a2 - a0) / 2.0);
    o.vy.n = (y - Y.g[1]) / (Y.g[2] - Y.g[1]);
    o.vy.m3 = (uint32_t)(vyx * ny + Y.s);
    return o;
}

// Interior points off the cell midlines (all but a thin set of markers)
// share weights pairwise: the basic stencil's x weight is the vx one and its
// y weight the vy one (clamps inactive), and the centre and vy x-weights
// coincide (both stencils start at grid[s-1] below the midline and at
// grid[s] above it), as do the centre and vx y-weights.  Such points need 4
// divisions instead of 8 with the same operands, so the same results; the
// other lanes take the full computation, in a branch the wavefront skips when
// none needs it.
__device__ __forceinline__ Stencil4 stencils4(const DevGrid &g, const Axis4 &X, const Axis4 &Y, double x, double y)
{
    const int ny = g.ynumy;
    const double xmid = (X.g[1] + X.g[2]) / 2.0, ymid = (Y.g[1] + Y.g[2]) / 2.0;
    const bool xup = x > xmid, yup = y > ymid;
    const bool rare = X.s == 0 || X.s == g.xnumx - 2 || Y.s == 0 || Y.s == ny - 2 || x == xmid || y == ymid ||
                      x != x || y != y;
    Stencil4 o;
    const double ex = div_rcp(x - X.g[1], X.g[2] - X.g[1], X.inv);
    const double ny_ = div_rcp(y - Y.g[1], Y.g[2] - Y.g[1], Y.inv);
    double a0, a1, a2, b0, b1, b2;
    sel3(X, xup ? 1 : 0, a0, a1, a2);
    sel3(Y, yup ? 1 : 0, b0, b1, b2);
    const double exc = div_rcp(x - (a0 + a1) / 2.0, (a2 - a0) / 2.0, xup ? X.rc[1] : X.rc[0]);
    const double nyc = div_rcp(y - (b0 + b1) / 2.0, (b2 - b0) / 2.0, yup ? Y.rc[1] : Y.rc[0]);
    const int cx = X.s + (xup ? 1 : 0), cy = Y.s + (yup ? 1 : 0);
    o.b.e = ex; o.b.n = ny_; o.b.m3 = (uint32_t)(X.s * ny + Y.s);
    o.c.e = exc; o.c.n = nyc; o.c.m3 = (uint32_t)(cx * ny + cy);
    o.vx.e = ex; o.vx.n = nyc; o.vx.m3 = (uint32_t)(X.s * ny + (cy - 1));
    o.vy.e = exc; o.vy.n = ny_; o.vy.m3 = (uint32_t)((cx - 1) * ny + Y.s);
    if (__ballot(rare) && rare) o = stencils4_full(g, X, Y, x, y);
    return o;
}

// Bilinear 4-point sum, reference term order.
struct W4 { double d0, d1, d2, d3; };
__device__ __forceinline__ W4 wts(const Stencil &s)
{
    W4 w;
    w.d0 = (1.0 - s.e) * (1.0 - s.n);
    w.d1 = (1.0 - s.e) * s.n;
    w.d2 = s.e * (1.0 - s.n);
    w.d3 = s.e * s.n;
    return w;
}
// 64-bit index arithmetic, so that the compiler merges each adjacent pair
// (m3, m3+1) into one 16-byte load.
__device__ __forceinline__ double bil(const double *__restrict__ f, uint32_t m3, int ny, const W4 &w)
{
    const long i = m3;
    return w.d0 * f[i] + w.d1 * f[i + 1] + w.d2 * f[i + ny] + w.d3 * f[i + ny + 1];
}

// Sub-cell sort key of a position (after clamping to the domain): cell in
// x-major order, times 4, plus the quadrant (2*qx + qy) with q = 1 past the
// cell midline.  All markers of one (cell, quadrant) share their scatter
// targets under interpolationmod == 2.
__device__ __forceinline__ uint32_t subcell_key(const DevGrid &g, double x, double y)
{
    x = clampd(x, 0, g.xsize);
    y = clampd(y, 0, g.ysize);
    const Axis4 X = find_x(g, x), Y = find_y(g, y);
    double cx = div_rcp(x - X.g[1], X.g[2] - X.g[1], X.inv);
    double cy = div_rcp(y - Y.g[1], Y.g[2] - Y.g[1], Y.inv);
    uint32_t q = (cx > 0.5 ? 2u : 0u) | (cy > 0.5 ? 1u : 0u);
    return (((uint32_t)X.s * (uint32_t)(g.ynumy - 1) + (uint32_t)Y.s) << 2) | q;
}

// ---- material laws -----------------------------------------------------------

// dencalcomp (density_ductilevisc_calc.c:2-27)
__device__ __forceinline__ void dencalc(const KArgs &a, double mtk, double mpb, int t,
                                        double &mro, double &mbb)
{
    const DevMat *M = a.mat;
    mbb = M->bb[t] / (1.0 - M->bb[t] * (mtk - 298.15));
    if (a.p.densimod == 0) mro = M->ro[t];
    else mro = M->ro[t] * (1.0 - M->bb[t] * (mtk - 298.15)) * (1.0 + M->aa[t] * (mpb - 1.0) * 1e-3);
}

// conductivity (recalc_marker_properties.c:51-54)
__device__ __forceinline__ double conductivity(const DevMat *M, int t, double mtk, double mpb)
{
    double k = (M->kt[t] + M->kf[t] / (mtk + 77.0)) * exp(M->kp[t] * mpb);
    if (k < 0) k = -k / (1.0 + M->kf[t] * (mtk - M->kp[t]) / M->kp[t]);
    return k;
}

__device__ __forceinline__ bool power_law_type(const DevMat *M, int t)
{
    return M->nu[t] > 0 && (M->dh[t] != 0 || M->dv[t] != 0) && M->ss[t] != 0 && M->mm[t] != 1.0;
}

__device__ __forceinline__ double clamp_visc(const KArgs &a, int t, double v)
{
    if (v < a.p.nubeg) v = a.p.nubeg;
    if (v > a.p.nuend) v = a.p.nuend;
    if (v < a.mat->n0[t]) v = a.mat->n0[t];
    if (v > a.mat->n1[t]) v = a.mat->n1[t];
    return v;
}

// The power-law part of ductilevisc (density_ductilevisc_calc.c:157-203):
// 1/eta of the quasi-Newton's last iterate, its iteration count in *iters.
// Fast: the loop's divisions, logarithms, exponentials and powers through
// quot, log_pos, exp_red and pow_pos (v2c_vmath.hpp, within about 2 ulp); otherwise IEEE division and
// the library log and pow.
template <bool Fast>
__device__ __forceinline__ double ductile_iterate(const KArgs &a, int t, double sigin, double e1, int *iters)
{
    const DevMat *M = a.mat;
    auto dv = [](double p, double q) { return Fast ? quot(p, q) : p / q; };
    auto ln = [](double x) { return Fast ? log_pos(x) : log(x); };
    auto ex = [](double x) { return Fast ? exp_red(x) : exp(x); };
    auto pw = [](double x, double y) { return Fast ? pow_pos(x, y) : pow(x, y); };
    const double n = M->mm[t];
    const double k1 = M->k1[t];
    double eta_disl = 0.5 * sigin * M->nu[t] * e1 / pw(sigin, n);
    double eta_new = eta_disl;
    double eta_diff = 0.5 * k1 * e1;
    double coeff = pw(eta_disl, M->rn[t]);
    double gdt = M->gg[t] * a.timestep;
    double mu_prev = 0, mu;
    // Loop-invariant pieces of the iteration, and two exact rewrites:
    // X^(-1/n) = X^((n-1)/n) / X, and eta_d^(2/(eta_d+eta)) evaluated as
    // exp(2/(eta_d+eta) * ln eta_d) with ln eta_d hoisted.  The stop test
    // |log10(mu_k/mu_{k-1})| < 1e-6 is applied as the equivalent bounds on
    // the ratio (10^-1e-6, 10^1e-6).
    const double ex1 = M->ex1[t], c1 = (n - 1) / (n * eta_diff);
    const double inv_ed = 1.0 / eta_diff, ln_ed = ln(eta_diff);
    const double r_lo = 0.999997697417558, r_hi = 1.000002302587744;
    int k = 0;
    for (;;) {
        if (k >= 50) break;
        k++;
        const double eta_old = eta_new;
        mu = dv(1.0, inv_ed + dv(1.0, eta_new));
        // (eta_diff + eta_old) / eta_diff as the exact division by the
        // correctly rounded reciprocal already at hand
        const double pref = div_rcp(eta_diff + eta_old, eta_diff, inv_ed);
        const double gm = gdt + mu;
        const double X = pref * gm;
        const double f1 = ex(ex1 * ln(X));
        const double f2 = eta_old - coeff * f1;
        // eta_d^(2/(eta_d+eta)): exactly 1.0 once the exponent's
        // magnitude is below 2^-54, which holds for any geological eta_d
        const double sum = eta_diff + eta_old;
        const double t3 = fabs(ln_ed) < 0x1p-56 * sum ? 1.0 : exp(2 / sum * ln_ed);
        const double df = 1 - c1 * coeff * dv(f1, X) * (gm + t3);
        eta_new = eta_old - dv(f2, df);
        if (k > 3) {
            const double r = dv(mu, mu_prev);
            if (r > r_lo && r < r_hi) break;
        }
        mu_prev = mu;
    }
    *iters = k;
    return 1 / eta_new;
}

// ductilevisc (density_ductilevisc_calc.c:29-223): sticky air, Newtonian, or
// diffusion + dislocation creep with the reference's quasi-Newton iteration
// (:157-203, stop on |log10(mu_k/mu_{k-1})| < 1e-6 after >3 iterations, at
// most 50).  The uninitialised branch (:55-65) returns NaN.  The iteration
// runs in the fast arithmetic; an iterate that turns NaN there (operands
// outside the fast forms' range) is recomputed with IEEE division and the
// library functions.
__device__ __forceinline__ double ductilevisc(const KArgs &a, int t, double sxy, double sxx,
                                              double mtk, double mpb, int *iters)
{
    const DevMat *M = a.mat;
    *iters = 0;
    if (t == 1) {
        double v = M->nu[1];
        if (a.p.stickymod == 1 && a.timesum > 4e7) {
            double cap = M->gg[2] * a.timestep / 100;
            if (v > cap) v = cap;
        }
        return v;
    }
    if (M->dh[t] == 0 && M->dv[t] == 0 && (M->ss[t] == 0 || M->mm[t] == 1.0)) return M->nu[t];
    if (!power_law_type(M, t)) return __builtin_nan("");
    // pow(s, 0.5) of a non-negative s is sqrt(s) up to pow's last-ulp error
    double sigin = sqrt(sxx * sxx + sxy * sxy);
    double e1 = (M->dh[t] + M->dv[t] * mpb) / (8.314 * mtk);
    if (e1 > 150.0) e1 = 150.0;
    e1 = exp(e1);
    double k1 = M->k1[t];
    double inv_newt = 1.0 / (0.5 * k1 * e1);
    double inv_powl = 0;
    if (sigin > 0) {
        inv_powl = ductile_iterate<true>(a, t, sigin, e1, iters);
        if (__builtin_expect(__builtin_isnan(inv_powl), 0)) inv_powl = ductile_iterate<false>(a, t, sigin, e1, iters);
    }
    return clamp_visc(a, t, 1.0 / (inv_newt + inv_powl));
}

// calc_frictionpara (calc_RSF.c:224-295): marker RSF parameters from the
// fault-band geometry.  Returns the (possibly reset) state phi; writes the
// parameter fields when requested.  Outside the band the fields are one of
// two constant sets (rock types > 0, type 0); fclass records which set a
// marker's fields already hold, so an unchanged set is not written again (the
// values in memory are the ones the reference writes every call).
enum { FC_NONE = 0, FC_OUT = 1, FC_TYPE0 = 2 };
__device__ __forceinline__ double friction_marker(const KArgs &a, long m, int t, double x, double y, double phi)
{
    const v2c_params &p = a.p;
    double radius = ref_abs(y - p.band_y);
    double hr2 = p.half_range;
    const int fc = t > 0 ? (radius <= p.band_halfwidth ? FC_NONE : FC_OUT) : FC_TYPE0;
    const bool fresh = fc != FC_NONE && a.write_friction && a.fclass[m] == fc;
    if (a.write_friction && !fresh) a.fclass[m] = (uint8_t)fc;
    if (t > 0) {
        if (radius <= p.band_halfwidth) {
            double A, B, L, V0;
            if (x > p.end_updip && x < p.start_downdip) {
                A = p.a_vw; B = p.b_vw; L = p.L_vw; V0 = p.V0_vw;
            } else if (x > p.start_updip && x < p.end_updip && p.half_range > 0) {
                double f = (x - (p.start_updip - p.xstpx / 2)) / (2 * hr2);
                A = p.a_vs + f * (p.a_vw - p.a_vs);
                B = p.b_vs + f * (p.b_vw - p.b_vs);
                L = p.L_vs + f * (p.L_vw - p.L_vs);
                V0 = p.V0_vs + f * (p.V0_vw - p.V0_vs);
            } else if (x >= p.start_downdip && x < p.end_downdip && p.half_range > 0) {
                double f = (x - (p.start_downdip - p.xstpx / 2)) / (2 * hr2);
                A = p.a_vw - f * (p.a_vw - p.a_vs);
                B = p.b_vw - f * (p.b_vw - p.b_vs);
                L = p.L_vw - f * (p.L_vw - p.L_vs);
                V0 = p.V0_vw - f * (p.V0_vw - p.V0_vs);
            } else {
                A = p.a_vs; B = p.b_vs; L = p.L_vs; V0 = p.V0_vs;
            }
            if (a.write_friction) {
                MK(a, MF_FS, m) = 0.2;
                MK(a, MF_A, m) = A; MK(a, MF_B, m) = B; MK(a, MF_L0, m) = L; MK(a, MF_L, m) = L;
                MK(a, MF_V0, m) = V0; MK(a, MF_C0, m) = 0.0;
            }
            if (a.timesum == a.timestep) phi = 0.0;
        } else {
            phi = 50.0;
            if (a.write_friction && !fresh) {
                MK(a, MF_FS, m) = 0.6;
                MK(a, MF_A, m) = p.a_hr; MK(a, MF_B, m) = p.b_hr; MK(a, MF_L0, m) = p.L_hr;
                MK(a, MF_L, m) = p.L_hr; MK(a, MF_V0, m) = p.V0_hr;
            }
        }
    } else if (a.write_friction && !fresh) {
        MK(a, MF_FS, m) = 0.6;
        MK(a, MF_A, m) = 1.0; MK(a, MF_B, m) = 0.01; MK(a, MF_L0, m) = 0.2; MK(a, MF_V0, m) = p.V0_hr;
    }
    return phi;
}

// calc_frictionpara_nodes (calc_RSF.c:298-350) as values: mus_s, as, bs,
// Ls, V0s of the node at (x, y).
__device__ __forceinline__ void friction_node_values(const KArgs &a, double x, double y, double &mus, double &A_,
                                                     double &B_, double &L_, double &V0_)
{
    const v2c_params &p = a.p;
    double radius = ref_abs(y - p.band_y);
    double hr2 = p.half_range;
    if (radius <= p.band_halfwidth) {
        if (x > p.end_updip && x < p.start_downdip) {
            A_ = p.a_vw; B_ = p.b_vw; L_ = p.L_vw; V0_ = p.V0_vw;
        } else if (x >= p.start_updip && x <= p.end_updip && p.half_range > 0) {
            double f = (x - (p.start_updip)) / (2 * hr2);
            A_ = p.a_vs + f * (p.a_vw - p.a_vs);
            B_ = p.b_vs + f * (p.b_vw - p.b_vs);
            L_ = p.L_vs + f * (p.L_vw - p.L_vs);
            V0_ = p.V0_vs + f * (p.V0_vw - p.V0_vs);
        } else if (x >= p.start_downdip && x < p.end_downdip && p.half_range > 0) {
            double f = (x - (p.start_downdip)) / (2 * hr2);
            A_ = p.a_vw - f * (p.a_vw - p.a_vs);
            B_ = p.b_vw - f * (p.b_vw - p.b_vs);
            L_ = p.L_vw - f * (p.L_vw - p.L_vs);
            V0_ = p.V0_vw - f * (p.V0_vw - p.V0_vs);
        } else {
            A_ = p.a_vs; B_ = p.b_vs; L_ = p.L_vs; V0_ = p.V0_vs;
        }
        mus = 0.2;
    } else {
        mus = 0.6;
        A_ = p.a_hr; B_ = p.b_hr; L_ = p.L_hr; V0_ = p.V0_hr;
    }
}

// allinteri (interpolation.c:1688-1758): velocity, spin and strain-rate
// invariant at a point.
// hint != nullptr: the cells of the clamped point are already known.
__device__ __forceinline__ void allinteri(const KArgs &a, double x, double y, double &VX, double &VY,
                                          double &ESP, double &EE, const Axis4 *hx = nullptr,
                                          const Axis4 *hy = nullptr)
{
    const DevGrid &g = a.g;
    const int ynumy = g.ynumy;
    x = clampd(x, 0, g.xsize);
    y = clampd(y, 0, g.ysize);
    double xrat = 2.0 / 3.0;
    if (x < g.xr_lo) xrat = 1.0;
    if (x > g.xr_hi) xrat = 1.0;
    if (y < g.yr_lo) xrat = 1.0;
    if (y > g.yr_hi) xrat = 1.0;
    const Axis4 X = hx ? *hx : find_x(g, x), Y = hy ? *hy : find_y(g, y);
    const double *vx = NOD(a, V2C_N_vx), *vy = NOD(a, V2C_N_vy);

    // the four stencils with their shared divisions (identical weights)
    const Stencil4 S4 = stencils4(g, X, Y, x, y);
    Stencil s = S4.b;
    W4 w = wts(s);
    double EXY = bil(NOD(a, V2C_N_exy), s.m3, ynumy, w);
    ESP = bil(NOD(a, V2C_N_esp), s.m3, ynumy, w);

    s = S4.c;
    w = wts(s);
    double EXX = bil(NOD(a, V2C_N_exx), s.m3, ynumy, w);
    // pressure-node averages of vx and vy (interpolation.c:1725-1736)
    const long m3 = s.m3;
    double vxa = (w.d0 * (vx[m3 - 1] + vx[m3 - ynumy - 1]) + w.d1 * (vx[m3] + vx[m3 - ynumy])
                  + w.d2 * (vx[m3 + ynumy - 1] + vx[m3 - 1]) + w.d3 * (vx[m3 + ynumy] + vx[m3])) * 0.5 * (1.0 - xrat);
    double vya = (w.d0 * (vy[m3 - ynumy] + vy[m3 - ynumy - 1]) + w.d1 * (vy[m3 - ynumy + 1] + vy[m3 - ynumy])
                  + w.d2 * (vy[m3] + vy[m3 - 1]) + w.d3 * (vy[m3 + 1] + vy[m3])) * 0.5 * (1.0 - xrat);
    EE = sqrt(EXX * EXX + EXY * EXY);

    s = S4.vx;
    VX = vxa + bil(vx, s.m3, ynumy, wts(s)) * xrat;
    s = S4.vy;
    VY = vya + bil(vy, s.m3, ynumy, wts(s)) * xrat;
}

// ---- allinteri over the RK4 stages, reusing one stencil configuration -------
// Inside an interior cell and off both cell midlines (the common case) the
// four stencils of allinteri are fixed by the cell and the quadrant: the basic
// stencil at the cell's corner node, the pressure node (cx, cy) of the
// quadrant, and vx- and vy-node stencils whose nodes are among the six vx and
// six vy nodes around that pressure node.  RkCache holds what allinteri reads
// there, in the form its sums use: the esp values, the pairwise sums of vx
// (vy) that the pressure-node average adds, and the four vx (vy) values of
// the vx- (vy-) node stencil; per axis the cell's gridlines, its width and
// reciprocal, and the pressure-node stencil's centre and half-spacing with
// their reciprocal.  Every later RK stage whose point lies in the same
// configuration evaluates allinteri from it with the same operands and the
// same operations, so the result is allinteri's bit for bit, without new
// gathers or cell searches.
struct RkAxis {
    double lo, hi, w, inv;              // cell [lo, hi), hi - lo, 1/(hi - lo)
    double mid;                         // (lo + hi) / 2: the cell midline
    double cm, hd, rc;                  // pressure-node stencil: (g[m-1] + g[m]) / 2, (g[m+1] - g[m-1]) / 2, 1/hd
};
struct RkCache {
    RkAxis X, Y;
    double e[4];                        // esp at the basic stencil, bil order
    double P[4], F[4];                  // vx: the average's pair sums, the vx-node stencil (bil order)
    double Q[4], H[4];                  // vy: the same
    int sx, sy;                         // the cell
    bool xup, yup, ok;
};

// Is the clamped point (x, y) in the cached cell (any quadrant)?
__device__ __forceinline__ bool rk_in_cell(const RkCache &c, double x, double y)
{
    return c.ok && x >= c.X.lo && x < c.X.hi && y >= c.Y.lo && y < c.Y.hi;
}

// Does the clamped point (x, y) lie in the cached configuration?
__device__ __forceinline__ bool rk_hit(const RkCache &c, double x, double y)
{
    return rk_in_cell(c, x, y) && (x > c.X.mid) == c.xup && (y > c.Y.mid) == c.yup && x != c.X.mid &&
           y != c.Y.mid;
}

__device__ __forceinline__ void rk_axis(const Axis4 &A, bool up, RkAxis &r)
{
    r.lo = A.g[1];
    r.hi = A.g[2];
    r.w = A.g[2] - A.g[1];
    r.inv = A.inv;
    r.mid = (A.g[1] + A.g[2]) / 2.0;
    // sel3 of the pressure-node stencil: g[m-1], g[m], g[m+1] with m = s + up
    const double a0 = up ? A.g[1] : A.g[0], a1 = up ? A.g[2] : A.g[1], a2 = up ? A.g[3] : A.g[2];
    r.cm = (a0 + a1) / 2.0;
    r.hd = (a2 - a0) / 2.0;
    r.rc = up ? A.rc[1] : A.rc[0];
}

// Fill the cache from the cells (X, Y) of the clamped point (x, y); returns
// false (cache left empty) where the general path applies.
__device__ __forceinline__ bool rk_fill(const KArgs &a, const Axis4 &X, const Axis4 &Y, double x, double y,
                                        RkCache &c)
{
    const DevGrid &g = a.g;
    const int ny = g.ynumy;
    const double xmid = (X.g[1] + X.g[2]) / 2.0, ymid = (Y.g[1] + Y.g[2]) / 2.0;
    const bool rare = X.s == 0 || X.s == g.xnumx - 2 || Y.s == 0 || Y.s == ny - 2 || x == xmid || y == ymid ||
                      x != x || y != y;
    c.ok = !rare;
    if (rare) return false;
    c.xup = x > xmid;
    c.yup = y > ymid;
    c.sx = X.s;
    c.sy = Y.s;
    rk_axis(X, c.xup, c.X);
    rk_axis(Y, c.yup, c.Y);
    const long b = (long)X.s * ny + Y.s;
    const long m3 = (long)(X.s + (c.xup ? 1 : 0)) * ny + Y.s + (c.yup ? 1 : 0);
    const double *esp = NOD(a, V2C_N_esp), *vx = NOD(a, V2C_N_vx), *vy = NOD(a, V2C_N_vy);
    c.e[0] = esp[b]; c.e[1] = esp[b + 1]; c.e[2] = esp[b + ny]; c.e[3] = esp[b + ny + 1];
    // vx at columns cx-1..cx+1 (u_c), rows cy-1, cy; vy at columns cx-1, cx,
    // rows cy-1..cy+1 (interpolation.c:1725-1736 reads these)
    const double u00 = vx[m3 - ny - 1], u01 = vx[m3 - ny], u10 = vx[m3 - 1], u11 = vx[m3];
    const double u20 = vx[m3 + ny - 1], u21 = vx[m3 + ny];
    const double v00 = vy[m3 - ny - 1], v01 = vy[m3 - ny], v02 = vy[m3 - ny + 1];
    const double v10 = vy[m3 - 1], v11 = vy[m3], v12 = vy[m3 + 1];
    c.P[0] = u10 + u00; c.P[1] = u11 + u01; c.P[2] = u20 + u10; c.P[3] = u21 + u11;
    c.Q[0] = v01 + v00; c.Q[1] = v02 + v01; c.Q[2] = v11 + v10; c.Q[3] = v12 + v11;
    // vx-node stencil: column cx - xup, rows cy-1, cy; vy-node stencil:
    // columns cx-1, cx, rows cy - yup, cy - yup + 1 (bil order)
    c.F[0] = c.xup ? u00 : u10; c.F[1] = c.xup ? u01 : u11;
    c.F[2] = c.xup ? u10 : u20; c.F[3] = c.xup ? u11 : u21;
    c.H[0] = c.yup ? v00 : v01; c.H[1] = c.yup ? v01 : v02;
    c.H[2] = c.yup ? v10 : v11; c.H[3] = c.yup ? v11 : v12;
    return true;
}

// The common-path weights of stencils4 at a point of the configuration:
// ex, ny the cell-local coordinates (basic-node stencil), exc, nyc the
// pressure-node stencil's.
__device__ __forceinline__ void rk_weights(const RkCache &c, double x, double y, double &ex, double &ny_,
                                           double &exc, double &nyc)
{
    ex = div_rcp(x - c.X.lo, c.X.w, c.X.inv);
    ny_ = div_rcp(y - c.Y.lo, c.Y.w, c.Y.inv);
    exc = div_rcp(x - c.X.cm, c.X.hd, c.X.rc);
    nyc = div_rcp(y - c.Y.cm, c.Y.hd, c.Y.rc);
}

// allinteri's velocity and spin (interpolation.c:1688-1758) at a clamped
// point of the cached configuration: allinteri's sums, term for term.
__device__ __forceinline__ void rk_eval(const KArgs &a, const RkCache &c, double x, double y, double &VX,
                                        double &VY, double &ESP)
{
    const DevGrid &g = a.g;
    double xrat = 2.0 / 3.0;
    if (x < g.xr_lo) xrat = 1.0;
    if (x > g.xr_hi) xrat = 1.0;
    if (y < g.yr_lo) xrat = 1.0;
    if (y > g.yr_hi) xrat = 1.0;
    double ex, ny_, exc, nyc;
    rk_weights(c, x, y, ex, ny_, exc, nyc);
    Stencil s;
    s.e = ex; s.n = ny_;
    W4 w = wts(s);
    ESP = w.d0 * c.e[0] + w.d1 * c.e[1] + w.d2 * c.e[2] + w.d3 * c.e[3];
    s.e = exc; s.n = nyc;
    w = wts(s);
    const double vxa = (w.d0 * c.P[0] + w.d1 * c.P[1] + w.d2 * c.P[2] + w.d3 * c.P[3]) * 0.5 * (1.0 - xrat);
    const double vya = (w.d0 * c.Q[0] + w.d1 * c.Q[1] + w.d2 * c.Q[2] + w.d3 * c.Q[3]) * 0.5 * (1.0 - xrat);
    s.e = ex; s.n = nyc;
    w = wts(s);
    VX = vxa + (w.d0 * c.F[0] + w.d1 * c.F[1] + w.d2 * c.F[2] + w.d3 * c.F[3]) * xrat;
    s.e = exc; s.n = ny_;
    w = wts(s);
    VY = vya + (w.d0 * c.H[0] + w.d1 * c.H[1] + w.d2 * c.H[2] + w.d3 * c.H[3]) * xrat;
}

// ---- build-defined per-marker local update (not reference behaviour) -------

// Ductile viscosity by Newton on the reference's residual
// (density_ductilevisc_calc.c:188-193): f(eta) = eta - coeff*X^((n-1)/n),
// X = (eta_d+eta)/eta_d*G dt + eta, exact f'(eta), to |f|/eta <= tol.
// Start: the reference's eta_disl0 (:159).  With a guess (> 0), the residual
// is evaluated at both starts and Newton starts from the smaller one; a guess
// whose residual grows after its first step falls back to eta_disl0
// (*fallback = 1).  *iters counts Newton updates, *evals residual evaluations.
template <bool Fast>
__device__ __forceinline__ double ductile_newton_impl(const KArgs &a, int t, double sxy, double sxx,
                                                      double mtk, double mpb, double guess, double tol,
                                                      int *iters, int *evals, int *fallback, double *root)
{
    auto pw = [](double x, double y) { return Fast ? pow_pos(x, y) : pow(x, y); };
    const DevMat *M = a.mat;
    *iters = 0;
    *evals = 0;
    *fallback = 0;
    *root = 0;
    if (!power_law_type(M, t) || t == 1) {
        int it;
        return ductilevisc(a, t, sxy, sxx, mtk, mpb, &it);
    }
    const double n = M->mm[t];
    const double sigin = sqrt(sxx * sxx + sxy * sxy);
    double e1 = (M->dh[t] + M->dv[t] * mpb) / (8.314 * mtk);
    if (e1 > 150.0) e1 = 150.0;
    e1 = exp(e1);
    const double k1 = M->k1[t];
    const double eta_diff = 0.5 * k1 * e1;
    if (!(sigin > 0)) return clamp_visc(a, t, eta_diff);
    const double eta_disl0 = 0.5 * sigin * M->nu[t] * e1 / pw(sigin, n);
    const double coeff = pw(eta_disl0, M->rn[t]);
    const double gdt = M->gg[t] * a.timestep;
    const double p1 = M->ex1[t], slope = gdt / eta_diff + 1;
    auto resid = [&](double eta, double &f, double &X, double &Xp) {
        X = (eta_diff + eta) / eta_diff * gdt + eta;
        Xp = pw(X, p1);
        f = eta - coeff * Xp;
        return fabs(f) / eta;
    };
    double eta = eta_disl0, f, X, Xp;
    double r = resid(eta, f, X, Xp);
    int ev = 1, k = 0;
    bool from_guess = false;
    if (r > tol && guess > 0) {
        double fg, Xg, Xpg;
        const double rg = resid(guess, fg, Xg, Xpg);
        ev++;
        if (rg < r) { eta = guess; f = fg; X = Xg; Xp = Xpg; r = rg; from_guess = true; }
    }
    double rprev = r;
    while (r > tol && k < 200) {
        const double df = 1 - p1 * coeff * (Xp / X) * slope;   // X^(-1/n) = X^((n-1)/n) / X
        double next;
        if (f < 0 && !(df > 0)) next = 2.0 * eta;               // left of a flat region
        else next = eta - f / df;
        if (!(next > 0)) next = 0.5 * eta;
        eta = next;
        k++;
        r = resid(eta, f, X, Xp);
        ev++;
        if (from_guess && k == 1 && r > rprev) {                 // warm start diverging
            from_guess = false;
            *fallback = 1;
            eta = eta_disl0;
            r = resid(eta, f, X, Xp);
            ev++;
        }
        rprev = r;
    }
    *iters = k;
    *evals = ev;
    *root = eta;
    return clamp_visc(a, t, 1.0 / (1.0 / eta_diff + 1.0 / eta));
}

// The solve in the fast arithmetic (pow_pos, within a few 1e-15); a result
// that turns NaN there is recomputed with the library pow.
__device__ __forceinline__ double ductile_newton(const KArgs &a, int t, double sxy, double sxx,
                                                 double mtk, double mpb, double guess, double tol,
                                                 int *iters, int *evals, int *fallback, double *root)
{
    const double v = ductile_newton_impl<true>(a, t, sxy, sxx, mtk, mpb, guess, tol, iters, evals, fallback, root);
    if (__builtin_expect(!__builtin_isnan(v), 1)) return v;
    return ductile_newton_impl<false>(a, t, sxy, sxx, mtk, mpb, guess, tol, iters, evals, fallback, root);
}

// Per-marker RSF slip rate with radiation damping at the aged state
// (aging law calc_RSF.c:96-103; yield law calc_RSF.c:129 solved for V by
// Newton in s = ln V; cap and floor calc_RSF.c:193-196).  The default start
// s_hi = ln min(undamped closed form calc_RSF.c:192, sig/eta_rad) bounds the
// root from above; g is increasing and convex in s, so Newton from s_hi
// converges monotonically.  Guess handling as in ductile_newton.
__device__ __forceinline__ int rsf_newton(double sig, double nP, double ra, double rb, double rL, double rV0,
                                          double mu0, double phi, double Vprev, double dt, double eta_rad,
                                          double EII, double W, double guessV, double tol, double &V_out,
                                          double &phi_out, int &fallback, double &root, int &evals)
{
    double phin;
    if (Vprev * dt / rL <= 1e-6) phin = log(exp(phi) * (1 - Vprev * dt / rL) + rV0 * dt / rL);
    else phin = log(rV0 / Vprev + (exp(phi) - rV0 / Vprev) * exp(-Vprev * dt / rL));
    phi_out = phin;
    fallback = 0;
    root = 0;
    evals = 0;
    if (!(sig > 0)) { V_out = 0; return 0; }
    const double K = exp((mu0 + rb * phin) / ra) / (2 * rV0);
    const double Vu = 2 * rV0 * sinh(sig / (ra * nP)) * exp(-(mu0 + rb * phin) / ra);
    double Vhi = Vu;
    if (eta_rad > 0 && sig / eta_rad < Vhi) Vhi = sig / eta_rad;
    double V;
    int k = 0;
    if (eta_rad == 0 || !(Vhi > 0) || !isfinite(Vhi)) {
        V = Vhi;
    } else {
        const double shi = log(Vhi);
        auto gfun = [&](double s, double &g, double &dg) {
            const double e = exp(s), z = e * K;
            g = nP * ra * asinh(z) + eta_rad * e - sig;
            dg = nP * ra * z / sqrt(1 + z * z) + eta_rad * e;
            return fabs(g) / sig;
        };
        double s = shi, g, dg;
        double r = gfun(s, g, dg);
        int ev = 1;
        bool from_guess = false;
        if (r > tol && guessV > 0) {
            double sg = log(guessV);
            if (sg > shi) sg = shi;
            double gg, dgg;
            const double rg = gfun(sg, gg, dgg);
            ev++;
            if (rg < r) { s = sg; g = gg; dg = dgg; r = rg; from_guess = true; }
        }
        double rprev = r;
        while (r > tol && k < 200) {
            double sn = s - g / dg;
            if (!(sn <= shi)) sn = shi;
            s = sn;
            k++;
            r = gfun(s, g, dg);
            ev++;
            if (from_guess && k == 1 && r > rprev) {
                from_guess = false;
                fallback = 1;
                s = shi;
                r = gfun(s, g, dg);
                ev++;
            }
            rprev = r;
        }
        evals = ev;
        V = exp(s);
    }
    root = V;
    V = ref_minv(V, EII * W);
    if (V < 1e-40) V = 1e-40;
    V_out = V;
    return k;
}

}  // namespace v2c
