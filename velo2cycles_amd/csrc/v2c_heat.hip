// v2c_heat.hip -- marker temperatures after the heat solve: the marker part
// of solve_temperature (temperature_iterate.c:181-419) on the device.
//
//   pass 1, per marker (physical order): fluid markers take the new nodal
//     temperature; rock markers get the reset of newly incoming markers, the
//     subgrid diffusion toward the old nodal temperature and the solved
//     increment, and leave their node weight w and w * dT in a scratch pair;
//   pass 2, per group of 64 cells (one wavefront, sorted order): the weights
//     of each cell's markers summed into its upper-left node (tol0, tol1),
//     then tol1 / tol0; the last column and row of nodes get zeros;
//   pass 3, per marker: the antidiffusion correction from that node.
// No atomics: each node's sums are a fixed sequence of lane scans over its
// cell's markers in (sub-cell, marker id) order, so runs are bitwise
// reproducible.
// Included at the end of v2c_api.hip: it uses the context internals.

namespace v2c {

constexpr int HEAT_BLOCK = 256;

__device__ __forceinline__ bool heat_inside(const DevGrid &g, double x, double y)
{
    return x > 0 && y > 0 && x < g.xsize && y < g.ysize;
}

// interpolate_temperature (interpolation.c:1492-1533) at an inside point of
// cell (X.s, Y.s): the new (tk) and old (tk2) nodal temperatures.
__device__ __forceinline__ void heat_tk(const KArgs &a, const Axis4 &X, const Axis4 &Y, double x, double y,
                                        double &TK, double &TK2)
{
    Stencil s;
    s.e = div_rcp(x - X.g[1], X.g[2] - X.g[1], X.inv);
    s.n = div_rcp(y - Y.g[1], Y.g[2] - Y.g[1], Y.inv);
    s.m3 = (uint32_t)(X.s * a.g.ynumy + Y.s);
    const W4 w = wts(s);
    TK = bil(NOD(a, V2C_N_tk), s.m3, a.g.ynumy, w);
    TK2 = bil(NOD(a, V2C_N_tk2), s.m3, a.g.ynumy, w);
}

// Pass 1 (temperature_iterate.c:196-208, 213-331).
__global__ __launch_bounds__(HEAT_BLOCK) void k_heat_markers(KArgs a, double heatdif)
{
    const long m = (long)blockIdx.x * HEAT_BLOCK + threadIdx.x;
    if (m >= a.n) return;
    const DevGrid &g = a.g;
    const double x = MK(a, MF_X, m), y = MK(a, MF_Y, m);
    double w = 0.0, wd = 0.0;
    const int t = a.t[m];
    if (heat_inside(g, x, y) && t < 100) {
        Axis4 X, Y;
        key_axes(a, a.key[m], X, Y);
        double TK, TK2;
        heat_tk(a, X, Y, x, y, TK, TK2);
        if (t >= 50) {
            MK(a, MF_K, m) = TK;
        } else {
            double mk = MK(a, MF_K, m);
            if (mk <= 0) {
                mk = TK2;
                if (mk < a.p.tdeep) mk = a.p.tdeep;
            }
            if (heatdif != 0) {
                const double dtk = TK2 - mk;
                const double dx = X.g[2] - X.g[1], dy = Y.g[2] - Y.g[1];
                const double swt1 = div_rcp(X.inv, dy, Y.inv);
                const double cx = div_rcp(x - X.g[1], dx, X.inv);
                const double cy = div_rcp(y - Y.g[1], dy, Y.inv);
                // k and rho*Cp at the marker, nodes in the reference's loop order
                const double *kt = NOD(a, V2C_N_kt), *cp = NOD(a, V2C_N_cp), *ro = NOD(a, V2C_N_ro);
                const long ny = g.ynumy, m4 = (long)X.s * ny + Y.s;
                const long n4[4] = {m4, m4 + 1, m4 + ny, m4 + ny + 1};
                const double iv[4] = {(1.0 - cx) * (1.0 - cy), (1.0 - cx) * cy, cx * (1.0 - cy), cx * cy};
                double mkt = 0, mrocp = 0;
                for (int k = 0; k < 4; k++) {
                    mkt += kt[n4[k]] * iv[k];
                    mrocp += cp[n4[k]] * ro[n4[k]] * iv[k];
                }
                double dtk1 = -heatdif * mkt / mrocp * 2.0 * (1.0 / dx / dx + 1.0 / dy / dy) * a.timestep;
                if (dtk1 < -150.0) dtk1 = -150.0;
                dtk1 = dtk * (1.0 - exp(dtk1));
                mk += dtk1;
                w = (1.0 - cx) * (1.0 - cy) * swt1;
                wd = dtk1 * w;
            }
            mk += (TK - TK2);
            MK(a, MF_K, m) = mk;
        }
    }
    a.prop[m] = w;
    a.prop[a.cap + m] = wd;
}

// Pass 2 (temperature_iterate.c:183-184, 325-364): node (m1, m2) collects the
// markers of cell (m1, m2).  One wavefront takes HEAT_CELLS consecutive cells
// and walks their sorted markers 64 at a time with coalesced loads: a
// segmented lane scan by cell sums each batch, the last lane of a cell's
// segment adds the batch sum to the cell's accumulator in LDS.
constexpr int HEAT_CELLS = 64;
static_assert(HEAT_CELLS == 64, "one accumulator per lane of a wavefront");

__global__ __launch_bounds__(64) void k_heat_cells(KArgs a)
{
    __shared__ double acc0[HEAT_CELLS], acc1[HEAT_CELLS];
    const int lane = threadIdx.x;
    const long c0 = (long)blockIdx.x * HEAT_CELLS;
    const long c1 = c0 + HEAT_CELLS < a.ncell ? c0 + HEAT_CELLS : a.ncell;
    acc0[lane] = 0.0;
    acc1[lane] = 0.0;
    __syncthreads();
    const uint32_t p0 = a.sub_start[4 * c0], p1 = a.sub_start[4 * c1];
    for (uint32_t base = p0; base < p1; base += 64) {
        const uint32_t p = base + lane;
        const bool ok = p < p1;
        int cell = -1;
        double w = 0.0, wd = 0.0;
        if (ok) {
            const uint32_t m = a.order[p];
            cell = (int)(a.skey[p] >> 2);
            w = a.prop[m];
            wd = a.prop[a.cap + m];
        }
        for (int off = 1; off < 64; off <<= 1) {
            const double w2 = __shfl_up(w, off), wd2 = __shfl_up(wd, off);
            const int c2 = __shfl_up(cell, off);
            if (lane >= off && c2 == cell) {
                w += w2;
                wd += wd2;
            }
        }
        const int next = __shfl_down(cell, 1);
        if (ok && (lane == 63 || next != cell)) {
            acc0[cell - c0] += w;
            acc1[cell - c0] += wd;
        }
        __syncthreads();
    }
    const long cell = c0 + lane;
    if (cell >= c1) return;
    const long ny = a.g.ynumy, m1 = cell / (ny - 1), m2 = cell % (ny - 1), m4 = m1 * ny + m2;
    const double t0 = acc0[lane];
    double t1 = acc1[lane];
    if (t0) t1 /= t0;
    NOD(a, V2C_N_tol0)[m4] = t0;
    NOD(a, V2C_N_tol1)[m4] = t1;
}

// Pass 2, nodes that are no cell's upper-left node (last column, last row).
__global__ __launch_bounds__(HEAT_BLOCK) void k_heat_border(KArgs a)
{
    const long i = (long)blockIdx.x * HEAT_BLOCK + threadIdx.x;
    const long nx = a.g.xnumx, ny = a.g.ynumy;
    if (i >= nx + ny - 1) return;
    const long m4 = i < ny ? (nx - 1) * ny + i : (i - ny) * ny + (ny - 1);
    NOD(a, V2C_N_tol0)[m4] = 0.0;
    NOD(a, V2C_N_tol1)[m4] = 0.0;
}

// Pass 3 (temperature_iterate.c:367-416).
__global__ __launch_bounds__(HEAT_BLOCK) void k_heat_antidiff(KArgs a)
{
    const long m = (long)blockIdx.x * HEAT_BLOCK + threadIdx.x;
    if (m >= a.n) return;
    const double x = MK(a, MF_X, m), y = MK(a, MF_Y, m);
    if (!(heat_inside(a.g, x, y) && a.t[m] < 50)) return;
    Axis4 X, Y;
    key_axes(a, a.key[m], X, Y);
    const double cx = div_rcp(x - X.g[1], X.g[2] - X.g[1], X.inv);
    const double cy = div_rcp(y - Y.g[1], Y.g[2] - Y.g[1], Y.inv);
    const double tol1 = NOD(a, V2C_N_tol1)[(long)X.s * a.g.ynumy + Y.s];
    if (tol1) MK(a, MF_K, m) -= tol1 * ((1.0 - cx) * (1.0 - cy));
}

}  // namespace v2c

extern "C" int v2c_temperature_markers(v2c_ctx *c, double heatdif)
{
    if (!c) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = ensure_order(c);
    if (rc) return rc;
    const KArgs a = make_args(c);
    rc = launch(c, V2C_K_POST, [&] {
        if (c->n > 0) hipLaunchKernelGGL(k_heat_markers, dim3(blocks_for(c->n, HEAT_BLOCK)), dim3(HEAT_BLOCK), 0,
                                         c->stream, a, heatdif);
        hipLaunchKernelGGL(k_heat_cells, dim3(blocks_for(c->ncell, HEAT_CELLS)), dim3(64), 0, c->stream, a);
        hipLaunchKernelGGL(k_heat_border, dim3(blocks_for(c->xnumx + c->ynumy - 1, HEAT_BLOCK)), dim3(HEAT_BLOCK), 0,
                           c->stream, a);
        if (heatdif != 0 && c->n > 0)
            hipLaunchKernelGGL(k_heat_antidiff, dim3(blocks_for(c->n, HEAT_BLOCK)), dim3(HEAT_BLOCK), 0, c->stream, a);
    });
    if (rc) return rc;
    return finish(c);
}
