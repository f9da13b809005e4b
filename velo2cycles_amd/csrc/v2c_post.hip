// v2c_post.hip -- the reference's per-node routines between the Stokes solve
// and the next marker2nodes, on the device-resident nodal fields:
// recompute_stress_strainrate, interpolate_center2basic(_old),
// interpolate_basic2center, copyarrays and calc_timestep, and the
// sticky-air marks of marker2nodes and statistics of calc_timestep.
//
// Every node is one thread in node order (m3 = m1*ynumy + m2).  No kernel
// writes a field it reads at another node, so each routine is one launch.
// The running minima and maxima of recompute_stress_strainrate and
// calc_timestep are reduced as compositions of per-node update functions,
// which reproduces a sequential MINV/MAXV fold over the nodes exactly (NaN
// and resets included) in any grouping.
// Included at the end of v2c_api.hip: it uses the context internals.

namespace v2c {

constexpr int POST_BLOCK = 256;

__device__ __forceinline__ double ref_maxv(double a, double b) { return a >= b ? a : b; }
__device__ __forceinline__ double ref_absv(double a) { return a >= 0 ? a : -a; }

#define NF_(f) NOD(a, V2C_N_##f)

// nueffcalc (momentum_continuity_addmatrix_checkerror.c:380-393)
__device__ __forceinline__ double nueff_calc(double nvp, double ggeff, double dt, double nubeg, double &Z)
{
    Z = ggeff * dt / (ggeff * dt + nvp);
    double nueff = Z * nvp;
    if (nueff < nubeg) nueff = nubeg;
    return nueff;
}

// ---- running minima and maxima -----------------------------------------------------
// A running minimum m <- MINV(m, x) is the function f_x(m) = m <= x ? m : x
// (f_x(NaN) = x; f_NaN = the constant NaN); a reset m <- r is the constant
// c_r.  Compositions stay in {identity, f_x with x not NaN, c_r}:
//   f_w . f_v = f_{f_w(v)},  f_w . c_r = c_{f_w(r)},  c_r . anything = c_r,
// so an ordered tree of compositions gives the sequential fold bit for bit.
// MAXV likewise with >=.
struct Fold {
    int kind;      // 0 identity, 1 running min/max with v, 2 constant v
    double v;
};

template <bool MAX>
__host__ __device__ inline double fold_step(double m, double x)
{
    return MAX ? (m >= x ? m : x) : (m <= x ? m : x);
}

template <bool MAX>
__host__ __device__ inline Fold fold_of(double x)
{
    return x != x ? Fold{2, x} : Fold{1, x};
}

// b after a
template <bool MAX>
__host__ __device__ inline Fold fold_then(Fold a, Fold b)
{
    if (b.kind == 0) return a;
    if (a.kind == 0 || b.kind == 2) return b;
    return Fold{a.kind, fold_step<MAX>(a.v, b.v)};
}

template <bool MAX>
__host__ __device__ inline double fold_apply(Fold f, double m)
{
    return f.kind == 0 ? m : (f.kind == 2 ? f.v : fold_step<MAX>(m, f.v));
}

// ---- recompute_stress_strainrate (momentum_continuity_iterate.c:214-263) ----
constexpr int NSTAT = 5;   // maxds, minnu, maxnu, maxdphi, maxnvp
__device__ constexpr bool STAT_MAX[NSTAT] = {true, false, true, true, false};

template <int K>
__device__ __forceinline__ Fold stat_then(Fold a, Fold b)
{
    return STAT_MAX[K] ? fold_then<true>(a, b) : fold_then<false>(a, b);
}

__device__ __forceinline__ void stat_compose(Fold *dst, const Fold *src)
{
    dst[0] = stat_then<0>(dst[0], src[0]);
    dst[1] = stat_then<1>(dst[1], src[1]);
    dst[2] = stat_then<2>(dst[2], src[2]);
    dst[3] = stat_then<3>(dst[3], src[3]);
    dst[4] = stat_then<4>(dst[4], src[4]);
}

// The routine's extrema as the sequential MINV/MAXV folds of its node loop
// (each node's updates in the routine's order), composed over nodes in order.
__global__ __launch_bounds__(POST_BLOCK) void k_stress(KArgs a, double dt, Fold *__restrict__ part)
{
    __shared__ Fold F[NSTAT][POST_BLOCK];
    const long m3 = (long)blockIdx.x * POST_BLOCK + threadIdx.x;
    Fold st[NSTAT];
    for (int k = 0; k < NSTAT; k++) st[k] = Fold{0, 0.0};
    if (m3 < a.nodenum) {
        const DevGrid &g = a.g;
        const long ny = g.ynumy, m2 = m3 % ny, m1 = m3 / ny;
        if (m1 >= 1 && m2 >= 1) {
            const double *vx = NF_(vx), *vy = NF_(vy);
            const long v0 = (m1 - 1) * ny + (m2 - 1), v1 = v0 + 1, v2 = v0 + ny, v3 = v2 + 1;
            // sxxcalc (:446-473)
            double Z;
            const double nux = nueff_calc(NF_(nd)[v3], NF_(gd)[v3], dt, a.p.nubeg, Z);
            const double exx = 0.5 * ((vx[v2] - vx[v0]) / (g.gx[m1] - g.gx[m1 - 1]) -
                                      (vy[v1] - vy[v0]) / (g.gy[m2] - g.gy[m2 - 1]));
            const double sxxe = NF_(sxxe)[v3];
            const double sxx = 2.0 * nux * exx + sxxe * (1.0 - Z);
            NF_(sxx)[m3] = sxx;
            NF_(exx)[m3] = exx;
            st[0] = fold_of<true>(ref_absv(sxx - sxxe));
            NF_(deltasxx)[m3] = (sxx - sxxe) / a.timestep;
            st[1] = fold_of<false>(nux);
            st[2] = fold_of<true>(nux);
            if (m1 < g.xnumx - 1 && m2 < ny - 1) {
                // sxycalc (:521-569)
                const double nuy = nueff_calc(NF_(nu)[v3], NF_(gg)[v3], dt, a.p.nubeg, Z);
                const double exy1 = (vx[v3] - vx[v2]) / (g.gy[m2 + 1] - g.gy[m2 - 1]);
                const double exy2 = (vy[v3] - vy[v1]) / (g.gx[m1 + 1] - g.gx[m1 - 1]);
                const double exy = exy1 + exy2;
                const double sxye = NF_(sxye)[v3];
                const double sxy = 2.0 * nuy * exy + sxye * (1.0 - Z);
                NF_(sxy)[m3] = sxy;
                NF_(exy)[m3] = exy;
                NF_(esp)[m3] = exy1 - exy2;
                NF_(deltasxy)[m3] = (sxy - sxye) / a.timestep;
                st[4] = fold_of<false>(NF_(nu)[m3] / NF_(gg)[m3]);
                st[0] = fold_then<true>(st[0], fold_of<true>(ref_absv(sxy - sxye)));
                st[3] = fold_of<true>(ref_absv(NF_(phis)[m3] - NF_(phis_it)[m3]));
                st[1] = fold_then<false>(st[1], fold_of<false>(nuy));
                st[2] = fold_then<true>(st[2], fold_of<true>(nuy));
            }
        }
    }
    for (int k = 0; k < NSTAT; k++) F[k][threadIdx.x] = st[k];
    __syncthreads();
    for (int s = 1; s < POST_BLOCK; s <<= 1) {           // ordered pairwise composition
        if ((threadIdx.x & (2 * s - 1)) == 0) {
            Fold l[NSTAT], r[NSTAT];
            for (int k = 0; k < NSTAT; k++) { l[k] = F[k][threadIdx.x]; r[k] = F[k][threadIdx.x + s]; }
            stat_compose(l, r);
            for (int k = 0; k < NSTAT; k++) F[k][threadIdx.x] = l[k];
        }
        __syncthreads();
    }
    if (threadIdx.x < NSTAT) part[(long)blockIdx.x * NSTAT + threadIdx.x] = F[threadIdx.x][0];
}

// ---- ordered composition of per-block folds --------------------------------------
// One workgroup of FIN_BLOCK threads: thread t composes, in order, the folds of
// a contiguous run of blocks, then the threads' results are composed pairwise
// in order in LDS.  Composition is associative, so the result is the
// sequential fold over every block whatever the grouping.
constexpr int FIN_BLOCK = 256;

// Bit k of MAXMASK: component k is a running maximum (else a minimum).
template <int NS, unsigned MAXMASK>
__device__ __forceinline__ void fin_compose(Fold (*F)[FIN_BLOCK], const Fold *__restrict__ part, int nblocks,
                                            Fold *acc)
{
    constexpr auto is_max = [](int k) { return ((MAXMASK >> k) & 1u) != 0; };
    const int per = (nblocks + FIN_BLOCK - 1) / FIN_BLOCK;
    const int b0 = threadIdx.x * per, b1 = min(b0 + per, nblocks);
    Fold f[NS];
    for (int k = 0; k < NS; k++) f[k] = Fold{0, 0.0};
    for (int b = b0; b < b1; b++)
        for (int k = 0; k < NS; k++) {
            const Fold x = part[(long)b * NS + k];
            f[k] = is_max(k) ? fold_then<true>(f[k], x) : fold_then<false>(f[k], x);
        }
    for (int k = 0; k < NS; k++) F[k][threadIdx.x] = f[k];
    __syncthreads();
    for (int s = 1; s < FIN_BLOCK; s <<= 1) {
        if ((threadIdx.x & (2 * s - 1)) == 0)
            for (int k = 0; k < NS; k++) {
                const Fold l = F[k][threadIdx.x], r = F[k][threadIdx.x + s];
                F[k][threadIdx.x] = is_max(k) ? fold_then<true>(l, r) : fold_then<false>(l, r);
            }
        __syncthreads();
    }
    for (int k = 0; k < NS; k++) acc[k] = F[k][0];
}

// Block folds composed in order, applied to the routine's initial values
// (0, -1e50, 1e50, 0, 1e30).
__global__ __launch_bounds__(FIN_BLOCK) void k_stress_final(const Fold *__restrict__ part, int nblocks,
                                                            double *__restrict__ out)
{
    __shared__ Fold F[NSTAT][FIN_BLOCK];
    Fold acc[NSTAT];
    fin_compose<NSTAT, 0b01101u>(F, part, nblocks, acc);     // maxds, maxnu, maxdphi
    if (threadIdx.x != 0) return;
    out[0] = fold_apply<true>(acc[0], 0.0);
    out[1] = fold_apply<false>(acc[1], -1e50);
    out[2] = fold_apply<true>(acc[2], 1e50);
    out[3] = fold_apply<true>(acc[3], 0.0);
    out[4] = fold_apply<false>(acc[4], 1e30);
}

// ---- interpolate_center2basic (interpolation.c:1600-1637) ----------------------
__global__ __launch_bounds__(POST_BLOCK) void k_center2basic(KArgs a, double dt)
{
    const long m3 = (long)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (m3 >= a.nodenum) return;
    const long ny = a.g.ynumy, m2 = m3 % ny, m1 = m3 / ny;
    if (!(m1 >= 1 && m2 >= 1 && m1 < a.g.xnumx - 1 && m2 < ny - 1)) return;
    const double *exx = NF_(exx), *sxx = NF_(sxx), *sxxe = NF_(sxxe), *gd = NF_(gd), *ndm = NF_(nd_m2c);
    const double *pr = NF_(pr);
    const long j[4] = {m3, m3 + 1, m3 + ny, m3 + ny + 1};
    double epl[4], el[4];
    for (int k = 0; k < 4; k++) {
        const long i = j[k];
        epl[k] = exx[i] - (sxx[i] - sxxe[i]) / (2 * gd[i] * dt) - sxx[i] / (2 * ndm[i]);
        el[k] = 2 * gd[i] * dt * exx[i] + sxxe[i];
    }
    NF_(exxs)[m3] = 0.25 * (exx[j[0]] + exx[j[1]] + exx[j[2]] + exx[j[3]]);
    NF_(exxpl)[m3] = 0.25 * (epl[0] + epl[1] + epl[2] + epl[3]);
    NF_(sxxs)[m3] = 0.25 * (sxx[j[0]] + sxx[j[1]] + sxx[j[2]] + sxx[j[3]]);
    NF_(prs)[m3] = 0.25 * (pr[j[0]] + pr[j[1]] + pr[j[2]] + pr[j[3]]);
    NF_(exxel)[m3] = 0.25 * (el[0] + el[1] + el[2] + el[3]);
}

// ---- interpolate_center2basic_old (interpolation.c:1575-1598) -----------------
__global__ __launch_bounds__(POST_BLOCK) void k_center2basic_old(KArgs a)
{
    const long m3 = (long)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (m3 >= a.nodenum) return;
    const long ny = a.g.ynumy, m2 = m3 % ny, m1 = m3 / ny;
    if (!(m1 >= 1 && m2 >= 1 && m1 < a.g.xnumx - 1 && m2 < ny - 1)) return;
    const double *exxe = NF_(exxe), *sxxe = NF_(sxxe), *sppe = NF_(sppe);
    NF_(exxes)[m3] = 0.25 * (exxe[m3] + exxe[m3 + 1] + exxe[m3 + ny] + exxe[m3 + ny + 1]);
    NF_(sxxes)[m3] = 0.25 * (sxxe[m3] + sxxe[m3 + 1] + sxxe[m3 + ny] + sxxe[m3 + ny + 1]);
    NF_(sppes)[m3] = 0.25 * (sppe[m3] + sppe[m3 + 1] + sppe[m3 + ny] + sppe[m3 + ny + 1]);
}

// ---- interpolate_basic2center (interpolation.c:1640-1685) ---------------------
// Pressure-node viscosity: harmonic bilinear average of the four basic nodes
// around the pressure point, found by the cell search of that point.
__global__ __launch_bounds__(POST_BLOCK) void k_basic2center(KArgs a)
{
    const long m3 = (long)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (m3 >= a.nodenum) return;
    const DevGrid &g = a.g;
    const long ny = g.ynumy, m2 = m3 % ny, m1 = m3 / ny;
    if (!(m1 >= 1 && m2 >= 1)) return;
    const double x = (g.gx[m1] + g.gx[m1 - 1]) / 2, y = (g.gy[m2] + g.gy[m2 - 1]) / 2;
    long m10 = find_x(g, x).s, m20 = find_y(g, y).s;
    if (m10 < 1) m10 = 1;
    if (m10 > g.xnumx - 3) m10 = g.xnumx - 3;
    if (m20 < 1) m20 = 1;
    if (m20 > ny - 3) m20 = ny - 3;
    double e = (x - g.gx[m10]) / (g.gx[m10 + 1] - g.gx[m10]);
    double n = (y - g.gy[m20]) / (g.gy[m20 + 1] - g.gy[m20]);
    if (e < 0) e = 0.0;
    if (n < 0) n = 0.0;
    if (e > 1) e = 1.0;
    if (n > 1) n = 1.0;
    const double dm0 = (1.0 - e) * (1.0 - n), dm1 = (1.0 - e) * n, dm2 = e * (1.0 - n), dm3 = e * n;
    const double *nu = NF_(nu);
    const long m4 = m10 * ny + m20;
    NF_(nd)[m3] = 1.0 / (dm0 * 1.0 / nu[m4] + dm1 * 1.0 / nu[m4 + 1] + dm2 * 1.0 / nu[m4 + ny] +
                         dm3 * 1.0 / nu[m4 + ny + 1]);
}

// ---- copyarrays (momentum_continuity_iterate.c:265-283) ------------------------
__device__ __forceinline__ void copyarrays_one(const KArgs &a, long m3)
{
    const long ny = a.g.ynumy, m2 = m3 % ny, m1 = m3 / ny;
    if (!(m1 >= 1 && m2 >= 1 && m1 < a.g.xnumx - 1 && m2 < ny - 1)) return;
    const double solved = NF_(RK_sliprs_solved)[m3];
    NF_(deltaVs)[m3] = (log(solved) - log(NF_(RK_sliprs_final)[m3])) / a.timestep;
    NF_(sliprs)[m3] = solved;
    NF_(RK_sliprs_final)[m3] = solved;
    NF_(RK_phidts_final)[m3] = NF_(RK_phidts_solved)[m3];
}

__global__ __launch_bounds__(POST_BLOCK) void k_copyarrays(KArgs a)
{
    const long m3 = (long)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (m3 < a.nodenum) copyarrays_one(a, m3);
}

// ---- calc_timestep (calc_RSF.c:352-572) ----------------------------------------
constexpr int NTS = 6;   // dtnvp, dtd, dtphi, dtL (min); maxvel, maxVx (max)

struct TsArgs {
    double maxtmstep, tcnvp, tcd, tcphih, dtfactor, lambda0;
    int oldornew;
};

// chi and the lambda it sets at an interior node (calc_RSF.c:464-481);
// returns 1 when chi sets lambda, 0 when it leaves it (chi NaN), -1 for
// chi == 0 (the reference exits there) and -2 for an unknown pmode.
__device__ __forceinline__ int ts_lambda(const KArgs &a, const TsArgs &T, long m3, double &lam)
{
    double nps;
    if (a.p.pmode == 0) nps = a.p.nstress;
    else if (a.p.pmode == 1) nps = T.oldornew == 0 ? NF_(sppes)[m3] : NF_(prs)[m3];
    else return -2;
    const double bm = NF_(bm)[m3], gg = NF_(gg)[m3], Ls = NF_(Ls)[m3], as = NF_(as)[m3], bs = NF_(bs)[m3];
    const double v = (3 * bm - 2 * gg) / (2 * (3 * bm + gg));
    const double k = 2.0 / 3.14 * (gg / (1 - v)) / a.p.xstpx;
    const double chi = 1.0 / 4.0 * pow(k * Ls / (as * nps) - (bs - as) / as, 2) - k * Ls / (as * nps);
    if (chi > 0) {
        lam = as * nps / (k * Ls - (bs - as) * nps);
        return 1;
    }
    if (chi < 0) {
        lam = 1 - (bs - as) * nps / (k * Ls);
        return 1;
    }
    if (chi == 0) return -1;
    return 0;
}

__device__ __forceinline__ bool ts_interior(const KArgs &a, long m3)
{
    const long ny = a.g.ynumy, m2 = m3 % ny, m1 = m3 / ny;
    return m1 >= 1 && m2 >= 1 && m1 < a.g.xnumx - 1 && m2 < ny - 1;
}

// Pass 1: the last lambda set in block b of the nodes (node order); error flags.
__device__ __forceinline__ void ts_lambda_block(const KArgs &a, const TsArgs &T, long b, int *__restrict__ bset,
                                                double *__restrict__ bval, int *__restrict__ err)
{
    __shared__ int last[POST_BLOCK];
    __shared__ double val[POST_BLOCK];
    const long m3 = b * POST_BLOCK + threadIdx.x;
    int set = 0;
    double lam = 0.0;
    if (m3 < a.nodenum && a.p.RSFmode == 1 && ts_interior(a, m3)) {
        const int r = ts_lambda(a, T, m3, lam);
        if (r < 0) atomicOr(err, r == -1 ? 1 : 2);
        set = r == 1;
    }
    last[threadIdx.x] = set ? (int)threadIdx.x : -1;
    val[threadIdx.x] = lam;
    __syncthreads();
    for (int s = POST_BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) last[threadIdx.x] = max(last[threadIdx.x], last[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bset[b] = last[0] >= 0;
        bval[b] = last[0] >= 0 ? val[last[0]] : 0.0;
    }
}

__global__ __launch_bounds__(POST_BLOCK) void k_ts_lambda(KArgs a, TsArgs T, int *__restrict__ bset,
                                                          double *__restrict__ bval, int *__restrict__ err)
{
    ts_lambda_block(a, T, blockIdx.x, bset, bval, err);
}

// Pass 2: lambda carried into every block, the last one set before it (or
// lambda0).  One workgroup: each thread takes a contiguous run of blocks; an
// inclusive scan of (set, value) with "the later set one wins" gives each
// run's carry-in.
__device__ __forceinline__ void ts_carry_block(const int *__restrict__ bset, const double *__restrict__ bval,
                                               int nblocks, double lambda0, double *__restrict__ carry)
{
    __shared__ int sset[FIN_BLOCK];
    __shared__ double sval[FIN_BLOCK];
    const int per = (nblocks + FIN_BLOCK - 1) / FIN_BLOCK;
    const int b0 = threadIdx.x * per, b1 = min(b0 + per, nblocks);
    int set = 0;
    double v = 0.0;
    for (int b = b0; b < b1; b++)
        if (bset[b]) {
            set = 1;
            v = bval[b];
        }
    sset[threadIdx.x] = set;
    sval[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < FIN_BLOCK; s <<= 1) {
        int ps = 0;
        double pv = 0.0;
        if ((int)threadIdx.x >= s) {
            ps = sset[threadIdx.x - s];
            pv = sval[threadIdx.x - s];
        }
        __syncthreads();
        if (ps && !sset[threadIdx.x]) {
            sset[threadIdx.x] = 1;
            sval[threadIdx.x] = pv;
        }
        __syncthreads();
    }
    double c = lambda0;
    if (threadIdx.x > 0 && sset[threadIdx.x - 1]) c = sval[threadIdx.x - 1];
    for (int b = b0; b < b1; b++) {
        carry[b] = c;
        if (bset[b]) c = bval[b];
    }
}

__global__ __launch_bounds__(FIN_BLOCK) void k_ts_carry(const int *__restrict__ bset,
                                                        const double *__restrict__ bval, int nblocks,
                                                        double lambda0, double *__restrict__ carry)
{
    ts_carry_block(bset, bval, nblocks, lambda0, carry);
}

// Pass 3: every interior node's update functions, composed in node order
// within block b.
__device__ __forceinline__ void ts_node_block(const KArgs &a, const TsArgs &T, long b,
                                              const double *__restrict__ carry, Fold *__restrict__ part)
{
    __shared__ int last[POST_BLOCK];
    __shared__ double val[POST_BLOCK];
    __shared__ Fold F[NTS][POST_BLOCK];
    const long m3 = b * POST_BLOCK + threadIdx.x;
    const bool in = m3 < a.nodenum && ts_interior(a, m3);
    const bool rsf = in && a.p.RSFmode == 1;
    int set = 0;
    double lam_set = 0.0;
    if (rsf) set = ts_lambda(a, T, m3, lam_set) == 1;
    // lambda in force at this node: the last one set at or before it
    last[threadIdx.x] = set ? (int)threadIdx.x : -1;
    val[threadIdx.x] = lam_set;
    __syncthreads();
    for (int s = 1; s < POST_BLOCK; s <<= 1) {          // inclusive max-scan
        const int v = threadIdx.x >= (unsigned)s ? last[threadIdx.x - s] : -1;
        __syncthreads();
        last[threadIdx.x] = max(last[threadIdx.x], v);
        __syncthreads();
    }
    Fold f[NTS];
    for (int k = 0; k < NTS; k++) f[k] = Fold{0, 0.0};
    if (in) {
        const DevGrid &g = a.g;
        const long ny = g.ynumy, m2 = m3 % ny, m1 = m3 / ny;
        const double dx = g.gx[m1 + 1] - g.gx[m1], dy = g.gy[m2 + 1] - g.gy[m2];
        const double fw = 2 * 0.5 * (dx + dy);
        const double Vf = NF_(RK_sliprs_final)[m3], Vs = NF_(RK_sliprs_solved)[m3];
        const double num2c = NF_(nu_m2c)[m3];
        double nuvp;
        if (T.oldornew == 0) {
            if (num2c > 1.00e+26) {
                const double sxxes = NF_(sxxes)[m3], sxye = NF_(sxye)[m3];
                const double sigin = sqrt(sxxes * sxxes + sxye * sxye);
                nuvp = (sigin > 0 && Vf > 0) ? num2c * sigin / (2 * num2c * Vf / fw + sigin) : num2c;
            } else {
                nuvp = 5.00e+26;
            }
        } else {
            nuvp = num2c > 1.00e+26 ? NF_(nu)[m3] : 5.00e+26;
        }
        f[0] = fold_of<false>(T.tcnvp * nuvp / NF_(gg)[m3]);
        const double ux = T.oldornew == 1 ? NF_(vx)[m3] : NF_(mvx)[m3];
        const double uy = T.oldornew == 1 ? NF_(vy)[m3] : NF_(mvy)[m3];
        if (ux != 0.0 && uy != 0.0) {
            f[5] = fold_of<true>(ref_maxv(ref_absv(ux), ref_absv(uy)));
            const double maxdr = ref_maxv(ref_absv(ux / dx), ref_absv(uy / dy));
            f[1] = fold_of<false>(T.tcd * 1.0 / maxdr);
        }
        if (rsf) {
            const int li = last[threadIdx.x];
            double lam = li >= 0 ? val[li] : carry[b];
            lam = ref_minv(lam, 1.0 / 5.0);
            const double bs = NF_(bs)[m3], Ls = NF_(Ls)[m3], V0s = NF_(V0s)[m3];
            if (bs > 0) {
                if (T.oldornew == 0 && Vf > 0) {
                    const double phisc = V0s / Ls * exp(-NF_(phis)[m3]);
                    f[2] = fold_of<false>(T.tcphih * 1.0 / phisc);
                    f[3] = fold_of<false>(lam * Ls / Vf);
                } else if (T.oldornew == 1 && Vs > 0) {
                    const double phisc = V0s / Ls * exp(-NF_(phis_it)[m3]);
                    f[2] = fold_of<false>(T.tcphih * 1.0 / phisc);
                    f[3] = fold_of<false>(lam * Ls / Vs);
                }
            } else {
                f[2] = Fold{2, T.maxtmstep};
                f[3] = Fold{2, T.maxtmstep};
            }
        }
        if (T.oldornew == 0 && Vf > 0) f[4] = fold_of<true>(Vf);
        if (T.oldornew == 1 && Vs > 0) f[4] = fold_of<true>(Vs);
    }
    for (int k = 0; k < NTS; k++) F[k][threadIdx.x] = f[k];
    __syncthreads();
    for (int s = 1; s < POST_BLOCK; s <<= 1) {           // ordered pairwise composition
        if ((threadIdx.x & (2 * s - 1)) == 0) {
            const int j = threadIdx.x + s;
            for (int k = 0; k < 4; k++) F[k][threadIdx.x] = fold_then<false>(F[k][threadIdx.x], F[k][j]);
            for (int k = 4; k < NTS; k++) F[k][threadIdx.x] = fold_then<true>(F[k][threadIdx.x], F[k][j]);
        }
        __syncthreads();
    }
    if (threadIdx.x < NTS) part[b * NTS + threadIdx.x] = F[threadIdx.x][0];
}

__global__ __launch_bounds__(POST_BLOCK) void k_ts_node(KArgs a, TsArgs T, const double *__restrict__ carry,
                                                        Fold *__restrict__ part)
{
    ts_node_block(a, T, blockIdx.x, carry, part);
}

// Pass 4: blocks composed in order, applied to the reference's initial
// values; out = dt, dtnvp, dtd, dtphi, dtL, maxvel, maxVx.
__device__ __forceinline__ void ts_final_block(const Fold *__restrict__ part, int nblocks, const TsArgs &T,
                                               double *__restrict__ out)
{
    __shared__ Fold F[NTS][FIN_BLOCK];
    Fold acc[NTS];
    fin_compose<NTS, 0b110000u>(F, part, nblocks, acc);      // maxvel, maxVx
    if (threadIdx.x != 0) return;
    const double dtnvp = fold_apply<false>(acc[0], 1e+30), dtd = fold_apply<false>(acc[1], 1e+30);
    const double dtphi = fold_apply<false>(acc[2], 1e+30), dtL = fold_apply<false>(acc[3], 1e+30);
    const double maxvel = fold_apply<true>(acc[4], 0.0), maxVx = fold_apply<true>(acc[5], 0.0);
    out[0] = T.dtfactor * ref_minv(T.maxtmstep, ref_minv(dtphi, ref_minv(ref_minv(dtL, dtnvp), dtd)));
    out[1] = dtnvp;
    out[2] = dtd;
    out[3] = dtphi;
    out[4] = dtL;
    out[5] = maxvel;
    out[6] = maxVx;
}

__global__ __launch_bounds__(FIN_BLOCK) void k_ts_final(const Fold *__restrict__ part, int nblocks, TsArgs T,
                                                        double *__restrict__ out)
{
    ts_final_block(part, nblocks, T, out);
}

__global__ void k_post_clear(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---- sticky-air marks (interpolation.c:861-916) ----------------------------------
// The reference's serial loop leaves at each node the type of the last marker
// (highest index) that marks it; win[node] = 1 + that index, from an atomic
// maximum over the markers, gives the same node values in any order.
__global__ __launch_bounds__(POST_BLOCK) void k_sticky_clear(uint32_t *__restrict__ win, long n)
{
    const long i = (long)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (i < n) win[i] = 0u;
}

// STK_MPT markers per thread: their types in one 8-byte load.
constexpr int STK_MPT = 8;

// The node that marker i marks, or -1 (a cell midline, or an index at or
// beyond nodenum).
__device__ __forceinline__ long sticky_node(const KArgs &a, long i)
{
    const double x = MK(a, MF_X, i), y = MK(a, MF_Y, i);
    const Axis4 X = find_x(a.g, x), Y = find_y(a.g, y);
    const double celdx = (x - X.g[1]) / (X.g[2] - X.g[1]);
    const double celdy = (y - Y.g[1]) / (Y.g[2] - Y.g[1]);
    const long ny = a.g.ynumy;
    long m3 = (long)X.s * ny + Y.s + 1;
    if (celdx < 0.5 && celdy < 0.5) {
    } else if (celdx < 0.5 && celdy > 0.5) {
        m3 += 1;
    } else if (celdx > 0.5 && celdy < 0.5) {
        m3 += ny;
    } else if (celdx > 0.5 && celdy > 0.5) {
        m3 += ny + 1;
    } else {
        return -1;
    }
    return m3 < a.nodenum ? m3 : -1;
}

// A thread's markers are consecutive, and neighbours often mark one node: one
// atomic per run of markers marking the same node, with the run's last
// (highest) index.
__global__ __launch_bounds__(POST_BLOCK) void k_sticky_marks(KArgs a, uint32_t *__restrict__ win)
{
    const long i0 = ((long)blockIdx.x * POST_BLOCK + threadIdx.x) * STK_MPT;
    if (i0 >= a.n) return;
    int8_t tt[STK_MPT];
    if (i0 + STK_MPT <= a.n) {
        const uint2 w = *reinterpret_cast<const uint2 *>(a.t + i0);   // i0 is a multiple of 8
        __builtin_memcpy(tt, &w, sizeof tt);
    } else {
        for (int k = 0; k < STK_MPT; k++) tt[k] = i0 + k < a.n ? a.t[i0 + k] : (int8_t)0;
    }
    long cur = -1;
    uint32_t top = 0;
#pragma unroll
    for (int k = 0; k < STK_MPT; k++) {
        if (tt[k] != 1 && tt[k] != 2) continue;
        const long m3 = sticky_node(a, i0 + k);
        if (m3 < 0) continue;
        if (m3 != cur) {
            if (cur >= 0) atomicMax(&win[cur], top);
            cur = m3;
        }
        top = (uint32_t)(i0 + k + 1);
    }
    if (cur >= 0) atomicMax(&win[cur], top);
}

__global__ __launch_bounds__(POST_BLOCK) void k_sticky_nodes(KArgs a, const uint32_t *__restrict__ win)
{
    const long m3 = (long)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (m3 >= a.nodenum) return;
    const uint32_t w = win[m3];
    NF_(sticky01)[m3] = w ? (double)a.t[w - 1] : 0.0;
}

// ---- calc_timestep's sticky-air statistics (calc_RSF.c:364-374, 505-550) ----------
// One node per thread; the per-node contributions are composed in node order:
// over the lanes of a wavefront (a shuffle tree), over the wavefronts of a
// block, then over the blocks.  Counts and sums add; extrema compose as
// running MAXV/MINV folds and the P1/P2 assignments as constants, so those
// come out as the serial loop's.  Set g = 0 gathers sticky01 == 1 (the
// reference's "_0" members), g = 1 sticky01 == 2 ("_1").
constexpr int STK_FIN = 1024;

struct StickyAcc {
    int nr[2];
    // f[4g + 0..3]: max sigin, min sigin, max nu, min nu of set g;
    // f[8..11]: sigin_P1, visco_P1, sigin_P2, visco_P2 (assignments)
    Fold f[12];
    // s[4g + 0..3]: sum sigin, sum sigin^2, sum log10 nu, sum (log10 nu)^2
    double s[8];
};

__device__ __forceinline__ void sticky_clear(StickyAcc &A)
{
    A.nr[0] = A.nr[1] = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) A.f[k] = Fold{0, 0.0};
#pragma unroll
    for (int k = 0; k < 8; k++) A.s[k] = 0.0;
}

// A <- B after A
__device__ __forceinline__ void sticky_then(StickyAcc &A, const StickyAcc &B)
{
    A.nr[0] += B.nr[0];
    A.nr[1] += B.nr[1];
#pragma unroll
    for (int k = 0; k < 12; k++)
        A.f[k] = (k < 8 && (k & 1) == 0) ? fold_then<true>(A.f[k], B.f[k]) : fold_then<false>(A.f[k], B.f[k]);
#pragma unroll
    for (int k = 0; k < 8; k++) A.s[k] += B.s[k];
}

// Lane 0 ends with the composition of the wavefront's lanes in lane order.
__device__ __forceinline__ void sticky_wave(StickyAcc &A, int lane)
{
    for (int d = 1; d < WAVE; d <<= 1) {
        StickyAcc B;
        B.nr[0] = __shfl_down(A.nr[0], d);
        B.nr[1] = __shfl_down(A.nr[1], d);
#pragma unroll
        for (int k = 0; k < 12; k++) {
            B.f[k].kind = __shfl_down(A.f[k].kind, d);
            B.f[k].v = __shfl_down(A.f[k].v, d);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) B.s[k] = __shfl_down(A.s[k], d);
        if ((lane & (2 * d - 1)) == 0) sticky_then(A, B);
    }
}

// One node of set G (constant indices keep the accumulator in registers).
template <int G>
__device__ __forceinline__ void sticky_node_in(StickyAcc &A, double sig, double v)
{
    A.nr[G] = 1;
    A.f[4 * G + 0] = fold_of<true>(sig);
    A.f[4 * G + 1] = fold_of<false>(sig);
    A.f[4 * G + 2] = fold_of<true>(v);
    A.f[4 * G + 3] = fold_of<false>(v);
    const double l = log10(v);
    A.s[4 * G + 0] = sig;
    A.s[4 * G + 1] = sig * sig;
    A.s[4 * G + 2] = l;
    A.s[4 * G + 3] = l * l;
}

__global__ __launch_bounds__(POST_BLOCK) void k_sticky_part(KArgs a, StickyAcc *__restrict__ part)
{
    __shared__ StickyAcc W[POST_BLOCK / WAVE];
    const long m3 = (long)blockIdx.x * POST_BLOCK + threadIdx.x;
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    StickyAcc A;
    sticky_clear(A);
    const long ny = a.g.ynumy;
    const long m2 = m3 % ny, m1 = m3 / ny;
    if (m3 < a.nodenum && m1 >= 1 && m2 >= 1 && m1 < a.g.xnumx - 1 && m2 < ny - 1) {
        const double sx = NF_(sxxs)[m3], sy = NF_(sxy)[m3];
        const double sig = sqrt(sx * sx + sy * sy);
        const double v = NF_(nu)[m3], mark = NF_(sticky01)[m3];
        if (mark == 1 && m1 == 50 && m2 == 50) { A.f[8] = Fold{2, sig}; A.f[9] = Fold{2, v}; }
        if (mark == 1 && m1 == 500 && m2 == 50) { A.f[10] = Fold{2, sig}; A.f[11] = Fold{2, v}; }
        if (mark == 2) sticky_node_in<1>(A, sig, v);
        else if (mark == 1) sticky_node_in<0>(A, sig, v);
    }
    sticky_wave(A, lane);
    if (lane == 0) W[wv] = A;
    __syncthreads();
    if (threadIdx.x != 0) return;
    StickyAcc R = W[0];
    for (int w = 1; w < POST_BLOCK / WAVE; w++) sticky_then(R, W[w]);
    part[blockIdx.x] = R;
}

__global__ __launch_bounds__(STK_FIN) void k_sticky_final(const StickyAcc *__restrict__ part, long np,
                                                          v2c_sticky_stats *__restrict__ out)
{
    __shared__ StickyAcc W[STK_FIN / WAVE];
    const int t = threadIdx.x, lane = t % WAVE;
    const long per = (np + STK_FIN - 1) / STK_FIN, b0 = t * per, b1 = min(b0 + per, np);
    StickyAcc A;
    sticky_clear(A);
    for (long b = b0; b < b1; b++) sticky_then(A, part[b]);
    sticky_wave(A, lane);
    if (lane == 0) W[t / WAVE] = A;
    __syncthreads();
    if (t != 0) return;
    StickyAcc R = W[0];
    for (int w = 1; w < STK_FIN / WAVE; w++) sticky_then(R, W[w]);
    v2c_sticky_stats o;
    o.nr_1 = R.nr[1];
    o.nr_0 = R.nr[0];
    o.max_sigin_1 = fold_apply<true>(R.f[4], 0.0);
    o.min_sigin_1 = fold_apply<false>(R.f[5], 1e30);
    o.max_visc_1 = fold_apply<true>(R.f[6], 0.0);
    o.min_visc_1 = fold_apply<false>(R.f[7], 1e30);
    o.max_sigin_0 = fold_apply<true>(R.f[0], 0.0);
    o.min_sigin_0 = fold_apply<false>(R.f[1], 1e30);
    o.max_visc_0 = fold_apply<true>(R.f[2], 0.0);
    o.min_visc_0 = fold_apply<false>(R.f[3], 1e30);
    o.sum_sigin_1 = R.s[4];
    o.sum2_sigin_1 = R.s[5];
    o.sum_visc_1 = R.s[6];
    o.sum2_visc_1 = R.s[7];
    o.sum_sigin_0 = R.s[0];
    o.sum2_sigin_0 = R.s[1];
    o.sum_visc_0 = R.s[2];
    o.sum2_visc_0 = R.s[3];
    o.sigin_P1 = fold_apply<true>(R.f[8], 0.0);
    o.visco_P1 = fold_apply<true>(R.f[9], 0.0);
    o.sigin_P2 = fold_apply<true>(R.f[10], 0.0);
    o.visco_P2 = fold_apply<true>(R.f[11], 0.0);
    *out = o;
}

#undef NF_

}  // namespace v2c

using namespace v2c;

// ---- host side -------------------------------------------------------------------

// Device scratch of the post-solve routines: per-block partials, sized for the
// context's node count on first use.
static int post_scratch(v2c_ctx *c, size_t bytes)
{
    if (bytes <= c->post_bytes) return 0;
    if (c->d_post) (void)hipFree(c->d_post);
    c->d_post = nullptr;
    c->post_bytes = 0;
    HIPCHK(c, hipMalloc(&c->d_post, bytes));
    c->post_bytes = bytes;
    return 0;
}

extern "C" {

int v2c_recompute_stress_strainrate(v2c_ctx *c, double dt, int rk, v2c_stress_stats *stats)
{
    (void)rk;
    if (!c) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const unsigned nb = blocks_for(c->nodenum, POST_BLOCK);
    int rc = post_scratch(c, sizeof(Fold) * NSTAT * (size_t)nb + sizeof(double) * NSTAT);
    if (rc) return rc;
    Fold *part = (Fold *)c->d_post;
    double *out = (double *)(part + (size_t)NSTAT * nb);
    const KArgs a = make_args(c);
    rc = launch(c, V2C_K_POST, [&] {
        hipLaunchKernelGGL(k_stress, dim3(nb), dim3(POST_BLOCK), 0, c->stream, a, dt, part);
        hipLaunchKernelGGL(k_stress_final, dim3(1), dim3(FIN_BLOCK), 0, c->stream, part, (int)nb, out);
    });
    if (rc) return rc;
    double h[NSTAT];
    HIPCHK(c, hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (stats) {
        stats->maxds = h[0];
        stats->minnu = h[1];
        stats->maxnu = h[2];
        stats->maxdphi = h[3];
        stats->maxnvp = h[4];
    }
    return 0;
}

#define POST_NODE_LAUNCH(kernel, ...)                                                               \
    do {                                                                                            \
        if (!c) return V2C_EINVAL;                                                                  \
        HIPCHK(c, hipSetDevice(c->device));                                                         \
        const KArgs a = make_args(c);                                                               \
        int rc_ = launch(c, V2C_K_POST, [&] {                                                       \
            hipLaunchKernelGGL(kernel, dim3(blocks_for(c->nodenum, POST_BLOCK)), dim3(POST_BLOCK), 0, \
                               c->stream, a, ##__VA_ARGS__);                                        \
        });                                                                                         \
        if (rc_) return rc_;                                                                        \
        return finish(c);                                                                           \
    } while (0)

int v2c_interpolate_center2basic(v2c_ctx *c, double dt) { POST_NODE_LAUNCH(k_center2basic, dt); }
int v2c_interpolate_center2basic_old(v2c_ctx *c) { POST_NODE_LAUNCH(k_center2basic_old); }
int v2c_interpolate_basic2center(v2c_ctx *c) { POST_NODE_LAUNCH(k_basic2center); }
int v2c_copyarrays(v2c_ctx *c, int rk)
{
    (void)rk;
    POST_NODE_LAUNCH(k_copyarrays);
}

#undef POST_NODE_LAUNCH

void v2c_timestep_params_default(v2c_timestep_params *p)
{
    if (!p) return;
    p->maxtmstep = 3.15576e7;
    p->tcnvp = 0.2;
    p->tcd = 1e-3;
    p->tcphih = 0.2;
    p->dtfactor = 1.0;
    p->lambda0 = NAN;
}

int v2c_calc_timestep(v2c_ctx *c, int oldornew, const v2c_timestep_params *p, v2c_timestep_result *r)
{
    if (!c || !p || !r || (oldornew != 0 && oldornew != 1)) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const unsigned nb = blocks_for(c->nodenum, POST_BLOCK);
    // layout: Fold part[NTS*nb] | double bval[nb], carry[nb], out[8] | int bset[nb], err
    const size_t bytes = sizeof(Fold) * NTS * nb + sizeof(double) * (2 * (size_t)nb + 8) + sizeof(int) * (nb + 1);
    int rc = post_scratch(c, bytes);
    if (rc) return rc;
    Fold *part = (Fold *)c->d_post;
    double *bval = (double *)(part + (size_t)NTS * nb), *carry = bval + nb, *out = carry + nb;
    int *bset = (int *)(out + 8), *err = bset + nb;
    TsArgs T{p->maxtmstep, p->tcnvp, p->tcd, p->tcphih, p->dtfactor, p->lambda0, oldornew};
    const KArgs a = make_args(c);
    rc = launch(c, V2C_K_POST, [&] {
        hipLaunchKernelGGL(k_post_clear, dim3(1), dim3(64), 0, c->stream, err, 1);
        hipLaunchKernelGGL(k_ts_lambda, dim3(nb), dim3(POST_BLOCK), 0, c->stream, a, T, bset, bval, err);
        hipLaunchKernelGGL(k_ts_carry, dim3(1), dim3(FIN_BLOCK), 0, c->stream, bset, bval, (int)nb, p->lambda0,
                           carry);
        hipLaunchKernelGGL(k_ts_node, dim3(nb), dim3(POST_BLOCK), 0, c->stream, a, T, carry, part);
        hipLaunchKernelGGL(k_ts_final, dim3(1), dim3(FIN_BLOCK), 0, c->stream, part, (int)nb, T, out);
    });
    if (rc) return rc;
    double h[7];
    int e = 0;
    HIPCHK(c, hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&e, err, sizeof e, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (e & 2) return set_err(c, V2C_EUNSUPPORTED, "calc_timestep: pmode %d (only 0 and 1 define the normal stress)",
                              c->prm.pmode);
    if (e & 1) return set_err(c, V2C_EUNSUPPORTED, "calc_timestep: chi is zero at a node (the reference exits)");
    r->dt = h[0];
    r->dtnvp = h[1];
    r->dtd = h[2];
    r->dtphi = h[3];
    r->dtL = h[4];
    r->maxvel = h[5];
    r->maxVx = h[6];
    return 0;
}

int v2c_sticky_marks(v2c_ctx *c)
{
    if (!c) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = post_scratch(c, sizeof(uint32_t) * (size_t)c->nodenum);
    if (rc) return rc;
    uint32_t *win = (uint32_t *)c->d_post;
    const KArgs a = make_args(c);
    const unsigned nbn = blocks_for(c->nodenum, POST_BLOCK);
    rc = launch(c, V2C_K_POST, [&] {
        hipLaunchKernelGGL(k_sticky_clear, dim3(nbn), dim3(POST_BLOCK), 0, c->stream, win, c->nodenum);
        if (c->n > 0)
            hipLaunchKernelGGL(k_sticky_marks, dim3(blocks_for((c->n + STK_MPT - 1) / STK_MPT, POST_BLOCK)),
                               dim3(POST_BLOCK), 0, c->stream, a, win);
        hipLaunchKernelGGL(k_sticky_nodes, dim3(nbn), dim3(POST_BLOCK), 0, c->stream, a, win);
    });
    if (rc) return rc;
    return finish(c);
}

int v2c_calc_sticky_stats(v2c_ctx *c, v2c_sticky_stats *out)
{
    if (!c || !out) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const long np = blocks_for(c->nodenum, POST_BLOCK);
    int rc = post_scratch(c, sizeof(StickyAcc) * (size_t)np + sizeof(v2c_sticky_stats));
    if (rc) return rc;
    StickyAcc *part = (StickyAcc *)c->d_post;
    v2c_sticky_stats *dout = (v2c_sticky_stats *)(part + np);
    const KArgs a = make_args(c);
    rc = launch(c, V2C_K_POST, [&] {
        hipLaunchKernelGGL(k_sticky_part, dim3((unsigned)np), dim3(POST_BLOCK), 0, c->stream, a, part);
        hipLaunchKernelGGL(k_sticky_final, dim3(1), dim3(STK_FIN), 0, c->stream, part, np, dout);
    });
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(out, dout, sizeof *out, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"
