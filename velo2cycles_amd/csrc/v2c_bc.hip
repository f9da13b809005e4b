// v2c_bc.hip -- boundary-condition tail of marker2nodes and the heat solve:
// vxvy_boundarycorrection, P_boundarycorrection (momentum_continuity_
// addmatrix_checkerror.c:1000-1184) and tkrecalc (temperature_iterate.c:
// 828-891) on the device.
//
// Each routine solves "identity plus boundary equations": a free parameter
// keeps its value, a bounded one equals CONST + sum KOEF_i * PAR_i.  The
// reference eliminates the rows in index order (gausevps.c:216-385); here
// every row is one thread and the system is solved by substitution to a
// fixed point (x <- b + C x until nothing changes), which ends after as many
// sweeps as the longest chain of boundary equations referring to each other
// and gives the elimination's value per row (up to the grouping of the sums
// along longer chains).  Equations that refer to each other in a cycle settle
// only if the cycle contracts; otherwise the sweep limit reports them as
// unsupported.  Couplings of a row are added in increasing parameter order,
// the order in which the elimination and back substitution add them.
// Included at the end of v2c_api.hip: it uses the context internals.

namespace v2c {

constexpr int BC_BLOCK = 256;
enum BcKind { BC_VXVY = 0, BC_P = 1, BC_T = 2 };
enum BcFlag { BCF_CHANGED = 1, BCF_ABSENT = 2, BCF_SELF = 4, BCF_RANGE = 8 };

struct BcArgs {
    V2C_GLOBAL const long *bondm;     // 4 * nodenum
    V2C_GLOBAL const long *bondn;     // nbond x 3
    V2C_GLOBAL const double *bondv;   // nbond x 4
    long nbond;
};

// Row u of the system: present?  its constant b and up to three couplings
// (j, c), sorted by j.  Returns false for an absent row.
template <int KIND>
__device__ __forceinline__ bool bc_row(const KArgs &a, const BcArgs &B, long u, double &b, int &nc, long j[3],
                                       double c[3], int &flags)
{
    const long nn = a.nodenum, ny = a.g.ynumy, nx = a.g.xnumx;
    long m3, p;
    const double *src;
    if (KIND == BC_VXVY) {
        const bool vy = u >= nn;
        m3 = vy ? u - nn : u;
        const long m2 = m3 % ny, m1 = m3 / ny;
        if (vy ? !(m1 < nx - 1) : !(m2 < ny - 1)) return false;
        p = 3 * m3 + (vy ? 2 : 1);
        src = vy ? NOD(a, V2C_N_mvy) : NOD(a, V2C_N_mvx);
    } else if (KIND == BC_P) {
        m3 = u;
        const long m2 = m3 % ny, m1 = m3 / ny;
        if (!(m1 && m2)) return false;
        p = 3 * m3;
        src = NOD(a, V2C_N_sppe);
    } else {
        m3 = u;
        p = 3 * nn + m3;
        src = NOD(a, V2C_N_tk);
    }
    const long k = B.bondm[p];
    nc = 0;
    if (k == 0 || (KIND == BC_P && a.p.compressiblemode == 1)) {
        b = src[m3];
        return true;
    }
    if (k < 0 || k >= B.nbond) {
        flags |= BCF_RANGE;
        b = 0.0;
        return true;
    }
    b = B.bondv[4 * k];
    const long N = KIND == BC_VXVY ? 2 * nn : nn;
    for (int i = 0; i < 3; i++) {
        const long q = B.bondn[3 * k + i];
        if (!q) continue;
        long jj;
        if (KIND == BC_VXVY) jj = u >= nn ? ((q - 1) - 2) / 3 + nn : ((q - 1) - 1) / 3;
        else if (KIND == BC_P) jj = (q - 1) / 3;
        else jj = q - 1 - 3 * nn;
        if (jj < 0 || jj >= N) { flags |= BCF_RANGE; continue; }
        if (jj == u) { flags |= BCF_SELF; continue; }
        j[nc] = jj;
        c[nc] = B.bondv[4 * k + i + 1];
        nc++;
    }
    for (int x = 1; x < nc; x++)                 // sort couplings by parameter
        for (int y = x; y > 0 && j[y] < j[y - 1]; y--) {
            const long tj = j[y]; j[y] = j[y - 1]; j[y - 1] = tj;
            const double tc = c[y]; c[y] = c[y - 1]; c[y - 1] = tc;
        }
    return true;
}

// x = b for present rows, 0 for absent ones; checks the couplings.
template <int KIND>
__global__ __launch_bounds__(BC_BLOCK) void k_bc_init(KArgs a, BcArgs B, long N, double *__restrict__ x,
                                                      int *__restrict__ flag)
{
    const long u = (long)blockIdx.x * BC_BLOCK + threadIdx.x;
    if (u >= N) return;
    double b, c[3];
    long j[3];
    int nc = 0, f = 0;
    const bool present = bc_row<KIND>(a, B, u, b, nc, j, c, f);
    x[u] = present ? b : 0.0;
    if (present)
        for (int i = 0; i < nc; i++) {
            // a row referring to an earlier row that does not exist: the
            // reference's elimination exits there
            if (j[i] < u) {
                double bj, cj[3];
                long jj[3];
                int ncj = 0, fj = 0;
                if (!bc_row<KIND>(a, B, j[i], bj, ncj, jj, cj, fj)) f |= BCF_ABSENT;
            }
        }
    if (f) atomicOr(flag, f);
}

// One substitution sweep: x_new = b + sum c x_old.
template <int KIND>
__global__ __launch_bounds__(BC_BLOCK) void k_bc_sweep(KArgs a, BcArgs B, long N, const double *__restrict__ xo,
                                                       double *__restrict__ xn, int *__restrict__ flag)
{
    const long u = (long)blockIdx.x * BC_BLOCK + threadIdx.x;
    if (u >= N) return;
    double b, c[3];
    long j[3];
    int nc = 0, f = 0;
    double v = 0.0;
    if (bc_row<KIND>(a, B, u, b, nc, j, c, f)) {
        v = b;
        for (int i = 0; i < nc; i++) v += c[i] * xo[j[i]];
    }
    xn[u] = v;
    if (__double_as_longlong(v) != __double_as_longlong(xo[u])) atomicOr(flag, BCF_CHANGED);
}

// The routine's reload: absent rows leave the reference's zero.
template <int KIND>
__global__ __launch_bounds__(BC_BLOCK) void k_bc_store(KArgs a, long N, const double *__restrict__ x)
{
    const long u = (long)blockIdx.x * BC_BLOCK + threadIdx.x;
    if (u >= N) return;
    const long nn = a.nodenum;
    if (KIND == BC_VXVY) {
        if (u < nn) NOD(a, V2C_N_mvx)[u] = x[u];
        else NOD(a, V2C_N_mvy)[u - nn] = x[u];
    } else if (KIND == BC_P) {
        NOD(a, V2C_N_sppe)[u] = x[u];
    } else {
        NOD(a, V2C_N_tk)[u] = x[u];
    }
}

__global__ void k_bc_clear(int *p)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

}  // namespace v2c

using namespace v2c;

constexpr int BC_MAX_SWEEPS = 64;

template <int KIND>
static int bc_solve(v2c_ctx *c, const char *what)
{
    if (!c) return V2C_EINVAL;
    if (!c->d_bondm) return set_err(c, V2C_EINVAL, "%s: no boundary equations (v2c_set_boundary)", what);
    HIPCHK(c, hipSetDevice(c->device));
    const long N = KIND == BC_VXVY ? 2 * c->nodenum : c->nodenum;
    const unsigned nb = blocks_for(N, BC_BLOCK);
    const KArgs a = make_args(c);
    BcArgs B;
    setp(B.bondm, c->d_bondm);
    setp(B.bondn, c->d_bondn);
    setp(B.bondv, c->d_bondv);
    B.nbond = c->nbond;
    double *x0 = c->d_bcx, *x1 = c->d_bcx + N;
    int *flag = c->d_bcflag;
    int h = 0;
    int rc = launch(c, V2C_K_POST, [&] {
        hipLaunchKernelGGL(k_bc_clear, dim3(1), dim3(64), 0, c->stream, flag);
        hipLaunchKernelGGL(k_bc_init<KIND>, dim3(nb), dim3(BC_BLOCK), 0, c->stream, a, B, N, x0, flag);
    });
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(&h, flag, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (h & BCF_RANGE) return set_err(c, V2C_EINVAL, "%s: boundary equation or parameter out of range", what);
    if (h & BCF_SELF) return set_err(c, V2C_EUNSUPPORTED, "%s: a boundary equation refers to its own parameter", what);
    if (h & BCF_ABSENT)
        return set_err(c, V2C_EUNSUPPORTED, "%s: a boundary equation refers to an absent earlier row "
                                            "(the reference exits)", what);
    int sweeps = 0;
    for (;;) {
        rc = launch(c, V2C_K_POST, [&] {
            hipLaunchKernelGGL(k_bc_clear, dim3(1), dim3(64), 0, c->stream, flag);
            hipLaunchKernelGGL(k_bc_sweep<KIND>, dim3(nb), dim3(BC_BLOCK), 0, c->stream, a, B, N, x0, x1, flag);
        });
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(&h, flag, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        std::swap(x0, x1);
        if (!(h & BCF_CHANGED)) break;
        if (++sweeps >= BC_MAX_SWEEPS)
            return set_err(c, V2C_EUNSUPPORTED, "%s: boundary equations refer to each other in a cycle "
                                                "that does not settle", what);
    }
    rc = launch(c, V2C_K_POST, [&] {
        hipLaunchKernelGGL(k_bc_store<KIND>, dim3(nb), dim3(BC_BLOCK), 0, c->stream, a, N, x0);
    });
    if (rc) return rc;
    return finish(c);
}

extern "C" {

int v2c_set_boundary(v2c_ctx *c, const long *bondm, const long *bondn, const double *bondv, long nbond)
{
    if (!c || !bondm || nbond < 1 || (nbond > 1 && (!bondn || !bondv))) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    if (nbond > c->bond_cap) {
        if (c->d_bondn) (void)hipFree(c->d_bondn);
        if (c->d_bondv) (void)hipFree(c->d_bondv);
        c->d_bondn = nullptr;
        c->d_bondv = nullptr;
        c->bond_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_bondn, sizeof(long) * 3 * nbond));
        HIPCHK(c, hipMalloc(&c->d_bondv, sizeof(double) * 4 * nbond));
        c->bond_cap = nbond;
    }
    if (!c->d_bondm) {
        HIPCHK(c, hipMalloc(&c->d_bondm, sizeof(long) * 4 * c->nodenum));
        HIPCHK(c, hipMalloc(&c->d_bcx, sizeof(double) * 4 * c->nodenum));
        HIPCHK(c, hipMalloc(&c->d_bcflag, sizeof(int) * 4));
    }
    HIPCHK(c, hipMemcpyAsync(c->d_bondm, bondm, sizeof(long) * 4 * c->nodenum, hipMemcpyHostToDevice, c->stream));
    if (bondn) HIPCHK(c, hipMemcpyAsync(c->d_bondn, bondn, sizeof(long) * 3 * nbond, hipMemcpyHostToDevice, c->stream));
    if (bondv) HIPCHK(c, hipMemcpyAsync(c->d_bondv, bondv, sizeof(double) * 4 * nbond, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nbond = nbond;
    return 0;
}

int v2c_vxvy_boundarycorrection(v2c_ctx *c) { return bc_solve<BC_VXVY>(c, "vxvy_boundarycorrection"); }
int v2c_P_boundarycorrection(v2c_ctx *c) { return bc_solve<BC_P>(c, "P_boundarycorrection"); }
int v2c_tkrecalc(v2c_ctx *c) { return bc_solve<BC_T>(c, "tkrecalc"); }

}  // extern "C"
