// v2c_api.hip -- C-ABI of libv2c: context, device memory, ordering and the
// hot-path entry points declared in include/v2c.h.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "v2c_kernels.hip"
#include "v2c_local.hip"

using namespace v2c;

struct TimedLaunch { int kernel; hipEvent_t start, stop; };

struct v2c_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // device address of this device's `scratch` array (noalloc staging)
    unsigned char *scratch = nullptr;
    // staging of the many-field nodal transfers (first use): pinned host and
    // device buffers of STAGE_FIELDS fields each
    double *h_stage = nullptr, *d_stage = nullptr;
    // marker2nodes runs its per-marker property pass here, beside the
    // re-ordering on `stream` (default priority, like `stream`)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int async_mode = 0;
    int write_friction = 1;
    v2c_params prm;
    v2c_materials mat;
    long xnumx = 0, ynumy = 0, nodenum = 0, ncell = 0;
    double xsize = 0, ysize = 0;
    long cap = 0, n = 0;
    double timesum = 0, timestep = 0;
    int key_bits = 1;
    bool order_valid = false;
    bool ever_sorted = false;
    long last_movers = 0, n_full_sorts = 0, n_incremental = 0;
    // device memory
    double *d_gx = nullptr, *d_gy = nullptr;
    Axis4 *d_bx = nullptr, *d_by = nullptr, *d_cellx = nullptr, *d_celly = nullptr;
    double bx0 = 0, by0 = 0, bxinv = 0, byinv = 0;
    double xr_lo = 0, xr_hi = 0, yr_lo = 0, yr_hi = 0;
    int nbx = 0, nby = 0;
    DevMat *d_mat = nullptr;
    double *d_mk = nullptr;
    int8_t *d_t = nullptr;
    uint32_t *d_key = nullptr, *d_skey = nullptr, *d_iota = nullptr, *d_order = nullptr, *d_cs = nullptr;
    // markers per sub-cell of the current order (4 ncell + 1 entries), valid
    // when cnt_valid; the incremental re-ordering keeps them and scans them
    // into the sub-cell starts
    uint32_t *d_cnt = nullptr;
    bool cnt_valid = false;
    void *d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    double *d_nod = nullptr, *d_part = nullptr, *d_red = nullptr, *d_prop = nullptr;
    int *d_counters = nullptr;
    int before_slot = 0;                // compute_before_solve's error-count slot (0 or BEFORE_SLOTS)
    long ord_max = 256;                 // workgroup cap of the re-ordering passes (0: one per tile)
    unsigned timing_mask = ~0u;         // kernels (1 << V2C_K_*) timed while timing is on
    // incremental re-ordering buffers
    uint32_t *d_order2 = nullptr, *d_skey2 = nullptr, *d_tile_count = nullptr, *d_tile_off = nullptr;
    unsigned long long *d_mask = nullptr;
    uint64_t *d_k64_in = nullptr, *d_k64_out = nullptr;
    long *d_nsel = nullptr;
    long *h_nsel = nullptr;   // pinned: mover count read back by the ordering
    void *d_post = nullptr;   // per-block partials of the post-solve routines
    size_t post_bytes = 0;
    // boundary equations (v2c_set_boundary) and the substitution buffers
    long *d_bondm = nullptr, *d_bondn = nullptr;
    double *d_bondv = nullptr, *d_bcx = nullptr;
    int *d_bcflag = nullptr;
    long nbond = 0, bond_cap = 0;
    // local update
    double *d_mx = nullptr;
    uint8_t *d_fault = nullptr;
    uint32_t *d_flist = nullptr;  // fault markers' ids, in increasing order
    long nfault = 0;
    uint32_t *d_defer = nullptr;  // markers the fast fused update hands to the general one
    uint8_t *d_fclass = nullptr;  // constant friction set each marker's fields hold (friction_marker)
    unsigned long long *d_lu = nullptr;
    // warm-start MLP
    float *d_mlp = nullptr, *d_mlp_m = nullptr, *d_mlp_v = nullptr, *d_mlp_norm = nullptr;
    long mlp_steps = 0;
    int mlp_ready = 0;
    void *d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    // timing
    int timing = 0;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    double ktime[V2C_K_COUNT] = {0};
    long klaunch[V2C_K_COUNT] = {0};
    char err[1024] = {0};
};

static int set_err(v2c_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(c, expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return set_err((c), V2C_EHIP, "%s failed: %s (%s:%d)", #expr,                 \
                           hipGetErrorString(e_), __FILE__, __LINE__);                    \
    } while (0)

static hipEvent_t take_event(v2c_ctx *c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Launch helper: records events around the launch when timing is on.
template <typename F>
static int launch(v2c_ctx *c, int kid, F &&fn, hipStream_t st = nullptr)
{
    if (!st) st = c->stream;
    TimedLaunch tl{kid, nullptr, nullptr};
    const bool timed = c->timing && ((c->timing_mask >> kid) & 1u);
    if (timed) {
        tl.start = take_event(c);
        tl.stop = take_event(c);
        if (tl.start) (void)hipEventRecord(tl.start, st);
    }
    hipError_t e = hipSuccess;
    if constexpr (std::is_same_v<decltype(fn()), hipError_t>) e = fn();
    else fn();
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, V2C_EHIP, "kernel %d launch failed: %s", kid, hipGetErrorString(e));
    if (timed && tl.start && tl.stop) {
        (void)hipEventRecord(tl.stop, st);
        c->pending.push_back(tl);
    }
    return 0;
}

static int collect_timing(v2c_ctx *c)
{
    if (c->pending.empty()) return 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &tl : c->pending) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, tl.start, tl.stop);
        c->ktime[tl.kernel] += ms;
        c->klaunch[tl.kernel] += 1;
        c->event_pool.push_back(tl.start);
        c->event_pool.push_back(tl.stop);
    }
    c->pending.clear();
    return 0;
}

// Device pointers into the argument struct (its pointer members are typed as
// global-memory pointers in the device pass).
template <class D, class S>
static inline void setp(D &dst, S *src) { dst = (D)src; }

static KArgs make_args(v2c_ctx *c)
{
    KArgs a;
    setp(a.g.gx, c->d_gx); setp(a.g.gy, c->d_gy);
    a.g.xnumx = (int)c->xnumx; a.g.ynumy = (int)c->ynumy;
    a.g.xsize = c->xsize; a.g.ysize = c->ysize;
    setp(a.g.bx, c->d_bx); setp(a.g.by, c->d_by); setp(a.g.cellx, c->d_cellx); setp(a.g.celly, c->d_celly);
    a.g.bx0 = c->bx0; a.g.by0 = c->by0; a.g.bxinv = c->bxinv; a.g.byinv = c->byinv;
    a.g.nbx = c->nbx; a.g.nby = c->nby;
    a.g.xr_lo = c->xr_lo; a.g.xr_hi = c->xr_hi; a.g.yr_lo = c->yr_lo; a.g.yr_hi = c->yr_hi;
    a.p = c->prm;
    setp(a.mat, c->d_mat);
    setp(a.mk, c->d_mk); a.cap = c->cap; setp(a.t, c->d_t); a.n = c->n;
    setp(a.order, c->d_order); setp(a.skey, c->d_skey); setp(a.sub_start, c->d_cs); setp(a.key, c->d_key);
    setp(a.nod, c->d_nod); a.nodenum = c->nodenum; a.ncell = c->ncell;
    setp(a.part, c->d_part);
    setp(a.prop, c->d_prop);
    a.timesum = c->timesum; a.timestep = c->timestep;
    setp(a.counters, c->d_counters);
    a.write_friction = c->write_friction;
    setp(a.fclass, c->d_fclass);
    setp(a.mx, c->d_mx);
    setp(a.fault, c->d_fault);
    setp(a.lu_stats, c->d_lu);
    setp(a.fault_list, c->d_flist);
    a.nfault = c->nfault;
    setp(a.mlp, c->d_mlp);
    setp(a.mlp_norm, c->d_mlp_norm);
    return a;
}

static int finish(v2c_ctx *c)
{
    if (!c->async_mode) HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Workgroups of the re-ordering passes (each loops over its share of the
// tiles): at most ord_max, so that beside the property pass they hold fewer
// of the CUs' wave slots (production step: 3.08 ms with one workgroup per
// tile, 3.05 with 256 or 512, 3.12 with 128 and 3.52 with 64, where the
// re-ordering outlasts the property pass; with the sub-cell starts scanned
// from counts: 3.002 ms with 512, 2.996 with 256, 3.042 with 128).
static inline unsigned ord_grid(const v2c_ctx *c, long ntiles)
{
    return (unsigned)(c->ord_max > 0 && ntiles > c->ord_max ? c->ord_max : ntiles);
}

// Sub-cell starts of the current order: from the keys in a pass over the
// order (and the per-sub-cell counts from them), or, where the counts are
// valid (kept by the incremental re-ordering), as their exclusive scan.
static int cell_starts(v2c_ctx *c)
{
    const long nsub = 4 * c->ncell;
    if (c->cnt_valid)
        return launch(c, V2C_K_SORT, [&] {
            return hipcub::DeviceScan::ExclusiveSum(c->d_scan_tmp, c->scan_tmp_bytes, c->d_cnt, c->d_cs,
                                                    (int)(nsub + 1), c->stream);
        });
    int rc = launch(c, V2C_K_SORT, [&] {
        hipLaunchKernelGGL(k_sub_start, dim3(ord_grid(c, blocks_for(c->n + 1, ORD_TILE))), dim3(ORD_BLOCK), 0, c->stream,
                           c->d_skey, c->n, nsub, c->d_cs);
        hipLaunchKernelGGL(k_sub_count, dim3(blocks_for(nsub + 1, 256)), dim3(256), 0, c->stream, c->d_cs, nsub,
                           c->d_cnt);
    });
    if (rc) return rc;
    c->cnt_valid = true;
    return 0;
}

// Full stable radix sort of (sub-cell key, marker id) -> order[], skey[].
static int full_sort(v2c_ctx *c)
{
    if (c->n > 0) {
        int rc = launch(c, V2C_K_SORT, [&] {
            return hipcub::DeviceRadixSort::SortPairs(c->d_sort_tmp, c->sort_tmp_bytes, c->d_key, c->d_skey,
                                                      c->d_iota, c->d_order, (int)c->n, 0, c->key_bits,
                                                      c->stream);
        });
        if (rc) return rc;
    }
    c->n_full_sorts++;
    c->ever_sorted = true;
    c->cnt_valid = false;
    return cell_starts(c);
}

// Bring order[]/skey[]/sub_start[] up to date with the current keys.  After
// an advection only markers whose key changed are re-sorted and merged in
// (k_order_flags finds them); the result is the same (key, id) order a full
// sort gives.  The work is split in two halves around the one read-back (the
// mover count), so that several contexts can have their first halves in
// flight together: order_begin launches the flagging and the count copy and
// sets *pending; order_finish, once the stream has reached the copy, launches
// the rest.
static int order_begin(v2c_ctx *c, bool *pending)
{
    *pending = false;
    if (c->order_valid) return 0;
    if (!c->ever_sorted || c->n == 0) {
        int rc = full_sort(c);
        if (rc) return rc;
        c->last_movers = c->n;
        c->order_valid = true;
        return 0;
    }
    const long n = c->n;
    const long ntiles = (n + ORD_TILE - 1) / ORD_TILE;
    int rc = launch(c, V2C_K_SORT, [&] {
        hipLaunchKernelGGL(k_order_flags, dim3(ord_grid(c, ntiles)), dim3(ORD_BLOCK), 0, c->stream, c->d_order,
                           c->d_key, c->d_skey, n, c->d_mask, c->d_tile_count);
        hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, c->stream, c->d_tile_count, ntiles,
                           c->d_tile_off, c->d_nsel);
    });
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->h_nsel, c->d_nsel, sizeof(long), hipMemcpyDeviceToHost, c->stream));
    *pending = true;
    return 0;
}

static int order_finish(v2c_ctx *c)
{
    const long n = c->n;
    const long ntiles = (n + ORD_TILE - 1) / ORD_TILE;
    const long nmov = *c->h_nsel;
    int rc;
    c->last_movers = nmov;
    if (nmov == 0) {
        c->order_valid = true;
        return 0;
    }
    if (nmov > n / 4) {
        rc = full_sort(c);
        if (rc) return rc;
        c->order_valid = true;
        return 0;
    }
    // the counts follow the movers only if they describe the old order
    uint32_t *cnt = c->cnt_valid ? c->d_cnt : nullptr;
    rc = launch(c, V2C_K_SORT, [&] {
        hipLaunchKernelGGL(k_order_movers, dim3(ord_grid(c, ntiles)), dim3(ORD_BLOCK), 0, c->stream, c->d_order,
                           c->d_key, n, c->d_mask, c->d_tile_off, c->d_k64_in, c->d_skey, cnt);
        hipError_t e = hipcub::DeviceRadixSort::SortKeys(c->d_sort_tmp, c->sort_tmp_bytes, c->d_k64_in,
                                                         c->d_k64_out, (int)nmov, 0, 32 + c->key_bits, c->stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_order_stayers, dim3(ord_grid(c, ntiles)), dim3(ORD_BLOCK), 0, c->stream, c->d_order,
                           c->d_skey, n, c->d_mask, c->d_tile_off, c->d_k64_out, nmov, c->d_order2, c->d_skey2);
        hipLaunchKernelGGL(k_order_place_movers, dim3(blocks_for(nmov, 256)), dim3(256), 0, c->stream,
                           c->d_k64_out, nmov, c->d_order, c->d_skey, n, c->d_cs, 4 * c->ncell, c->d_mask,
                           c->d_tile_off, c->d_order2, c->d_skey2);
        return hipSuccess;
    });
    if (rc) return rc;
    std::swap(c->d_order, c->d_order2);
    std::swap(c->d_skey, c->d_skey2);
    c->n_incremental++;
    rc = cell_starts(c);
    if (rc) return rc;
    c->order_valid = true;
    return 0;
}

static int ensure_order(v2c_ctx *c)
{
    bool pending = false;
    int rc = order_begin(c, &pending);
    if (rc || !pending) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return order_finish(c);
}

extern "C" {

void v2c_params_default(v2c_params *p)
{
    memset(p, 0, sizeof *p);
    p->markmod = 2; p->stoksmod = 1; p->outgrid = 2; p->densimod = 0; p->interpolationmod = 2;
    p->antidiffinter = 1; p->rkmod = 0; p->RSFmode = 1; p->power_law_RSF = 0; p->reloadfromDP = 0;
    p->compressiblemode = 0; p->stickymod = 1; p->inertyn = 0; p->pmode = 0; p->tempmod = 0;
    p->n2m_phantom = 1;
    p->nubeg = 1e1; p->nuend = 1.667e27; p->tdeep = 1873.0; p->zdeep = 3000000.0; p->incoming_T = 0.0;
    p->nstress = 10.0e6; p->Dphis = 0.0; p->half_range = 2.0e3;
    p->start_updip = 10.0e3 - p->half_range; p->end_updip = 10.0e3 + p->half_range;
    p->start_downdip = 40.0e3 - p->half_range; p->end_downdip = 40.0e3 + p->half_range;
    p->a_vw = 0.009; p->b_vw = 0.017; p->L_vw = 0.0100; p->V0_vw = 4.0e-09;
    p->a_vs = 0.020; p->b_vs = 0.001; p->L_vs = 0.0100; p->V0_vs = 4.0e-09;
    p->a_hr = 0.009; p->b_hr = 0.017; p->L_hr = 0.0100; p->V0_hr = 4.0e-09;
    p->xstpx = 0.0;
    p->faultwidth = 400.0; p->rad_damping = 4.3303e+06; p->bulk_modulus = 50e9;
    p->band_y = 10.000e3; p->band_halfwidth = 100.0;
}

static const char *const kNodalNames[V2C_NF] = {
    "pr", "vx", "vy", "exx", "exy", "esp", "sxx", "sxy", "sxxs", "sppes", "exxpl", "exxel",
    "bm", "ro", "et", "nu", "nd", "gg", "gd", "sppe", "phis", "phin", "sxyen", "sxxen", "exyen",
    "exxen", "sppen", "mus_s", "as", "bs", "Ls", "V0s", "sxxe", "sxye", "exxe", "exye", "cohesion",
    "dro", "drp", "cp", "kt", "ht", "tk", "mrx", "mry", "mvx", "mvy", "pe", "pen", "nu_m2c",
    "nd_m2c", "mvxn", "mvyn",
    "phie", "sxyee", "exyee", "pee", "sppee", "sxxee", "exxee", "mvxe", "mvye",
    "wtphisum", "phisum", "sxysum", "exysum", "esum", "wtsxxsum", "psum", "sxxsum", "exxsum",
    "wtvxsum", "vxsum", "wtvysum", "vysum", "mrx0", "mry0",
    "pes", "phis_it", "syields", "nuvp_it", "sliprs", "RK_sliprs_final", "RK_phidts_final",
    "RK_sliprs_solved", "RK_phidts_solved", "RK_deltas", "RK_sliprs0", "RK_sliprs1",
    "RK_phidts0", "RK_phidts1",
    "deltasxx", "deltasxy", "deltaVs", "exxs", "exxes", "sxxes", "prs",
    "tk2", "tol0", "tol1", "sticky01"};

const char *v2c_nodal_name(int field)
{
    return (field >= 0 && field < V2C_NF) ? kNodalNames[field] : nullptr;
}

const char *v2c_last_error(const v2c_ctx *c) { return c ? c->err : "null context"; }

// ---- static staging of v2c_ductilevisc_batch_noalloc ------------------------
// Five double columns (sxy, sxx, tk, pb, result), the iteration counts and
// the types, each column V2C_DUCTILEVISC_NOALLOC_MAX entries long: 45 bytes
// per input, 11.25 MiB per device, held until the process ends.  In host
// code `scratch` names a host-side stand-in, not device memory: the device
// address of each device's copy comes from hipGetSymbolAddress, taken once
// per device at context creation (scratch_address) and kept in
// g_scratch[device]; only that address reaches kernels and copies.
constexpr size_t NOALLOC_COL = (size_t)V2C_DUCTILEVISC_NOALLOC_MAX * sizeof(double);
constexpr size_t NOALLOC_BYTES = 5 * NOALLOC_COL + (size_t)V2C_DUCTILEVISC_NOALLOC_MAX * (sizeof(int) + 1);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on `device`, which must be current.
static hipError_t scratch_address(int device, unsigned char **out)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[device] = (unsigned char *)p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

static void free_all(v2c_ctx *c)
{
    void *ptrs[] = {c->d_gx, c->d_gy, c->d_bx, c->d_by, c->d_cellx, c->d_celly, c->d_mat, c->d_mk, c->d_t, c->d_key, c->d_skey, c->d_iota,
                    c->d_order, c->d_cs, c->d_nod, c->d_part, c->d_prop, c->d_red, c->d_counters, c->d_sort_tmp,
                    c->d_order2, c->d_skey2, c->d_tile_count, c->d_tile_off, c->d_mask,
                    c->d_k64_in, c->d_k64_out, c->d_nsel, c->d_mx, c->d_fault, c->d_lu, c->d_mlp,
                    c->d_mlp_m, c->d_mlp_v, c->d_mlp_norm, c->d_post, c->d_bondm, c->d_bondn,
                    c->d_bondv, c->d_bcx, c->d_bcflag, c->d_defer, c->d_fclass, c->d_flist, c->d_cnt,
                    c->d_scan_tmp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &tl : c->pending) { (void)hipEventDestroy(tl.start); (void)hipEventDestroy(tl.stop); }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    if (c->h_nsel) (void)hipHostFree(c->h_nsel);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->d_stage) (void)hipFree(c->d_stage);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

int v2c_create(v2c_ctx **out, const v2c_grid *g, const v2c_params *prm, const v2c_materials *mat,
               long capacity, int device)
{
    if (!out || !g || !prm || !mat) return V2C_EINVAL;
    *out = nullptr;
    // per-field arrays are addressed with 32-bit byte offsets on the device
    if (g->xnumx < 4 || g->ynumy < 4 || !g->gx || !g->gy || capacity < 0 || capacity >= (1L << 29) ||
        g->xnumx * g->ynumy >= (1L << 29))
        return V2C_EINVAL;
    if (prm->interpolationmod != 2) return V2C_EUNSUPPORTED;
    if (mat->ntypes < 0 || mat->ntypes > V2C_MAX_TYPES) return V2C_EINVAL;
    for (long i = 1; i < g->xnumx; i++) if (!(g->gx[i] > g->gx[i - 1])) return V2C_EINVAL;
    for (long i = 1; i < g->ynumy; i++) if (!(g->gy[i] > g->gy[i - 1])) return V2C_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return V2C_ENOGPU;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return V2C_ENOGPU;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return V2C_ENOGPU;

    v2c_ctx *c = new v2c_ctx();
    c->device = device;
    c->prm = *prm;
    c->mat = *mat;
    c->xnumx = g->xnumx; c->ynumy = g->ynumy; c->nodenum = g->xnumx * g->ynumy;
    c->ncell = (g->xnumx - 1) * (g->ynumy - 1);
    c->xsize = g->xsize; c->ysize = g->ysize;
    c->cap = capacity > 0 ? capacity : 1;
    while ((1L << c->key_bits) < 4 * c->ncell) c->key_bits++;
    if (c->prm.xstpx == 0.0) c->prm.xstpx = g->xsize / (double)(g->xnumx - 1);
#define CK(expr)                                                                      \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            free_all(c);                                                              \
            delete c;                                                                 \
            return e_ == hipErrorOutOfMemory ? V2C_ENOMEM : V2C_EHIP;                 \
        }                                                                             \
    } while (0)
    CK(hipSetDevice(device));
    CK(scratch_address(device, &c->scratch));
    CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    CK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    CK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    CK(hipMalloc(&c->d_gx, sizeof(double) * g->xnumx));
    CK(hipMalloc(&c->d_gy, sizeof(double) * g->ynumy));
    CK(hipMalloc(&c->d_mat, sizeof(DevMat)));
    CK(hipMalloc(&c->d_mk, sizeof(double) * NMF * c->cap));
    CK(hipMalloc(&c->d_t, c->cap));
    CK(hipMalloc(&c->d_key, sizeof(uint32_t) * c->cap));
    CK(hipMalloc(&c->d_skey, sizeof(uint32_t) * c->cap));
    CK(hipMalloc(&c->d_iota, sizeof(uint32_t) * c->cap));
    CK(hipMalloc(&c->d_order, sizeof(uint32_t) * c->cap));
    CK(hipMalloc(&c->d_cs, sizeof(uint32_t) * (4 * c->ncell + 1)));
    CK(hipMalloc(&c->d_cnt, sizeof(uint32_t) * (4 * c->ncell + 1)));
    CK(hipMalloc(&c->d_nod, sizeof(double) * V2C_NF * c->nodenum));
    CK(hipMalloc(&c->d_part, sizeof(double) * 4 * M_NF * c->nodenum));
    CK(hipMalloc(&c->d_prop, sizeof(double) * NPROP * c->cap));
    CK(hipMalloc(&c->d_red, sizeof(double) * (2 * blocks_for(c->nodenum, RED_BLOCK) + 8)));
    CK(hipMalloc(&c->d_counters, sizeof(int) * 8));
    CK(hipMalloc(&c->d_order2, sizeof(uint32_t) * c->cap));
    CK(hipMalloc(&c->d_skey2, sizeof(uint32_t) * c->cap));
    {
        const long ntiles = (c->cap + ORD_TILE - 1) / ORD_TILE;
        CK(hipMalloc(&c->d_tile_count, sizeof(uint32_t) * (ntiles + 1)));
        CK(hipMalloc(&c->d_tile_off, sizeof(uint32_t) * (ntiles + 1)));
        CK(hipMalloc(&c->d_mask, sizeof(unsigned long long) * (ntiles * ORD_MASKS + 1)));
    }
    CK(hipMalloc(&c->d_k64_in, sizeof(uint64_t) * c->cap));
    CK(hipMalloc(&c->d_k64_out, sizeof(uint64_t) * c->cap));
    CK(hipMalloc(&c->d_nsel, sizeof(long) * 2));
    CK(hipHostMalloc((void **)&c->h_nsel, sizeof(long), hipHostMallocDefault));
    CK(hipMalloc(&c->d_mx, sizeof(double) * NMX * c->cap));
    CK(hipMalloc(&c->d_fault, c->cap));
    CK(hipMalloc(&c->d_defer, sizeof(uint32_t) * c->cap));
    CK(hipMalloc(&c->d_flist, sizeof(uint32_t) * c->cap));
    CK(hipMalloc(&c->d_fclass, c->cap));
    CK(hipMemsetAsync(c->d_fclass, 0, c->cap, c->stream));
    CK(hipMalloc(&c->d_lu, sizeof(unsigned long long) * LS_COUNT));
    CK(hipMemsetAsync(c->d_mx, 0, sizeof(double) * NMX * c->cap, c->stream));
    CK(hipMemsetAsync(c->d_fault, 0, c->cap, c->stream));
    CK(hipMalloc(&c->d_mlp, sizeof(float) * MLP_NPARAM));
    CK(hipMalloc(&c->d_mlp_m, sizeof(float) * MLP_NPARAM));
    CK(hipMalloc(&c->d_mlp_v, sizeof(float) * MLP_NPARAM));
    CK(hipMalloc(&c->d_mlp_norm, sizeof(float) * MLP_NNORM));
    CK(hipMemsetAsync(c->d_mlp, 0, sizeof(float) * MLP_NPARAM, c->stream));
    CK(hipMemsetAsync(c->d_mlp_m, 0, sizeof(float) * MLP_NPARAM, c->stream));
    CK(hipMemsetAsync(c->d_mlp_v, 0, sizeof(float) * MLP_NPARAM, c->stream));
    CK(hipMemsetAsync(c->d_nod, 0, sizeof(double) * V2C_NF * c->nodenum, c->stream));
    CK(hipMemsetAsync(c->d_mk, 0, sizeof(double) * NMF * c->cap, c->stream));
    CK(hipMemsetAsync(c->d_t, 0, c->cap, c->stream));
    CK(hipMemsetAsync(c->d_counters, 0, sizeof(int) * 8, c->stream));
    CK(hipMemcpyAsync(c->d_gx, g->gx, sizeof(double) * g->xnumx, hipMemcpyHostToDevice, c->stream));
    CK(hipMemcpyAsync(c->d_gy, g->gy, sizeof(double) * g->ynumy, hipMemcpyHostToDevice, c->stream));
    {
        std::vector<Axis4> bx, by, cx, cy;
        make_tables(g->gx, (int)g->xnumx, bx, cx, c->bx0, c->bxinv);
        make_tables(g->gy, (int)g->ynumy, by, cy, c->by0, c->byinv);
        c->nbx = (int)bx.size();
        c->nby = (int)by.size();
        const long nx = g->xnumx, ny = g->ynumy;
        c->xr_lo = (g->gx[0] + g->gx[1]) / 2.0; c->xr_hi = (g->gx[nx - 2] + g->gx[nx - 1]) / 2.0;
        c->yr_lo = (g->gy[0] + g->gy[1]) / 2.0; c->yr_hi = (g->gy[ny - 2] + g->gy[ny - 1]) / 2.0;
        CK(hipMalloc(&c->d_bx, sizeof(Axis4) * bx.size()));
        CK(hipMalloc(&c->d_by, sizeof(Axis4) * by.size()));
        CK(hipMalloc(&c->d_cellx, sizeof(Axis4) * cx.size()));
        CK(hipMalloc(&c->d_celly, sizeof(Axis4) * cy.size()));
        CK(hipMemcpyAsync(c->d_bx, bx.data(), sizeof(Axis4) * bx.size(), hipMemcpyHostToDevice, c->stream));
        CK(hipMemcpyAsync(c->d_by, by.data(), sizeof(Axis4) * by.size(), hipMemcpyHostToDevice, c->stream));
        CK(hipMemcpyAsync(c->d_cellx, cx.data(), sizeof(Axis4) * cx.size(), hipMemcpyHostToDevice, c->stream));
        CK(hipMemcpyAsync(c->d_celly, cy.data(), sizeof(Axis4) * cy.size(), hipMemcpyHostToDevice, c->stream));
        CK(hipStreamSynchronize(c->stream));
    }
    {
        DevMat dm;
        memcpy(dm.nu, mat->nu, sizeof dm.nu); memcpy(dm.dh, mat->dh, sizeof dm.dh);
        memcpy(dm.dv, mat->dv, sizeof dm.dv); memcpy(dm.ss, mat->ss, sizeof dm.ss);
        memcpy(dm.mm, mat->mm, sizeof dm.mm); memcpy(dm.gg, mat->gg, sizeof dm.gg);
        memcpy(dm.n0, mat->n0, sizeof dm.n0); memcpy(dm.n1, mat->n1, sizeof dm.n1);
        memcpy(dm.ro, mat->ro, sizeof dm.ro); memcpy(dm.bb, mat->bb, sizeof dm.bb);
        memcpy(dm.aa, mat->aa, sizeof dm.aa); memcpy(dm.cp, mat->cp, sizeof dm.cp);
        memcpy(dm.kt, mat->kt, sizeof dm.kt); memcpy(dm.kf, mat->kf, sizeof dm.kf);
        memcpy(dm.kp, mat->kp, sizeof dm.kp); memcpy(dm.ht, mat->ht, sizeof dm.ht);
        memcpy(dm.im, mat->im, sizeof dm.im);
        for (int t = 0; t < V2C_MAX_TYPES; t++) {
            dm.ssn1[t] = pow(mat->ss[t], mat->mm[t] - 1.0);
            dm.k1[t] = mat->nu[t] / dm.ssn1[t];
            dm.rn[t] = 1.0 / mat->mm[t];
            dm.ex1[t] = (mat->mm[t] - 1) / mat->mm[t];
        }
        CK(hipMemcpyAsync(c->d_mat, &dm, sizeof dm, hipMemcpyHostToDevice, c->stream));
        CK(hipStreamSynchronize(c->stream));
    }
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(c->cap, 256)), dim3(256), 0, c->stream, c->d_iota, c->cap);
    CK(hipGetLastError());
    {
        size_t b1 = 0, b3 = 0;
        CK(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, c->d_key, c->d_skey, c->d_iota, c->d_order,
                                              (int)c->cap, 0, c->key_bits, c->stream));
        CK(hipcub::DeviceRadixSort::SortKeys(nullptr, b3, c->d_k64_in, c->d_k64_out, (int)c->cap, 0,
                                             32 + c->key_bits, c->stream));
        c->sort_tmp_bytes = std::max(b1, b3);
        CK(hipcub::DeviceScan::ExclusiveSum(nullptr, c->scan_tmp_bytes, c->d_cnt, c->d_cs, (int)(4 * c->ncell + 1),
                                            c->stream));
    }
    CK(hipMalloc(&c->d_sort_tmp, c->sort_tmp_bytes > 0 ? c->sort_tmp_bytes : 16));
    CK(hipMalloc(&c->d_scan_tmp, c->scan_tmp_bytes > 0 ? c->scan_tmp_bytes : 16));
    CK(hipStreamSynchronize(c->stream));
#undef CK
    *out = c;
    return 0;
}

void v2c_destroy(v2c_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
}

int v2c_set_async(v2c_ctx *c, int on)
{
    if (!c) return V2C_EINVAL;
    c->async_mode = on ? 1 : 0;
    return 0;
}

int v2c_set_friction_fields(v2c_ctx *c, int on)
{
    if (!c) return V2C_EINVAL;
    c->write_friction = on ? 1 : 0;
    return 0;
}

int v2c_synchronize(v2c_ctx *c)
{
    if (!c) return V2C_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int v2c_set_time(v2c_ctx *c, double timesum, double timestep)
{
    if (!c) return V2C_EINVAL;
    c->timesum = timesum;
    c->timestep = timestep;
    return 0;
}

long v2c_marker_count(const v2c_ctx *c) { return c ? c->n : -1; }

static double *host_field(const v2c_markers *m, int f)
{
    double *const p[NMF] = {m->x, m->y, m->k, m->p, m->xx, m->xy, m->exx, m->exy, m->e, m->phi, m->vx,
                            m->vy, m->C, m->d, m->w, m->a, m->b, m->L, m->L0, m->V0, m->fric_static, m->C0};
    return p[f];
}

int v2c_markers_upload(v2c_ctx *c, const v2c_markers *m, long n)
{
    if (!c || !m || n < 0) return V2C_EINVAL;
    if (n > c->cap) return set_err(c, V2C_ECAPACITY, "marker count %ld above capacity %ld", n, c->cap);
    for (int f = 0; f < NMF; f++) if (!host_field(m, f)) return set_err(c, V2C_EINVAL, "marker field %d is NULL", f);
    if (!m->t) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    for (int f = 0; f < NMF; f++)
        HIPCHK(c, hipMemcpyAsync(c->d_mk + (size_t)f * c->cap, host_field(m, f), sizeof(double) * n,
                                 hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_t, m->t, n, hipMemcpyHostToDevice, c->stream));
    // the friction fields came from the host: none is known to hold a constant set
    HIPCHK(c, hipMemsetAsync(c->d_fclass, 0, n, c->stream));
    c->n = n;
    KArgs a = make_args(c);
    if (n > 0) {
        hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, a);
        HIPCHK(c, hipGetLastError());
    }
    c->order_valid = false;
    c->ever_sorted = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int v2c_order_stats(v2c_ctx *c, long *last_movers, long *full_sorts, long *incremental)
{
    if (!c) return V2C_EINVAL;
    if (last_movers) *last_movers = c->last_movers;
    if (full_sorts) *full_sorts = c->n_full_sorts;
    if (incremental) *incremental = c->n_incremental;
    return 0;
}

int v2c_order_get(v2c_ctx *c, unsigned int *order, unsigned int *skey, long n)
{
    if (!c || n < 0 || n > c->n) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = ensure_order(c);
    if (rc) return rc;
    if (n > 0 && order)
        HIPCHK(c, hipMemcpyAsync(order, c->d_order, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    if (n > 0 && skey)
        HIPCHK(c, hipMemcpyAsync(skey, c->d_skey, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int v2c_markers_download(v2c_ctx *c, v2c_markers *m, long n)
{
    if (!c || !m || n < 0) return V2C_EINVAL;
    if (n > c->n) return set_err(c, V2C_EINVAL, "download of %ld markers, %ld resident", n, c->n);
    HIPCHK(c, hipSetDevice(c->device));
    for (int f = 0; f < NMF; f++) {
        double *h = host_field(m, f);
        if (!h) return V2C_EINVAL;
        HIPCHK(c, hipMemcpyAsync(h, c->d_mk + (size_t)f * c->cap, sizeof(double) * n, hipMemcpyDeviceToHost,
                                 c->stream));
    }
    if (m->t) HIPCHK(c, hipMemcpyAsync(m->t, c->d_t, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int v2c_nodal_set(v2c_ctx *c, int field, const double *host)
{
    if (!c || !host || field < 0 || field >= V2C_NF) return V2C_EINVAL;
    HIPCHK(c, hipMemcpyAsync(c->d_nod + (size_t)field * c->nodenum, host, sizeof(double) * c->nodenum,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int v2c_nodal_get(v2c_ctx *c, int field, double *host)
{
    if (!c || !host || field < 0 || field >= V2C_NF) return V2C_EINVAL;
    HIPCHK(c, hipMemcpyAsync(host, c->d_nod + (size_t)field * c->nodenum, sizeof(double) * c->nodenum,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// ---- many nodal fields per transfer ----------------------------------------
// The host arrays are packed into one pinned buffer by a few host threads,
// cross PCIe in one copy, and a kernel scatters the rows into the nodal
// fields (and the reverse for reads), STAGE_FIELDS fields at a time.
constexpr int STAGE_FIELDS = 24;
struct FieldIds { int f[STAGE_FIELDS]; };

__global__ void k_stage_to_nodal(double *__restrict__ nod, const double *__restrict__ stage, FieldIds ids, int k,
                                 long nodenum)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (r < k && i < nodenum) nod[(size_t)ids.f[r] * nodenum + i] = stage[(size_t)r * nodenum + i];
}

__global__ void k_nodal_to_stage(const double *__restrict__ nod, double *__restrict__ stage, FieldIds ids, int k,
                                 long nodenum)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (r < k && i < nodenum) stage[(size_t)r * nodenum + i] = nod[(size_t)ids.f[r] * nodenum + i];
}

static int stage_ready(v2c_ctx *c)
{
    if (c->h_stage) return 0;
    const size_t bytes = sizeof(double) * STAGE_FIELDS * c->nodenum;
    HIPCHK(c, hipHostMalloc((void **)&c->h_stage, bytes, hipHostMallocDefault));
    const hipError_t e = hipMalloc(&c->d_stage, bytes);
    if (e != hipSuccess) {
        (void)hipHostFree(c->h_stage);
        c->h_stage = nullptr;
        return set_err(c, e == hipErrorOutOfMemory ? V2C_ENOMEM : V2C_EHIP, "staging: %s", hipGetErrorString(e));
    }
    return 0;
}

// Row copies between caller arrays and the pinned buffer on up to 8 threads.
static void stage_copy(double *const *dst, const double *const *src, int k, long nodenum)
{
    const size_t bytes = sizeof(double) * nodenum;
    const int nt = std::max(1, std::min({k, 8, (int)std::thread::hardware_concurrency()}));
    if (nt == 1) {
        for (int r = 0; r < k; r++) memcpy(dst[r], src[r], bytes);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
        pool.emplace_back([=] {
            for (int r = t; r < k; r += nt) memcpy(dst[r], src[r], bytes);
        });
    for (auto &th : pool) th.join();
}

static int check_fields(const int *fields, int n)
{
    if (!fields || n < 0) return V2C_EINVAL;
    for (int i = 0; i < n; i++)
        if (fields[i] < 0 || fields[i] >= V2C_NF) return V2C_EINVAL;
    return 0;
}

int v2c_nodal_set_many(v2c_ctx *c, const int *fields, const double *const *host, int n)
{
    if (!c || !host || check_fields(fields, n)) return V2C_EINVAL;
    for (int i = 0; i < n; i++) if (!host[i]) return V2C_EINVAL;
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = stage_ready(c);
    if (rc) return rc;
    const long nn = c->nodenum;
    for (int b = 0; b < n; b += STAGE_FIELDS) {
        const int k = std::min(STAGE_FIELDS, n - b);
        // the previous chunk's copy out of the pinned buffer has finished
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double *rows[STAGE_FIELDS];
        FieldIds ids{};
        for (int r = 0; r < k; r++) {
            rows[r] = c->h_stage + (size_t)r * nn;
            ids.f[r] = fields[b + r];
        }
        stage_copy(rows, host + b, k, nn);
        HIPCHK(c, hipMemcpyAsync(c->d_stage, c->h_stage, sizeof(double) * k * nn, hipMemcpyHostToDevice,
                                 c->stream));
        hipLaunchKernelGGL(k_stage_to_nodal, dim3((unsigned)blocks_for(nn, 256), (unsigned)k), dim3(256), 0,
                           c->stream, c->d_nod, c->d_stage, ids, k, nn);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int v2c_nodal_get_many(v2c_ctx *c, const int *fields, double *const *host, int n)
{
    if (!c || !host || check_fields(fields, n)) return V2C_EINVAL;
    for (int i = 0; i < n; i++) if (!host[i]) return V2C_EINVAL;
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = stage_ready(c);
    if (rc) return rc;
    const long nn = c->nodenum;
    for (int b = 0; b < n; b += STAGE_FIELDS) {
        const int k = std::min(STAGE_FIELDS, n - b);
        FieldIds ids{};
        const double *rows[STAGE_FIELDS];
        for (int r = 0; r < k; r++) {
            rows[r] = c->h_stage + (size_t)r * nn;
            ids.f[r] = fields[b + r];
        }
        hipLaunchKernelGGL(k_nodal_to_stage, dim3((unsigned)blocks_for(nn, 256), (unsigned)k), dim3(256), 0,
                           c->stream, c->d_nod, c->d_stage, ids, k, nn);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(c->h_stage, c->d_stage, sizeof(double) * k * nn, hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        stage_copy(host + b, rows, k, nn);
    }
    return 0;
}

int v2c_marker2nodes(v2c_ctx *c)
{
    if (!c) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    // The property pass reads markers in physical order and writes only the
    // property hand-over and marker phi, which the re-ordering never touches:
    // it runs on the side stream while `stream` brings the order up to date,
    // and the scatter waits for both.
    int rc = 0;
    if (c->n > 0) {
        const KArgs a = make_args(c);
        const int nb = (int)blocks_for(c->n, 256);
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
        rc = launch(c, V2C_K_M2N_PROPS, [&] {
            hipLaunchKernelGGL(k_m2n_props, dim3(nb), dim3(256), 0, c->side, a, nb);
        }, c->side);
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(c->ev_join, c->side));
    }
    rc = ensure_order(c);
    if (c->n > 0) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    if (rc) return rc;
    KArgs a = make_args(c);
    const int nwg = (int)((c->ncell + CPW - 1) / CPW);
    rc = launch(c, V2C_K_M2N_CELLS, [&] {
        hipLaunchKernelGGL(k_m2n_cells, dim3(nwg), dim3(WAVE), 0, c->stream, a, nwg);
    });
    if (rc) return rc;
    rc = launch(c, V2C_K_M2N_NODES, [&] {
        hipLaunchKernelGGL(k_m2n_nodes, dim3(blocks_for(c->nodenum, 256)), dim3(256), 0, c->stream, a);
    });
    if (rc) return rc;
    if (c->prm.Dphis > 0) {
        rc = launch(c, V2C_K_M2N_NODES, [&] {
            hipLaunchKernelGGL(k_phis_bc, dim3(blocks_for(c->nodenum, 256)), dim3(256), 0, c->stream, a);
        });
        if (rc) return rc;
    }
    return finish(c);
}

static int read_counter(v2c_ctx *c, int idx, int *val)
{
    HIPCHK(c, hipMemcpyAsync(val, c->d_counters + idx, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int n2m_impl(v2c_ctx *c, bool advect)
{
    if (!c) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const bool anti = c->prm.antidiffinter == 1;
    int rc = 0;
    if (anti) {
        rc = ensure_order(c);
        if (rc) return rc;
    }
    KArgs a = make_args(c);
    // with antidiffusion the scatter clears the update's counters itself
    if (advect && !anti) HIPCHK(c, hipMemsetAsync(c->d_counters + 1, 0, sizeof(int), c->stream));
    // The antidiffusion scatter reads marker positions and types and the
    // solve's nodal fields; its node pass writes the sums and the *ee fields.
    // The marker update reads neither and, without advection, writes no
    // position, so then the two run side by side; with advection the scatter
    // goes first (it reads positions before any marker moves).
    const bool side = anti && !advect && c->n > 0;
    hipStream_t st = side ? c->side : c->stream;
    if (side) {
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
    }
    // on an error after the fork the main stream is still ordered after
    // whatever reached the side stream (it writes the antidiffusion sums)
    auto join_side = [&](int code) {
        if (side) {
            (void)hipEventRecord(c->ev_join, c->side);
            (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
        }
        return code;
    };
    if (anti) {
        const int nwg = (int)((c->ncell + CPW - 1) / CPW);
        const int pre = (c->prm.n2m_phantom ? N2M_PRE_PHANTOM : 0) | (advect ? N2M_PRE_CLEAR : 0);
        rc = launch(c, V2C_K_N2M_CELLS, [&] {
            hipLaunchKernelGGL(k_n2m_scatter, dim3(nwg), dim3(WAVE), 0, st, a, nwg, pre);
        }, st);
        if (rc) return join_side(rc);
        rc = launch(c, V2C_K_N2M_NODES, [&] {
            hipLaunchKernelGGL(k_n2m_nodes, dim3(blocks_for(c->nodenum, 256)), dim3(256), 0, st, a);
        }, st);
        if (rc) return join_side(rc);
    }
    if (side) {
        const hipError_t je = hipEventRecord(c->ev_join, c->side);
        if (je != hipSuccess) {
            (void)hipStreamSynchronize(c->side);
            return set_err(c, V2C_EHIP, "hipEventRecord: %s", hipGetErrorString(je));
        }
    }
    if (c->n > 0) {
        const int nb = (int)blocks_for(c->n, 256);
        // With antidiffusion and advection the fast form takes every marker
        // whose RK stages stay in one stencil configuration and lists the
        // others for the general form.
        // (the scatter cleared its deferred-marker count)
        const bool fast = advect && anti;
        if (!rc)
            rc = launch(c, V2C_K_N2M_UPDATE, [&] {
                if (fast) {
                    hipLaunchKernelGGL(k_n2m_update_fast, dim3(nb), dim3(256), 0, c->stream, a, nb, c->d_defer,
                                       c->d_counters + 2);
                    hipLaunchKernelGGL(k_n2m_update_list, dim3(N2M_LIST_BLOCKS), dim3(256), 0, c->stream, a,
                                       c->d_defer, c->d_counters + 2);
                } else if (advect) {
                    hipLaunchKernelGGL(k_n2m_update<true>, dim3(nb), dim3(256), 0, c->stream, a, nb);
                } else {
                    hipLaunchKernelGGL(k_n2m_update<false>, dim3(nb), dim3(256), 0, c->stream, a, nb);
                }
            });
    }
    if (side) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    if (rc) return rc;
    if (advect) c->order_valid = false;
    rc = finish(c);
    if (rc) return rc;
    if (advect && !c->async_mode) {
        int uns = 0;
        rc = read_counter(c, 1, &uns);
        if (rc) return rc;
        if (uns) return set_err(c, V2C_EUNSUPPORTED, "%d markers of type >= 100 (immobile spawning is out of scope)", uns);
    }
    return 0;
}

int v2c_nodes2markers(v2c_ctx *c) { return n2m_impl(c, false); }
int v2c_nodes2markers_advect(v2c_ctx *c) { return n2m_impl(c, true); }

int v2c_moveandrotatemarkers(v2c_ctx *c)
{
    if (!c) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    KArgs a = make_args(c);
    HIPCHK(c, hipMemsetAsync(c->d_counters + 1, 0, sizeof(int), c->stream));
    int rc = 0;
    if (c->n > 0) {
        rc = launch(c, V2C_K_ADVECT, [&] {
            hipLaunchKernelGGL(k_advect, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, a);
        });
        if (rc) return rc;
    }
    c->order_valid = false;
    rc = finish(c);
    if (rc) return rc;
    if (!c->async_mode) {
        int uns = 0;
        rc = read_counter(c, 1, &uns);
        if (rc) return rc;
        if (uns) return set_err(c, V2C_EUNSUPPORTED, "%d markers of type >= 100 (immobile spawning is out of scope)", uns);
    }
    return 0;
}

int v2c_compute_before_solve(v2c_ctx *c, double dt, int rk, long *nerr)
{
    if (!c) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    KArgs a = make_args(c);
    // the error count alternates between two counter slots; the kernel clears
    // the other one for the next call
    const int slot = c->before_slot;
    c->before_slot ^= BEFORE_SLOTS;
    int rc = launch(c, V2C_K_BEFORE, [&] {
        hipLaunchKernelGGL(k_before, dim3(blocks_for(c->nodenum, 256)), dim3(256), 0, c->stream, a, dt, rk, slot);
    });
    if (rc) return rc;
    rc = finish(c);
    if (rc) return rc;
    if (nerr) {
        int v = 0;
        rc = read_counter(c, slot, &v);
        if (rc) return rc;
        *nerr = v;
    }
    return 0;
}

int v2c_compute_after_solve(v2c_ctx *c, double dt, int rk, double *deltas, long *nerr)
{
    if (!c) return V2C_EINVAL;
    (void)rk;
    HIPCHK(c, hipSetDevice(c->device));
    KArgs a = make_args(c);
    const int nb = (int)blocks_for(c->nodenum, RED_BLOCK);
    double *out = c->d_red + 2 * nb;
    int rc = launch(c, V2C_K_AFTER, [&] {
        hipLaunchKernelGGL(k_after, dim3(nb), dim3(RED_BLOCK), 0, c->stream, a, dt, c->d_red);
        hipLaunchKernelGGL(k_after_final, dim3(1), dim3(RED_BLOCK), 0, c->stream, c->d_red, nb, out);
    });
    if (rc) return rc;
    rc = finish(c);
    if (rc) return rc;
    if (deltas || nerr) {
        double h[2];
        HIPCHK(c, hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (deltas) *deltas = h[0];
        if (nerr) *nerr = h[0] != h[0] ? 1 : 0;
    }
    return 0;
}

// Inputs to the device, k_ductile_batch, results back, on the context's
// stream; d_* are device columns of at least n entries.
static hipError_t ductile_staged(v2c_ctx *c, long n, const char *type, const double *sxy, const double *sxx,
                                 const double *tk, const double *pb, double *nu_out, int *iters, double *d_sxy,
                                 double *d_sxx, double *d_tk, double *d_pb, double *d_out, int *d_it, int8_t *d_t)
{
    hipError_t e = hipMemcpyAsync(d_sxy, sxy, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sxx, sxx, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tk, tk, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_pb, pb, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_t, type, n, hipMemcpyHostToDevice, c->stream);
    KArgs a = make_args(c);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ductile_batch, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, a, n, d_t, d_sxy,
                           d_sxx, d_tk, d_pb, d_out, d_it);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(nu_out, d_out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && iters) e = hipMemcpyAsync(iters, d_it, n * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->stream);
    return e;
}

int v2c_ductilevisc_batch(v2c_ctx *c, long n, const char *type, const double *sxy, const double *sxx,
                          const double *exy, const double *exx, const double *tk, const double *pb,
                          double *nu_out, int *iters)
{
    (void)exy; (void)exx;   // accepted but unused, as in density_ductilevisc_calc.c:114
    if (!c || n < 0 || !type || !sxy || !sxx || !tk || !pb || !nu_out) return V2C_EINVAL;
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->device));
    char *buf = nullptr;
    const size_t bytes = (size_t)n * (5 * sizeof(double) + sizeof(int) + 1);
    HIPCHK(c, hipMalloc(&buf, bytes));
    double *d_sxy = (double *)buf, *d_sxx = d_sxy + n, *d_tk = d_sxx + n, *d_pb = d_tk + n, *d_out = d_pb + n;
    int *d_it = (int *)(d_out + n);
    int8_t *d_t = (int8_t *)(d_it + n);
    const hipError_t e = ductile_staged(c, n, type, sxy, sxx, tk, pb, nu_out, iters, d_sxy, d_sxx, d_tk, d_pb,
                                        d_out, d_it, d_t);
    (void)hipFree(buf);
    if (e != hipSuccess) return set_err(c, V2C_EHIP, "ductilevisc_batch: %s", hipGetErrorString(e));
    return 0;
}

int v2c_ductilevisc_batch_noalloc(v2c_ctx *c, long n, const char *type, const double *sxy, const double *sxx,
                                  const double *exy, const double *exx, const double *tk, const double *pb,
                                  double *nu_out, int *iters)
{
    (void)exy; (void)exx;
    if (!c || n < 0 || !type || !sxy || !sxx || !tk || !pb || !nu_out) return V2C_EINVAL;
    if (n > V2C_DUCTILEVISC_NOALLOC_MAX)
        return set_err(c, V2C_EINVAL,
                       "ductilevisc_batch_noalloc: n = %ld exceeds the static staging limit of %ld inputs "
                       "(V2C_DUCTILEVISC_NOALLOC_MAX); use v2c_ductilevisc_batch or split the batch",
                       n, (long)V2C_DUCTILEVISC_NOALLOC_MAX);
    if (n == 0) return 0;
    if (!c->scratch) return set_err(c, V2C_EHIP, "ductilevisc_batch_noalloc: no staging address for device %d",
                                    c->device);
    HIPCHK(c, hipSetDevice(c->device));
    // fixed column layout: every column starts on a 256-byte boundary
    unsigned char *b = c->scratch;
    double *d_sxy = (double *)b, *d_sxx = (double *)(b + NOALLOC_COL), *d_tk = (double *)(b + 2 * NOALLOC_COL);
    double *d_pb = (double *)(b + 3 * NOALLOC_COL), *d_out = (double *)(b + 4 * NOALLOC_COL);
    int *d_it = (int *)(b + 5 * NOALLOC_COL);
    int8_t *d_t = (int8_t *)(b + 5 * NOALLOC_COL + (size_t)V2C_DUCTILEVISC_NOALLOC_MAX * sizeof(int));
    const hipError_t e = ductile_staged(c, n, type, sxy, sxx, tk, pb, nu_out, iters, d_sxy, d_sxx, d_tk, d_pb,
                                        d_out, d_it, d_t);
    if (e != hipSuccess) return set_err(c, V2C_EHIP, "ductilevisc_batch_noalloc: %s", hipGetErrorString(e));
    return 0;
}

int v2c_rsf_upload(v2c_ctx *c, const unsigned char *fault, const double *V, const double *phi,
                   const double *a, const double *b, const double *L, const double *V0,
                   const double *mu0, long n)
{
    if (!c || n < 0 || n > c->cap) return V2C_EINVAL;
    if (!fault || !V || !phi || !a || !b || !L || !V0 || !mu0) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const double *src[7] = {V, phi, a, b, L, V0, mu0};
    const int dst[7] = {MX_V, MX_PHI, MX_A, MX_B, MX_L, MX_V0, MX_MU0};
    for (int k = 0; k < 7; k++)
        HIPCHK(c, hipMemcpyAsync(c->d_mx + (size_t)dst[k] * c->cap, src[k], sizeof(double) * n,
                                 hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_fault, fault, n, hipMemcpyHostToDevice, c->stream));
    std::vector<uint32_t> list;
    for (long i = 0; i < n; i++)
        if (fault[i]) list.push_back((uint32_t)i);
    if (!list.empty())
        HIPCHK(c, hipMemcpyAsync(c->d_flist, list.data(), sizeof(uint32_t) * list.size(), hipMemcpyHostToDevice,
                                 c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nfault = (long)list.size();
    return 0;
}

int v2c_local_set_guess(v2c_ctx *c, const double *eta, const double *V, long n)
{
    if (!c || n < 0 || n > c->cap) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    if (eta) HIPCHK(c, hipMemcpyAsync(c->d_mx + (size_t)MX_GETA * c->cap, eta, sizeof(double) * n,
                                      hipMemcpyHostToDevice, c->stream));
    if (V) HIPCHK(c, hipMemcpyAsync(c->d_mx + (size_t)MX_GV * c->cap, V, sizeof(double) * n,
                                    hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int v2c_local_get_guess(v2c_ctx *c, double *eta, double *V, long n)
{
    if (!c || n < 0 || n > c->n) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    if (eta) HIPCHK(c, hipMemcpyAsync(eta, c->d_mx + (size_t)MX_GETA * c->cap, sizeof(double) * n,
                                      hipMemcpyDeviceToHost, c->stream));
    if (V) HIPCHK(c, hipMemcpyAsync(V, c->d_mx + (size_t)MX_GV * c->cap, sizeof(double) * n,
                                    hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int v2c_local_update(v2c_ctx *c, double dt, int warm, double tol, v2c_lu_stats *st)
{
    if (!c || !(tol > 0) || warm < 0 || warm > 2) return V2C_EINVAL;
    if (warm == 2 && !c->mlp_ready) return set_err(c, V2C_EINVAL, "no MLP parameters loaded (v2c_mlp_set)");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemsetAsync(c->d_lu, 0, sizeof(unsigned long long) * LS_COUNT, c->stream));
    KArgs a = make_args(c);
    int rc = 0;
    // The slip-rate solve (fault markers, with the network for warm == 2)
    // runs on the side stream beside the ductile solve of every marker; the
    // visco-plastic viscosity that needs both follows the join.  The region
    // from the fork to the last kernel is timed as V2C_K_LOCAL_WALL.
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (c->timing) {
        t0 = take_event(c);
        t1 = take_event(c);
        if (t0) (void)hipEventRecord(t0, c->stream);
    }
    const bool side = c->nfault > 0;
    if (side) {
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
        const unsigned nb = blocks_for(c->nfault, LU_BLOCK);
        rc = launch(c, V2C_K_LOCAL_RSF, [&] {
            if (warm == 2) hipLaunchKernelGGL(k_local_rsf<2>, dim3(nb), dim3(LU_BLOCK), 0, c->side, a, dt, tol);
            else if (warm == 1) hipLaunchKernelGGL(k_local_rsf<1>, dim3(nb), dim3(LU_BLOCK), 0, c->side, a, dt, tol);
            else hipLaunchKernelGGL(k_local_rsf<0>, dim3(nb), dim3(LU_BLOCK), 0, c->side, a, dt, tol);
        }, c->side);
        const hipError_t ej = hipEventRecord(c->ev_join, c->side);
        if (!rc && ej != hipSuccess) rc = set_err(c, V2C_EHIP, "hipEventRecord: %s", hipGetErrorString(ej));
    }
    if (!rc && c->n > 0)
        rc = launch(c, V2C_K_LOCAL, [&] {
            hipLaunchKernelGGL(k_local_ductile, dim3(blocks_for(c->n, LU_BLOCK)), dim3(LU_BLOCK), 0, c->stream, a,
                               warm, tol);
        });
    // join before any return, so no side-stream work is left unordered
    if (side) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    if (rc) return rc;
    if (side) {
        rc = launch(c, V2C_K_LOCAL, [&] {
            hipLaunchKernelGGL(k_local_etavp, dim3(blocks_for(c->nfault, LU_BLOCK)), dim3(LU_BLOCK), 0, c->stream, a);
        });
        if (rc) return rc;
    }
    if (t0 && t1) {
        (void)hipEventRecord(t1, c->stream);
        c->pending.push_back(TimedLaunch{V2C_K_LOCAL_WALL, t0, t1});
    }
    rc = finish(c);
    if (rc) return rc;
    if (st) {
        unsigned long long h[LS_COUNT];
        HIPCHK(c, hipMemcpyAsync(h, c->d_lu, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        st->markers = (long)h[LS_N]; st->ductile_markers = (long)h[LS_DUCT];
        st->ductile_iters = (long)h[LS_DUCT_IT]; st->fault_markers = (long)h[LS_FAULT];
        st->rsf_iters = (long)h[LS_RSF_IT]; st->max_ductile_iters = (long)h[LS_DUCT_MAX];
        st->max_rsf_iters = (long)h[LS_RSF_MAX]; st->fallbacks = (long)h[LS_FALLBACK];
        st->nonfinite = (long)h[LS_NONFINITE];
        st->ductile_evals = (long)h[LS_DUCT_EV];
        st->rsf_evals = (long)h[LS_RSF_EV];
        for (int k = 0; k < LU_HIST; k++) {
            st->hist_ductile[k] = (long)h[LS_HIST_D + k];
            st->hist_rsf[k] = (long)h[LS_HIST_R + k];
        }
    }
    return 0;
}

int v2c_local_download(v2c_ctx *c, double *eta, double *V, double *phi, double *eta_vp, double *syield, long n)
{
    if (!c || n < 0 || n > c->n) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    double *dst[5] = {eta, V, phi, eta_vp, syield};
    const int src[5] = {MX_ETA, MX_V, MX_PHI, MX_ETAVP, MX_SY};
    for (int k = 0; k < 5; k++)
        if (dst[k])
            HIPCHK(c, hipMemcpyAsync(dst[k], c->d_mx + (size_t)src[k] * c->cap, sizeof(double) * n,
                                     hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// ---- learned warm start --------------------------------------------------

int v2c_mlp_set(v2c_ctx *c, const float *params, const float *norm)
{
    if (!c || !params || !norm) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->d_mlp, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mlp_norm, norm, sizeof(float) * MLP_NNORM, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_mlp_m, 0, sizeof(float) * MLP_NPARAM, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_mlp_v, 0, sizeof(float) * MLP_NPARAM, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->mlp_steps = 0;
    c->mlp_ready = 1;
    return 0;
}

int v2c_mlp_get(v2c_ctx *c, float *params, float *norm)
{
    if (!c) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    if (params)
        HIPCHK(c, hipMemcpyAsync(params, c->d_mlp, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
    if (norm)
        HIPCHK(c, hipMemcpyAsync(norm, c->d_mlp_norm, sizeof(float) * MLP_NNORM, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int v2c_mlp_predict(v2c_ctx *c, double dt) { return v2c_mlp_predict_outputs(c, dt, 3); }

int v2c_mlp_predict_outputs(v2c_ctx *c, double dt, int outputs)
{
    if (!c || outputs < 0 || outputs > 3) return V2C_EINVAL;
    if (!c->mlp_ready) return set_err(c, V2C_EINVAL, "no MLP parameters loaded (v2c_mlp_set)");
    HIPCHK(c, hipSetDevice(c->device));
    KArgs a = make_args(c);
    int rc = 0;
    if (c->n > 0) {
        rc = launch(c, V2C_K_MLP, [&] {
            hipLaunchKernelGGL(k_mlp_guess, dim3(blocks_for(c->n, 64 * MI_WAVES)), dim3(64 * MI_WAVES), 0, c->stream,
                               a, dt, c->d_mlp, c->d_mlp_norm, outputs);
        });
        if (rc) return rc;
    }
    return finish(c);
}

int v2c_mlp_samples(v2c_ctx *c, double dt, double tol, float *X, float *mask)
{
    if (!c || !X || !mask) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const long n = c->n;
    if (n == 0) return 0;
    float *dX = nullptr, *dM = nullptr;
    HIPCHK(c, hipMalloc(&dX, sizeof(float) * MLP_ROW * n));
    hipError_t e = hipMalloc(&dM, sizeof(float) * 2 * n);
    if (e != hipSuccess) { (void)hipFree(dX); return set_err(c, V2C_ENOMEM, "mlp_samples"); }
    KArgs a = make_args(c);
    hipLaunchKernelGGL(k_mlp_features, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, a, dt, dX, dM);
    e = hipGetLastError();
    int rc = 0;
    if (e == hipSuccess) rc = v2c_local_update(c, dt, 0, tol, nullptr);
    if (e == hipSuccess && rc == 0) {
        hipLaunchKernelGGL(k_mlp_labels, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, a, dX, dM);
        e = hipGetLastError();
    }
    if (e == hipSuccess && rc == 0) e = hipMemcpyAsync(X, dX, sizeof(float) * MLP_ROW * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && rc == 0) e = hipMemcpyAsync(mask, dM, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dX);
    (void)hipFree(dM);
    if (rc) return rc;
    if (e != hipSuccess) return set_err(c, V2C_EHIP, "mlp_samples: %s", hipGetErrorString(e));
    return 0;
}

// Deterministic mini-batch training: a fixed xorshift permutation per epoch
// (seeded), per-wave gradient partials summed in wave order, Adam on FP32
// master weights.  Xs/Ts are already standardised.  loss_out[epoch] receives
// the epoch's mean masked squared error.
namespace {
struct TrainBufs {
    float *X = nullptr, *T = nullptr, *M = nullptr, *G = nullptr, *L = nullptr;
    uint32_t *idx = nullptr;
    ~TrainBufs()
    {
        void *p[] = {X, T, M, G, L, idx};
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
};
}  // namespace

static int mlp_upload(v2c_ctx *c, TrainBufs &b, const float *Xs, const float *Ts, const float *Ms, long ns,
                      long batch)
{
    const long nw = (batch + 63) / 64;
    HIPCHK(c, hipMalloc(&b.X, sizeof(float) * MLP_IN * ns));
    HIPCHK(c, hipMalloc(&b.T, sizeof(float) * 2 * ns));
    HIPCHK(c, hipMalloc(&b.M, sizeof(float) * 2 * ns));
    HIPCHK(c, hipMalloc(&b.G, sizeof(float) * MLP_NPARAM * nw));
    HIPCHK(c, hipMalloc(&b.L, sizeof(float) * nw));
    HIPCHK(c, hipMalloc(&b.idx, sizeof(uint32_t) * ns));
    HIPCHK(c, hipMemcpyAsync(b.X, Xs, sizeof(float) * MLP_IN * ns, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.T, Ts, sizeof(float) * 2 * ns, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.M, Ms, sizeof(float) * 2 * ns, hipMemcpyHostToDevice, c->stream));
    return 0;
}

int v2c_mlp_train(v2c_ctx *c, const float *Xs, const float *Ts, const float *Ms, long ns, int epochs, int batch,
                  float lr, unsigned seed, float *loss_out)
{
    if (!c || !Xs || !Ts || !Ms || ns <= 0 || epochs <= 0 || batch < 64 || batch % 64) return V2C_EINVAL;
    if (!c->mlp_ready) return set_err(c, V2C_EINVAL, "initialise the MLP with v2c_mlp_set first");
    HIPCHK(c, hipSetDevice(c->device));
    TrainBufs b;
    int rc = mlp_upload(c, b, Xs, Ts, Ms, ns, batch);
    if (rc) return rc;
    std::vector<uint32_t> perm(ns);
    std::vector<float> lhost((batch + 63) / 64);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
    auto rnd = [&]() {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return st;
    };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ep++) {
        for (long i = 0; i < ns; i++) perm[i] = (uint32_t)i;
        for (long i = ns - 1; i > 0; i--) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
        HIPCHK(c, hipMemcpyAsync(b.idx, perm.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice, c->stream));
        double ep_loss = 0, ep_cnt = 0;
        for (long off = 0; off < ns; off += batch) {
            const long nb = std::min<long>(batch, ns - off);
            double cnt = 0;
            for (long i = 0; i < nb; i++) cnt += Ms[2 * perm[off + i]] + Ms[2 * perm[off + i] + 1];
            if (cnt == 0) continue;
            const int nw = (int)((nb + 63) / 64);
            c->mlp_steps++;
            const float bc1 = 1.f - powf(b1, (float)c->mlp_steps), bc2 = 1.f - powf(b2, (float)c->mlp_steps);
            rc = launch(c, V2C_K_MLP, [&] {
                hipLaunchKernelGGL(k_mlp_train, dim3(nw), dim3(64), 0, c->stream, c->d_mlp, b.X, b.T, b.M,
                                   b.idx + off, nb, (float)(1.0 / cnt), b.G, b.L);
                hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(MLP_NPARAM, 256)), dim3(256), 0, c->stream, c->d_mlp,
                                   c->d_mlp_m, c->d_mlp_v, b.G, nw, lr, b1, b2, eps, bc1, bc2);
            });
            if (rc) return rc;
            HIPCHK(c, hipMemcpyAsync(lhost.data(), b.L, sizeof(float) * nw, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (int w = 0; w < nw; w++) ep_loss += lhost[w];
            ep_cnt += cnt;
        }
        if (loss_out) loss_out[ep] = (float)(ep_cnt > 0 ? ep_loss / ep_cnt : 0.0);
    }
    return 0;
}

int v2c_mlp_grad(v2c_ctx *c, const float *Xs, const float *Ts, const float *Ms, long ns, float *grad, float *loss)
{
    if (!c || !Xs || !Ts || !Ms || ns <= 0 || !grad) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const long batch = ((ns + 63) / 64) * 64;
    TrainBufs b;
    int rc = mlp_upload(c, b, Xs, Ts, Ms, ns, batch);
    if (rc) return rc;
    std::vector<uint32_t> idx(ns);
    double cnt = 0;
    for (long i = 0; i < ns; i++) { idx[i] = (uint32_t)i; cnt += Ms[2 * i] + Ms[2 * i + 1]; }
    if (cnt == 0) cnt = 1;
    HIPCHK(c, hipMemcpyAsync(b.idx, idx.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice, c->stream));
    const int nw = (int)((ns + 63) / 64);
    float *dgrad = b.T;   // reuse: T is not needed after the kernel reads it? keep separate
    float *dg = nullptr;
    HIPCHK(c, hipMalloc(&dg, sizeof(float) * MLP_NPARAM));
    hipLaunchKernelGGL(k_mlp_train, dim3(nw), dim3(64), 0, c->stream, c->d_mlp, b.X, b.T, b.M, b.idx, ns,
                       (float)(1.0 / cnt), b.G, b.L);
    hipLaunchKernelGGL(k_mlp_gradsum, dim3(blocks_for(MLP_NPARAM, 256)), dim3(256), 0, c->stream, b.G, nw, dg);
    (void)dgrad;
    hipError_t e = hipGetLastError();
    std::vector<float> lh(nw);
    if (e == hipSuccess) e = hipMemcpyAsync(grad, dg, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(lh.data(), b.L, sizeof(float) * nw, hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dg);
    if (e != hipSuccess) return set_err(c, V2C_EHIP, "mlp_grad: %s", hipGetErrorString(e));
    if (loss) {
        double t = 0;
        for (float v : lh) t += v;
        *loss = (float)(t / cnt);
    }
    return 0;
}

int v2c_mlp_forward(v2c_ctx *c, const float *Xs, long ns, float *Ys)
{
    if (!c || !Xs || !Ys || ns <= 0) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    float *dX = nullptr, *dY = nullptr;
    HIPCHK(c, hipMalloc(&dX, sizeof(float) * MLP_IN * ns));
    hipError_t e = hipMalloc(&dY, sizeof(float) * 2 * ns);
    if (e == hipSuccess) e = hipMemcpyAsync(dX, Xs, sizeof(float) * MLP_IN * ns, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_for(ns, 64)), dim3(64), 0, c->stream, c->d_mlp, dX, ns, dY);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(Ys, dY, sizeof(float) * 2 * ns, hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    if (dX) (void)hipFree(dX);
    if (dY) (void)hipFree(dY);
    if (e != hipSuccess) return set_err(c, V2C_EHIP, "mlp_forward: %s", hipGetErrorString(e));
    return 0;
}

int v2c_copy_calibration(v2c_ctx *c, long bytes, double *ms)
{
    if (!c || bytes <= 0) return V2C_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const long n = bytes / 8;
    double *src = nullptr, *dst = nullptr;
    HIPCHK(c, hipMalloc(&src, n * 8));
    hipError_t e = hipMalloc(&dst, n * 8);
    if (e == hipSuccess) e = hipMemsetAsync(src, 0, n * 8, c->stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_copy_calib, dim3(8192), dim3(256), 0, c->stream, src, dst, n);   // warm
        (void)hipEventRecord(e0, c->stream);
        hipLaunchKernelGGL(k_copy_calib, dim3(8192), dim3(256), 0, c->stream, src, dst, n);
        (void)hipEventRecord(e1, c->stream);
        e = hipStreamSynchronize(c->stream);
    }
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e != hipSuccess) return set_err(c, V2C_EHIP, "copy_calibration: %s", hipGetErrorString(e));
    return 0;
}

int v2c_enable_timing(v2c_ctx *c, int on)
{
    if (!c) return V2C_EINVAL;
    c->timing = on ? 1 : 0;
    return 0;
}

int v2c_timing_select(v2c_ctx *c, unsigned int kernel_mask)
{
    if (!c) return V2C_EINVAL;
    c->timing_mask = kernel_mask;
    return 0;
}

int v2c_reset_timing(v2c_ctx *c)
{
    if (!c) return V2C_EINVAL;
    int rc = collect_timing(c);
    for (int k = 0; k < V2C_K_COUNT; k++) { c->ktime[k] = 0; c->klaunch[k] = 0; }
    return rc;
}

int v2c_kernel_time(v2c_ctx *c, int kernel, double *ms, long *launches)
{
    if (!c || kernel < 0 || kernel >= V2C_K_COUNT) return V2C_EINVAL;
    int rc = collect_timing(c);
    if (rc) return rc;
    if (ms) *ms = c->ktime[kernel];
    if (launches) *launches = c->klaunch[kernel];
    return 0;
}

}  // extern "C"

#include "v2c_post.hip"
#include "v2c_ensemble.hip"
#include "v2c_heat.hip"
#include "v2c_bc.hip"
#include "v2c_prn.cpp"
