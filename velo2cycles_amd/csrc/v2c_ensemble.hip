// v2c_ensemble.hip -- batched ensembles of independent models.
//
// Every model is an ordinary context (its own grid, marker cloud and nodal
// state).  A step of the whole ensemble runs each stage of the per-model step
// (marker2nodes, compute_before_solve, compute_after_solve, nodes2markers +
// advection) as ONE launch over a work list of (model, range) entries, and
// every entry calls the same device function the single-model kernel calls,
// on that model's arguments.  The stage sequence is captured once into a HIP
// graph and replayed each step; only the cell ordering (which decides between
// an incremental merge and a full sort on the host) and the argument upload
// run outside the graph.
//
// Two launch forms:
//  * item stages (per marker, per node): one 256-thread block per entry of at
//    most 256 items of one model, one item per thread and no loop, so each
//    model's arguments are loaded where they are used (a loop over entries
//    would keep them live across the whole per-marker body).  The host orders
//    the entry list so that runs of consecutive entries of one model land on
//    one XCD (blocks are dealt to the 8 XCDs round-robin), which keeps a
//    model's nodal rows in that XCD's L2.
//  * scatter stages (per cell group, one wavefront each): a persistent grid
//    strides over the list of cell groups.
//
// Shapes (cells per scatter group; markers and nodes per XCD run) are chosen
// per model size class; v2c_ens_autotune times candidates on the ensemble
// itself and keeps the fastest.  Results do not depend on the shapes: every
// sub-cell sum is the same left-to-right sum over its sorted markers whatever
// the batch and group boundaries, and every item is computed alone, so a
// batched step equals the per-model steps bit for bit.
// Included at the end of v2c_api.hip: it uses the context internals.

namespace v2c {

// The model's arguments, read from global memory (their pointer members are
// typed as global pointers, so loads through them stay global loads).
__device__ __forceinline__ const KArgs &model_args(const KArgs *__restrict__ A, int mod) { return A[mod]; }

struct EnsTile {
    int model;
    int pad;
    long begin, end;            // item range: markers, cells or nodes
};

constexpr int ENS_BLOCK = 256;  // items per item-stage entry (= RED_BLOCK)
static_assert(ENS_BLOCK == RED_BLOCK, "node entries double as reduction blocks");

// Item stages: block b takes entry b of the list; thread i its i-th item.
template <class F>
__device__ __forceinline__ void ens_item(const KArgs *__restrict__ A, const EnsTile *__restrict__ tiles, F &&f)
{
    const EnsTile T = tiles[blockIdx.x];
    const long i = T.begin + threadIdx.x;
    if (i < T.end) f(model_args(A, T.model), i);
}

__global__ __launch_bounds__(ENS_BLOCK) void k_ens_m2n_props(const KArgs *__restrict__ A,
                                                             const EnsTile *__restrict__ tiles)
{
    ens_item(A, tiles, [&](const KArgs &a, long m) { m2n_props_one(a, m); });
}

// The persistent scatter grids deal the cell-group list to the XCDs in chunks
// of ENS_CHUNK consecutive groups, round-robin (chunk q to XCD q % 8), and the
// workgroups of an XCD (workgroup b runs on XCD b % 8) stride over that XCD's
// chunks: neighbouring groups, which share marker cache lines and node
// records, mostly meet in one L2, while every XCD gets chunks of every part
// of the list (models of all sizes).  With a grid that is not a multiple of 8
// the list is simply strided.
constexpr int ENS_CHUNK = 32;
struct EnsBand {
    int x, r, step;
    bool chunked;
};
__device__ __forceinline__ EnsBand ens_band(int ntiles)
{
    const int G = gridDim.x;
    if (G % 8 != 0 || ntiles < G) return EnsBand{0, (int)blockIdx.x, G, false};
    return EnsBand{(int)blockIdx.x % 8, (int)blockIdx.x / 8, G / 8, true};
}
// The j-th group this workgroup takes (j = r, r + step, ...), or past the end.
__device__ __forceinline__ int ens_band_tile(const EnsBand &bd, int j)
{
    if (!bd.chunked) return j;
    return ((j / ENS_CHUNK) * 8 + bd.x) * ENS_CHUNK + j % ENS_CHUNK;
}

__global__ __launch_bounds__(WAVE) void k_ens_m2n_cells(const KArgs *__restrict__ A,
                                                        const EnsTile *__restrict__ tiles, int ntiles)
{
    __shared__ M2nLds L;
    const EnsBand bd = ens_band(ntiles);
    for (int j = bd.r;; j += bd.step) {
        const int t = ens_band_tile(bd, j);
        if (t >= ntiles) break;
        const EnsTile T = tiles[t];
        m2n_cells_group(model_args(A, T.model), T.begin, T.end, threadIdx.x, L);
        __syncthreads();
    }
}

__global__ __launch_bounds__(ENS_BLOCK) void k_ens_m2n_nodes(const KArgs *__restrict__ A,
                                                             const EnsTile *__restrict__ tiles)
{
    ens_item(A, tiles, [&](const KArgs &a, long m3) { m2n_nodes_one(a, m3); });
}

__global__ __launch_bounds__(ENS_BLOCK) void k_ens_phis_bc(const KArgs *__restrict__ A,
                                                           const EnsTile *__restrict__ tiles)
{
    ens_item(A, tiles, [&](const KArgs &a, long m3) {
        if (a.p.Dphis > 0) phis_bc_one(a, m3);
    });
}

__global__ __launch_bounds__(ENS_BLOCK) void k_ens_before(const KArgs *__restrict__ A,
                                                          const EnsTile *__restrict__ tiles)
{
    ens_item(A, tiles, [&](const KArgs &a, long m3) { before_one(a, m3, a.timestep, 0, 0); });
}

// Node entries are whole RED_BLOCK chunks of one model (they start at
// multiples of RED_BLOCK), so each chunk's fixed-order sum lands in the
// model's partial slot begin / RED_BLOCK, as in the single-model launch.
__global__ __launch_bounds__(RED_BLOCK) void k_ens_after(const KArgs *__restrict__ A, double *const *__restrict__ red,
                                                         const EnsTile *__restrict__ tiles)
{
    const EnsTile T = tiles[blockIdx.x];
    const long m3 = T.begin + threadIdx.x;
    double dsq = 0.0, cnt = 0.0;
    if (m3 < T.end) {
        const KArgs &a = model_args(A, T.model);
        after_one(a, m3, a.timestep, dsq, cnt);
    }
    after_block_sum(dsq, cnt, red[T.model], T.begin / RED_BLOCK);
}

// One workgroup per model: fixed-order sum of the model's partials.
__global__ __launch_bounds__(RED_BLOCK) void k_ens_after_final(double *const *__restrict__ red,
                                                               const int *__restrict__ nblocks, double *out, int k0)
{
    __shared__ double ssum[RED_BLOCK];
    __shared__ double scnt[RED_BLOCK];
    const int mod = k0 + (int)blockIdx.x;
    const double *r = red[mod];
    double s = 0, c = 0;
    for (int b = threadIdx.x; b < nblocks[mod]; b += RED_BLOCK) { s += r[2 * b]; c += r[2 * b + 1]; }
    ssum[threadIdx.x] = s;
    scnt[threadIdx.x] = c;
    __syncthreads();
    for (int k = RED_BLOCK / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) { ssum[threadIdx.x] += ssum[threadIdx.x + k]; scnt[threadIdx.x] += scnt[threadIdx.x + k]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * mod] = scnt[0] > 0 ? pow(ssum[0] / scnt[0], 0.5) : 0.0;
        out[2 * mod + 1] = scnt[0];
    }
}

__global__ void k_ens_phantom(const KArgs *__restrict__ A, int k0, int k1)
{
    const int mod = k0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (mod < k1 && A[mod].p.antidiffinter == 1 && A[mod].p.n2m_phantom) n2m_phantom_one(model_args(A, mod));
}

__global__ __launch_bounds__(WAVE) void k_ens_n2m_scatter(const KArgs *__restrict__ A,
                                                          const EnsTile *__restrict__ tiles, int ntiles)
{
    __shared__ N2mLds L;
    const EnsBand bd = ens_band(ntiles);
    for (int j = bd.r;; j += bd.step) {
        const int t = ens_band_tile(bd, j);
        if (t >= ntiles) break;
        const EnsTile T = tiles[t];
        const KArgs &a = model_args(A, T.model);
        if (a.p.antidiffinter == 1) n2m_scatter_group(a, T.begin, T.end, threadIdx.x, L);
        __syncthreads();
    }
}

__global__ __launch_bounds__(ENS_BLOCK) void k_ens_n2m_nodes(const KArgs *__restrict__ A,
                                                             const EnsTile *__restrict__ tiles)
{
    ens_item(A, tiles, [&](const KArgs &a, long m3) {
        if (a.p.antidiffinter == 1) n2m_nodes_one(a, m3);
    });
}

// Append the lanes' flagged (model, marker) items to the ensemble's list.
__device__ __forceinline__ void list_append64(bool flag, uint64_t item, uint64_t *list, int *count)
{
    const unsigned long long mask = __ballot(flag);
    if (!mask) return;
    const int lane = threadIdx.x & (WAVE - 1);
    int base = 0;
    if (lane == 0) base = atomicAdd(count, (int)__popcll(mask));
    base = __shfl(base, 0);
    if (flag) list[base + (int)__popcll(mask & ((1ull << lane) - 1ull))] = item;
}

// Fused node-to-marker update and advection, fast form (n2m_update_fast);
// the markers it cannot take, and every marker of a model without
// antidiffusion, go to the deferred list for the general form.
__global__ __launch_bounds__(ENS_BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void k_ens_n2m_update_fast(const KArgs *__restrict__ A,
                                                                   const EnsTile *__restrict__ tiles,
                                                                   uint64_t *defer, int *ndefer)
{
    const EnsTile T = tiles[blockIdx.x];
    const long m = T.begin + threadIdx.x;
    bool def = false;
    if (m < T.end) {
        const KArgs &a = model_args(A, T.model);
        def = a.p.antidiffinter != 1 || !n2m_update_fast(a, m);
    }
    list_append64(def, ((uint64_t)T.model << 32) | (uint64_t)(uint32_t)m, defer, ndefer);
}

__global__ __launch_bounds__(256) void k_ens_n2m_update_list(const KArgs *__restrict__ A,
                                                             const uint64_t *__restrict__ list,
                                                             const int *__restrict__ count, int *unsupported)
{
    const int n = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t v = list[i];
        if (n2m_update_one<true>(model_args(A, (int)(v >> 32)), (long)(uint32_t)v)) atomicAdd(unsupported, 1);
    }
}

// ---- per-marker local update of every model (v2c_local.hip) ----------------
__global__ void k_ens_lu_clear(const KArgs *__restrict__ A, int k0)
{
    unsigned long long *st = model_args(A, k0 + (int)blockIdx.x).lu_stats;
    for (int i = threadIdx.x; i < LS_COUNT; i += blockDim.x) st[i] = 0;
}

__global__ __launch_bounds__(LU_BLOCK) void k_ens_local_ductile(const KArgs *__restrict__ A,
                                                                const EnsTile *__restrict__ tiles, int warm,
                                                                double tol)
{
    __shared__ unsigned long long sh[LS_COUNT];
    lu_clear(sh);
    const EnsTile T = tiles[blockIdx.x];
    const KArgs &a = model_args(A, T.model);
    const long m = T.begin + threadIdx.x;
    LuAcc acc;
    if (m < T.end) lu_ductile_one(a, m, warm == 1, tol, acc);
    lu_wave_add(acc, sh);
    lu_flush(sh, a.lu_stats);
}

// Fault-list entries of the models (index ranges of each model's list).
template <int WARM>
__global__ __launch_bounds__(LU_BLOCK) void k_ens_local_rsf(const KArgs *__restrict__ A,
                                                            const EnsTile *__restrict__ tiles, double tol)
{
    __shared__ unsigned long long sh[LS_COUNT];
    lu_clear(sh);
    const EnsTile T = tiles[blockIdx.x];
    const KArgs &a = model_args(A, T.model);
    const long i = T.begin + threadIdx.x;
    if constexpr (WARM == 2) {
        __shared__ LuRsfLds L;
        lu_rsf_block<2>(a, i, T.end, a.timestep, tol, sh, &L);
    } else {
        lu_rsf_block<WARM>(a, i, T.end, a.timestep, tol, sh, nullptr);
    }
    lu_flush(sh, a.lu_stats);
}

__global__ __launch_bounds__(LU_BLOCK) void k_ens_local_etavp(const KArgs *__restrict__ A,
                                                              const EnsTile *__restrict__ tiles)
{
    const EnsTile T = tiles[blockIdx.x];
    const long i = T.begin + threadIdx.x;
    if (i < T.end) lu_etavp_one(model_args(A, T.model), i);
}

// ---- copyarrays and calc_timestep of every model (v2c_post.hip) ------------
// Per-model scratch of the batched calc_timestep; out (8 doubles) and err
// point into ensemble-wide arrays read back after the step.
struct EnsTs {
    V2C_GLOBAL Fold *part;
    V2C_GLOBAL double *bval, *carry, *out;
    V2C_GLOBAL int *bset, *err;
    int nblocks, pad;
};

__global__ __launch_bounds__(ENS_BLOCK) void k_ens_copyarrays(const KArgs *__restrict__ A,
                                                              const EnsTile *__restrict__ tiles)
{
    ens_item(A, tiles, [&](const KArgs &a, long m3) { copyarrays_one(a, m3); });
}

__global__ void k_ens_ts_clear(const EnsTs *__restrict__ S, int k0, int k1)
{
    const int k = k0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < k1) S[k].err[0] = 0;
}

// Node entries start at multiples of POST_BLOCK: entry = block begin / POST_BLOCK.
static_assert(ENS_BLOCK == POST_BLOCK, "node entries double as calc_timestep blocks");
__global__ __launch_bounds__(POST_BLOCK) void k_ens_ts_lambda(const KArgs *__restrict__ A,
                                                              const EnsTile *__restrict__ tiles,
                                                              const EnsTs *__restrict__ S, TsArgs T)
{
    const EnsTile E = tiles[blockIdx.x];
    const EnsTs &s = S[E.model];
    ts_lambda_block(model_args(A, E.model), T, E.begin / POST_BLOCK, s.bset, s.bval, s.err);
}

__global__ __launch_bounds__(FIN_BLOCK) void k_ens_ts_carry(const EnsTs *__restrict__ S, TsArgs T, int k0)
{
    const EnsTs &s = S[k0 + (int)blockIdx.x];
    ts_carry_block(s.bset, s.bval, s.nblocks, T.lambda0, s.carry);
}

__global__ __launch_bounds__(POST_BLOCK) void k_ens_ts_node(const KArgs *__restrict__ A,
                                                            const EnsTile *__restrict__ tiles,
                                                            const EnsTs *__restrict__ S, TsArgs T)
{
    const EnsTile E = tiles[blockIdx.x];
    const EnsTs &s = S[E.model];
    ts_node_block(model_args(A, E.model), T, E.begin / POST_BLOCK, s.carry, s.part);
}

__global__ __launch_bounds__(FIN_BLOCK) void k_ens_ts_final(const EnsTs *__restrict__ S, TsArgs T, int k0)
{
    const EnsTs &s = S[k0 + (int)blockIdx.x];
    ts_final_block(s.part, s.nblocks, T, s.out);
}

__global__ void k_ens_clear(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Batched incremental re-ordering: the passes of the single-model
// re-ordering, each one launch over all models, with ONE radix sort of the
// movers of every model (model number in the bits above the key, so each
// model's movers come out contiguous and in their own (key, id) order) and
// ONE read-back of the mover counts.  A model takes this path when it was
// sorted before and 0 < movers <= n/4, the rule of the single-model path;
// above n/4 it gets a full sort of its own.
// ---------------------------------------------------------------------------
struct EnsOrd {
    V2C_GLOBAL const uint32_t *key;                 // physical key per marker
    V2C_GLOBAL const uint32_t *order, *skey;        // current order
    V2C_GLOBAL uint32_t *order2, *skey2;            // the new order
    V2C_GLOBAL unsigned long long *mask;
    V2C_GLOBAL uint32_t *tile_count, *tile_off, *sub_start;
    long n, ntiles, nsub;
    int active;                                     // re-ordered this step
    int pad;
};

__host__ __device__ inline bool eo_incremental(const EnsOrd &o, long nmov)
{
    return o.active && nmov > 0 && nmov <= o.n / 4;
}

__global__ __launch_bounds__(ORD_BLOCK) void k_eo_flags(const EnsOrd *__restrict__ O, const EnsTile *__restrict__ tiles)
{
    const EnsTile T = tiles[blockIdx.x];
    const EnsOrd &o = O[T.model];
    if (o.active) order_flags_tile(o.order, o.key, o.skey, o.n, o.mask, o.tile_count, T.begin);
}

__global__ __launch_bounds__(1024) void k_eo_scan(const EnsOrd *__restrict__ O, long *__restrict__ nmov)
{
    const EnsOrd &o = O[blockIdx.x];
    if (!o.active) {
        if (threadIdx.x == 0) nmov[blockIdx.x] = 0;
        return;
    }
    order_scan_block(o.tile_count, o.ntiles, o.tile_off, &nmov[blockIdx.x]);
}

// base[k] = first slot of model k's movers; base[nm] = all movers sorted.
__global__ void k_eo_bases(const EnsOrd *__restrict__ O, const long *__restrict__ nmov, int nm, long *__restrict__ base)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    long run = 0;
    for (int k = 0; k < nm; k++) {
        base[k] = run;
        if (eo_incremental(O[k], nmov[k])) run += nmov[k];
    }
    base[nm] = run;
}

__global__ __launch_bounds__(ORD_BLOCK) void k_eo_movers(const EnsOrd *__restrict__ O, const EnsTile *__restrict__ tiles,
                                                         const long *__restrict__ nmov, const long *__restrict__ base,
                                                         uint64_t *__restrict__ mov, int shift)
{
    const EnsTile T = tiles[blockIdx.x];
    const EnsOrd &o = O[T.model];
    if (eo_incremental(o, nmov[T.model]))
        order_movers_tile(o.order, o.key, o.n, o.mask, o.tile_off, mov + base[T.model], (uint64_t)T.model << shift,
                          T.begin);
}

__global__ __launch_bounds__(ORD_BLOCK) void k_eo_stayers(const EnsOrd *__restrict__ O, const EnsTile *__restrict__ tiles,
                                                          const long *__restrict__ nmov, const long *__restrict__ base,
                                                          const uint64_t *__restrict__ mov, int shift)
{
    const EnsTile T = tiles[blockIdx.x];
    const EnsOrd &o = O[T.model];
    const long nm_k = nmov[T.model];
    if (eo_incremental(o, nm_k))
        order_stayers_tile(o.order, o.skey, o.n, o.mask, o.tile_off, mov + base[T.model], nm_k,
                           (uint64_t)T.model << shift, o.order2, o.skey2, T.begin);
}

__global__ void k_eo_place(const EnsOrd *__restrict__ O, const uint64_t *__restrict__ mov,
                           const long *__restrict__ base, long total, int shift)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= total) return;
    const uint64_t v = mov[r];
    const int k = (int)(v >> shift);
    const EnsOrd &o = O[k];
    order_place_one(v & ((1ull << shift) - 1ull), r - base[k], o.order, o.skey, o.n, o.sub_start, o.nsub, o.mask,
                    o.tile_off, o.order2, o.skey2);
}

// Sub-cell starts of the new order over the marker entries (sorted
// positions); the last position also writes the tail up to nsub.
__global__ __launch_bounds__(ENS_BLOCK) void k_eo_sub_start(const EnsOrd *__restrict__ O,
                                                            const EnsTile *__restrict__ tiles,
                                                            const long *__restrict__ nmov)
{
    const EnsTile T = tiles[blockIdx.x];
    const EnsOrd &o = O[T.model];
    if (!eo_incremental(o, nmov[T.model])) return;
    const long i = T.begin + threadIdx.x;
    if (i < T.end) {
        sub_start_at(o.skey2, o.n, o.nsub, o.sub_start, i);
        if (i == o.n - 1) sub_start_at(o.skey2, o.n, o.nsub, o.sub_start, o.n);
    }
}

}  // namespace v2c

using namespace v2c;

// Work lists of one shape assignment.
struct EnsPlan {
    std::vector<EnsTile> markers, cells, nodes, otiles, faults;
};

// A branch of the step graph: the models [k0, k1) and their entries in the
// marker, cell, node and fault lists, which hold the branches one after the
// other; defer_off is the branch's part of the deferred-update list.
constexpr int ENS_MAXBR = 4;
struct EnsBranch {
    int k0 = 0, k1 = 0;
    int mark_off = 0, n_mark = 0, cell_off = 0, n_cell = 0, node_off = 0, n_node = 0, fault_off = 0,
        n_fault = 0;
    long defer_off = 0;
};

struct v2c_ens {
    int device = 0;
    hipStream_t stream = nullptr;
    // the property pass of every model runs here, beside the re-ordering
    // (default priority, like `stream`), with its own copy of the arguments
    hipStream_t side = nullptr;
    hipEvent_t ev_props = nullptr;
    KArgs *d_args_p = nullptr, *h_args_p = nullptr;
    std::vector<v2c_ctx *> models;
    std::vector<int> size_class;
    v2c_tile_shape shape[V2C_ENS_CLASSES];
    int blocks_per_cu = 8, ncu = 256;
    // device buffers
    KArgs *d_args = nullptr;
    KArgs *h_args = nullptr;                 // pinned staging
    EnsTile *d_tiles = nullptr;
    long tile_cap = 0;
    int n_mark = 0, n_cell = 0, n_node = 0;
    EnsTile *t_mark = nullptr, *t_cell = nullptr, *t_node = nullptr;
    double **d_red = nullptr;
    int *d_redn = nullptr;
    double *d_rms = nullptr;
    int *d_flags = nullptr;                  // [0] unsupported marker types
    std::vector<hipEvent_t> ordered;         // per model: its ordering is done
    std::vector<long> plan_n, plan_nf;       // marker and fault counts the work lists were cut for
    int n_fault = 0;
    EnsTile *t_fault = nullptr;              // fault-list entries (local update)
    uint64_t *d_defer = nullptr;             // (model, marker) the fast fused update hands on
    long defer_cap = 0;
    // branches: contiguous groups of models whose stage sequences run on
    // their own streams inside the one graph, so that different stages of
    // different groups overlap (a VALU-bound stage of one group beside a
    // memory-bound stage of another); flags hold 4 ints per branch
    int nbranch = 1, nbranch_req = 1;
    EnsBranch br[ENS_MAXBR];
    hipStream_t bstream[ENS_MAXBR] = {};
    hipEvent_t bfork = nullptr, bjoin[ENS_MAXBR] = {};
    // optional stages: the local update (warm mode, tolerance) and adaptive steps
    int local_warm = -1;
    double local_tol = 1e-10;
    hipEvent_t ev_lfork = nullptr, ev_ljoin = nullptr;
    hipEvent_t ev_pre = nullptr;             // property pass after queued work
    int adaptive = 0;
    TsArgs ts{};
    double dt_min = 0, dt_max = 0;
    EnsTs *d_ts = nullptr;
    std::vector<void *> ts_mem;              // per-model calc_timestep scratch
    double *d_tsout = nullptr, *h_tsout = nullptr;   // 8 per model
    int *d_tserr = nullptr;
    // batched re-ordering
    bool batch_order = false;                // packed mover values fit in 64 bits
    int mov_shift = 0, mov_bits = 0;         // model number at bit mov_shift; sort bits
    int n_otile = 0;
    EnsTile *t_otile = nullptr;
    EnsOrd *d_ord = nullptr, *h_ord = nullptr;
    long *d_nmov = nullptr, *d_base = nullptr, *h_counts = nullptr;
    uint64_t *d_mov_in = nullptr, *d_mov_out = nullptr;
    void *d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    char err[512] = {0};
};

static int ens_err(v2c_ens *e, int code, const char *fmt, ...)
{
    if (e) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(e->err, sizeof e->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define ENSCHK(e, expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ens_err((e), V2C_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

static int ens_class_of(const v2c_ctx *c)
{
    if (c->n < 500000) return 0;
    if (c->n < 2000000) return 1;
    return 2;
}

static void ens_default_shape(v2c_tile_shape *s)
{
    s->cells_per_group = CPW;
    s->markers_per_tile = 16384;
    s->nodes_per_tile = 16384;
}

// Item-stage list: every model's range cut into entries of ENS_BLOCK items,
// grouped into runs of `run` consecutive entries of one model; each run goes
// to the XCD queue with the fewest entries so far, and the queues are
// interleaved so that list position b sits in queue b % 8, the XCD that the
// round-robin block dispatch sends block b to.
static void ens_item_list(std::vector<EnsTile> &out, const std::vector<std::vector<EnsTile>> &runs)
{
    constexpr int NXCD = 8;
    std::vector<EnsTile> q[NXCD];
    size_t qn[NXCD] = {0};
    for (const auto &r : runs) {
        int best = 0;
        for (int x = 1; x < NXCD; x++)
            if (qn[x] < qn[best]) best = x;
        q[best].insert(q[best].end(), r.begin(), r.end());
        qn[best] += r.size();
    }
    size_t pos[NXCD] = {0};
    for (bool more = true; more;) {
        more = false;
        for (int x = 0; x < NXCD; x++)
            if (pos[x] < q[x].size()) {
                out.push_back(q[x][pos[x]++]);
                more = true;
            }
    }
}

static void ens_cut_runs(std::vector<std::vector<EnsTile>> &runs, int mod, long n, int run)
{
    for (long b = 0; b < n; b += (long)ENS_BLOCK * run) {
        std::vector<EnsTile> r;
        const long e = std::min(n, b + (long)ENS_BLOCK * run);
        for (long t = b; t < e; t += ENS_BLOCK) r.push_back(EnsTile{mod, 0, t, std::min(e, t + ENS_BLOCK)});
        runs.push_back(std::move(r));
    }
}

// Contiguous model ranges of about equal marker counts, one per branch.
static void ens_branches(v2c_ens *e)
{
    const int nm = (int)e->models.size();
    const int nb = std::max(1, std::min(e->nbranch_req, nm));
    long total = 0;
    for (auto *c : e->models) total += c->n;
    int k = 0;
    long acc = 0;
    for (int b = 0; b < nb; b++) {
        e->br[b] = EnsBranch{};
        e->br[b].k0 = k;
        const long target = total * (b + 1) / nb;
        // at least one model per branch, and leave one for each later branch
        do {
            acc += e->models[k]->n;
            k++;
        } while (k < nm - (nb - 1 - b) && acc < target);
        if (b == nb - 1) k = nm;
        e->br[b].k1 = k;
    }
    e->nbranch = nb;
}

static EnsPlan ens_plan(v2c_ens *e)
{
    EnsPlan p;
    long defer = 0;
    for (int b = 0; b < e->nbranch; b++) {
        EnsBranch &B = e->br[b];
        std::vector<std::vector<EnsTile>> mruns, nruns;
        B.mark_off = (int)p.markers.size();
        B.cell_off = (int)p.cells.size();
        B.node_off = (int)p.nodes.size();
        B.fault_off = (int)p.faults.size();
        B.defer_off = defer;
        for (int k = B.k0; k < B.k1; k++) {
            const v2c_ctx *c = e->models[k];
            const v2c_tile_shape &s = e->shape[e->size_class[k]];
            ens_cut_runs(mruns, k, c->n, s.markers_per_tile / ENS_BLOCK);
            ens_cut_runs(nruns, k, c->nodenum, s.nodes_per_tile / ENS_BLOCK);
            for (long q = 0; q < c->ncell; q += s.cells_per_group)
                p.cells.push_back(EnsTile{k, 0, q, std::min(c->ncell, q + (long)s.cells_per_group)});
            for (long q = 0; q < c->nfault; q += LU_BLOCK)
                p.faults.push_back(EnsTile{k, 0, q, std::min(c->nfault, q + (long)LU_BLOCK)});
            defer += c->cap;
        }
        ens_item_list(p.markers, mruns);
        ens_item_list(p.nodes, nruns);
        B.n_mark = (int)p.markers.size() - B.mark_off;
        B.n_cell = (int)p.cells.size() - B.cell_off;
        B.n_node = (int)p.nodes.size() - B.node_off;
        B.n_fault = (int)p.faults.size() - B.fault_off;
    }
    for (size_t k = 0; k < e->models.size(); k++) {
        const v2c_ctx *c = e->models[k];
        for (long t = 0; t < (c->n + ORD_TILE - 1) / ORD_TILE; t++) p.otiles.push_back(EnsTile{(int)k, 0, t, t + 1});
    }
    return p;
}

static int ens_upload_plan(v2c_ens *e)
{
    ens_branches(e);
    const EnsPlan p = ens_plan(e);
    const long total = (long)(p.markers.size() + p.cells.size() + p.nodes.size() + p.otiles.size() +
                              p.faults.size());
    if (total > e->tile_cap) {
        if (e->d_tiles) (void)hipFree(e->d_tiles);
        e->d_tiles = nullptr;
        ENSCHK(e, hipMalloc(&e->d_tiles, sizeof(EnsTile) * std::max(total, 1L)));
        e->tile_cap = total;
    }
    std::vector<EnsTile> all;
    all.reserve(total);
    all.insert(all.end(), p.markers.begin(), p.markers.end());
    all.insert(all.end(), p.cells.begin(), p.cells.end());
    all.insert(all.end(), p.nodes.begin(), p.nodes.end());
    all.insert(all.end(), p.otiles.begin(), p.otiles.end());
    all.insert(all.end(), p.faults.begin(), p.faults.end());
    if (total)
        ENSCHK(e, hipMemcpyAsync(e->d_tiles, all.data(), sizeof(EnsTile) * total, hipMemcpyHostToDevice, e->stream));
    ENSCHK(e, hipStreamSynchronize(e->stream));
    e->n_mark = (int)p.markers.size();
    e->n_cell = (int)p.cells.size();
    e->n_node = (int)p.nodes.size();
    e->n_otile = (int)p.otiles.size();
    e->n_fault = (int)p.faults.size();
    e->t_mark = e->d_tiles;
    e->t_cell = e->t_mark + e->n_mark;
    e->t_node = e->t_cell + e->n_cell;
    e->t_otile = e->t_node + e->n_node;
    e->t_fault = e->t_otile + e->n_otile;
    for (size_t k = 0; k < e->models.size(); k++) {
        e->plan_n[k] = e->models[k]->n;
        e->plan_nf[k] = e->models[k]->nfault;
    }
    e->graph_valid = false;
    return 0;
}

// Persistent grid of the scatter stages (one wavefront per block).
static unsigned ens_grid(const v2c_ens *e, int ntiles)
{
    const long g = (long)e->ncu * e->blocks_per_cu * 4;
    return (unsigned)std::max(1L, std::min((long)ntiles, g));
}

// The stage sequence of branch b on stream st.  side: the slip-rate solves
// may fork to the side stream (one branch only); otherwise they run before
// the ductile solves on st.
static void ens_branch_stages(v2c_ens *e, int b, hipStream_t st, bool side)
{
    const EnsBranch &B = e->br[b];
    const int k0 = B.k0, nmb = B.k1 - B.k0;
    const EnsTile *t_mark = e->t_mark + B.mark_off, *t_cell = e->t_cell + B.cell_off;
    const EnsTile *t_node = e->t_node + B.node_off, *t_fault = e->t_fault + B.fault_off;
    int *flags = e->d_flags + 4 * b;
    uint64_t *defer = e->d_defer + B.defer_off;
    const dim3 mg((unsigned)B.n_mark), ng((unsigned)B.n_node);
    auto grid = [&](int ntiles) {
        // a multiple of 8 where it can be (ens_band's XCD bands)
        const long g = std::max(8L, ((long)e->ncu * e->blocks_per_cu * 4 / e->nbranch) / 8 * 8);
        return (unsigned)std::max(1L, std::min((long)ntiles, g));
    };
    hipLaunchKernelGGL(k_ens_clear, dim3(1), dim3(64), 0, st, flags, 4);
    // (the property pass ran on the side stream beside the re-ordering:
    // ens_props.  Inside the branches, all of them or all but the first,
    // it measured 2-6 % slower at 24 production models, where it no longer
    // overlaps the re-ordering)
    if (B.n_cell)
        hipLaunchKernelGGL(k_ens_m2n_cells, dim3(grid(B.n_cell)), dim3(WAVE), 0, st, e->d_args, t_cell, B.n_cell);
    if (B.n_node) {
        hipLaunchKernelGGL(k_ens_m2n_nodes, ng, dim3(ENS_BLOCK), 0, st, e->d_args, t_node);
        bool dphis = false;
        for (int k = B.k0; k < B.k1; k++) dphis = dphis || e->models[k]->prm.Dphis > 0;
        if (dphis) hipLaunchKernelGGL(k_ens_phis_bc, ng, dim3(ENS_BLOCK), 0, st, e->d_args, t_node);
        hipLaunchKernelGGL(k_ens_before, ng, dim3(ENS_BLOCK), 0, st, e->d_args, t_node);
        hipLaunchKernelGGL(k_ens_after, ng, dim3(RED_BLOCK), 0, st, e->d_args, e->d_red, t_node);
    }
    hipLaunchKernelGGL(k_ens_after_final, dim3(nmb), dim3(RED_BLOCK), 0, st, e->d_red, e->d_redn, e->d_rms, k0);
    if (e->local_warm >= 0) {
        // local update of every model: the slip-rate solves (fault-list
        // entries), beside the ductile solves (marker entries) on the side
        // stream when there is one branch, then the visco-plastic viscosity
        // that needs both
        hipLaunchKernelGGL(k_ens_lu_clear, dim3(nmb), dim3(64), 0, st, e->d_args, k0);
        hipStream_t rs = st;
        if (B.n_fault && side) {
            (void)hipEventRecord(e->ev_lfork, st);
            (void)hipStreamWaitEvent(e->side, e->ev_lfork, 0);
            rs = e->side;
        }
        if (B.n_fault) {
            const dim3 fg((unsigned)B.n_fault);
            if (e->local_warm == 2)
                hipLaunchKernelGGL(k_ens_local_rsf<2>, fg, dim3(LU_BLOCK), 0, rs, e->d_args, t_fault, e->local_tol);
            else if (e->local_warm == 1)
                hipLaunchKernelGGL(k_ens_local_rsf<1>, fg, dim3(LU_BLOCK), 0, rs, e->d_args, t_fault, e->local_tol);
            else
                hipLaunchKernelGGL(k_ens_local_rsf<0>, fg, dim3(LU_BLOCK), 0, rs, e->d_args, t_fault, e->local_tol);
            if (side) (void)hipEventRecord(e->ev_ljoin, e->side);
        }
        if (B.n_mark)
            hipLaunchKernelGGL(k_ens_local_ductile, mg, dim3(LU_BLOCK), 0, st, e->d_args, t_mark, e->local_warm,
                               e->local_tol);
        if (B.n_fault) {
            if (side) (void)hipStreamWaitEvent(st, e->ev_ljoin, 0);
            hipLaunchKernelGGL(k_ens_local_etavp, dim3((unsigned)B.n_fault), dim3(LU_BLOCK), 0, st, e->d_args,
                               t_fault);
        }
    }
    hipLaunchKernelGGL(k_ens_phantom, dim3((nmb + 63) / 64), dim3(64), 0, st, e->d_args, B.k0, B.k1);
    if (B.n_cell)
        hipLaunchKernelGGL(k_ens_n2m_scatter, dim3(grid(B.n_cell)), dim3(WAVE), 0, st, e->d_args, t_cell, B.n_cell);
    if (B.n_node) hipLaunchKernelGGL(k_ens_n2m_nodes, ng, dim3(ENS_BLOCK), 0, st, e->d_args, t_node);
    if (B.n_mark) {
        hipLaunchKernelGGL(k_ens_n2m_update_fast, mg, dim3(ENS_BLOCK), 0, st, e->d_args, t_mark, defer, flags + 1);
        hipLaunchKernelGGL(k_ens_n2m_update_list, dim3(N2M_LIST_BLOCKS), dim3(256), 0, st, e->d_args, defer,
                           flags + 1, flags);
    }
    if (e->adaptive && B.n_node) {
        // the next time step of every model: copyarrays, then calc_timestep
        hipLaunchKernelGGL(k_ens_copyarrays, ng, dim3(ENS_BLOCK), 0, st, e->d_args, t_node);
        hipLaunchKernelGGL(k_ens_ts_clear, dim3((nmb + 63) / 64), dim3(64), 0, st, e->d_ts, B.k0, B.k1);
        hipLaunchKernelGGL(k_ens_ts_lambda, ng, dim3(POST_BLOCK), 0, st, e->d_args, t_node, e->d_ts, e->ts);
        hipLaunchKernelGGL(k_ens_ts_carry, dim3(nmb), dim3(FIN_BLOCK), 0, st, e->d_ts, e->ts, k0);
        hipLaunchKernelGGL(k_ens_ts_node, ng, dim3(POST_BLOCK), 0, st, e->d_args, t_node, e->d_ts, e->ts);
        hipLaunchKernelGGL(k_ens_ts_final, dim3(nmb), dim3(FIN_BLOCK), 0, st, e->d_ts, e->ts, k0);
    }
}

// The graph-captured stage sequence of one step: every branch on its own
// stream, forked from and joined back to the ensemble stream.
static void ens_stages(v2c_ens *e)
{
    if (e->nbranch == 1) {
        ens_branch_stages(e, 0, e->stream, true);
        return;
    }
    (void)hipEventRecord(e->bfork, e->stream);
    for (int b = 1; b < e->nbranch; b++) (void)hipStreamWaitEvent(e->bstream[b], e->bfork, 0);
    for (int b = 0; b < e->nbranch; b++) ens_branch_stages(e, b, b ? e->bstream[b] : e->stream, false);
    for (int b = 1; b < e->nbranch; b++) {
        (void)hipEventRecord(e->bjoin[b], e->bstream[b]);
        (void)hipStreamWaitEvent(e->stream, e->bjoin[b], 0);
    }
}

// Per-step preparation outside the graph: cell ordering of every model and
// the argument array.  Without batched re-ordering every model orders on its
// own stream; the first halves (flagging and the mover count) of all models
// are launched before any count is read back, so the small ordering kernels
// of different models overlap.  The ensemble stream then waits for each
// model's ordering through an event.
static int ens_order_streams(v2c_ens *e)
{
    const size_t nm = e->models.size();
    std::vector<char> pending(nm, 0);
    for (size_t k = 0; k < nm; k++) {
        bool p = false;
        int rc = order_begin(e->models[k], &p);
        if (rc) return ens_err(e, rc, "model %zu: ordering failed: %s", k, e->models[k]->err);
        pending[k] = p;
    }
    for (size_t k = 0; k < nm; k++) {
        v2c_ctx *c = e->models[k];
        if (pending[k]) {
            ENSCHK(e, hipStreamSynchronize(c->stream));
            int rc = order_finish(c);
            if (rc) return ens_err(e, rc, "model %zu: ordering failed: %s", k, c->err);
        }
        ENSCHK(e, hipEventRecord(e->ordered[k], c->stream));
        ENSCHK(e, hipStreamWaitEvent(e->stream, e->ordered[k], 0));
    }
    return 0;
}

// Batched re-ordering on the ensemble stream (see k_eo_flags and the
// following kernels); models sorted for the first time, empty ones and those
// with more than n/4 movers get a full sort on their own stream.
static int ens_order_batched(v2c_ens *e)
{
    const int nm = (int)e->models.size();
    auto full = [&](int k) -> int {
        v2c_ctx *c = e->models[k];
        int rc = full_sort(c);
        if (rc) return ens_err(e, rc, "model %d: sort failed: %s", k, c->err);
        c->order_valid = true;
        ENSCHK(e, hipEventRecord(e->ordered[k], c->stream));
        ENSCHK(e, hipStreamWaitEvent(e->stream, e->ordered[k], 0));
        return 0;
    };
    bool any = false;
    for (int k = 0; k < nm; k++) {
        v2c_ctx *c = e->models[k];
        EnsOrd &o = e->h_ord[k];
        setp(o.key, c->d_key);
        setp(o.order, c->d_order); setp(o.skey, c->d_skey);
        setp(o.order2, c->d_order2); setp(o.skey2, c->d_skey2);
        setp(o.mask, c->d_mask);
        setp(o.tile_count, c->d_tile_count); setp(o.tile_off, c->d_tile_off); setp(o.sub_start, c->d_cs);
        o.n = c->n;
        o.ntiles = (c->n + ORD_TILE - 1) / ORD_TILE;
        o.nsub = 4 * c->ncell;
        o.active = 0;
        if (c->order_valid) continue;
        if (!c->ever_sorted || c->n == 0) {
            c->last_movers = c->n;
            int rc = full(k);
            if (rc) return rc;
            continue;
        }
        o.active = 1;
        any = true;
    }
    if (!any) return 0;
    ENSCHK(e, hipMemcpyAsync(e->d_ord, e->h_ord, sizeof(EnsOrd) * nm, hipMemcpyHostToDevice, e->stream));
    if (e->n_otile) hipLaunchKernelGGL(k_eo_flags, dim3(e->n_otile), dim3(ORD_BLOCK), 0, e->stream, e->d_ord, e->t_otile);
    hipLaunchKernelGGL(k_eo_scan, dim3(nm), dim3(1024), 0, e->stream, e->d_ord, e->d_nmov);
    hipLaunchKernelGGL(k_eo_bases, dim3(1), dim3(64), 0, e->stream, e->d_ord, e->d_nmov, nm, e->d_base);
    ENSCHK(e, hipGetLastError());
    ENSCHK(e, hipMemcpyAsync(e->h_counts, e->d_nmov, sizeof(long) * nm, hipMemcpyDeviceToHost, e->stream));
    ENSCHK(e, hipMemcpyAsync(e->h_counts + nm, e->d_base + nm, sizeof(long), hipMemcpyDeviceToHost, e->stream));
    ENSCHK(e, hipStreamSynchronize(e->stream));
    const long total = e->h_counts[nm];
    if (total > 0) {
        hipLaunchKernelGGL(k_eo_movers, dim3(e->n_otile), dim3(ORD_BLOCK), 0, e->stream, e->d_ord, e->t_otile,
                           e->d_nmov, e->d_base, e->d_mov_in, e->mov_shift);
        ENSCHK(e, hipcub::DeviceRadixSort::SortKeys(e->d_sort_tmp, e->sort_tmp_bytes, e->d_mov_in, e->d_mov_out,
                                                    (int)total, 0, e->mov_bits, e->stream));
        hipLaunchKernelGGL(k_eo_stayers, dim3(e->n_otile), dim3(ORD_BLOCK), 0, e->stream, e->d_ord, e->t_otile,
                           e->d_nmov, e->d_base, e->d_mov_out, e->mov_shift);
        hipLaunchKernelGGL(k_eo_place, dim3(blocks_for(total, 256)), dim3(256), 0, e->stream, e->d_ord, e->d_mov_out,
                           e->d_base, total, e->mov_shift);
        if (e->n_mark)
            hipLaunchKernelGGL(k_eo_sub_start, dim3(e->n_mark), dim3(ENS_BLOCK), 0, e->stream, e->d_ord, e->t_mark,
                               e->d_nmov);
        ENSCHK(e, hipGetLastError());
    }
    for (int k = 0; k < nm; k++) {
        v2c_ctx *c = e->models[k];
        const EnsOrd &o = e->h_ord[k];
        if (!o.active) continue;
        const long nmov = e->h_counts[k];
        c->last_movers = nmov;
        if (eo_incremental(o, nmov)) {
            std::swap(c->d_order, c->d_order2);
            std::swap(c->d_skey, c->d_skey2);
            c->n_incremental++;
            c->order_valid = true;
            c->cnt_valid = false;       // the batched passes recompute the starts, not the counts
        } else if (nmov == 0) {
            c->order_valid = true;
        } else {
            int rc = full(k);
            if (rc) return rc;
        }
    }
    return 0;
}

static int ens_plan_check(v2c_ens *e)
{
    const size_t nm = e->models.size();
    for (size_t k = 0; k < nm; k++)
        if (e->models[k]->n != e->plan_n[k] || e->models[k]->nfault != e->plan_nf[k])
            return ens_upload_plan(e);   // new marker or fault counts: new work lists
    return 0;
}

// The property pass of every model, on the side stream: it reads markers in
// physical order and writes only the property hand-over and marker phi,
// which the re-ordering never touches, so it runs while the ordering does.
static int ens_props(v2c_ens *e)
{
    if (!e->n_mark) return 0;
    const size_t nm = e->models.size();
    // ordered after all work queued so far on the ensemble stream and on the
    // streams of models in asynchronous mode (the only ones a caller can have
    // left work on: every other entry point returns with its stream drained,
    // and the previous step's per-model ordering work was joined into the
    // ensemble stream); the events are re-recorded by the ordering
    ENSCHK(e, hipEventRecord(e->ev_pre, e->stream));
    ENSCHK(e, hipStreamWaitEvent(e->side, e->ev_pre, 0));
    for (size_t k = 0; k < nm; k++) {
        if (!e->models[k]->async_mode) continue;
        ENSCHK(e, hipEventRecord(e->ordered[k], e->models[k]->stream));
        ENSCHK(e, hipStreamWaitEvent(e->side, e->ordered[k], 0));
    }
    for (size_t k = 0; k < nm; k++) e->h_args_p[k] = make_args(e->models[k]);
    ENSCHK(e, hipMemcpyAsync(e->d_args_p, e->h_args_p, sizeof(KArgs) * nm, hipMemcpyHostToDevice, e->side));
    hipLaunchKernelGGL(k_ens_m2n_props, dim3((unsigned)e->n_mark), dim3(ENS_BLOCK), 0, e->side, e->d_args_p,
                       e->t_mark);
    ENSCHK(e, hipGetLastError());
    ENSCHK(e, hipEventRecord(e->ev_props, e->side));
    return 0;
}

static int ens_prepare(v2c_ens *e)
{
    const size_t nm = e->models.size();
    int rc = e->batch_order ? ens_order_batched(e) : ens_order_streams(e);
    if (rc) return rc;
    for (size_t k = 0; k < nm; k++) e->h_args[k] = make_args(e->models[k]);
    ENSCHK(e, hipMemcpyAsync(e->d_args, e->h_args, sizeof(KArgs) * nm, hipMemcpyHostToDevice, e->stream));
    return 0;
}

// After a step: markers moved; every model's time advances by its step and,
// with adaptive steps, takes the next step from its calc_timestep (read back
// with the step's flags), clamped to [dt_min, dt_max].
static int ens_finish_step(v2c_ens *e)
{
    for (size_t k = 0; k < e->models.size(); k++) {
        v2c_ctx *c = e->models[k];
        c->order_valid = false;
        c->timesum += c->timestep;
        if (e->adaptive) c->timestep = std::min(e->dt_max, std::max(e->dt_min, e->h_tsout[8 * k]));
    }
    return 0;
}

// The step's read-back: unsupported-type count, and with adaptive steps the
// calc_timestep results and error flags of every model.
static int ens_readback(v2c_ens *e, int *bad)
{
    int fl[4 * ENS_MAXBR];
    ENSCHK(e, hipMemcpyAsync(fl, e->d_flags, sizeof(int) * 4 * e->nbranch, hipMemcpyDeviceToHost, e->stream));
    const size_t nm = e->models.size();
    std::vector<int> err(nm, 0);
    if (e->adaptive) {
        ENSCHK(e, hipMemcpyAsync(e->h_tsout, e->d_tsout, sizeof(double) * 8 * nm, hipMemcpyDeviceToHost, e->stream));
        ENSCHK(e, hipMemcpyAsync(err.data(), e->d_tserr, sizeof(int) * nm, hipMemcpyDeviceToHost, e->stream));
    }
    ENSCHK(e, hipStreamSynchronize(e->stream));
    *bad = 0;
    for (int b = 0; b < e->nbranch; b++) *bad += fl[4 * b];
    for (size_t k = 0; k < nm; k++)
        if (err[k])
            return ens_err(e, V2C_EUNSUPPORTED, "model %zu: calc_timestep: %s", k,
                           (err[k] & 2) ? "pmode other than 0 and 1" : "chi is zero at a node (the reference exits)");
    return 0;
}

extern "C" {

int v2c_ens_create(v2c_ens **out, v2c_ctx *const *models, int nmodels)
{
    if (!out || !models || nmodels <= 0) return V2C_EINVAL;
    *out = nullptr;
    const int dev = models[0]->device;
    for (int k = 0; k < nmodels; k++) {
        if (!models[k] || models[k]->device != dev) return V2C_EINVAL;
        for (int j = 0; j < k; j++)
            if (models[j] == models[k]) return V2C_EINVAL;      // each context once
    }
    v2c_ens *e = new v2c_ens();
    e->device = dev;
    for (int k = 0; k < nmodels; k++) {
        e->models.push_back(models[k]);
        e->size_class.push_back(ens_class_of(models[k]));
    }
    for (int s = 0; s < V2C_ENS_CLASSES; s++) ens_default_shape(&e->shape[s]);
    auto fail = [&](int code) {
        v2c_ens_destroy(e);
        return code;
    };
    if (hipSetDevice(dev) != hipSuccess) return fail(V2C_EHIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) e->ncu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) return fail(V2C_EHIP);
    if (hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) != hipSuccess) return fail(V2C_EHIP);
    if (hipEventCreateWithFlags(&e->ev_props, hipEventDisableTiming) != hipSuccess) return fail(V2C_EHIP);
    e->ordered.assign(nmodels, nullptr);
    for (int k = 0; k < nmodels; k++)
        if (hipEventCreateWithFlags(&e->ordered[k], hipEventDisableTiming) != hipSuccess) return fail(V2C_EHIP);
    if (hipMalloc(&e->d_args, sizeof(KArgs) * nmodels) != hipSuccess) return fail(V2C_ENOMEM);
    if (hipHostMalloc((void **)&e->h_args, sizeof(KArgs) * nmodels, hipHostMallocDefault) != hipSuccess)
        return fail(V2C_ENOMEM);
    if (hipMalloc(&e->d_args_p, sizeof(KArgs) * nmodels) != hipSuccess) return fail(V2C_ENOMEM);
    if (hipHostMalloc((void **)&e->h_args_p, sizeof(KArgs) * nmodels, hipHostMallocDefault) != hipSuccess)
        return fail(V2C_ENOMEM);
    std::vector<double *> reds(nmodels);
    std::vector<int> nred(nmodels);
    for (int k = 0; k < nmodels; k++) {
        reds[k] = models[k]->d_red;
        nred[k] = (int)blocks_for(models[k]->nodenum, RED_BLOCK);
    }
    if (hipMalloc(&e->d_red, sizeof(double *) * nmodels) != hipSuccess) return fail(V2C_ENOMEM);
    if (hipMalloc(&e->d_redn, sizeof(int) * nmodels) != hipSuccess) return fail(V2C_ENOMEM);
    if (hipMalloc(&e->d_rms, sizeof(double) * 2 * nmodels) != hipSuccess) return fail(V2C_ENOMEM);
    if (hipMalloc(&e->d_flags, sizeof(int) * 4 * ENS_MAXBR) != hipSuccess) return fail(V2C_ENOMEM);
    if (hipEventCreateWithFlags(&e->bfork, hipEventDisableTiming) != hipSuccess) return fail(V2C_EHIP);
    for (int b = 1; b < ENS_MAXBR; b++)
        if (hipStreamCreateWithFlags(&e->bstream[b], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e->bjoin[b], hipEventDisableTiming) != hipSuccess)
            return fail(V2C_EHIP);
    if (hipMemcpy(e->d_red, reds.data(), sizeof(double *) * nmodels, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->d_redn, nred.data(), sizeof(int) * nmodels, hipMemcpyHostToDevice) != hipSuccess)
        return fail(V2C_EHIP);
    // batched re-ordering: model number above the widest sub-cell key
    int kb = 1, mb = 0;
    long capsum = 0;
    for (int k = 0; k < nmodels; k++) {
        kb = std::max(kb, models[k]->key_bits);
        capsum += models[k]->cap;
    }
    while ((1 << mb) < nmodels) mb++;
    e->mov_shift = 32 + kb;
    e->mov_bits = e->mov_shift + mb;
    e->batch_order = e->mov_bits <= 64 && capsum < (1L << 31);
    if (e->batch_order) {
        if (hipMalloc(&e->d_ord, sizeof(EnsOrd) * nmodels) != hipSuccess ||
            hipHostMalloc((void **)&e->h_ord, sizeof(EnsOrd) * nmodels, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&e->d_nmov, sizeof(long) * nmodels) != hipSuccess ||
            hipMalloc(&e->d_base, sizeof(long) * (nmodels + 1)) != hipSuccess ||
            hipHostMalloc((void **)&e->h_counts, sizeof(long) * (nmodels + 1), hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&e->d_mov_in, sizeof(uint64_t) * std::max(capsum, 1L)) != hipSuccess ||
            hipMalloc(&e->d_mov_out, sizeof(uint64_t) * std::max(capsum, 1L)) != hipSuccess)
            return fail(V2C_ENOMEM);
        if (hipcub::DeviceRadixSort::SortKeys(nullptr, e->sort_tmp_bytes, e->d_mov_in, e->d_mov_out,
                                              (int)std::max(capsum, 1L), 0, e->mov_bits, e->stream) != hipSuccess)
            return fail(V2C_EHIP);
        if (hipMalloc(&e->d_sort_tmp, std::max(e->sort_tmp_bytes, (size_t)16)) != hipSuccess) return fail(V2C_ENOMEM);
    }
    if (hipMalloc(&e->d_defer, sizeof(uint64_t) * std::max(capsum, 1L)) != hipSuccess) return fail(V2C_ENOMEM);
    e->defer_cap = capsum;
    if (hipEventCreateWithFlags(&e->ev_lfork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_ljoin, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_pre, hipEventDisableTiming) != hipSuccess)
        return fail(V2C_EHIP);
    e->plan_n.assign(nmodels, -1);
    e->plan_nf.assign(nmodels, -1);
    if (ens_upload_plan(e)) return fail(V2C_EHIP);
    *out = e;
    return 0;
}

void v2c_ens_destroy(v2c_ens *e)
{
    if (!e) return;
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->side) (void)hipStreamSynchronize(e->side);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    void *ptrs[] = {e->d_args, e->d_args_p, e->d_tiles, e->d_red, e->d_redn, e->d_rms, e->d_flags, e->d_ord,
                    e->d_nmov, e->d_base, e->d_mov_in, e->d_mov_out, e->d_sort_tmp, e->d_defer, e->d_ts,
                    e->d_tsout, e->d_tserr};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (void *p : e->ts_mem) (void)hipFree(p);
    if (e->h_tsout) (void)hipHostFree(e->h_tsout);
    if (e->ev_lfork) (void)hipEventDestroy(e->ev_lfork);
    if (e->ev_ljoin) (void)hipEventDestroy(e->ev_ljoin);
    if (e->ev_pre) (void)hipEventDestroy(e->ev_pre);
    if (e->bfork) (void)hipEventDestroy(e->bfork);
    for (int b = 1; b < ENS_MAXBR; b++) {
        if (e->bstream[b]) {
            (void)hipStreamSynchronize(e->bstream[b]);
            (void)hipStreamDestroy(e->bstream[b]);
        }
        if (e->bjoin[b]) (void)hipEventDestroy(e->bjoin[b]);
    }
    if (e->h_args) (void)hipHostFree(e->h_args);
    if (e->h_args_p) (void)hipHostFree(e->h_args_p);
    if (e->ev_props) (void)hipEventDestroy(e->ev_props);
    if (e->side) (void)hipStreamDestroy(e->side);
    if (e->h_ord) (void)hipHostFree(e->h_ord);
    if (e->h_counts) (void)hipHostFree(e->h_counts);
    for (auto ev : e->ordered)
        if (ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *v2c_ens_last_error(const v2c_ens *e) { return e ? e->err : "null ensemble"; }

int v2c_ens_size_class(const v2c_ens *e, int model)
{
    if (!e || model < 0 || model >= (int)e->models.size()) return V2C_EINVAL;
    return e->size_class[model];
}

int v2c_ens_set_shape(v2c_ens *e, int size_class, const v2c_tile_shape *s)
{
    if (!e || !s || size_class < 0 || size_class >= V2C_ENS_CLASSES) return V2C_EINVAL;
    if (s->cells_per_group < 1 || s->cells_per_group > 1024 || s->markers_per_tile < ENS_BLOCK ||
        s->markers_per_tile % ENS_BLOCK || s->markers_per_tile > (1 << 24) || s->nodes_per_tile < ENS_BLOCK ||
        s->nodes_per_tile % ENS_BLOCK || s->nodes_per_tile > (1 << 24))
        return ens_err(e, V2C_EINVAL, "tile shape out of range");
    e->shape[size_class] = *s;
    ENSCHK(e, hipSetDevice(e->device));
    return ens_upload_plan(e);
}

int v2c_ens_get_shape(const v2c_ens *e, int size_class, v2c_tile_shape *s)
{
    if (!e || !s || size_class < 0 || size_class >= V2C_ENS_CLASSES) return V2C_EINVAL;
    *s = e->shape[size_class];
    return 0;
}

int v2c_ens_set_blocks_per_cu(v2c_ens *e, int blocks)
{
    if (!e || blocks < 1 || blocks > 64) return V2C_EINVAL;
    e->blocks_per_cu = blocks;
    e->graph_valid = false;
    return 0;
}

int v2c_ens_set_branches(v2c_ens *e, int branches)
{
    if (!e || branches < 1 || branches > ENS_MAXBR) return V2C_EINVAL;
    ENSCHK(e, hipSetDevice(e->device));
    e->nbranch_req = branches;
    return ens_upload_plan(e);
}

int v2c_ens_get_branches(const v2c_ens *e) { return e ? e->nbranch : V2C_EINVAL; }

int v2c_ens_step(v2c_ens *e, int use_graph)
{
    if (!e) return V2C_EINVAL;
    ENSCHK(e, hipSetDevice(e->device));
    int rc = ens_plan_check(e);
    if (!rc) rc = ens_props(e);
    if (rc) return rc;
    rc = ens_prepare(e);
    if (e->n_mark) ENSCHK(e, hipStreamWaitEvent(e->stream, e->ev_props, 0));
    if (rc) return rc;
    if (!use_graph) {
        ens_stages(e);
        ENSCHK(e, hipGetLastError());
    } else {
        if (!e->graph_valid) {
            if (e->exec) (void)hipGraphExecDestroy(e->exec);
            if (e->graph) (void)hipGraphDestroy(e->graph);
            e->exec = nullptr;
            e->graph = nullptr;
            ENSCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            ens_stages(e);
            hipError_t ce = hipStreamEndCapture(e->stream, &e->graph);
            if (ce != hipSuccess) return ens_err(e, V2C_EHIP, "graph capture failed: %s", hipGetErrorString(ce));
            ENSCHK(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
            e->graph_valid = true;
        }
        ENSCHK(e, hipGraphLaunch(e->exec, e->stream));
    }
    int bad = 0;
    rc = ens_readback(e, &bad);
    if (rc) return rc;
    ens_finish_step(e);
    if (bad) return ens_err(e, V2C_EUNSUPPORTED, "%d markers of type >= 100 (immobile spawning is out of scope)", bad);
    return 0;
}

int v2c_ens_order(v2c_ens *e)
{
    if (!e) return V2C_EINVAL;
    ENSCHK(e, hipSetDevice(e->device));
    int rc = e->batch_order ? ens_order_batched(e) : ens_order_streams(e);
    if (rc) return rc;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    for (auto *c : e->models) ENSCHK(e, hipStreamSynchronize(c->stream));
    return 0;
}

// The same step as per-model launches of the single-model entry points.
int v2c_ens_step_looped(v2c_ens *e)
{
    if (!e) return V2C_EINVAL;
    for (size_t k = 0; k < e->models.size(); k++) {
        v2c_ctx *c = e->models[k];
        const int was_async = c->async_mode;
        c->async_mode = 1;
        const double dt = c->timestep;
        int rc = v2c_marker2nodes(c);
        if (!rc) rc = v2c_compute_before_solve(c, dt, 0, nullptr);
        if (!rc) rc = v2c_compute_after_solve(c, dt, 0, nullptr, nullptr);
        if (!rc && e->local_warm >= 0) rc = v2c_local_update(c, dt, e->local_warm, e->local_tol, nullptr);
        if (!rc) rc = v2c_nodes2markers_advect(c);
        c->async_mode = was_async;
        if (rc) return ens_err(e, rc, "model %zu: %s", k, c->err);
    }
    for (size_t k = 0; k < e->models.size(); k++) {
        v2c_ctx *c = e->models[k];
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double next = 0;
        if (e->adaptive) {
            int rc = v2c_copyarrays(c, 0);
            v2c_timestep_params tp{e->ts.maxtmstep, e->ts.tcnvp, e->ts.tcd, e->ts.tcphih, e->ts.dtfactor,
                                   e->ts.lambda0};
            v2c_timestep_result r;
            if (!rc) rc = v2c_calc_timestep(c, e->ts.oldornew, &tp, &r);
            if (rc) return ens_err(e, rc, "model %zu: %s", k, c->err);
            next = std::min(e->dt_max, std::max(e->dt_min, r.dt));
        }
        c->timesum += c->timestep;
        if (e->adaptive) c->timestep = next;
    }
    return 0;
}

// Adds the per-marker local update to every batched step (after
// compute_after_solve); warm < 0 removes it.
int v2c_ens_set_local(v2c_ens *e, int warm, double tol)
{
    if (!e || warm > 2 || (warm >= 0 && !(tol > 0))) return V2C_EINVAL;
    if (warm == 2)
        for (size_t k = 0; k < e->models.size(); k++)
            if (!e->models[k]->mlp_ready)
                return ens_err(e, V2C_EINVAL, "model %zu: no MLP parameters loaded (v2c_mlp_set)", k);
    e->local_warm = warm < 0 ? -1 : warm;
    e->local_tol = tol;
    e->graph_valid = false;
    return 0;
}

int v2c_ens_local_stats(v2c_ens *e, int model, v2c_lu_stats *st)
{
    if (!e || !st || model < 0 || model >= (int)e->models.size()) return V2C_EINVAL;
    v2c_ctx *c = e->models[model];
    unsigned long long h[LS_COUNT];
    ENSCHK(e, hipMemcpy(h, c->d_lu, sizeof h, hipMemcpyDeviceToHost));
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
    return 0;
}

int v2c_ens_set_adaptive(v2c_ens *e, const v2c_timestep_params *p, int oldornew, double dt_min, double dt_max)
{
    if (!e || (p && (oldornew < 0 || oldornew > 1 || !(dt_min > 0) || !(dt_max >= dt_min)))) return V2C_EINVAL;
    ENSCHK(e, hipSetDevice(e->device));
    e->graph_valid = false;
    if (!p) {
        e->adaptive = 0;
        return 0;
    }
    const size_t nm = e->models.size();
    if (!e->d_ts) {
        std::vector<EnsTs> h(nm);
        ENSCHK(e, hipMalloc(&e->d_ts, sizeof(EnsTs) * nm));
        ENSCHK(e, hipMalloc(&e->d_tsout, sizeof(double) * 8 * nm));
        ENSCHK(e, hipMalloc(&e->d_tserr, sizeof(int) * nm));
        ENSCHK(e, hipHostMalloc((void **)&e->h_tsout, sizeof(double) * 8 * nm, hipHostMallocDefault));
        for (size_t k = 0; k < nm; k++) {
            const int nb = (int)blocks_for(e->models[k]->nodenum, POST_BLOCK);
            void *mem = nullptr;
            ENSCHK(e, hipMalloc(&mem, sizeof(Fold) * NTS * nb + sizeof(double) * 2 * nb + sizeof(int) * nb));
            e->ts_mem.push_back(mem);
            Fold *part = (Fold *)mem;
            double *bval = (double *)(part + (size_t)NTS * nb);
            setp(h[k].part, part);
            setp(h[k].bval, bval);
            setp(h[k].carry, bval + nb);
            setp(h[k].bset, (int *)(bval + 2 * (size_t)nb));
            setp(h[k].out, e->d_tsout + 8 * k);
            setp(h[k].err, e->d_tserr + k);
            h[k].nblocks = nb;
            h[k].pad = 0;
        }
        ENSCHK(e, hipMemcpy(e->d_ts, h.data(), sizeof(EnsTs) * nm, hipMemcpyHostToDevice));
    }
    e->ts = TsArgs{p->maxtmstep, p->tcnvp, p->tcd, p->tcphih, p->dtfactor, p->lambda0, oldornew};
    e->dt_min = dt_min;
    e->dt_max = dt_max;
    e->adaptive = 1;
    return 0;
}

int v2c_ens_times(v2c_ens *e, double *timesum_timestep)
{
    if (!e || !timesum_timestep) return V2C_EINVAL;
    for (size_t k = 0; k < e->models.size(); k++) {
        timesum_timestep[2 * k] = e->models[k]->timesum;
        timesum_timestep[2 * k + 1] = e->models[k]->timestep;
    }
    return 0;
}

int v2c_ens_rms(v2c_ens *e, double *rms_count)
{
    if (!e || !rms_count) return V2C_EINVAL;
    ENSCHK(e, hipMemcpy(rms_count, e->d_rms, sizeof(double) * 2 * e->models.size(), hipMemcpyDeviceToHost));
    return 0;
}

// Coordinate descent over the tile shapes of each size class present, then
// the persistent grid size: each candidate is timed over `reps` graph-replayed
// steps (after one untimed step) and the fastest kept.  The time is the
// graph's: the re-ordering and the property pass beside it run before.  The ensemble's state
// advances by the steps run.
int v2c_ens_autotune(v2c_ens *e, int reps, double *best_ms)
{
    if (!e || reps < 1) return V2C_EINVAL;
    hipEvent_t t0, t1;
    ENSCHK(e, hipEventCreate(&t0));
    ENSCHK(e, hipEventCreate(&t1));
    auto measure = [&](double &ms) -> int {
        int rc = v2c_ens_step(e, 1);
        if (rc) return rc;
        double tot = 0;
        for (int r = 0; r < reps; r++) {
            rc = ens_plan_check(e);
            if (!rc) rc = ens_props(e);
            if (!rc) rc = ens_prepare(e);
            if (e->n_mark) ENSCHK(e, hipStreamWaitEvent(e->stream, e->ev_props, 0));
            if (rc) return rc;
            ENSCHK(e, hipEventSynchronize(e->ev_props));    // time the graph alone
            ENSCHK(e, hipEventRecord(t0, e->stream));
            ENSCHK(e, hipGraphLaunch(e->exec, e->stream));
            ENSCHK(e, hipEventRecord(t1, e->stream));
            ENSCHK(e, hipEventSynchronize(t1));
            float m = 0;
            ENSCHK(e, hipEventElapsedTime(&m, t0, t1));
            tot += m;
            int bad = 0;
            rc = ens_readback(e, &bad);
            if (rc) return rc;
            ens_finish_step(e);
        }
        ms = tot / reps;
        return 0;
    };
    bool present[V2C_ENS_CLASSES] = {false};
    for (int c : e->size_class) present[c] = true;
    double best = 0;
    int rc = measure(best);
    const int cpg[] = {8, 16, 32, 64};
    const int mpt[] = {256, 4096, 65536};
    const int npt[] = {256, 4096, 65536};
    for (int cls = 0; cls < V2C_ENS_CLASSES && !rc; cls++) {
        if (!present[cls]) continue;
        for (int pass = 0; pass < 3 && !rc; pass++) {
            const int *cand = pass == 0 ? cpg : (pass == 1 ? mpt : npt);
            const int ncand = pass == 0 ? 4 : 3;
            v2c_tile_shape keep = e->shape[cls];
            for (int q = 0; q < ncand && !rc; q++) {
                v2c_tile_shape s = keep;
                if (pass == 0) s.cells_per_group = cand[q];
                else if (pass == 1) s.markers_per_tile = cand[q];
                else s.nodes_per_tile = cand[q];
                if (memcmp(&s, &keep, sizeof s) == 0) continue;
                rc = v2c_ens_set_shape(e, cls, &s);
                double ms = 0;
                if (!rc) rc = measure(ms);
                if (!rc && ms < best) { best = ms; keep = s; }
            }
            if (!rc) rc = v2c_ens_set_shape(e, cls, &keep);
        }
    }
    const int bpc[] = {2, 4, 8, 16};
    int keepb = e->blocks_per_cu;
    for (int q = 0; q < 4 && !rc; q++) {
        if (bpc[q] == keepb) continue;
        v2c_ens_set_blocks_per_cu(e, bpc[q]);
        double ms = 0;
        rc = measure(ms);
        if (!rc && ms < best) { best = ms; keepb = bpc[q]; }
    }
    if (!rc) v2c_ens_set_blocks_per_cu(e, keepb);
    // concurrent branches of model groups
    int keepbr = e->nbranch_req;
    for (int b = 1; b <= ENS_MAXBR && !rc; b++) {
        if (b == keepbr || b > (int)e->models.size()) continue;
        rc = v2c_ens_set_branches(e, b);
        double ms = 0;
        if (!rc) rc = measure(ms);
        if (!rc && ms < best) { best = ms; keepbr = b; }
    }
    if (!rc) rc = v2c_ens_set_branches(e, keepbr);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (best_ms) *best_ms = best;
    return rc;
}

}  // extern "C"
