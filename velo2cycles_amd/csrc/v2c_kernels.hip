// v2c_kernels.hip -- HIP kernels of the per-marker hot path for gfx950.
//
// Marker state is an FP64 structure-of-arrays resident in HBM.  Markers are
// ordered by sub-cell key (cell, quadrant) with a stable radix sort of
// (key, marker id); under interpolationmod == 2 every marker of one
// (cell, quadrant) scatters to the same basic, vx, vy and centre node, so the
// marker->node reductions run as deterministic segmented sums.  Each pass is
// split by access pattern:
//
//   per-marker kernels  (m2n_props, n2m_update) one thread per marker in
//                    physical order: material laws, gathers, RK4 advection;
//                    coalesced, no LDS, occupancy set by registers only;
//   scatter kernels  (m2n_cells, n2m_scatter) one 64-lane wavefront per group
//                    of CPW consecutive cells; lanes take 64 markers at a time
//                    in sorted order and put their contributions in LDS; every
//                    (cell, quadrant) run is summed in lane order by work items
//                    (run, field) and written once per sub-cell; a run that
//                    continues into the next batch is carried in LDS;
//   node kernels     one thread per node adds the <= 4 sub-cell sums that map
//                    to it in a fixed order and normalises.
//
// No floating-point atomics are used anywhere, so results are bitwise
// reproducible run to run.
#include "v2c_device.hpp"

namespace v2c {

constexpr int WAVE = 64;
constexpr int CPW = 16;          // cells per wavefront group

// Contribution layout of one marker, m2n (interpolation.c:2-52 accumulators):
//   basic  [0,14): nu, nu/G, e, sxy, exy, C, phi, ro, et, ht, cp, kt, T, w
//   vx     [14,17): v*ro*w, ro*w, w      vy [17,20): same
//   centre [20,27): bm, nu, nu/G, sxx, exx, P, w
constexpr int M_NB = 14, M_NV = 3, M_NC = 7;
constexpr int M_VX = M_NB, M_VY = M_NB + M_NV, M_C = M_NB + 2 * M_NV, M_NF = M_C + M_NC;

// n2m antidiffusion sums (interpolation.c:1159-1209, 1360-1374, 1405-1489):
//   basic [0,5): PHIS*w, SXY*w, EXY*w, ES*w, w
//   vx [5,8): VX*ro*w, ro*w, w   vy [8,11)   centre [11,15): PR*w, SXX*w, EXX*w, w
constexpr int N_NB = 5, N_NV = 3, N_NC = 4;
constexpr int N_VX = N_NB, N_VY = N_NB + N_NV, N_C = N_NB + 2 * N_NV, N_NF = N_C + N_NC;

// Map workgroup ids so each of the 8 XCDs gets a contiguous range of cells
// (dispatch assigns workgroups round-robin over XCDs; neighbouring cells share
// nodal data, which then stays in one XCD's L2).
__device__ __forceinline__ int xcd_remap(int b, int nwg)
{
    int full = (nwg / 8) * 8;
    if (b >= full) return b;
    return (b % 8) * (full / 8) + b / 8;
}

// ---------------------------------------------------------------------------
// Keys, ordering
// ---------------------------------------------------------------------------
__global__ void k_keys(KArgs a)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    a.key[i] = subcell_key(a.g, MK(a, MF_X, i), MK(a, MF_Y, i));
}

__global__ void k_iota(uint32_t *v, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

// ss[k] = first sorted position whose sub-cell key is >= k; ss[nsub] = n.
// Position i (0 <= i <= n) writes the entries between its predecessor's key
// and its own.
__device__ __forceinline__ void sub_start_at(const uint32_t *__restrict__ skey, long n, long nsub, uint32_t *ss, long i)
{
    long cur = i < n ? (long)skey[i] : nsub;
    long prev = i > 0 ? (long)skey[i - 1] : -1;
    for (long k = prev + 1; k <= cur && k <= nsub; k++) ss[k] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// Incremental re-ordering after advection.  The order is always the (key, id)
// lexicographic order (a stable radix sort of keys over ids produces exactly
// that).  Markers whose key did not change ("stayers") keep that order among
// themselves, so the new order is the merge of the stayers, already sorted,
// with the few movers sorted by (key, id): a stayer lands at its rank among
// stayers plus the number of movers before it, a mover at its rank among
// movers plus the number of stayers before it.
//
// The old sorted positions are cut into tiles of ORD_TILE; a 64-bit ballot
// mask per 64 positions marks the movers, and an exclusive scan of the tile
// counts gives every position's mover prefix without a global compaction.
//
// Movers are sorted as 64-bit values (key << 32 | id), optionally with
// constant high bits `hi` above the key (a batched launch sorts the movers
// of several models at once, with the model number there); comparisons
// against stayers carry the same bits, and `low` masks them off again.
// ---------------------------------------------------------------------------
constexpr int ORD_BLOCK = 256, ORD_ITEMS = 16, ORD_TILE = ORD_BLOCK * ORD_ITEMS;
constexpr int ORD_MASKS = ORD_TILE / WAVE;        // masks per tile

__device__ __forceinline__ uint64_t kid(uint32_t k, uint32_t id) { return ((uint64_t)k << 32) | id; }

// Sub-cell starts over positions [t0, t0 + ORD_TILE) (and n, in the last
// tile), ORD_ITEMS positions per thread with every key load issued before
// the first store: a workgroup's waves stay resident for a few memory
// latencies rather than one per position.
__global__ __launch_bounds__(ORD_BLOCK) void k_sub_start(const uint32_t *__restrict__ skey, long n, long nsub,
                                                         uint32_t *ss)
{
    if (n == 0) {
        if (blockIdx.x == 0)
            for (long q = threadIdx.x; q <= nsub; q += ORD_BLOCK) ss[q] = 0u;
        return;
    }
    const long ntiles = (n + 1 + ORD_TILE - 1) / ORD_TILE;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long t0 = tile * ORD_TILE;
    // unconditional loads at clamped positions, the tile's out-of-range
    // cases resolved after all of them are in flight
    uint32_t rc[ORD_ITEMS], rp[ORD_ITEMS];
#pragma unroll
    for (int k = 0; k < ORD_ITEMS; k++) {
        const long i = t0 + (long)k * ORD_BLOCK + threadIdx.x;
        rc[k] = skey[i < n ? i : n - 1];
        rp[k] = skey[i <= 0 ? 0 : (i <= n ? i - 1 : n - 1)];
    }
#pragma unroll
    for (int k = 0; k < ORD_ITEMS; k++) {
        const long i = t0 + (long)k * ORD_BLOCK + threadIdx.x;
        if (i > n) continue;
        const long cur = i < n ? (long)rc[k] : nsub;
        const long prev = i > 0 ? (long)rp[k] : -1;
        for (long q = prev + 1; q <= cur && q <= nsub; q++) ss[q] = (uint32_t)i;
    }
    }
}

// Markers per sub-cell from the sub-cell starts: cnt[k] = ss[k+1] - ss[k],
// cnt[nsub] = 0 (so that an exclusive scan of cnt gives ss back).
__global__ void k_sub_count(const uint32_t *__restrict__ ss, long nsub, uint32_t *__restrict__ cnt)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nsub) cnt[k] = ss[k + 1] - ss[k];
    else if (k == nsub) cnt[k] = 0u;
}

// Pass 1: mover masks and mover count of one tile.
__device__ __forceinline__ void order_flags_tile(const uint32_t *__restrict__ order, const uint32_t *__restrict__ key,
                                                 const uint32_t *__restrict__ skey, long n,
                                                 unsigned long long *__restrict__ mask,
                                                 uint32_t *__restrict__ tile_count, long tile)
{
    __shared__ uint32_t wsum[ORD_BLOCK / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const long t0 = tile * ORD_TILE;
    // all loads of the tile first (ids, then their keys), so a wave waits
    // two memory latencies rather than two per position
    uint32_t id[ORD_ITEMS], sk[ORD_ITEMS], kk[ORD_ITEMS];
#pragma unroll
    for (int k = 0; k < ORD_ITEMS; k++) {
        const long i = t0 + (long)k * ORD_BLOCK + threadIdx.x;
        id[k] = i < n ? order[i] : 0u;
        sk[k] = i < n ? skey[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < ORD_ITEMS; k++) kk[k] = t0 + (long)k * ORD_BLOCK + threadIdx.x < n ? key[id[k]] : 0u;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < ORD_ITEMS; k++) {
        const long i = t0 + (long)k * ORD_BLOCK + threadIdx.x;
        const bool mv = i < n && kk[k] != sk[k];
        const unsigned long long b = __ballot(mv);
        if (lane == 0 && t0 + (long)k * ORD_BLOCK + wv * WAVE < n) mask[(t0 + (long)k * ORD_BLOCK) / WAVE + wv] = b;
        cnt += (uint32_t)__popcll(b);
    }
    if (lane == 0) wsum[wv] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < ORD_BLOCK / WAVE; w++) t += wsum[w];
        tile_count[tile] = t;
    }
}

__global__ __launch_bounds__(ORD_BLOCK) void k_order_flags(const uint32_t *__restrict__ order,
                                                           const uint32_t *__restrict__ key,
                                                           const uint32_t *__restrict__ skey, long n,
                                                           unsigned long long *__restrict__ mask,
                                                           uint32_t *__restrict__ tile_count)
{
    const long ntiles = (n + ORD_TILE - 1) / ORD_TILE;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();
        order_flags_tile(order, key, skey, n, mask, tile_count, tile);
    }
}

// Pass 2: exclusive scan of the tile counts (one 1024-thread workgroup);
// off[ntiles] = total.
__device__ __forceinline__ void order_scan_block(const uint32_t *__restrict__ cnt, long ntiles,
                                                 uint32_t *__restrict__ off, long *__restrict__ total)
{
    __shared__ uint32_t part[1024];
    const long per = (ntiles + 1023) / 1024;
    const long b = threadIdx.x * per, e = b + per < ntiles ? b + per : ntiles;
    uint32_t s = 0;
    for (long i = b; i < e; i++) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {          // inclusive Hillis-Steele scan
        const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0u;
    for (long i = b; i < e; i++) { off[i] = run; run += cnt[i]; }
    if (threadIdx.x == 1023) { off[ntiles] = part[1023]; *total = part[1023]; }
}

__global__ __launch_bounds__(1024) void k_order_scan(const uint32_t *__restrict__ cnt, long ntiles,
                                                     uint32_t *__restrict__ off, long *__restrict__ total)
{
    order_scan_block(cnt, ntiles, off, total);
}

// Movers before old position i: tile offset + masks of the tile before i's
// mask + the bits of i's mask below i.
__device__ __forceinline__ uint32_t movers_before(const unsigned long long *__restrict__ mask,
                                                  const uint32_t *__restrict__ off, long i)
{
    const long tile = i / ORD_TILE, w = i / WAVE;
    uint32_t c = off[tile];
    for (long q = tile * ORD_MASKS; q < w; q++) c += (uint32_t)__popcll(mask[q]);
    const int bit = (int)(i & (WAVE - 1));
    return c + (uint32_t)__popcll(mask[w] & ((1ull << bit) - 1ull));
}

// Pass 3: the tile's movers' (key, id) | hi, compacted in old-position order.
__device__ __forceinline__ void order_movers_tile(const uint32_t *__restrict__ order,
                                                  const uint32_t *__restrict__ key, long n,
                                                  const unsigned long long *__restrict__ mask,
                                                  const uint32_t *__restrict__ off, uint64_t *__restrict__ mov,
                                                  uint64_t hi, long tile,
                                                  const uint32_t *__restrict__ skey = nullptr,
                                                  uint32_t *__restrict__ cnt = nullptr)
{
    __shared__ uint32_t wpre[ORD_BLOCK / WAVE + 1];
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const long t0 = tile * ORD_TILE;
    uint32_t base = off[tile];
    if (off[tile + 1] == base) return;            // no movers in this tile
    for (int k = 0; k < ORD_ITEMS; k++) {
        const long c0 = t0 + (long)k * ORD_BLOCK;
        if (c0 >= n) break;
        const long i = c0 + threadIdx.x;
        const unsigned long long b = c0 + wv * WAVE < n ? mask[c0 / WAVE + wv] : 0ull;
        if (lane == 0) wpre[wv] = (uint32_t)__popcll(b);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t r = 0;
            for (int w = 0; w < ORD_BLOCK / WAVE; w++) { const uint32_t t = wpre[w]; wpre[w] = r; r += t; }
            wpre[ORD_BLOCK / WAVE] = r;
        }
        __syncthreads();
        if ((b >> lane) & 1ull) {
            const uint32_t id = order[i];
            const uint32_t k = key[id];
            mov[base + wpre[wv] + (uint32_t)__popcll(b & ((1ull << lane) - 1ull))] = kid(k, id) | hi;
            if (cnt) {
                // the sub-cell counts of the new order: one marker fewer in its
                // old sub-cell, one more in its new one (integer atomics: the
                // totals do not depend on their order)
                atomicSub(&cnt[skey[i]], 1u);
                atomicAdd(&cnt[k], 1u);
            }
        }
        base += wpre[ORD_BLOCK / WAVE];
        __syncthreads();
    }
}

__global__ __launch_bounds__(ORD_BLOCK) void k_order_movers(const uint32_t *__restrict__ order,
                                                            const uint32_t *__restrict__ key, long n,
                                                            const unsigned long long *__restrict__ mask,
                                                            const uint32_t *__restrict__ off,
                                                            uint64_t *__restrict__ mov,
                                                            const uint32_t *__restrict__ skey, uint32_t *__restrict__ cnt)
{
    const long ntiles = (n + ORD_TILE - 1) / ORD_TILE;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();
        order_movers_tile(order, key, n, mask, off, mov, 0ull, tile, skey, cnt);
    }
}

__device__ __forceinline__ long lower_bound64(const uint64_t *__restrict__ a, long lo, long hi, uint64_t c)
{
    while (lo < hi) {
        const long mid = (lo + hi) >> 1;
        if (a[mid] < c) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Pass 5: the tile's stayers to their new positions.  The tile's masks and
// their prefix counts go to LDS once; the movers that can precede any of the
// tile's stayers are confined to mov[lo, hi), found once per tile and staged
// in LDS when there are at most ORD_WIN of them; every id and key of the
// tile is loaded before the first placement.
constexpr int ORD_WIN = 256;
static_assert(ORD_MASKS == WAVE, "one wave scans a tile's masks");

__device__ __forceinline__ long lower_bound_lds(const uint64_t *a, long m, uint64_t c)
{
    long lo = 0, hi = m;
    while (lo < hi) {
        const long mid = (lo + hi) >> 1;
        if (a[mid] < c) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void order_stayers_tile(const uint32_t *__restrict__ order,
                                                   const uint32_t *__restrict__ skey, long n,
                                                   const unsigned long long *__restrict__ mask,
                                                   const uint32_t *__restrict__ off,
                                                   const uint64_t *__restrict__ mov, long nmov, uint64_t hib,
                                                   uint32_t *__restrict__ order_out,
                                                   uint32_t *__restrict__ skey_out, long tile)
{
    __shared__ long win[2];
    __shared__ unsigned long long msk[ORD_MASKS];
    __shared__ uint32_t mpre[ORD_MASKS];
    __shared__ uint64_t wmov[ORD_WIN];
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const long t0 = tile * ORD_TILE;
    const long t1 = t0 + ORD_TILE < n ? t0 + ORD_TILE : n;
    if (threadIdx.x == 0) {
        // stayer keys of this tile lie in [skey[t0], skey[t1-1]]
        win[0] = lower_bound64(mov, 0, nmov, kid(skey[t0], 0u) | hib);
        win[1] = lower_bound64(mov, win[0], nmov, (kid(skey[t1 - 1], 0xffffffffu) | hib) + 1ull);
        if (skey[t1 - 1] == 0xffffffffu) win[1] = nmov;
    }
    if (wv == 0) {                                  // the tile's masks, exclusive popcount prefix
        const unsigned long long b = t0 + (long)lane * WAVE < n ? mask[t0 / WAVE + lane] : 0ull;
        const uint32_t own = (uint32_t)__popcll(b);
        uint32_t c = own;
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint32_t v = __shfl_up(c, d);
            if (lane >= d) c += v;
        }
        msk[lane] = b;
        mpre[lane] = c - own;
    }
    uint32_t id[ORD_ITEMS], sk[ORD_ITEMS];
#pragma unroll
    for (int k = 0; k < ORD_ITEMS; k++) {
        const long i = t0 + (long)k * ORD_BLOCK + threadIdx.x;
        id[k] = i < n ? order[i] : 0u;
        sk[k] = i < n ? skey[i] : 0u;
    }
    __syncthreads();
    const long lo = win[0], hi = win[1];
    const bool staged = hi - lo <= ORD_WIN;
    if (staged)
        for (long j = lo + threadIdx.x; j < hi; j += ORD_BLOCK) wmov[j - lo] = mov[j];
    __syncthreads();
    const uint32_t moved = off[tile];               // movers before the tile
#pragma unroll
    for (int k = 0; k < ORD_ITEMS; k++) {
        const long i = t0 + (long)k * ORD_BLOCK + threadIdx.x;
        const int q = k * (ORD_BLOCK / WAVE) + wv;
        const unsigned long long b = msk[q];
        if (i < n && !((b >> lane) & 1ull)) {
            const uint64_t c = kid(sk[k], id[k]) | hib;
            const uint32_t mb = moved + mpre[q] + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
            const long before = staged ? lo + lower_bound_lds(wmov, hi - lo, c) : lower_bound64(mov, lo, hi, c);
            const long pos = (i - (long)mb) + before;
            order_out[pos] = id[k];
            skey_out[pos] = sk[k];
        }
    }
}

__global__ __launch_bounds__(ORD_BLOCK) void k_order_stayers(const uint32_t *__restrict__ order,
                                                             const uint32_t *__restrict__ skey, long n,
                                                             const unsigned long long *__restrict__ mask,
                                                             const uint32_t *__restrict__ off,
                                                             const uint64_t *__restrict__ mov, long nmov,
                                                             uint32_t *__restrict__ order_out,
                                                             uint32_t *__restrict__ skey_out)
{
    const long ntiles = (n + ORD_TILE - 1) / ORD_TILE;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();
        order_stayers_tile(order, skey, n, mask, off, mov, nmov, 0ull, order_out, skey_out, tile);
    }
}

// Pass 6: mover r (value c, high bits removed) to its new position.  Stayers
// before a mover = entries of the old order below it minus the movers among
// them.  The old order is sorted by the keys it was built with (stayers kept
// theirs), and ss holds its sub-cell starts: the entries below c are those of
// the sub-cells before c's and those of c's own sub-cell with a smaller id,
// found by a search confined to that sub-cell.
__device__ __forceinline__ void order_place_one(uint64_t c, long r, const uint32_t *__restrict__ order,
                                                const uint32_t *__restrict__ skey, long n,
                                                const uint32_t *__restrict__ ss, long nsub,
                                                const unsigned long long *__restrict__ mask,
                                                const uint32_t *__restrict__ off, uint32_t *__restrict__ order_out,
                                                uint32_t *__restrict__ skey_out)
{
    const uint32_t kc = (uint32_t)(c >> 32);
    long lo = kc < (uint64_t)nsub ? (long)ss[kc] : (long)ss[nsub];
    long hi = kc < (uint64_t)nsub ? (long)ss[kc + 1] : n;
    while (lo < hi) {
        const long mid = (lo + hi) >> 1;
        if (kid(skey[mid], order[mid]) < c) lo = mid + 1; else hi = mid;
    }
    const long stay_before = lo - (long)(lo < n ? movers_before(mask, off, lo) : off[(n + ORD_TILE - 1) / ORD_TILE]);
    const long pos = r + stay_before;
    order_out[pos] = (uint32_t)(c & 0xffffffffu);
    skey_out[pos] = (uint32_t)(c >> 32);
}

__global__ void k_order_place_movers(const uint64_t *__restrict__ mov, long nmov,
                                     const uint32_t *__restrict__ order, const uint32_t *__restrict__ skey, long n,
                                     const uint32_t *__restrict__ ss, long nsub,
                                     const unsigned long long *__restrict__ mask,
                                     const uint32_t *__restrict__ off, uint32_t *__restrict__ order_out,
                                     uint32_t *__restrict__ skey_out)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nmov) order_place_one(mov[r], r, order, skey, n, ss, nsub, mask, off, order_out, skey_out);
}

// ---------------------------------------------------------------------------
// Node records of the scatter passes.  Every sub-cell (cell (cx, cy),
// quadrant q = 2*(cx > .5) + (cy > .5)) feeds exactly one basic, one vx, one
// vy and one centre node, and every node is fed by four sub-cells, so each
// sub-cell sum has a fixed slot in its node's record:
//   basic  node (cx+q/2, cy+q%2), slot 3-q          (slot k: neighbour cells
//          (m11-1,m22-1), (m11-1,m22), (m11,m22-1), (m11,m22) for k = 0..3)
//   vx     q 2,3 -> node (cx+1, cy) slots 0,1;  q 0,1 -> node (cx, cy) slots 2,3
//   vy     q 1,3 -> node (cx, cy+1) slots 0,1;  q 0,2 -> node (cx, cy) slots 2,3
//   centre node (cx+1, cy+1), slot q
// Records are node-major (four blocks: basic NB, vx NV, vy NV, centre NC
// fields per slot), so a node thread reads one contiguous record per block.
// ---------------------------------------------------------------------------
template <int NB, int NV, int NC>
struct NodeRec {
    static constexpr int NF = NB + 2 * NV + NC;
    __device__ static long index(uint32_t key, int f, int ynumy, long nodenum)
    {
        const uint32_t cell = key >> 2, ncy = (uint32_t)(ynumy - 1);
        const int q = (int)(key & 3u);
        const long cx = (long)(cell / ncy), cy = (long)cell - cx * (long)ncy;
        if (f < NB) {
            const long node = (cx + (q >> 1)) * ynumy + cy + (q & 1);
            return node * 4 * NB + (3 - q) * NB + f;
        }
        if (f < NB + NV) {
            const long node = (cx + (q >> 1)) * ynumy + cy;
            const int slot = (q >> 1) ? (q & 1) : 2 + (q & 1);
            return nodenum * 4 * NB + node * 4 * NV + slot * NV + (f - NB);
        }
        if (f < NB + 2 * NV) {
            const long node = cx * ynumy + cy + (q & 1);
            const int slot = (q & 1) ? (q >> 1) : 2 + (q >> 1);
            return nodenum * 4 * (NB + NV) + node * 4 * NV + slot * NV + (f - NB - NV);
        }
        const long node = (cx + 1) * ynumy + cy + 1;
        return nodenum * 4 * (NB + 2 * NV) + node * 4 * NC + q * NC + (f - NB - 2 * NV);
    }
    __device__ static const double *basic(const double *P, long m3) { return P + m3 * 4 * NB; }
    __device__ static const double *vx(const double *P, long m3, long nodenum)
    {
        return P + nodenum * 4 * NB + m3 * 4 * NV;
    }
    __device__ static const double *vy(const double *P, long m3, long nodenum)
    {
        return P + nodenum * 4 * (NB + NV) + m3 * 4 * NV;
    }
    __device__ static const double *centre(const double *P, long m3, long nodenum)
    {
        return P + nodenum * 4 * (NB + 2 * NV) + m3 * 4 * NC;
    }
};

using MRec = NodeRec<M_NB, M_NV, M_NC>;
using NRec = NodeRec<N_NB, N_NV, N_NC>;

// ---------------------------------------------------------------------------
// Segmented sums of one batch of 64 sorted markers.  Lanes [0, nact) hold NF
// contributions in registers and their sub-cell key in skey[lane]; equal
// keys are contiguous.  The contributions pass through LDS CH fields at a
// time (con[lane][f]), so a wavefront holds 64 x CH doubles of LDS rather
// than 64 x NF and more wavefronts fit on a CU.  Per chunk, work items
// (run, field), run-major, are spread over the lanes: each sums its run
// serially in lane order (the result depends neither on timing nor on the
// chunking), and consecutive lanes read consecutive LDS words and write
// consecutive fields of a node record.  A sub-cell whose markers continue
// into the next batch (cont) is carried in LDS, double-buffered by batch
// parity, and finished there; every other run is complete and written once.
// ---------------------------------------------------------------------------
template <int NF, int CH = NF>
struct ScatterLds {
    static constexpr int CHUNK = CH;
    double con[WAVE][CH];
    double carry[2][NF];
    uint32_t skey[WAVE];
    int runs[WAVE + 1];
};

// Run heads of the batch into runs[0..nr], runs[nr] = nact; returns nr.  The
// caller synchronises before the runs are read.
__device__ __forceinline__ int scatter_runs(const uint32_t *skey, int *runs, int lane, int nact)
{
    const bool head = lane < nact && (lane == 0 || skey[lane - 1] != skey[lane]);
    const unsigned long long hm = __ballot(head);
    const int nr = __popcll(hm);
    if (head) runs[__popcll(hm & ((1ull << lane) - 1ull))] = lane;
    if (lane == 0) runs[nr] = nact;
    return nr;
}

// Fields [F0, F0 + W) of the batch, staged in con (row stride CH).
template <class R, int F0, int W, int CH>
__device__ __forceinline__ void scatter_fields(const double (*con)[CH], const uint32_t *skey, const int *runs,
                                               int nr, double (*carry)[R::NF], int par, bool have_carry,
                                               bool cont, int lane, const KArgs &a)
{
    for (int w = lane; w < nr * W; w += WAVE) {
        const int r = w / W, f = w - r * W;
        const int b = runs[r], e = runs[r + 1];
        double sum = (r == 0 && have_carry) ? carry[par][F0 + f] : 0.0;
        for (int j = b; j < e; j++) sum += con[j][f];
        if (r == nr - 1 && cont) carry[par ^ 1][F0 + f] = sum;
        else a.part[R::index(skey[b], F0 + f, a.g.ynumy, a.nodenum)] = sum;
    }
}

// One batch: the lane's key and NF contributions v in, sums out.  Starts
// and ends with the wavefront synchronised.
template <class R, int CH, int F0 = 0>
__device__ __forceinline__ void scatter_chunks(ScatterLds<R::NF, CH> &L, const double *v, int nr, int par,
                                               bool have_carry, bool cont, int lane, const KArgs &a)
{
    constexpr int W = R::NF - F0 < CH ? R::NF - F0 : CH;
#pragma unroll
    for (int f = 0; f < W; f++) L.con[lane][f] = v[F0 + f];
    __syncthreads();
    scatter_fields<R, F0, W, CH>(L.con, L.skey, L.runs, nr, L.carry, par, have_carry, cont, lane, a);
    __syncthreads();
    if constexpr (F0 + CH < R::NF) scatter_chunks<R, CH, F0 + CH>(L, v, nr, par, have_carry, cont, lane, a);
}

template <class R, int CH>
__device__ __forceinline__ void scatter_batch(ScatterLds<R::NF, CH> &L, const double *v, uint32_t key, int par,
                                              bool have_carry, bool cont, int lane, int nact, const KArgs &a)
{
    L.skey[lane] = key;
    __syncthreads();
    const int nr = scatter_runs(L.skey, L.runs, lane, nact);
    scatter_chunks<R, CH>(L, v, nr, par, have_carry, cont, lane, a);
}

// Zero the slots fed by the group's empty sub-cells (node passes read every
// slot whose source cell exists).
template <class R>
__device__ __forceinline__ void zero_empty(const KArgs &a, long c0, long c1, int lane)
{
    for (long k = c0 * 4 + lane; k < c1 * 4; k += WAVE)
        if (a.sub_start[k] == a.sub_start[k + 1])
            for (int f = 0; f < R::NF; f++) a.part[R::index((uint32_t)k, f, a.g.ynumy, a.nodenum)] = 0.0;
}

// Cells of a marker from its sub-cell key (the key is computed by the same
// search from the same clamped position), with their gridline records.
__device__ __forceinline__ void key_axes(const KArgs &a, uint32_t key, Axis4 &X, Axis4 &Y)
{
    const uint32_t cell = key >> 2, ncy = (uint32_t)(a.g.ynumy - 1);
    const uint32_t nx = cell / ncy, ny = cell - nx * ncy;
    X = a.g.cellx[nx];
    Y = a.g.celly[ny];
}

// ---------------------------------------------------------------------------
// marker2nodes, per-marker properties (interpolation.c:262-330) in physical
// marker order: pressure at the marker, density, expansivity, conductivity,
// viscosity (the FP64 quasi-Newton of ductilevisc) and the friction reset of
// calc_frictionpara.  Separate from the scatter so the transcendental-heavy
// part runs at full occupancy with coalesced access.
// ---------------------------------------------------------------------------
// Hand-over of the property pass to the scatter: the viscosity (ductilevisc)
// and the conductivity (it needs the pressure at the marker); the density
// only where it depends on that pressure (densimod != 0).  The expansivity
// depends on T and type alone and is recomputed by the scatter.
enum MProp { P_NU = 0, P_KT, P_RO, NPROP };

__device__ __forceinline__ void m2n_props_one(const KArgs &a, long m)
{
    const DevGrid &g = a.g;
    const double x = MK(a, MF_X, m), y = MK(a, MF_Y, m);
    if (!(x > 0 && y > 0 && x < g.xsize && y < g.ysize)) return;
    Axis4 X, Y;
    key_axes(a, a.key[m], X, Y);
    const int t = a.t[m];
    const double mtk = MK(a, MF_K, m);
    const Stencil ps = centre_stencil(g, X, Y, x, y);   // allinterpomp (interpolation.c:1923-1958)
    const double mpb = 1e-5 * bil(NOD(a, V2C_N_pr), ps.m3, g.ynumy, wts(ps));
    double mro, mbb;
    dencalc(a, mtk, mpb, t, mro, mbb);
    const double mkt = conductivity(a.mat, t, mtk, mpb);
    int iters;
    const double mnu = ductilevisc(a, t, MK(a, MF_XY, m), MK(a, MF_XX, m), mtk, mpb, &iters);
    const double phi0 = MK(a, MF_PHI, m);
    const double phi = friction_marker(a, m, t, x, y, phi0);
    if (__double_as_longlong(phi) != __double_as_longlong(phi0)) MK(a, MF_PHI, m) = phi;
    double *P = a.prop;
    (void)mbb;
    P[(long)P_NU * a.cap + m] = mnu;
    P[(long)P_KT * a.cap + m] = mkt;
    if (a.p.densimod != 0) P[(long)P_RO * a.cap + m] = mro;
}

__global__ __launch_bounds__(256) void k_m2n_props(KArgs a, int nwg)
{
    const long m = (long)xcd_remap(blockIdx.x, nwg) * 256 + threadIdx.x;
    if (m < a.n) m2n_props_one(a, m);
}

// ---------------------------------------------------------------------------
// marker2nodes, scatter pass (interpolation.c:215-629) over the sorted order:
// weights and contributions of every marker inside the grid, reduced per
// (cell, quadrant) and written as per-cell partials.
// ---------------------------------------------------------------------------
// Cells [c0, c1) of one wavefront group.
// Contributions staged through LDS 9 fields at a time (3 chunks of the 27):
// 5.6 KB of LDS per wavefront instead of 14.8 KB.
constexpr int M_CH = 9;
using M2nLds = ScatterLds<M_NF, M_CH>;

__device__ __forceinline__ void m2n_cells_group(const KArgs &a, long c0, long c1, int lane, M2nLds &L)
{
    const uint32_t beg = a.sub_start[c0 * 4], end = a.sub_start[c1 * 4];
    const DevMat *M = a.mat;
    const DevGrid &g = a.g;
    const double *P = a.prop;
    bool have_carry = false;
    int par = 0;
    // the next batch's marker ids and keys are loaded one batch ahead
    uint32_t m_next = beg + lane < end ? a.order[beg + lane] : 0u;
    uint32_t k_next = beg + lane < end ? a.skey[beg + lane] : 0xffffffffu;
    for (uint32_t base = beg; base < end; base += WAVE, par ^= 1) {
        const uint32_t i = base + lane;
        const bool active = i < end;
        const int nact = end - base < (uint32_t)WAVE ? (int)(end - base) : WAVE;
        const long m = m_next;
        uint32_t key = active ? k_next : 0xffffffffu;
        if (i + WAVE < end) { m_next = a.order[i + WAVE]; k_next = a.skey[i + WAVE]; }
        double v[M_NF];
#pragma unroll
        for (int f = 0; f < M_NF; f++) v[f] = 0.0;
        if (active) {
            const double x = MK(a, MF_X, m), y = MK(a, MF_Y, m);
            if (x > 0 && y > 0 && x < g.xsize && y < g.ysize) {
                Axis4 X, Y;
                key_axes(a, key, X, Y);
                const int t = a.t[m];
                const double mbm = a.p.bulk_modulus;
                const double mgg = M->gg[t];
                const double mtk = MK(a, MF_K, m);
                const double mnu = P[(long)P_NU * a.cap + m], mkt = P[(long)P_KT * a.cap + m];
                const double mro = a.p.densimod == 0 ? M->ro[t] : P[(long)P_RO * a.cap + m];
                // dencalcomp's expansivity (density_ductilevisc_calc.c:2-27)
                const double mbb = M->bb[t] / (1.0 - M->bb[t] * (mtk - 298.15));
                const double sxy = MK(a, MF_XY, m), sxx = MK(a, MF_XX, m);
                const double phi = MK(a, MF_PHI, m);
                const double me = MK(a, MF_E, m);
                // 1/dx/dy and the cell-local coordinates, exact quotients
                const double swt1 = div_rcp(X.inv, Y.g[2] - Y.g[1], Y.inv);
                const double cx = div_rcp(x - X.g[1], X.g[2] - X.g[1], X.inv);
                const double cy = div_rcp(y - Y.g[1], Y.g[2] - Y.g[1], Y.inv);
                double bw = 0.0;
                bool hasb = true;
                if (cx < 0.5 && cy < 0.5) bw = swt1 * (1.0 - 2.0 * cx) * (1.0 - 2.0 * cy);
                else if (cx < 0.5 && cy > 0.5) bw = swt1 * (1.0 - 2.0 * cx) * (2.0 * cy - 1.0);
                else if (cx > 0.5 && cy < 0.5) bw = swt1 * (2.0 * cx - 1.0) * (1.0 - 2.0 * cy);
                else if (cx > 0.5 && cy > 0.5) bw = swt1 * (2.0 * cx - 1.0) * (2.0 * cy - 1.0);
                else hasb = false;
                if (hasb) {
                    v[0] = mnu * bw; v[1] = mnu / mgg * bw; v[2] = me * bw;
                    v[3] = sxy * bw; v[4] = MK(a, MF_EXY, m) * bw; v[5] = MK(a, MF_C, m) * bw;
                    v[6] = phi * bw; v[7] = mro * bw; v[8] = mbb * bw; v[9] = M->ht[t] * bw;
                    v[10] = M->cp[t] * bw; v[11] = mkt * bw; v[12] = mtk * bw; v[13] = bw;
                }
                double vw = 0.0;
                bool hasv = true;
                if (cx < 0.5) vw = swt1 * (1.0 - cx) * (1.0 - ref_abs(cy - 0.5));
                else if (cx > 0.5) vw = swt1 * cx * (1.0 - ref_abs(cy - 0.5));
                else hasv = false;
                if (hasv) {
                    v[M_VX] = MK(a, MF_VX, m) * mro * vw; v[M_VX + 1] = mro * vw; v[M_VX + 2] = vw;
                }
                hasv = true;
                if (cy < 0.5) vw = swt1 * (1.0 - ref_abs(cx - 0.5)) * (1.0 - cy);
                else if (cy > 0.5) vw = swt1 * (1.0 - ref_abs(cx - 0.5)) * cy;
                else hasv = false;
                if (hasv) {
                    v[M_VY] = MK(a, MF_VY, m) * mro * vw; v[M_VY + 1] = mro * vw; v[M_VY + 2] = vw;
                }
                const double cw = swt1 * (1.0 - 2.0 * ref_abs(cx - 0.5)) * (1.0 - 2.0 * ref_abs(cy - 0.5));
                v[M_C + 0] = mbm * cw; v[M_C + 1] = mnu * cw; v[M_C + 2] = mnu / mgg * cw;
                v[M_C + 3] = sxx * cw; v[M_C + 4] = MK(a, MF_EXX, m) * cw; v[M_C + 5] = MK(a, MF_P, m) * cw;
                v[M_C + 6] = cw;
            }
        }
        const bool cont = base + WAVE < end && a.skey[base + WAVE] == a.skey[base + WAVE - 1];
        scatter_batch<MRec, M_CH>(L, v, key, par, have_carry, cont, lane, nact, a);
        have_carry = cont;
    }
    zero_empty<MRec>(a, c0, c1, lane);
}

__global__ __launch_bounds__(WAVE) void k_m2n_cells(KArgs a, int nwg)
{
    __shared__ M2nLds L;
    const long c0 = (long)xcd_remap(blockIdx.x, nwg) * CPW;
    m2n_cells_group(a, c0, c0 + CPW < a.ncell ? c0 + CPW : a.ncell, threadIdx.x, L);
}

// ---------------------------------------------------------------------------
// marker2nodes, node pass: combine cell partials and normalise
// (interpolation.c:58-153 zeroing, :714-824 normalisation, :827-859 BC copy
// runs as k_phis_bc)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void m2n_nodes_one(const KArgs &a, long m3)
{
    const int ynumy = a.g.ynumy, xnumx = a.g.xnumx;
    const int m22 = (int)(m3 % ynumy), m11 = (int)(m3 / ynumy);
    const double *P = a.part;
    // The node's record slots, summed in slot order; a slot counts only if
    // its source cell exists.
    const bool c00 = m11 >= 1 && m22 >= 1, c01 = m11 >= 1 && m22 < ynumy - 1;
    const bool c10 = m11 < xnumx - 1 && m22 >= 1, c11 = m11 < xnumx - 1 && m22 < ynumy - 1;
    double B[M_NB];
#pragma unroll
    for (int f = 0; f < M_NB; f++) B[f] = 0.0;
    {
        const double *R = MRec::basic(P, m3);
        const bool has[4] = {c00, c01, c10, c11};
        for (int k = 0; k < 4; k++) {
            if (!has[k]) continue;
#pragma unroll
            for (int f = 0; f < M_NB; f++) B[f] += R[k * M_NB + f];
        }
    }
    double VXs[M_NV] = {0, 0, 0}, VYs[M_NV] = {0, 0, 0}, Cs[M_NC];
#pragma unroll
    for (int f = 0; f < M_NC; f++) Cs[f] = 0.0;
    {
        const double *R = MRec::vx(P, m3, a.nodenum);
        for (int f = 0; f < M_NV; f++) {
            if (c01) VXs[f] += R[0 * M_NV + f] + R[1 * M_NV + f];
            if (c11) VXs[f] += R[2 * M_NV + f] + R[3 * M_NV + f];
        }
        R = MRec::vy(P, m3, a.nodenum);
        for (int f = 0; f < M_NV; f++) {
            if (c10) VYs[f] += R[0 * M_NV + f] + R[1 * M_NV + f];
            if (c11) VYs[f] += R[2 * M_NV + f] + R[3 * M_NV + f];
        }
        R = MRec::centre(P, m3, a.nodenum);
        if (c00)
            for (int f = 0; f < M_NC; f++)
                Cs[f] = R[0 * M_NC + f] + R[1 * M_NC + f] + R[2 * M_NC + f] + R[3 * M_NC + f];
    }

    // Every field the reference's initialisation loop zeroes
    // (interpolation.c:58-153) is written exactly once below, with its value
    // or with the zero that loop left; the others only where the reference
    // writes them.  The antidiffusion inputs are read before any output.
    const bool anti = a.timesum > 0 && a.p.antidiffinter == 1;
    const bool inA = B[13] != 0 && m11 >= 1 && m22 >= 1 && m11 < xnumx - 1 && m22 < ynumy - 1;
    const bool inC = Cs[6] != 0 && m11 >= 1 && m22 >= 1;
    double in_phis_it = 0, in_phie = 0, in_sxy = 0, in_sxyee = 0, in_exy = 0, in_exyee = 0,
           in_pes = 0, in_pee = 0, in_sxx = 0, in_sxxee = 0, in_pr = 0, in_sppee = 0,
           in_exx = 0, in_exxee = 0, in_vx = 0, in_mvxe = 0, in_vy = 0, in_mvye = 0;
    if (anti) {
        if (inA) {
            in_phis_it = NOD(a, V2C_N_phis_it)[m3]; in_phie = NOD(a, V2C_N_phie)[m3];
            in_sxy = NOD(a, V2C_N_sxy)[m3]; in_sxyee = NOD(a, V2C_N_sxyee)[m3];
            in_exy = NOD(a, V2C_N_exy)[m3]; in_exyee = NOD(a, V2C_N_exyee)[m3];
            in_pes = NOD(a, V2C_N_pes)[m3]; in_pee = NOD(a, V2C_N_pee)[m3];
        }
        if (inC) {
            in_sxx = NOD(a, V2C_N_sxx)[m3]; in_sxxee = NOD(a, V2C_N_sxxee)[m3];
            in_pr = NOD(a, V2C_N_pr)[m3]; in_sppee = NOD(a, V2C_N_sppee)[m3];
            in_exx = NOD(a, V2C_N_exx)[m3]; in_exxee = NOD(a, V2C_N_exxee)[m3];
        }
        if (VXs[2] != 0) { in_vx = NOD(a, V2C_N_vx)[m3]; in_mvxe = NOD(a, V2C_N_mvxe)[m3]; }
        if (VYs[2] != 0) { in_vy = NOD(a, V2C_N_vy)[m3]; in_mvye = NOD(a, V2C_N_mvye)[m3]; }
    }
    const double w = B[13];
    // basic nodes (interpolation.c:714-767)
    double phin = 0, sxyen = 0, exyen = 0, phis = 0, sxye = 0, exye = 0, pe = 0, coh = 0, gg = 0;
    double mus = 0, as_ = 0, bs = 0, Ls = 0, V0s = 0;
    if (inA) {
        phin = B[6] / w; sxyen = B[3] / w; exyen = B[4] / w;
        const double pen = B[2] / w;
        NOD(a, V2C_N_pen)[m3] = pen;
        if (anti) {
            phis = in_phis_it + (phin - in_phie);
            sxye = in_sxy + (sxyen - in_sxyee);
            exye = in_exy + (exyen - in_exyee);
            pe = in_pes + (pen - in_pee);
        } else {
            phis = phin; sxye = sxyen; exye = exyen; pe = pen;
        }
        if (pe < 0) pe = 0;
        friction_node_values(a, a.g.gx[m11], a.g.gy[m22], mus, as_, bs, Ls, V0s);
        coh = B[5] / w;
        const double nu_m2c = B[0] / w;
        NOD(a, V2C_N_nu_m2c)[m3] = nu_m2c;
        gg = nu_m2c / (B[1] / w);
    }
    NOD(a, V2C_N_phin)[m3] = phin; NOD(a, V2C_N_sxyen)[m3] = sxyen; NOD(a, V2C_N_exyen)[m3] = exyen;
    NOD(a, V2C_N_phis)[m3] = phis; NOD(a, V2C_N_sxye)[m3] = sxye; NOD(a, V2C_N_exye)[m3] = exye;
    NOD(a, V2C_N_pe)[m3] = pe; NOD(a, V2C_N_cohesion)[m3] = coh; NOD(a, V2C_N_gg)[m3] = gg;
    NOD(a, V2C_N_mus_s)[m3] = mus; NOD(a, V2C_N_as)[m3] = as_; NOD(a, V2C_N_bs)[m3] = bs;
    NOD(a, V2C_N_Ls)[m3] = Ls; NOD(a, V2C_N_V0s)[m3] = V0s;
    // pressure nodes (interpolation.c:769-801)
    double sxxen = 0, sppen = 0, exxen = 0, sxxe = 0, sppe = 0, exxe = 0, gd = 0, bm = 0, dro = 0, drp = 0;
    if (inC) {
        const double wc = Cs[6];
        sxxen = Cs[3] / wc; sppen = Cs[5] / wc; exxen = Cs[4] / wc;
        if (anti) {
            sxxe = in_sxx + (sxxen - in_sxxee);
            sppe = in_pr + (sppen - in_sppee);
            exxe = in_exx + (exxen - in_exxee);
        } else {
            sxxe = sxxen; sppe = sppen; exxe = exxen;
        }
        if (a.timesum == 0 && a.p.reloadfromDP == 0 && a.p.compressiblemode == 1) sppe = a.p.nstress;
        const double nd = Cs[1] / wc;
        NOD(a, V2C_N_nd_m2c)[m3] = nd;
        gd = nd / (Cs[2] / wc);
        bm = Cs[0] / wc;
        dro = 0 / wc;
        drp = 0 / wc;
    }
    NOD(a, V2C_N_sxxen)[m3] = sxxen; NOD(a, V2C_N_sppen)[m3] = sppen; NOD(a, V2C_N_exxen)[m3] = exxen;
    NOD(a, V2C_N_sxxe)[m3] = sxxe; NOD(a, V2C_N_sppe)[m3] = sppe; NOD(a, V2C_N_exxe)[m3] = exxe;
    NOD(a, V2C_N_gd)[m3] = gd; NOD(a, V2C_N_bm)[m3] = bm; NOD(a, V2C_N_dro)[m3] = dro;
    NOD(a, V2C_N_drp)[m3] = drp;
    // velocity nodes (interpolation.c:803-824)
    double mvx = 0, mrx = 0, mvy = 0, mry = 0;
    if (VXs[2] != 0) {
        const double mvxn = VXs[0] / VXs[1];
        NOD(a, V2C_N_mvxn)[m3] = mvxn;
        mvx = anti ? in_vx + (mvxn - in_mvxe) : mvxn;
        mrx = VXs[1] / VXs[2];
    }
    if (VYs[2] != 0) {
        const double mvyn = VYs[0] / VYs[1];
        NOD(a, V2C_N_mvyn)[m3] = mvyn;
        mvy = anti ? in_vy + (mvyn - in_mvye) : mvyn;
        mry = VYs[1] / VYs[2];
    }
    NOD(a, V2C_N_mvx)[m3] = mvx; NOD(a, V2C_N_mrx)[m3] = mrx;
    NOD(a, V2C_N_mvy)[m3] = mvy; NOD(a, V2C_N_mry)[m3] = mry;
    // thermal and density fields of the basic node
    double ro = 0, et = 0, ht = 0, kt = 0, cp = 0, tk = 0;
    if (w != 0) {
        ro = B[7] / w; et = B[8] / w; ht = B[9] / w; kt = B[11] / w; cp = B[10] / w; tk = B[12] / w;
    }
    NOD(a, V2C_N_ro)[m3] = ro; NOD(a, V2C_N_et)[m3] = et; NOD(a, V2C_N_ht)[m3] = ht;
    NOD(a, V2C_N_kt)[m3] = kt; NOD(a, V2C_N_cp)[m3] = cp; NOD(a, V2C_N_tk)[m3] = tk;
    // fields zeroed and not set again here: nu, nd, and nodes2markers' sums
    static constexpr int zero_only[] = {
        V2C_N_nu, V2C_N_nd, V2C_N_wtphisum, V2C_N_phisum, V2C_N_sxysum, V2C_N_exysum, V2C_N_esum,
        V2C_N_wtsxxsum, V2C_N_psum, V2C_N_sxxsum, V2C_N_exxsum, V2C_N_wtvxsum, V2C_N_vxsum,
        V2C_N_wtvysum, V2C_N_vysum, V2C_N_mrx0, V2C_N_mry0};
    for (int z = 0; z < (int)(sizeof(zero_only) / sizeof(zero_only[0])); z++) NOD(a, zero_only[z])[m3] = 0.0;
}

__global__ __launch_bounds__(256) void k_m2n_nodes(KArgs a)
{
    const long m3 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m3 < a.nodenum) m2n_nodes_one(a, m3);
}

// State-variable boundary copy when its diffusion is on (interpolation.c:827-859).
// Reads interior values only and writes boundary nodes only, so one pass is race free.
__device__ __forceinline__ void phis_bc_one(const KArgs &a, long m3)
{
    const int ynumy = a.g.ynumy, xnumx = a.g.xnumx;
    const int m22 = (int)(m3 % ynumy), m11 = (int)(m3 / ynumy);
    double *phis = NOD(a, V2C_N_phis);
    if (m11 == 0 && m22 >= 1 && m22 <= ynumy - 2) phis[m3] = phis[m3 + ynumy];
    else if (m11 == xnumx - 1 && m22 >= 1 && m22 <= ynumy - 2) phis[m3] = phis[m3 - ynumy];
    else if (m22 == 0 && m11 >= 1 && m11 <= xnumx - 2) phis[m3] = phis[m3 + 1];
    else if (m22 == ynumy - 1 && m11 > 0 && m11 <= xnumx - 2) phis[m3] = phis[m3 - 1];
    else if ((m11 == 0 && m22 == 0) || (m11 == 0 && m22 == ynumy - 1) ||
             (m11 == xnumx - 1 && m22 == 0) || (m11 == xnumx - 1 && m22 == ynumy - 1))
        phis[m3] = 0;
}

__global__ __launch_bounds__(256) void k_phis_bc(KArgs a)
{
    const long m3 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m3 < a.nodenum) phis_bc_one(a, m3);
}

// ---------------------------------------------------------------------------
// nodes2markers (interpolation.c:929-1048) and moveandrotatemarkers
// (moveandrotatemarkers.c:1-138)
// ---------------------------------------------------------------------------
struct Mech {
    double EXY, EXYE, SXY, SXYE, PHIS, PHISE, SXX, SXXE, EXX, EXXE, PR, SPPE, VX, MVX, VY, MVY, ES, ESE;
};

// interpolate_mechanical gathers (interpolation.c:1060-1092, 1211-1237,
// 1376-1397, 1433-1454) at a clamped position in cell (X.s, Y.s).
// WITH_E = false gathers only the nine fields the antidiffusion branch of the
// update reads (the *E fields are left unset).
template <bool WITH_E>
__device__ __forceinline__ void mech_gather(const KArgs &a, double x, double y, const Axis4 &X, const Axis4 &Y,
                                            Mech &o)
{
    const DevGrid &g = a.g;
    const int ynumy = g.ynumy;
    Stencil S = basic_stencil(g, X, Y, x, y);
    W4 w = wts(S);
    uint32_t m3 = S.m3;
    o.EXY = bil(NOD(a, V2C_N_exy), m3, ynumy, w);
    o.SXY = bil(NOD(a, V2C_N_sxy), m3, ynumy, w);
    o.ES = bil(NOD(a, V2C_N_pes), m3, ynumy, w);
    o.PHIS = bil(NOD(a, V2C_N_phis_it), m3, ynumy, w);
    if (WITH_E) {
        o.EXYE = bil(NOD(a, V2C_N_exye), m3, ynumy, w);
        o.SXYE = bil(NOD(a, V2C_N_sxye), m3, ynumy, w);
        o.ESE = bil(NOD(a, V2C_N_pe), m3, ynumy, w);
        o.PHISE = bil(NOD(a, V2C_N_phis), m3, ynumy, w);
    }
    S = centre_stencil(g, X, Y, x, y);
    w = wts(S);
    m3 = S.m3;
    o.EXX = bil(NOD(a, V2C_N_exx), m3, ynumy, w);
    o.SXX = bil(NOD(a, V2C_N_sxx), m3, ynumy, w);
    o.PR = bil(NOD(a, V2C_N_pr), m3, ynumy, w);
    if (WITH_E) {
        o.EXXE = bil(NOD(a, V2C_N_exxe), m3, ynumy, w);
        o.SXXE = bil(NOD(a, V2C_N_sxxe), m3, ynumy, w);
        o.SPPE = bil(NOD(a, V2C_N_sppe), m3, ynumy, w);
    }
    S = vx_stencil(g, X, Y, x, y);
    w = wts(S);
    o.VX = bil(NOD(a, V2C_N_vx), S.m3, ynumy, w);
    if (WITH_E) o.MVX = bil(NOD(a, V2C_N_mvx), S.m3, ynumy, w);
    S = vy_stencil(g, X, Y, x, y);
    w = wts(S);
    o.VY = bil(NOD(a, V2C_N_vy), S.m3, ynumy, w);
    if (WITH_E) o.MVY = bil(NOD(a, V2C_N_mvy), S.m3, ynumy, w);
}

// The nine interpolants of mech_gather<false> at the clamped point (x, y) of
// a filled RK cache (the same point, so the same stencils): the basic- and
// pressure-node fields are gathered, vx and vy come from the cached nodes the
// vx- and vy-node stencils read; weights and sums as mech_gather's.
__device__ __forceinline__ void mech_from_cache(const KArgs &a, const RkCache &c, double x, double y, Mech &o)
{
    const int ynumy = a.g.ynumy;
    double ex, ny_, exc, nyc;
    rk_weights(c, x, y, ex, ny_, exc, nyc);
    Stencil S;
    S.e = ex; S.n = ny_;
    W4 w = wts(S);
    uint32_t m3 = (uint32_t)(c.sx * ynumy + c.sy);
    o.EXY = bil(NOD(a, V2C_N_exy), m3, ynumy, w);
    o.SXY = bil(NOD(a, V2C_N_sxy), m3, ynumy, w);
    o.ES = bil(NOD(a, V2C_N_pes), m3, ynumy, w);
    o.PHIS = bil(NOD(a, V2C_N_phis_it), m3, ynumy, w);
    S.e = exc; S.n = nyc;
    w = wts(S);
    m3 = (uint32_t)((c.sx + (c.xup ? 1 : 0)) * ynumy + c.sy + (c.yup ? 1 : 0));
    o.EXX = bil(NOD(a, V2C_N_exx), m3, ynumy, w);
    o.SXX = bil(NOD(a, V2C_N_sxx), m3, ynumy, w);
    o.PR = bil(NOD(a, V2C_N_pr), m3, ynumy, w);
    S.e = ex; S.n = nyc;
    w = wts(S);
    o.VX = w.d0 * c.F[0] + w.d1 * c.F[1] + w.d2 * c.F[2] + w.d3 * c.F[3];
    S.e = exc; S.n = ny_;
    w = wts(S);
    o.VY = w.d0 * c.H[0] + w.d1 * c.H[1] + w.d2 * c.H[2] + w.d3 * c.H[3];
}

// Sub-cell key of a clamped point in the cached cell: subcell_key's cell and
// quadrant without the search.
__device__ __forceinline__ uint32_t key_from_cache(const KArgs &a, const RkCache &c, double x, double y)
{
    const double cx = div_rcp(x - c.X.lo, c.X.w, c.X.inv);
    const double cy = div_rcp(y - c.Y.lo, c.Y.w, c.Y.inv);
    const uint32_t q = (cx > 0.5 ? 2u : 0u) | (cy > 0.5 ? 1u : 0u);
    return (((uint32_t)c.sx * (uint32_t)(a.g.ynumy - 1) + (uint32_t)c.sy) << 2) | q;
}

// Antidiffusion scatter-back contributions of one marker (interpolation.c:
// 1159-1209, 1360-1374, 1405-1489): the interpolated PHIS, SXY, EXY, ES, PR,
// SXX, EXX, VX, VY times the node weights, at the clamped position in cell
// (X.s, Y.s).  Only these nine interpolants are gathered.
__device__ __forceinline__ void anti_contrib(const KArgs &a, double x, double y, const Axis4 &X, const Axis4 &Y,
                                             double mro, double *v /* N_NF, zeroed */)
{
    const DevGrid &g = a.g;
    const int ynumy = g.ynumy;
    const Stencil4 S = stencils4(g, X, Y, x, y);
    W4 w = wts(S.b);
    const double EXY = bil(NOD(a, V2C_N_exy), S.b.m3, ynumy, w);
    const double SXY = bil(NOD(a, V2C_N_sxy), S.b.m3, ynumy, w);
    const double ES = bil(NOD(a, V2C_N_pes), S.b.m3, ynumy, w);
    const double PHIS = bil(NOD(a, V2C_N_phis_it), S.b.m3, ynumy, w);
    w = wts(S.c);
    const double EXX = bil(NOD(a, V2C_N_exx), S.c.m3, ynumy, w);
    const double SXX = bil(NOD(a, V2C_N_sxx), S.c.m3, ynumy, w);
    const double PR = bil(NOD(a, V2C_N_pr), S.c.m3, ynumy, w);
    const double VX = bil(NOD(a, V2C_N_vx), S.vx.m3, ynumy, wts(S.vx));
    const double VY = bil(NOD(a, V2C_N_vy), S.vy.m3, ynumy, wts(S.vy));
    // cell-local coordinates (interpolation.c:1140-1150) are the vx x- and
    // vy y-weights: the same formula on the same cell
    const double swt1 = div_rcp(X.inv, Y.g[2] - Y.g[1], Y.inv);
    const double cx = S.vx.e, cy = S.vy.n;
    double bw = 0.0;
    bool hasb = true;
    if (cx < 0.5 && cy < 0.5) bw = swt1 * (1.0 - 2 * cx) * (1.0 - 2 * cy);
    else if (cx < 0.5 && cy > 0.5) bw = swt1 * (1.0 - 2.0 * cx) * (2.0 * cy - 1.0);
    else if (cx > 0.5 && cy < 0.5) bw = swt1 * (2.0 * cx - 1.0) * (1.0 - 2.0 * cy);
    else if (cx > 0.5 && cy > 0.5) bw = swt1 * (2.0 * cx - 1.0) * (2.0 * cy - 1.0);
    else hasb = false;
    if (hasb) { v[0] = PHIS * bw; v[1] = SXY * bw; v[2] = EXY * bw; v[3] = ES * bw; v[4] = bw; }
    const double cw = swt1 * (1.0 - 2 * ref_abs(cx - 0.5)) * (1.0 - 2 * ref_abs(cy - 0.5));
    v[N_C] = PR * cw; v[N_C + 1] = SXX * cw; v[N_C + 2] = EXX * cw; v[N_C + 3] = cw;
    double vw = 0.0;
    bool hasv = true;
    if (cx < 0.5) vw = swt1 * (1.0 - cx) * (1.0 - ref_abs(cy - 0.5));
    else if (cx > 0.5) vw = swt1 * cx * (1.0 - ref_abs(cy - 0.5));
    else hasv = false;
    if (hasv) { v[N_VX] = VX * mro * vw; v[N_VX + 1] = mro * vw; v[N_VX + 2] = vw; }
    hasv = true;
    if (cy < 0.5) vw = swt1 * (1.0 - ref_abs(cx - 0.5)) * (1.0 - cy);
    else if (cy > 0.5) vw = swt1 * (1.0 - ref_abs(cx - 0.5)) * cy;
    else hasv = false;
    if (hasv) { v[N_VY] = VY * mro * vw; v[N_VY + 1] = mro * vw; v[N_VY + 2] = vw; }
}

// Marker state of one marker, loaded once and written once.
struct MState {
    double x, y, k, p, xx, xy, exx, exy, e, phi, vx, vy;
    bool dirty_state, dirty_k;
};

// Marker update of nodes2markers (interpolation.c:943-1011).
__device__ __forceinline__ void n2m_update(const KArgs &a, int t, MState &s, const Mech &v)
{
    const DevGrid &g = a.g;
    const bool anti = a.p.antidiffinter == 1;
    if (s.x > 0 && s.y > 0 && s.x < g.xsize && s.y < g.ysize && t < 50 && s.k <= 0) {
        if (anti) {
            s.xx = v.SXX; s.xy = v.SXY; s.p = v.PR; s.exx = v.EXX; s.exy = v.EXY;
            s.vx = v.VX; s.vy = v.VY; s.phi = v.PHIS; s.e = v.ES;
        } else {
            s.xx = v.SXXE; s.xy = v.SXYE; s.p = v.SPPE; s.exx = v.EXXE; s.exy = v.EXYE;
            s.vx = v.MVX; s.vy = v.MVY; s.phi = v.PHISE; s.e = v.ESE;
        }
        s.dirty_state = true;
        if (!a.p.tempmod) { s.k = a.p.incoming_T; s.dirty_k = true; }
    }
    if (s.x >= 0 && s.y >= 0 && s.x <= g.xsize && s.y <= g.ysize && t < 50 && s.k > 0) {
        if (anti) {
            s.phi = v.PHIS; s.exx = v.EXX; s.exy = v.EXY; s.xx = v.SXX; s.xy = v.SXY;
            s.p = v.PR; s.vx = v.VX; s.vy = v.VY; s.e = v.ES;
        } else {
            s.phi += (v.PHIS - v.PHISE); s.exx += (v.EXX - v.EXXE); s.exy += (v.EXY - v.EXYE);
            s.xx += (v.SXX - v.SXXE); s.xy += (v.SXY - v.SXYE); s.p += (v.PR - v.SPPE);
            s.vx += (v.VX - v.MVX); s.vy += (v.VY - v.MVY); s.e += (v.ES - v.ESE);
        }
        s.dirty_state = true;
    }
}

// The end of moveandrotatemarkers for one marker (moveandrotatemarkers.c:
// 45-87) given its velocity and spin (vx0, vy0, sp0): orthogonal-only motion
// outside (outgrid == 2), the move, the stress rotation, the deep-temperature
// and out-of-grid resets.
__device__ __forceinline__ void advect_apply(const KArgs &a, MState &s, double vx0, double vy0, double sp0,
                                             bool &reset)
{
    const DevGrid &g = a.g;
    const v2c_params &p = a.p;
    const double dt = a.timestep;
    if (p.outgrid == 2) {
        if (s.x < 0 || s.x > g.xsize) vy0 = 0;
        if (s.y < 0 || s.y > g.ysize) vx0 = 0;
    }
    s.x += dt * vx0;
    s.y += dt * vy0;
    const double th = sp0 * dt;
    const double sxx = s.xx, sxy = s.xy;
    if (p.stoksmod == 1) {
        // sin 2th and cos 2th from the one sincos (double-angle identities,
        // a few ulp from sin(2 th), cos(2 th)); stresses carry the 1e-12 bar,
        // positions do not depend on them
        double c, sn;
        sincos(th, &sn, &c);
        const double s2 = 2.0 * sn * c, c2 = fma(-2.0 * sn, sn, 1.0);
        s.xx = sxx * c * c - sxx * sn * sn + sxy * s2;
        s.xy = 0.5 * (-sxx - sxx) * s2 + sxy * c2;
    } else if (p.stoksmod == 2) {
        s.xx = sxx + sxy * 2.0 * th;
        s.xy = sxy + 0.5 * (-sxx - sxx) * 2.0 * th;
    }
    if (s.y > p.zdeep && vy0 < 0 && s.k < p.tdeep) { s.k = p.tdeep; s.dirty_k = true; }
    if (s.x < 0 || s.y < 0 || s.x > g.xsize || s.y > g.ysize) {
        s.k = 0; s.dirty_k = true; s.e = 0; reset = true;
    }
}

// Does the marker move at all (moveandrotatemarkers.c:12-16)?
__device__ __forceinline__ bool advect_moves(const KArgs &a, int t, const MState &s)
{
    const DevGrid &g = a.g;
    const bool inside = s.x >= 0 && s.y >= 0 && s.x <= g.xsize && s.y <= g.ysize;
    return (inside || a.p.outgrid != 1) && !a.mat->im[t];
}

// The same with the type's immobility flag already loaded.
__device__ __forceinline__ bool advect_moves_flag(const KArgs &a, bool immobile, const MState &s)
{
    const DevGrid &g = a.g;
    const bool inside = s.x >= 0 && s.y >= 0 && s.x <= g.xsize && s.y <= g.ysize;
    return (inside || a.p.outgrid != 1) && !immobile;
}

// (k1 + 2 k2 + 2 k3 + k4) / 6, summed stage by stage: the same left-to-right
// additions as the reference, one accumulator live instead of every stage's
// rates; the strain-rate invariant is not used.  The spin is the RK average
// (markmod == 2) or the first stage's.  interp(x, y, vx, vy, esp) evaluates
// allinteri; it returns false to abandon the marker.
template <class F>
__device__ __forceinline__ bool rk4_velocity(const KArgs &a, const MState &s, double &vx0, double &vy0,
                                             double &sp0, F &&interp)
{
    const double dt = a.timestep;
    double vx, vy, sp;
    if (!interp(s.x, s.y, vx, vy, sp)) return false;
    const double sp1 = sp;
    double ax = vx, ay = vy, as = sp;
    if (!interp(s.x + vx * dt / 2.0, s.y + vy * dt / 2.0, vx, vy, sp)) return false;
    ax = ax + 2.0 * vx; ay = ay + 2.0 * vy; as = as + 2.0 * sp;
    if (!interp(s.x + vx * dt / 2.0, s.y + vy * dt / 2.0, vx, vy, sp)) return false;
    ax = ax + 2.0 * vx; ay = ay + 2.0 * vy; as = as + 2.0 * sp;
    if (!interp(s.x + vx * dt, s.y + vy * dt, vx, vy, sp)) return false;
    ax = ax + vx; ay = ay + vy; as = as + sp;
    vx0 = div_rcp(ax, 6.0, 1.0 / 6.0);
    vy0 = div_rcp(ay, 6.0, 1.0 / 6.0);
    sp0 = a.p.markmod == 2 ? div_rcp(as, 6.0, 1.0 / 6.0) : sp1;
    return true;
}

// moveandrotatemarkers body for one marker of type < 100 (moveandrotatemarkers.c:12-88),
// general form (any cell, any stage point).  Returns false (and changes
// nothing) for types >= 100.  (hx, hy): the cells of the marker's clamped
// position, or null to search.
__device__ __forceinline__ bool advect_one(const KArgs &a, int t, MState &s, bool &moved, bool &reset,
                                           const Axis4 *hx = nullptr, const Axis4 *hy = nullptr)
{
    moved = false;
    reset = false;
    if (!advect_moves(a, t, s)) return true;
    if (t >= 100) return false;
    double vx0, vy0, sp0;
    if (a.p.markmod == 1) {
        double ee;
        allinteri(a, s.x, s.y, vx0, vy0, sp0, ee, hx, hy);
    } else {
        int st = 0;
        rk4_velocity(a, s, vx0, vy0, sp0, [&](double x, double y, double &vx, double &vy, double &sp) {
            double ee;
            if (st++ == 0) allinteri(a, x, y, vx, vy, sp, ee, hx, hy);
            else allinteri(a, x, y, vx, vy, sp, ee);
            return true;
        });
    }
    advect_apply(a, s, vx0, vy0, sp0, reset);
    moved = true;
    return true;
}

// nodes2markers antidiffusion scatter (interpolation.c:1159-1209, 1360-1374,
// 1405-1489) over the sorted order, before any marker moves: contributions of
// every marker at its clamped position, reduced per (cell, quadrant) into
// per-cell partials.
// Cells [c0, c1) of one wavefront group.
using N2mLds = ScatterLds<N_NF>;

__device__ __forceinline__ void n2m_scatter_group(const KArgs &a, long c0, long c1, int lane, N2mLds &L)
{
    const DevGrid &g = a.g;
    const uint32_t beg = a.sub_start[c0 * 4], end = a.sub_start[c1 * 4];
    bool have_carry = false;
    int par = 0;
    uint32_t m_next = beg + lane < end ? a.order[beg + lane] : 0u;
    uint32_t k_next = beg + lane < end ? a.skey[beg + lane] : 0xffffffffu;
    for (uint32_t base = beg; base < end; base += WAVE, par ^= 1) {
        const uint32_t i = base + lane;
        const int nact = end - base < (uint32_t)WAVE ? (int)(end - base) : WAVE;
        const long m = m_next;
        uint32_t key = i < end ? k_next : 0xffffffffu;
        if (i + WAVE < end) { m_next = a.order[i + WAVE]; k_next = a.skey[i + WAVE]; }
        double v[N_NF];
#pragma unroll
        for (int f = 0; f < N_NF; f++) v[f] = 0.0;
        if (i < end) {
            Axis4 X, Y;
            key_axes(a, key, X, Y);
            const double x = clampd(MK(a, MF_X, m), 0, g.xsize), y = clampd(MK(a, MF_Y, m), 0, g.ysize);
            anti_contrib(a, x, y, X, Y, a.mat->ro[a.t[m]], v);
        }
        const bool cont = base + WAVE < end && a.skey[base + WAVE] == a.skey[base + WAVE - 1];
        scatter_batch<NRec, N_NF>(L, v, key, par, have_carry, cont, lane, nact, a);
        have_carry = cont;
    }
    zero_empty<NRec>(a, c0, c1, lane);
}


// nodes2markers marker update (interpolation.c:943-1011), fused with
// moveandrotatemarkers (moveandrotatemarkers.c:1-138) when ADVECT, in physical
// marker order: every marker reads only its own state and nodal fields.
// Returns true when the marker's type is not supported by the advection.
template <bool ADVECT>
__device__ __forceinline__ bool n2m_update_one(const KArgs &a, long m)
{
    bool unsupported = false;
    const DevGrid &g = a.g;
    const bool anti = a.p.antidiffinter == 1;
    const int t = a.t[m];
    Axis4 X, Y;
    key_axes(a, a.key[m], X, Y);
    MState s;
    s.x = MK(a, MF_X, m); s.y = MK(a, MF_Y, m); s.k = MK(a, MF_K, m);
    s.dirty_state = false; s.dirty_k = false;
    if (!anti) {
        s.xx = MK(a, MF_XX, m); s.xy = MK(a, MF_XY, m); s.p = MK(a, MF_P, m);
        s.exx = MK(a, MF_EXX, m); s.exy = MK(a, MF_EXY, m); s.vx = MK(a, MF_VX, m);
        s.vy = MK(a, MF_VY, m); s.phi = MK(a, MF_PHI, m); s.e = MK(a, MF_E, m);
    }
    {
        Mech mv;
        const double cxp = clampd(s.x, 0, g.xsize), cyp = clampd(s.y, 0, g.ysize);
        if (anti) mech_gather<false>(a, cxp, cyp, X, Y, mv);
        else mech_gather<true>(a, cxp, cyp, X, Y, mv);
        n2m_update(a, t, s, mv);
    }
    if (s.dirty_state) {
        MK(a, MF_PHI, m) = s.phi; MK(a, MF_EXX, m) = s.exx; MK(a, MF_EXY, m) = s.exy;
        MK(a, MF_P, m) = s.p; MK(a, MF_VX, m) = s.vx; MK(a, MF_VY, m) = s.vy; MK(a, MF_E, m) = s.e;
    }
    bool moved = false, reset = false;
    if (ADVECT) {
        if (!s.dirty_state) { s.xx = MK(a, MF_XX, m); s.xy = MK(a, MF_XY, m); }
        if (!advect_one(a, t, s, moved, reset, &X, &Y)) unsupported = true;
    }
    if (s.dirty_state || moved) { MK(a, MF_XX, m) = s.xx; MK(a, MF_XY, m) = s.xy; }
    if (s.dirty_k) MK(a, MF_K, m) = s.k;
    if (ADVECT) {
        if (moved) {
            MK(a, MF_X, m) = s.x; MK(a, MF_Y, m) = s.y;
            a.key[m] = subcell_key(a.g, s.x, s.y);
        }
        if (reset) { MK(a, MF_D, m) = -1.0; MK(a, MF_W, m) = -1.0; MK(a, MF_E, m) = 0; }
    }
    return unsupported;
}

template <bool ADVECT>
__global__ __launch_bounds__(256) void k_n2m_update(KArgs a, int nwg)
{
    const long m = (long)xcd_remap(blockIdx.x, nwg) * 256 + threadIdx.x;
    const bool unsupported = m < a.n && n2m_update_one<ADVECT>(a, m);
    if (ADVECT) {
        const unsigned long long bad = __ballot(unsupported);
        if (bad && (threadIdx.x & (WAVE - 1)) == 0) atomicAdd(&a.counters[1], (int)__popcll(bad));
    }
}

// Nodal values of the fast form, addressed from the marker's sub-cell key
// alone.  The key fixes the cell (sx, sy) and the quadrant; rk_fill takes the
// quadrant from the cell's midlines instead.  Here the key's quadrant is
// checked against the midline tests once the gridline records arrive (a
// marker where the two disagree, a position within rounding of a midline,
// takes the general form), and the stencil cache's and the update's nodal
// values are gathered as one batch of loads at the key's stencils.
struct FastNodes {
    double e[4];                            // esp at the basic stencil (bil order)
    double u[6], v[6];                      // rk_fill's u00 u01 u10 u11 u20 u21 and v00 v01 v02 v10 v11 v12
    double exy[4], sxy[4], pes[4], phs[4];  // basic stencil
    double exx[4], sxx[4], pr[4];           // pressure-node stencil
};

// b: the basic stencil's first node (sx, sy); m3: the pressure node
// (sx + xup, sy + yup).  Interior cells only: every index is in range.
__device__ __forceinline__ void fast_gather(const KArgs &a, long b, long m3, FastNodes &f)
{
    const long ny = a.g.ynumy;
    const double *esp = NOD(a, V2C_N_esp), *vx = NOD(a, V2C_N_vx), *vy = NOD(a, V2C_N_vy);
    f.e[0] = esp[b]; f.e[1] = esp[b + 1]; f.e[2] = esp[b + ny]; f.e[3] = esp[b + ny + 1];
    f.u[0] = vx[m3 - ny - 1]; f.u[1] = vx[m3 - ny]; f.u[2] = vx[m3 - 1]; f.u[3] = vx[m3];
    f.u[4] = vx[m3 + ny - 1]; f.u[5] = vx[m3 + ny];
    f.v[0] = vy[m3 - ny - 1]; f.v[1] = vy[m3 - ny]; f.v[2] = vy[m3 - ny + 1];
    f.v[3] = vy[m3 - 1]; f.v[4] = vy[m3]; f.v[5] = vy[m3 + 1];
    auto quad = [&](int fld, long i, double *o) {
        const double *q = NOD(a, fld);
        o[0] = q[i]; o[1] = q[i + 1]; o[2] = q[i + ny]; o[3] = q[i + ny + 1];
    };
    quad(V2C_N_exy, b, f.exy); quad(V2C_N_sxy, b, f.sxy); quad(V2C_N_pes, b, f.pes);
    quad(V2C_N_phis_it, b, f.phs);
    quad(V2C_N_exx, m3, f.exx); quad(V2C_N_sxx, m3, f.sxx); quad(V2C_N_pr, m3, f.pr);
}

// rk_fill's cache from gathered values, for an interior cell (sx, sy) with
// the configuration (xup, yup) already checked against the midlines.
__device__ __forceinline__ void rk_fill_from(const Axis4 &X, const Axis4 &Y, bool xup, bool yup,
                                             const FastNodes &f, RkCache &c)
{
    c.ok = true;
    c.xup = xup;
    c.yup = yup;
    c.sx = X.s;
    c.sy = Y.s;
    rk_axis(X, xup, c.X);
    rk_axis(Y, yup, c.Y);
    for (int k = 0; k < 4; k++) c.e[k] = f.e[k];
    const double u00 = f.u[0], u01 = f.u[1], u10 = f.u[2], u11 = f.u[3], u20 = f.u[4], u21 = f.u[5];
    const double v00 = f.v[0], v01 = f.v[1], v02 = f.v[2], v10 = f.v[3], v11 = f.v[4], v12 = f.v[5];
    c.P[0] = u10 + u00; c.P[1] = u11 + u01; c.P[2] = u20 + u10; c.P[3] = u21 + u11;
    c.Q[0] = v01 + v00; c.Q[1] = v02 + v01; c.Q[2] = v11 + v10; c.Q[3] = v12 + v11;
    c.F[0] = xup ? u00 : u10; c.F[1] = xup ? u01 : u11;
    c.F[2] = xup ? u10 : u20; c.F[3] = xup ? u11 : u21;
    c.H[0] = yup ? v00 : v01; c.H[1] = yup ? v01 : v02;
    c.H[2] = yup ? v10 : v11; c.H[3] = yup ? v11 : v12;
}

__device__ __forceinline__ double bil4(const W4 &w, const double *f)
{
    return w.d0 * f[0] + w.d1 * f[1] + w.d2 * f[2] + w.d3 * f[3];
}

// mech_from_cache with the basic- and pressure-node fields from FastNodes
// (the same values bil would load, summed in bil's order).
__device__ __forceinline__ void mech_from_fast(const RkCache &c, const FastNodes &f, double x, double y, Mech &o)
{
    double ex, ny_, exc, nyc;
    rk_weights(c, x, y, ex, ny_, exc, nyc);
    Stencil S;
    S.e = ex; S.n = ny_;
    W4 w = wts(S);
    o.EXY = bil4(w, f.exy);
    o.SXY = bil4(w, f.sxy);
    o.ES = bil4(w, f.pes);
    o.PHIS = bil4(w, f.phs);
    S.e = exc; S.n = nyc;
    w = wts(S);
    o.EXX = bil4(w, f.exx);
    o.SXX = bil4(w, f.sxx);
    o.PR = bil4(w, f.pr);
    S.e = ex; S.n = nyc;
    w = wts(S);
    o.VX = w.d0 * c.F[0] + w.d1 * c.F[1] + w.d2 * c.F[2] + w.d3 * c.F[3];
    S.e = exc; S.n = ny_;
    w = wts(S);
    o.VY = w.d0 * c.H[0] + w.d1 * c.H[1] + w.d2 * c.H[2] + w.d3 * c.H[3];
}

// nodes2markers + moveandrotatemarkers of one marker through one stencil
// configuration (antidiffinter == 1): the marker's clamped position lies in
// an interior cell off both midlines, every RK stage point in the same
// configuration and the new position in the same cell.  Then the update's
// gather, all RK stages and the new key come from the one RkCache fill, with
// the same operands and operations as the general path.  Returns false,
// having written nothing, for any other marker (the general path takes it).
// (2.5 % faster than filling the cache from the records: fewer dependent
// loads.  Issuing the nodal loads before the midline test as well, with the
// loads of boundary cells redirected to an interior cell, needs 8 more
// registers than the 3-waves budget holds and spills: 7 % slower.)
__device__ __forceinline__ bool n2m_update_fast(const KArgs &a, long m)
{
    const DevGrid &g = a.g;
    const int t = a.t[m];
    const uint32_t key = a.key[m];
    MState s;
    s.x = MK(a, MF_X, m); s.y = MK(a, MF_Y, m); s.k = MK(a, MF_K, m);
    s.dirty_state = false; s.dirty_k = false;
    if (t >= 100) return false;
    const uint32_t ncy = (uint32_t)(g.ynumy - 1), cell = key >> 2;
    const uint32_t sx = cell / ncy, sy = cell - sx * ncy;
    if (sx == 0 || sx == (uint32_t)g.xnumx - 2 || sy == 0 || sy == ncy - 1) return false;
    const bool kxup = (key & 2u) != 0, kyup = (key & 1u) != 0;
    const long b = (long)sx * g.ynumy + sy;
    const long m3 = b + (kxup ? (long)g.ynumy : 0L) + (kyup ? 1L : 0L);
    const Axis4 X = g.cellx[sx], Y = g.celly[sy];
    const bool immobile = a.mat->im[t] != 0;
    FastNodes f;
    fast_gather(a, b, m3, f);
    const double cxp = clampd(s.x, 0, g.xsize), cyp = clampd(s.y, 0, g.ysize);
    const double xmid = (X.g[1] + X.g[2]) / 2.0, ymid = (Y.g[1] + Y.g[2]) / 2.0;
    if (cxp == xmid || cyp == ymid || cxp != cxp || cyp != cyp || (cxp > xmid) != kxup || (cyp > ymid) != kyup)
        return false;
    RkCache c;
    rk_fill_from(X, Y, kxup, kyup, f, c);
    Mech mv;
    mech_from_fast(c, f, cxp, cyp, mv);
    n2m_update(a, t, s, mv);
    // The update's values are written now (a marker handed on to the general
    // path gets the same values written again: with antidiffinter == 1 they
    // are the interpolants at the unchanged position, whatever the marker's
    // previous state; an incoming marker's new temperature makes the general
    // path take the other branch of the update, which writes the same
    // interpolants).  Only the stresses, which the rotation still changes,
    // stay in registers.
    if (s.dirty_state) {
        MK(a, MF_PHI, m) = s.phi; MK(a, MF_EXX, m) = s.exx; MK(a, MF_EXY, m) = s.exy;
        MK(a, MF_P, m) = s.p; MK(a, MF_VX, m) = s.vx; MK(a, MF_VY, m) = s.vy; MK(a, MF_E, m) = s.e;
    }
    if (s.dirty_k) MK(a, MF_K, m) = s.k;
    s.dirty_k = false;
    if (!s.dirty_state) { s.xx = MK(a, MF_XX, m); s.xy = MK(a, MF_XY, m); }
    bool moved = false, reset = false;
    uint32_t nkey = 0;
    if (advect_moves_flag(a, immobile, s)) {
        double vx0, vy0, sp0;
        auto interp = [&](double x, double y, double &vx, double &vy, double &sp) {
            const double xc = clampd(x, 0, g.xsize), yc = clampd(y, 0, g.ysize);
            if (!rk_hit(c, xc, yc)) return false;
            rk_eval(a, c, xc, yc, vx, vy, sp);
            return true;
        };
        if (a.p.markmod == 1) {
            if (!interp(s.x, s.y, vx0, vy0, sp0)) return false;
        } else if (!rk4_velocity(a, s, vx0, vy0, sp0, interp)) {
            return false;
        }
        advect_apply(a, s, vx0, vy0, sp0, reset);
        const double xc = clampd(s.x, 0, g.xsize), yc = clampd(s.y, 0, g.ysize);
        if (!rk_in_cell(c, xc, yc)) return false;
        nkey = key_from_cache(a, c, xc, yc);
        moved = true;
    }
    if (s.dirty_state || moved) { MK(a, MF_XX, m) = s.xx; MK(a, MF_XY, m) = s.xy; }
    if (s.dirty_k) MK(a, MF_K, m) = s.k;
    if (moved) {
        MK(a, MF_X, m) = s.x; MK(a, MF_Y, m) = s.y;
        a.key[m] = nkey;
    }
    if (reset) { MK(a, MF_D, m) = -1.0; MK(a, MF_W, m) = -1.0; MK(a, MF_E, m) = 0; }
    return true;
}

// Append the lanes' flagged items to a list, one atomic per wavefront (the
// list order varies from run to run; every item is processed on its own).
__device__ __forceinline__ void list_append(bool flag, uint32_t item, uint32_t *list, int *count)
{
    const unsigned long long mask = __ballot(flag);
    if (!mask) return;
    const int lane = threadIdx.x & (WAVE - 1);
    int base = 0;
    if (lane == 0) base = atomicAdd(count, (int)__popcll(mask));
    base = __shfl(base, 0);
    if (flag) list[base + (int)__popcll(mask & ((1ull << lane) - 1ull))] = item;
}

// Fused update, fast form: markers the fast path cannot take go to `defer`.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_n2m_update_fast(KArgs a, int nwg, uint32_t *defer, int *ndefer)
{
    const long m = (long)xcd_remap(blockIdx.x, nwg) * 256 + threadIdx.x;
    const bool def = m < a.n && !n2m_update_fast(a, m);
    list_append(def, (uint32_t)m, defer, ndefer);
}

// The deferred markers through the general fused update (grid-stride over
// the device-side count, a fixed grid of N2M_LIST_BLOCKS blocks).
constexpr int N2M_LIST_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_n2m_update_list(KArgs a, const uint32_t *list, const int *count)
{
    const int n = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (n2m_update_one<true>(a, list[i])) atomicAdd(&a.counters[1], 1);
    }
}

// The extra iteration of the reference's `m3<=marknum` loop
// (interpolation.c:937): a zero marker of type 0 at (0,0).  Only its
// antidiffusion scatter has an effect; added into the persistent sums.
__device__ __forceinline__ void n2m_phantom_one(const KArgs &a)
{
    double v[N_NF];
    for (int f = 0; f < N_NF; f++) v[f] = 0.0;
    anti_contrib(a, 0.0, 0.0, find_x(a.g, 0.0), find_y(a.g, 0.0), a.mat->ro[0], v);
    const int ynumy = a.g.ynumy;
    // at (0,0): cx = cy = 0 -> basic node 0, vx node 0, vy node 0, centre (1,1)
    const long cn = (long)ynumy + 1;
    NOD(a, V2C_N_phisum)[0] += v[0]; NOD(a, V2C_N_sxysum)[0] += v[1]; NOD(a, V2C_N_exysum)[0] += v[2];
    NOD(a, V2C_N_esum)[0] += v[3]; NOD(a, V2C_N_wtphisum)[0] += v[4];
    NOD(a, V2C_N_vxsum)[0] += v[N_VX]; NOD(a, V2C_N_mrx0)[0] += v[N_VX + 1]; NOD(a, V2C_N_wtvxsum)[0] += v[N_VX + 2];
    NOD(a, V2C_N_vysum)[0] += v[N_VY]; NOD(a, V2C_N_mry0)[0] += v[N_VY + 1]; NOD(a, V2C_N_wtvysum)[0] += v[N_VY + 2];
    NOD(a, V2C_N_psum)[cn] += v[N_C]; NOD(a, V2C_N_sxxsum)[cn] += v[N_C + 1];
    NOD(a, V2C_N_exxsum)[cn] += v[N_C + 2]; NOD(a, V2C_N_wtsxxsum)[cn] += v[N_C + 3];
}

__global__ void k_n2m_phantom(KArgs a)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && a.p.antidiffinter == 1) n2m_phantom_one(a);
}

// The antidiffusion scatter.  Its first workgroup also does what would
// otherwise be separate one-thread launches before it: with N2M_PRE_PHANTOM
// the phantom marker's contribution (which goes straight to the nodal sums
// the node pass completes, not to the records this kernel writes), with
// N2M_PRE_CLEAR the clearing of the update's counters (unsupported markers,
// deferred markers), which no kernel before the update reads.
enum { N2M_PRE_PHANTOM = 1, N2M_PRE_CLEAR = 2 };
__global__ __launch_bounds__(WAVE) void k_n2m_scatter(KArgs a, int nwg, int pre)
{
    __shared__ N2mLds L;
    if (pre && blockIdx.x == 0 && threadIdx.x == 0) {
        if (pre & N2M_PRE_PHANTOM) n2m_phantom_one(a);
        if (pre & N2M_PRE_CLEAR) { a.counters[1] = 0; a.counters[2] = 0; }
    }
    const long c0 = (long)xcd_remap(blockIdx.x, nwg) * CPW;
    n2m_scatter_group(a, c0, c0 + CPW < a.ncell ? c0 + CPW : a.ncell, threadIdx.x, L);
}

// nodes2markers node pass: add the cell partials to the persistent sums and
// normalise (interpolation.c:1014-1046).
__device__ __forceinline__ void n2m_nodes_one(const KArgs &a, long m3)
{
    const int ynumy = a.g.ynumy, xnumx = a.g.xnumx;
    const int m22 = (int)(m3 % ynumy), m11 = (int)(m3 / ynumy);
    const double *P = a.part;
    const bool c00 = m11 >= 1 && m22 >= 1, c01 = m11 >= 1 && m22 < ynumy - 1;
    const bool c10 = m11 < xnumx - 1 && m22 >= 1, c11 = m11 < xnumx - 1 && m22 < ynumy - 1;
    double B[N_NB] = {0, 0, 0, 0, 0};
    {
        const double *R = NRec::basic(P, m3);
        const bool has[4] = {c00, c01, c10, c11};
        for (int k = 0; k < 4; k++) {
            if (!has[k]) continue;
            for (int f = 0; f < N_NB; f++) B[f] += R[k * N_NB + f];
        }
    }
    double VXs[N_NV] = {0, 0, 0}, VYs[N_NV] = {0, 0, 0}, Cs[N_NC] = {0, 0, 0, 0};
    {
        const double *R = NRec::vx(P, m3, a.nodenum);
        for (int f = 0; f < N_NV; f++) {
            if (c01) VXs[f] += R[0 * N_NV + f] + R[1 * N_NV + f];
            if (c11) VXs[f] += R[2 * N_NV + f] + R[3 * N_NV + f];
        }
        R = NRec::vy(P, m3, a.nodenum);
        for (int f = 0; f < N_NV; f++) {
            if (c10) VYs[f] += R[0 * N_NV + f] + R[1 * N_NV + f];
            if (c11) VYs[f] += R[2 * N_NV + f] + R[3 * N_NV + f];
        }
        R = NRec::centre(P, m3, a.nodenum);
        if (c00)
            for (int f = 0; f < N_NC; f++)
                Cs[f] = R[0 * N_NC + f] + R[1 * N_NC + f] + R[2 * N_NC + f] + R[3 * N_NC + f];
    }
    double *phisum = NOD(a, V2C_N_phisum), *sxysum = NOD(a, V2C_N_sxysum), *exysum = NOD(a, V2C_N_exysum);
    double *esum = NOD(a, V2C_N_esum), *wtphisum = NOD(a, V2C_N_wtphisum);
    double *vxsum = NOD(a, V2C_N_vxsum), *mrx0 = NOD(a, V2C_N_mrx0), *wtvxsum = NOD(a, V2C_N_wtvxsum);
    double *vysum = NOD(a, V2C_N_vysum), *mry0 = NOD(a, V2C_N_mry0), *wtvysum = NOD(a, V2C_N_wtvysum);
    double *psum = NOD(a, V2C_N_psum), *sxxsum = NOD(a, V2C_N_sxxsum), *exxsum = NOD(a, V2C_N_exxsum);
    double *wtsxxsum = NOD(a, V2C_N_wtsxxsum);
    const double s_phi = phisum[m3] + B[0], s_sxy = sxysum[m3] + B[1], s_exy = exysum[m3] + B[2];
    const double s_e = esum[m3] + B[3], s_w = wtphisum[m3] + B[4];
    const double s_vx = vxsum[m3] + VXs[0], s_mrx = mrx0[m3] + VXs[1], s_wvx = wtvxsum[m3] + VXs[2];
    const double s_vy = vysum[m3] + VYs[0], s_mry = mry0[m3] + VYs[1], s_wvy = wtvysum[m3] + VYs[2];
    const double s_p = psum[m3] + Cs[0], s_sxx = sxxsum[m3] + Cs[1], s_exx = exxsum[m3] + Cs[2];
    const double s_wc = wtsxxsum[m3] + Cs[3];
    phisum[m3] = s_phi; sxysum[m3] = s_sxy; exysum[m3] = s_exy; esum[m3] = s_e; wtphisum[m3] = s_w;
    vxsum[m3] = s_vx; mrx0[m3] = s_mrx; wtvxsum[m3] = s_wvx;
    vysum[m3] = s_vy; mry0[m3] = s_mry; wtvysum[m3] = s_wvy;
    psum[m3] = s_p; sxxsum[m3] = s_sxx; exxsum[m3] = s_exx; wtsxxsum[m3] = s_wc;
    if (s_w != 0 && m11 >= 1 && m22 >= 1 && m11 < xnumx - 1 && m22 < ynumy - 1) {
        NOD(a, V2C_N_phie)[m3] = s_phi / s_w;
        NOD(a, V2C_N_sxyee)[m3] = s_sxy / s_w;
        NOD(a, V2C_N_exyee)[m3] = s_exy / s_w;
        NOD(a, V2C_N_pee)[m3] = s_e / s_w;
    }
    if (s_wc != 0 && m11 >= 1 && m22 >= 1) {
        NOD(a, V2C_N_sppee)[m3] = s_p / s_wc;
        NOD(a, V2C_N_sxxee)[m3] = s_sxx / s_wc;
        NOD(a, V2C_N_exxee)[m3] = s_exx / s_wc;
    }
    if (s_wvx != 0) NOD(a, V2C_N_mvxe)[m3] = s_vx / s_mrx;
    if (s_wvy != 0) NOD(a, V2C_N_mvye)[m3] = s_vy / s_mry;
}

__global__ __launch_bounds__(256) void k_n2m_nodes(KArgs a)
{
    const long m3 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m3 < a.nodenum) n2m_nodes_one(a, m3);
}

// Standalone moveandrotatemarkers over physical marker order.
__global__ __launch_bounds__(256) void k_advect(KArgs a)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int unsupported = 0;
    if (m < a.n) {
        const int t = a.t[m];
        MState s;
        s.x = MK(a, MF_X, m); s.y = MK(a, MF_Y, m); s.k = MK(a, MF_K, m);
        s.xx = MK(a, MF_XX, m); s.xy = MK(a, MF_XY, m);
        s.dirty_k = false;
        Axis4 X, Y;
        key_axes(a, a.key[m], X, Y);
        bool moved, reset;
        if (!advect_one(a, t, s, moved, reset, &X, &Y)) unsupported = 1;
        if (moved) {
            MK(a, MF_X, m) = s.x; MK(a, MF_Y, m) = s.y;
            MK(a, MF_XX, m) = s.xx; MK(a, MF_XY, m) = s.xy;
            a.key[m] = subcell_key(a.g, s.x, s.y);
        }
        if (s.dirty_k) MK(a, MF_K, m) = s.k;
        if (reset) { MK(a, MF_D, m) = -1.0; MK(a, MF_W, m) = -1.0; MK(a, MF_E, m) = 0; }
    }
    if (unsupported) atomicAdd(&a.counters[1], 1);
}

// ---------------------------------------------------------------------------
// compute_before_solve / calc_vpvisc (plastic.c:2-56, calc_RSF.c:33-169)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void before_one(const KArgs &a, long m3, double dt, int rk, int slot)
{
    const int ynumy = a.g.ynumy;
    const int m2 = (int)(m3 % ynumy), m1 = (int)(m3 / ynumy);
    if (!(m1 >= 1 && m2 >= 1 && m1 < a.g.xnumx - 1 && m2 < ynumy - 1)) return;
    const v2c_params &p = a.p;
    const double W = p.faultwidth;
    const double nps = (p.pmode == 0 || a.timesum == 0) ? p.nstress : NOD(a, V2C_N_sppes)[m3];
    const double sxxs = NOD(a, V2C_N_sxxs)[m3], sxy = NOD(a, V2C_N_sxy)[m3];
    const double sigin = pow(sxxs * sxxs + sxy * sxy, 0.5);
    int err = 0;
    double nu;
    if (sigin > 0 || p.RSFmode == 0) {
        double V = 0;
        if (rk == 0) {
            if (a.timesum > 0) V = NOD(a, V2C_N_RK_sliprs_final)[m3];
            else if (p.RSFmode == 1) V = NOD(a, V2C_N_RK_sliprs_final)[m3];
        }
        if ((p.rkmod == 0 && rk > 0) || (rk > 0 && rk < 3)) {
            if (p.rkmod == 1) {
                V = NOD(a, V2C_N_RK_sliprs_solved)[m3];
            } else {
                if (p.rkmod == 3) {
                    NOD(a, rk - 1 == 0 ? V2C_N_RK_sliprs0 : V2C_N_RK_sliprs1)[m3] = NOD(a, V2C_N_RK_sliprs_solved)[m3];
                    if (p.RSFmode == 1)
                        NOD(a, rk - 1 == 0 ? V2C_N_RK_phidts0 : V2C_N_RK_phidts1)[m3] = NOD(a, V2C_N_RK_phidts_solved)[m3];
                }
                V = NOD(a, V2C_N_RK_sliprs_solved)[m3];
            }
        }
        if (rk == 3 && p.rkmod == 3) V = NOD(a, V2C_N_RK_sliprs_solved)[m3];
        NOD(a, V2C_N_pes)[m3] = NOD(a, V2C_N_pe)[m3] + V / W * dt;
        double newstate = 0, syield, vp = 0;
        const double as = NOD(a, V2C_N_as)[m3], bs = NOD(a, V2C_N_bs)[m3], Ls = NOD(a, V2C_N_Ls)[m3];
        const double V0s = NOD(a, V2C_N_V0s)[m3], mus = NOD(a, V2C_N_mus_s)[m3];
        const double nu_m2c = NOD(a, V2C_N_nu_m2c)[m3];
        if (p.RSFmode == 1) {
            const double phis = NOD(a, V2C_N_phis)[m3];
            if (V * dt / Ls <= 1e-6) newstate = log(exp(phis) * (1 - V * dt / Ls) + V0s * dt / Ls);
            else newstate = log(V0s / V + (exp(phis) - V0s / V) * exp(-V * dt / Ls));
            const double rad = p.inertyn == 0 ? p.rad_damping * V : 0.0;
            if (p.power_law_RSF > 0) {
                const double sy0 = nps * mus;
                syield = sy0 * pow(V / V0s, as / mus) * exp((newstate * bs) / mus) + rad;
            } else {
                syield = nps * as * asinh(V / (2 * V0s) * exp((mus + newstate * bs) / as)) + rad;
            }
        } else {
            syield = nps * mus + NOD(a, V2C_N_cohesion)[m3];
        }
        if (p.RSFmode == 1) vp = syield > 0 ? nu_m2c * syield / (2 * nu_m2c * V / W + syield) : nu_m2c;
        else if (p.RSFmode == 0) vp = nu_m2c * syield / (2 * nu_m2c * V / W + syield);
        NOD(a, V2C_N_phis_it)[m3] = newstate;
        NOD(a, V2C_N_syields)[m3] = syield;
        NOD(a, V2C_N_nuvp_it)[m3] = vp;
        NOD(a, V2C_N_sliprs)[m3] = V;
        if ((a.timesum > 0 && (isnan(newstate) || isnan(syield) || isnan(vp) || vp == 0 || isnan(V)))
            || V < 0 || syield < 0)
            err = 1;
        nu = vp;
    } else {
        nu = NOD(a, V2C_N_nu_m2c)[m3];
    }
    NOD(a, V2C_N_nu)[m3] = nu;
    if (isnan(nu) || nu <= 0) err = 1;
    if (err) atomicAdd(&a.counters[slot], 1);
}

// The error count goes to counters[slot]; the other slot of the pair
// (slot ^ BEFORE_SLOTS), the next call's, is cleared here, so no separate
// clearing launch precedes the kernel.
constexpr int BEFORE_SLOTS = 4;
__global__ __launch_bounds__(256) void k_before(KArgs a, double dt, int rk, int slot)
{
    const long m3 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m3 == 0) a.counters[slot ^ BEFORE_SLOTS] = 0;
    if (m3 < a.nodenum) before_one(a, m3, dt, rk, slot);
}

// compute_after_solve / calc_sliprate_after_solve (plastic.c:61-129,
// calc_RSF.c:172-221).  Per-block partial sums of the misfit (fixed tree
// order) go to red[2*block], red[2*block+1].
constexpr int RED_BLOCK = 256;
// One node of compute_after_solve; adds its misfit term to (dsq, cnt).
__device__ __forceinline__ void after_one(const KArgs &a, long m3, double dt, double &dsq, double &cnt)
{
    const int ynumy = a.g.ynumy;
    const int m2 = (int)(m3 % ynumy), m1 = (int)(m3 / ynumy);
    if (m1 >= 1 && m2 >= 1 && m1 < a.g.xnumx - 1 && m2 < ynumy - 1) {
        const v2c_params &p = a.p;
        const double W = p.faultwidth;
        const double nP = (p.pmode == 0 || a.timesum == 0) ? p.nstress : NOD(a, V2C_N_sppes)[m3];
        const double sxxs = NOD(a, V2C_N_sxxs)[m3], sxy = NOD(a, V2C_N_sxy)[m3];
        const double sigin = pow(sxxs * sxxs + sxy * sxy, 0.5);
        double *Vs = NOD(a, V2C_N_RK_sliprs_solved);
        const double as = NOD(a, V2C_N_as)[m3], bs = NOD(a, V2C_N_bs)[m3], Ls = NOD(a, V2C_N_Ls)[m3];
        const double V0s = NOD(a, V2C_N_V0s)[m3], mus = NOD(a, V2C_N_mus_s)[m3];
        const double phi = NOD(a, V2C_N_phis_it)[m3];
        double Vout;
        if (p.RSFmode == 1) {
            if (sigin > 0) {
                double V;
                if (p.power_law_RSF > 0) {
                    const double sy0 = nP * mus;
                    V = V0s * pow(sigin / sy0, (mus / as)) * exp((-phi * bs) / as);
                } else {
                    V = 2 * V0s * sinh(sigin / (as * nP)) * exp(-(mus + bs * phi) / as);
                }
                const double EPXX = NOD(a, V2C_N_exxpl)[m3], EXY = NOD(a, V2C_N_exy)[m3];
                const double EII = pow(EPXX * EPXX + EXY * EXY, 0.5);
                V = ref_minv(V, EII * W);
                if (V < 1e-40) V = 1e-40;
                Vout = V;
            } else {
                Vout = 0;
            }
            Vs[m3] = Vout;
            NOD(a, V2C_N_RK_phidts_solved)[m3] = V0s / Ls * exp(-phi) - Vout / Ls;
        } else {
            const double G = NOD(a, V2C_N_gg)[m3], NUVP = NOD(a, V2C_N_nu)[m3];
            const double NU = NOD(a, V2C_N_nu_m2c)[m3];
            const double E2EL = sigin / (NUVP / (G * dt + NUVP));
            const double A = NOD(a, V2C_N_syields)[m3] / E2EL;
            if (A < 1) {
                const double nup = G * dt * A / (1.0 - A);
                Vout = W * (sigin / 2.0 * (1.0 / (nup) - 1.0 / NU));
            } else {
                Vout = 0;
            }
            Vs[m3] = Vout;
        }
        if (Vout > 0) {
            const double d = NOD(a, V2C_N_syields)[m3] - sigin;
            NOD(a, V2C_N_RK_deltas)[m3] = d;
            dsq = pow(d, 2);
            cnt = 1.0;
        } else {
            NOD(a, V2C_N_RK_deltas)[m3] = 0;
        }
    }
}

// Fixed-order tree sum of the block's (dsq, cnt) into red[2*slot], red[2*slot+1].
__device__ __forceinline__ void after_block_sum(double dsq, double cnt, double *red, long slot)
{
    __shared__ double ssum[RED_BLOCK];
    __shared__ double scnt[RED_BLOCK];
    ssum[threadIdx.x] = dsq;
    scnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = RED_BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
            scnt[threadIdx.x] += scnt[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        red[2 * slot] = ssum[0];
        red[2 * slot + 1] = scnt[0];
    }
}

__global__ __launch_bounds__(RED_BLOCK) void k_after(KArgs a, double dt, double *red)
{
    const long m3 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double dsq = 0.0, cnt = 0.0;
    if (m3 < a.nodenum) after_one(a, m3, dt, dsq, cnt);
    after_block_sum(dsq, cnt, red, blockIdx.x);
}

// Single-block fixed-order reduction of the per-block partials.
__global__ __launch_bounds__(RED_BLOCK) void k_after_final(const double *red, int nblocks, double *out)
{
    __shared__ double ssum[RED_BLOCK];
    __shared__ double scnt[RED_BLOCK];
    double s = 0, c = 0;
    for (int b = threadIdx.x; b < nblocks; b += RED_BLOCK) { s += red[2 * b]; c += red[2 * b + 1]; }
    ssum[threadIdx.x] = s;
    scnt[threadIdx.x] = c;
    __syncthreads();
    for (int k = RED_BLOCK / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) { ssum[threadIdx.x] += ssum[threadIdx.x + k]; scnt[threadIdx.x] += scnt[threadIdx.x + k]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double count = scnt[0];
        out[0] = count > 0 ? pow(ssum[0] / count, 0.5) : 0.0;
        out[1] = count;
    }
}

// ductilevisc on independent inputs (the host-callable helper).
__global__ void k_ductile_batch(KArgs a, long n, const int8_t *t, const double *sxy, const double *sxx,
                                const double *tk, const double *pb, double *out, int *iters)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int it;
    out[i] = ductilevisc(a, t[i], sxy[i], sxx[i], tk[i], pb[i], &it);
    iters[i] = it;
}

// Plain streaming copy of known size (calibrates the HBM byte counters).
__global__ void k_copy_calib(const double *__restrict__ src, double *__restrict__ dst, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace v2c
