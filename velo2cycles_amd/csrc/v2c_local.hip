// v2c_local.hip -- the build-defined per-marker local update (no reference
// counterpart): the ductile viscosity of every marker by Newton on the
// reference's residual (density_ductilevisc_calc.c:188-193), and the slip
// rate of every fault marker by Newton on the rate-and-state yield law with
// radiation damping at the aged state (calc_RSF.c:96-129, 193-196), each to
// the stated relative residual.
//
// Two kernels: the ductile solve over all markers in physical order, and the
// slip-rate solve over the context's list of fault markers (built when the
// RSF inputs are uploaded), so the slip-rate work and its warm start follow
// the fault markers alone.  Warm starts:
//   warm == 1  the guesses stored in MX_GETA / MX_GV (v2c_mlp_predict or the
//              host) for both solves;
//   warm == 2  the slip-rate guess from the network, evaluated inside the
//              slip-rate kernel (BF16 MFMA, one wavefront per 64 fault
//              markers); the ductile solve starts from its default, where it
//              converges in about one iteration.
// Statistics go to lu_stats with integer atomics (order independent), summed
// per wavefront first.
#pragma once
#include "v2c_mlp.hip"

namespace v2c {

constexpr int LU_BLOCK = 256;
constexpr int LU_HIST = 32;
enum { LS_N = 0, LS_DUCT, LS_DUCT_IT, LS_FAULT, LS_RSF_IT, LS_DUCT_MAX, LS_RSF_MAX, LS_FALLBACK,
       LS_NONFINITE, LS_DUCT_EV, LS_RSF_EV, LS_HIST_D, LS_HIST_R = LS_HIST_D + LU_HIST,
       LS_COUNT = LS_HIST_R + LU_HIST };

__device__ __forceinline__ void lu_clear(unsigned long long *sh)
{
    for (int i = threadIdx.x; i < LS_COUNT; i += blockDim.x) sh[i] = 0;
    __syncthreads();
}

__device__ __forceinline__ void lu_flush(const unsigned long long *sh, unsigned long long *out)
{
    __syncthreads();
    for (int i = threadIdx.x; i < LS_COUNT; i += blockDim.x) {
        if (!sh[i]) continue;
        if (i == LS_DUCT_MAX || i == LS_RSF_MAX) atomicMax(&out[i], sh[i]);
        else atomicAdd(&out[i], sh[i]);
    }
}

// One marker's contribution to the statistics (hist < 0: none).
struct LuAcc {
    int n = 0, duct = 0, duct_it = 0, duct_ev = 0, duct_hist = -1;
    int fault = 0, rsf_it = 0, rsf_ev = 0, rsf_hist = -1;
    int fallback = 0, nonfinite = 0;
};

__device__ __forceinline__ int wave_sum(int v)
{
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
    for (int o = WAVE / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, WAVE));
    return v;
}

// Histogram bins of a wavefront: one LDS add per distinct bin.
__device__ __forceinline__ void wave_hist(int bin, unsigned long long *sh, int base, int lane)
{
    unsigned long long pending = __ballot(bin >= 0);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int b = __shfl(bin, leader, WAVE);
        const unsigned long long same = __ballot(bin == b);
        if (lane == leader) atomicAdd(&sh[base + b], (unsigned long long)__popcll(same));
        pending &= ~same;
    }
}

// Every lane of the wavefront calls this with its marker's contribution
// (zeros for lanes without one); lane 0 adds the wavefront's totals to the
// workgroup's counters, so the LDS sees a handful of atomics per wavefront
// instead of several per marker.
__device__ __forceinline__ void lu_wave_add(const LuAcc &c, unsigned long long *sh)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int n = (int)__popcll(__ballot(c.n != 0)), duct = (int)__popcll(__ballot(c.duct != 0));
    const int fault = (int)__popcll(__ballot(c.fault != 0));
    const int fb = (int)__popcll(__ballot(c.fallback != 0)), nf = (int)__popcll(__ballot(c.nonfinite != 0));
    const int dit = wave_sum(c.duct_it), dev = wave_sum(c.duct_ev), dmax = wave_max(c.duct_it);
    const int rit = wave_sum(c.rsf_it), rev = wave_sum(c.rsf_ev), rmax = wave_max(c.rsf_it);
    wave_hist(c.duct_hist, sh, LS_HIST_D, lane);
    wave_hist(c.rsf_hist, sh, LS_HIST_R, lane);
    if (lane == 0) {
        if (n) atomicAdd(&sh[LS_N], (unsigned long long)n);
        if (duct) {
            atomicAdd(&sh[LS_DUCT], (unsigned long long)duct);
            atomicAdd(&sh[LS_DUCT_IT], (unsigned long long)dit);
            atomicAdd(&sh[LS_DUCT_EV], (unsigned long long)dev);
            atomicMax(&sh[LS_DUCT_MAX], (unsigned long long)dmax);
        }
        if (fault) {
            atomicAdd(&sh[LS_FAULT], (unsigned long long)fault);
            atomicAdd(&sh[LS_RSF_IT], (unsigned long long)rit);
            atomicAdd(&sh[LS_RSF_EV], (unsigned long long)rev);
            atomicMax(&sh[LS_RSF_MAX], (unsigned long long)rmax);
        }
        if (fb) atomicAdd(&sh[LS_FALLBACK], (unsigned long long)fb);
        if (nf) atomicAdd(&sh[LS_NONFINITE], (unsigned long long)nf);
    }
}

// Ductile solve of marker m (all markers).
__device__ __forceinline__ void lu_ductile_one(const KArgs &a, long m, bool warm, double tol, LuAcc &acc)
{
    const int t = a.t[m];
    const double sxx = MK(a, MF_XX, m), sxy = MK(a, MF_XY, m);
    const double tk = MK(a, MF_K, m), p = MK(a, MF_P, m);
    int it = 0, fb = 0, ev = 0;
    double root_eta;
    const double eta = ductile_newton(a, t, sxy, sxx, tk, 1e-5 * p, warm ? MXF(a, MX_GETA, m) : 0.0, tol, &it, &ev,
                                      &fb, &root_eta);
    MXF(a, MX_ETA, m) = eta;
    MXF(a, MX_ROOT_ETA, m) = root_eta;
    acc.n = 1;
    if (power_law_type(a.mat, t) && t != 1) {
        acc.duct = 1;
        acc.duct_it = it;
        acc.duct_ev = ev;
        acc.duct_hist = it < LU_HIST ? it : LU_HIST - 1;
    }
    if (fb) acc.fallback = 1;
    if (!isfinite(eta)) acc.nonfinite = 1;
}

// Slip-rate solve of fault marker m from the guess guessV (0: default start).
// It does not read the ductile solve's result, so the two run side by side;
// k_local_etavp combines them.
__device__ __forceinline__ void lu_rsf_one(const KArgs &a, long m, double dt, double guessV, double tol,
                                           LuAcc &acc)
{
    const double sxx = MK(a, MF_XX, m), sxy = MK(a, MF_XY, m), p = MK(a, MF_P, m);
    const double exx = MK(a, MF_EXX, m), exy = MK(a, MF_EXY, m);
    const double sig = pow(sxx * sxx + sxy * sxy, 0.5);
    const double nP = a.p.pmode == 0 ? a.p.nstress : p;
    const double EII = pow(exx * exx + exy * exy, 0.5);
    const double eta_rad = a.p.inertyn == 0 ? a.p.rad_damping : 0.0;
    const double W = a.p.faultwidth;
    double V, phin, root_v;
    int rfb = 0, rev = 0;
    const int rit = rsf_newton(sig, nP, MXF(a, MX_A, m), MXF(a, MX_B, m), MXF(a, MX_L, m), MXF(a, MX_V0, m),
                               MXF(a, MX_MU0, m), MXF(a, MX_PHI, m), MXF(a, MX_V, m), dt, eta_rad, EII, W, guessV,
                               tol, V, phin, rfb, root_v, rev);
    MXF(a, MX_ROOT_V, m) = root_v;
    MXF(a, MX_V, m) = V;
    MXF(a, MX_PHI, m) = phin;
    MXF(a, MX_SY, m) = sig;
    acc.fault = 1;
    acc.rsf_it = rit;
    acc.rsf_ev = rev;
    acc.rsf_hist = rit < LU_HIST ? rit : LU_HIST - 1;
    if (rfb) acc.fallback = 1;
    if (!isfinite(V) || !isfinite(phin)) acc.nonfinite = 1;
}

__global__ __launch_bounds__(LU_BLOCK) void k_local_ductile(KArgs a, int warm, double tol)
{
    __shared__ unsigned long long sh[LS_COUNT];
    lu_clear(sh);
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    LuAcc acc;
    if (m < a.n) lu_ductile_one(a, m, warm == 1, tol, acc);
    lu_wave_add(acc, sh);
    lu_flush(sh, a.lu_stats);
}

// Fault markers i0 + blockIdx.x * LU_BLOCK + threadIdx.x < nf of the list.
// WARM == 2: the workgroup stages the network's weights once and every
// wavefront evaluates it for its 64 fault markers.
struct LuRsfLds {
    MlpW w;
    short x[LU_BLOCK / 64][64][16];
    short h[LU_BLOCK / 64][64][MLP_H];
};

// Entry i (< iend) of the fault list.
template <int WARM>
__device__ __forceinline__ void lu_rsf_block(const KArgs &a, long i, long iend, double dt, double tol,
                                             unsigned long long *sh, LuRsfLds *L)
{
    LuAcc acc;
    const bool live = i < iend;
    const long m = live ? (long)a.fault_list[i] : 0;
    double guess = 0.0;
    if (WARM == 1) {
        if (live) guess = MXF(a, MX_GV, m);
    } else if (WARM == 2) {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        MlpFeat ft;
        ft.rsf = false;
        ft.vhi = 0;
        if (live) mlp_features<false>(a, m, dt, ft);
        mlp_stage_weights(a.mlp, L->w, threadIdx.x, LU_BLOCK);
        mlp_load_row(ft, a.mlp_norm, live, L->x[w][lane]);
        __syncthreads();
        float y0, y1;
        mlp_infer_wave(L->w, L->x[w], L->h[w], y0, y1);
        (void)y0;
        if (ft.rsf) {
            const double r1 = y1 * a.mlp_norm[2 * MLP_IN + 3] + a.mlp_norm[2 * MLP_IN + 1];
            guess = ft.vhi * feat_exp(r1 * LN10);
        }
    }
    if (live) lu_rsf_one(a, m, dt, guess, tol, acc);
    lu_wave_add(acc, sh);
}

template <int WARM>
__global__ __launch_bounds__(LU_BLOCK) void k_local_rsf(KArgs a, double dt, double tol)
{
    __shared__ unsigned long long sh[LS_COUNT];
    lu_clear(sh);
    const long i = (long)blockIdx.x * LU_BLOCK + threadIdx.x;
    if constexpr (WARM == 2) {
        __shared__ LuRsfLds L;     // only the network form holds the staging
        lu_rsf_block<2>(a, i, a.nfault, dt, tol, sh, &L);
    } else {
        lu_rsf_block<WARM>(a, i, a.nfault, dt, tol, sh, nullptr);
    }
    lu_flush(sh, a.lu_stats);
}

// Visco-plastic viscosity of the fault markers from both solves
// (calc_RSF.c:138-152 with the marker's eta, V and sigII).
__device__ __forceinline__ void lu_etavp_one(const KArgs &a, long i)
{
    const long m = a.fault_list[i];
    const double eta = MXF(a, MX_ETA, m), V = MXF(a, MX_V, m), sig = MXF(a, MX_SY, m);
    const double W = a.p.faultwidth;
    MXF(a, MX_ETAVP, m) = sig > 0 ? eta * sig / (2 * eta * V / W + sig) : eta;
}

__global__ __launch_bounds__(LU_BLOCK) void k_local_etavp(KArgs a)
{
    const long i = (long)blockIdx.x * LU_BLOCK + threadIdx.x;
    if (i < a.nfault) lu_etavp_one(a, i);
}

}  // namespace v2c
