// v2c_mlp.hip -- learned warm start of the per-marker local update.
//
// A 9 -> 64 -> 64 -> 2 tanh MLP proposes, per marker, the log10 ratio of the
// Newton root to the analytic default start for (i) the ductile unknown
// (dislocation viscosity; default eta_disl0 of density_ductilevisc_calc.c:159)
// and (ii) the fault slip rate (default min(undamped closed form
// calc_RSF.c:192, sigII/eta_rad)).  Inputs: log10 strain-rate invariant, T, P,
// log10 stress invariant, log10 theta, a-b, and for fault markers the log10 of
// the slip-rate default start, log10 L and log10 of the previous slip rate,
// standardised.  (The last three fix where the default start sits relative to
// the root: it follows the aged state, which depends on L and the previous
// rate.)
//
// Forward and backward run on the matrix cores: v_mfma_f32_16x16x16_bf16 with
// BF16 operands and FP32 accumulation.  One wavefront owns 64 rows (markers or
// training samples); activations are staged in LDS between layers.  Training
// writes per-wavefront gradient partials that a second kernel sums in a fixed
// order before an Adam step on FP32 master weights, so training is
// deterministic.
#pragma once
#include "v2c_device.hpp"

namespace v2c {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MLP_IN = 9, MLP_H = 64, MLP_OUT = 2;
// flat FP32 parameter layout, weights stored [out][in]
constexpr int P_W1 = 0, P_B1 = P_W1 + MLP_H * MLP_IN, P_W2 = P_B1 + MLP_H, P_B2 = P_W2 + MLP_H * MLP_H,
              P_W3 = P_B2 + MLP_H, P_B3 = P_W3 + MLP_OUT * MLP_H, MLP_NPARAM = P_B3 + MLP_OUT;
static_assert(MLP_NPARAM == 4930, "parameter count");
static_assert(MLP_IN <= 16, "layer 1 pads the inputs to one 16-wide MFMA step");
// normalisation block: in_mean[MLP_IN] in_std[MLP_IN] out_mean[2] out_std[2]
constexpr int MLP_NNORM = 2 * MLP_IN + 4;

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ short f2bf(float f) { return __builtin_bit_cast(short, (__bf16)f); }
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

// tanh(x) = 1 - 2 / (exp(2x) + 1): one exponential and one hardware
// reciprocal (about 1 ulp); saturates to +-1 through exp's overflow and
// underflow.  Absolute error near 1e-7, ample for a guess network (its
// training uses the same function).
__device__ __forceinline__ float fast_tanh(float x)
{
    return fmaf(-2.0f, __builtin_amdgcn_rcpf(__expf(2.0f * x) + 1.0f), 1.0f);
}

// One 16x16 output tile, K = 16*ksteps.  A(i,k) and B(k,n) are fetched by
// callables; lane l supplies A rows i = l%16 with k = 4*(l/16)+j and B
// columns n = l%16 with the same k; the result element j of lane l is
// D(4*(l/16)+j, l%16).
template <typename FA, typename FB>
__device__ __forceinline__ f32x4 mfma_tile(int ksteps, FA &&fa, FB &&fb)
{
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, kb = 4 * (lane >> 4);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < ksteps; ks++) {
        bf16x4 av, bv;
        for (int j = 0; j < 4; j++) {
            const int k = ks * 16 + kb + j;
            av[j] = fa(i, k);
            bv[j] = fb(k, i);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// Activation staging of one wavefront (64 rows).
struct MlpLds {
    short x[64][16];          // inputs, bf16, padded to K = 16
    float h1[64][MLP_H];      // tanh(Z1)
    float h2[64][MLP_H];      // tanh(Z2)
    float g[64][MLP_H];       // scratch: dZ2 then dZ1 (FP32)
    float y[64][MLP_OUT];     // outputs / dY
};

// Forward pass of the wave's 64 rows: x must be filled; fills h1, h2, y.
__device__ __forceinline__ void mlp_forward(const float *__restrict__ P, MlpLds &s)
{
    const int lane = threadIdx.x & 63;
    const int n = lane & 15, ib = 4 * (lane >> 4);
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++) {
            f32x4 d = mfma_tile(1, [&](int i, int k) { return s.x[mt * 16 + i][k]; },
                                [&](int k, int c) { return k < MLP_IN ? f2bf(P[P_W1 + (nt * 16 + c) * MLP_IN + k]) : (short)0; });
            for (int j = 0; j < 4; j++)
                s.h1[mt * 16 + ib + j][nt * 16 + n] = fast_tanh(d[j] + P[P_B1 + nt * 16 + n]);
        }
    __syncthreads();
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++) {
            f32x4 d = mfma_tile(4, [&](int i, int k) { return f2bf(s.h1[mt * 16 + i][k]); },
                                [&](int k, int c) { return f2bf(P[P_W2 + (nt * 16 + c) * MLP_H + k]); });
            for (int j = 0; j < 4; j++)
                s.h2[mt * 16 + ib + j][nt * 16 + n] = fast_tanh(d[j] + P[P_B2 + nt * 16 + n]);
        }
    __syncthreads();
    for (int mt = 0; mt < 4; mt++) {
        f32x4 d = mfma_tile(4, [&](int i, int k) { return f2bf(s.h2[mt * 16 + i][k]); },
                            [&](int k, int c) { return c < MLP_OUT ? f2bf(P[P_W3 + c * MLP_H + k]) : (short)0; });
        if (n < MLP_OUT)
            for (int j = 0; j < 4; j++) s.y[mt * 16 + ib + j][n] = d[j] + P[P_B3 + n];
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Features and analytic default starts of one marker (shared by inference and
// training-set extraction).  Returns false when neither solve is active.
// ---------------------------------------------------------------------------
struct MlpFeat {
    float f[MLP_IN];
    bool duct, rsf;           // which outputs apply
    double eta0, vhi;         // default starts
};

// Logarithm and exponential of the features: the FP64 forms of the viscosity
// solves (a few ulp, far below what BF16 inputs resolve), the library where
// those forms' range ends.  Training rows and inference share them.
__device__ __forceinline__ double feat_log(double x)
{
    const double r = log_pos(x);
    return r == r ? r : log(x);
}
__device__ __forceinline__ double feat_exp(double x)
{
    const double r = exp_red(x);
    return r == r ? r : exp(x);
}
constexpr double INV_LN10 = 0.43429448190325182765, LN10 = 2.30258509299404568402;

// DUCT = false: the ductile default start is not needed (the slip-rate solve
// asks for the slip-rate guess only) and is left 0; the features do not
// depend on it.
template <bool DUCT = true>
__device__ __forceinline__ void mlp_features(const KArgs &a, long m, double dt, MlpFeat &o)
{
    const int t = a.t[m];
    const DevMat *M = a.mat;
    const double sxx = MK(a, MF_XX, m), sxy = MK(a, MF_XY, m);
    const double exx = MK(a, MF_EXX, m), exy = MK(a, MF_EXY, m);
    const double tk = MK(a, MF_K, m), p = MK(a, MF_P, m);
    const double sig = sqrt(sxx * sxx + sxy * sxy);
    const double eii = sqrt(exx * exx + exy * exy);
    o.duct = false;
    o.rsf = false;
    o.eta0 = 0;
    o.vhi = 0;
    if (DUCT && power_law_type(M, t) && t != 1 && sig > 0) {
        const double n = M->mm[t], mpb = 1e-5 * p;
        double e1 = (M->dh[t] + M->dv[t] * mpb) / (8.314 * tk);
        if (e1 > 150.0) e1 = 150.0;
        e1 = exp(e1);
        o.eta0 = 0.5 * sig * M->nu[t] * e1 / pow(sig, n);
        o.duct = o.eta0 > 0 && isfinite(o.eta0);
    }
    double theta_l10 = 0.0, amb = 0.0, vhi_l10 = 0.0, L_l10 = 0.0, vp_l10 = 0.0;
    if (a.fault[m]) {
        const double ra = MXF(a, MX_A, m), rb = MXF(a, MX_B, m), rL = MXF(a, MX_L, m), rV0 = MXF(a, MX_V0, m);
        const double mu0 = MXF(a, MX_MU0, m), phi = MXF(a, MX_PHI, m), Vp = MXF(a, MX_V, m);
        const double nP = a.p.pmode == 0 ? a.p.nstress : p;
        const double eta_rad = a.p.inertyn == 0 ? a.p.rad_damping : 0.0;
        const double ephi = feat_exp(phi);
        double phin;
        if (Vp * dt / rL <= 1e-6) phin = feat_log(ephi * (1 - Vp * dt / rL) + rV0 * dt / rL);
        else phin = feat_log(rV0 / Vp + (ephi - rV0 / Vp) * feat_exp(-Vp * dt / rL));
        if (sig > 0 && eta_rad > 0) {
            const double x = sig / (ra * nP);
            const double ex = feat_exp(x);
            double vhi = 2 * rV0 * (0.5 * (ex - 1.0 / ex)) * feat_exp(-(mu0 + rb * phin) / ra);
            if (sig / eta_rad < vhi) vhi = sig / eta_rad;
            o.vhi = vhi;
            o.rsf = vhi > 0 && isfinite(vhi);
            if (o.rsf) vhi_l10 = feat_log(vhi) * INV_LN10;
        }
        theta_l10 = (phi + feat_log(rL / rV0)) * INV_LN10;   // theta = L e^phi / V0
        amb = ra - rb;
        L_l10 = feat_log(rL) * INV_LN10;
        vp_l10 = feat_log(Vp > 1e-40 ? Vp : 1e-40) * INV_LN10;
    }
    o.f[0] = (float)(feat_log(eii > 1e-30 ? eii : 1e-30) * INV_LN10);
    o.f[1] = (float)tk;
    o.f[2] = (float)p;
    o.f[3] = (float)(feat_log(sig > 1e-30 ? sig : 1e-30) * INV_LN10);
    o.f[4] = (float)theta_l10;
    o.f[5] = (float)amb;
    o.f[6] = (float)vhi_l10;
    o.f[7] = (float)L_l10;
    o.f[8] = (float)vp_l10;
}

// BF16 weights of one workgroup, staged in LDS in the layout the MFMA B
// operand reads (4 consecutive k per lane), and the FP32 biases.
struct MlpW {
    short w1[MLP_H][16];          // [out][in], in padded to 16
    short w2[MLP_H][MLP_H];       // [out][in]
    short w3[16][MLP_H];          // [out][in], out padded to 16
    float b1[MLP_H], b2[MLP_H], b3[MLP_OUT];
};

// All threads of the workgroup (nthreads) convert the FP32 master weights;
// the caller synchronises before use.
__device__ __forceinline__ void mlp_stage_weights(const float *__restrict__ P, MlpW &s, int tid, int nthreads)
{
    for (int e = tid; e < MLP_H * 16; e += nthreads) {
        const int o = e >> 4, k = e & 15;
        s.w1[o][k] = k < MLP_IN ? f2bf(P[P_W1 + o * MLP_IN + k]) : (short)0;
    }
    for (int e = tid; e < MLP_H * MLP_H; e += nthreads) (&s.w2[0][0])[e] = f2bf(P[P_W2 + e]);
    for (int e = tid; e < 16 * MLP_H; e += nthreads) {
        const int o = e / MLP_H;
        (&s.w3[0][0])[e] = o < MLP_OUT ? f2bf(P[P_W3 + e]) : (short)0;
    }
    for (int e = tid; e < MLP_H; e += nthreads) { s.b1[e] = P[P_B1 + e]; s.b2[e] = P[P_B2 + e]; }
    if (tid < MLP_OUT) s.b3[tid] = P[P_B3 + tid];
}

// Four consecutive BF16 values (8-byte aligned) as one LDS read.
__device__ __forceinline__ bf16x4 lds_bf16x4(const short *p) { return *(const bf16x4 *)p; }

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward pass of one wavefront's 64 rows: x[lane][0..15] holds the lane's
// standardised inputs as BF16 (zero padded); on return (y0, y1) are the lane's
// two standardised outputs.  x and h are the wavefront's own LDS rows; the
// weights must be staged.  The arithmetic is the training forward's (same
// roundings, same k order).
__device__ __forceinline__ void mlp_infer_wave(const MlpW &s, short (*x)[16], short (*h)[MLP_H], float &y0,
                                               float &y1)
{
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, kb = 4 * (lane >> 4);
    // layer 1: h = tanh(x W1^T + b1)
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++) {
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_bf16x4(&x[mt * 16 + i][kb]),
                                                          lds_bf16x4(&s.w1[nt * 16 + i][kb]), d, 0, 0, 0);
            for (int j = 0; j < 4; j++)
                h[mt * 16 + kb + j][nt * 16 + i] = f2bf(fast_tanh(d[j] + s.b1[nt * 16 + i]));
        }
    wave_lds_sync();
    // layer 2: all 16 tiles read layer 1 before any is written back
    f32x4 d2[4][4];
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++) {
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
            for (int ks = 0; ks < 4; ks++)
                d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_bf16x4(&h[mt * 16 + i][ks * 16 + kb]),
                                                              lds_bf16x4(&s.w2[nt * 16 + i][ks * 16 + kb]), d, 0, 0, 0);
            d2[mt][nt] = d;
        }
    wave_lds_sync();
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++)
            for (int j = 0; j < 4; j++)
                h[mt * 16 + kb + j][nt * 16 + i] = f2bf(fast_tanh(d2[mt][nt][j] + s.b2[nt * 16 + i]));
    wave_lds_sync();
    // layer 3: y = h W3^T + b3; row mt*16 + kb + j, output i (< 2 used), handed
    // to the lane owning that row through x (no longer needed)
    for (int mt = 0; mt < 4; mt++) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 4; ks++)
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_bf16x4(&h[mt * 16 + i][ks * 16 + kb]),
                                                          lds_bf16x4(&s.w3[i][ks * 16 + kb]), d, 0, 0, 0);
        for (int j = 0; j < 4; j++)
            if (i < MLP_OUT) *(float *)&x[mt * 16 + kb + j][2 * i] = d[j] + s.b3[i];
    }
    wave_lds_sync();
    y0 = *(const float *)&x[lane][0];
    y1 = *(const float *)&x[lane][2];
}

// Standardised BF16 inputs of one lane into its LDS row.
__device__ __forceinline__ void mlp_load_row(const MlpFeat &ft, const float *__restrict__ norm, bool live,
                                             short *row)
{
    for (int k = 0; k < 16; k++) {
        float v = 0.f;
        if (k < MLP_IN && live) v = (ft.f[k] - norm[k]) / norm[MLP_IN + k];
        row[k] = f2bf(v);
    }
}

// Inference: standardised features -> guesses written to MX_GETA / MX_GV
// (0 where a solve does not apply, which the local update treats as "no
// guess").  Four wavefronts per block share the weights staged in LDS; each
// wavefront keeps its 64 rows of activations as BF16 in LDS.
constexpr int MI_WAVES = 4;
struct MlpInferLds {
    MlpW w;
    short x[MI_WAVES][64][16];
    short h[MI_WAVES][64][MLP_H];
};

// outputs: bit 0 the ductile guess, bit 1 the slip-rate guess; the other is
// written as 0 (default start).  Markers that can need a selected guess are
// found first from the type and fault flag alone: a workgroup with none
// writes zeros and stops before the features and the weights, and a
// wavefront with none stops after the weights are staged.
__global__ __launch_bounds__(64 * MI_WAVES) void k_mlp_guess(KArgs a, double dt, const float *__restrict__ P,
                                                             const float *__restrict__ norm, int outputs)
{
    __shared__ MlpInferLds s;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const long m = ((long)blockIdx.x * MI_WAVES + w) * 64 + lane;
    bool need = false;
    if (m < a.n) {
        const int t = a.t[m];
        need = ((outputs & 2) && a.fault[m]) || ((outputs & 1) && power_law_type(a.mat, t) && t != 1);
    }
    if (!__syncthreads_or(need)) {
        if (m < a.n) {
            MXF(a, MX_GETA, m) = 0.0;
            MXF(a, MX_GV, m) = 0.0;
        }
        return;
    }
    const bool wave_need = __ballot(need) != 0;
    MlpFeat ft;
    ft.duct = ft.rsf = false;
    if (m < a.n && wave_need) mlp_features(a, m, dt, ft);
    ft.duct = ft.duct && (outputs & 1);
    ft.rsf = ft.rsf && (outputs & 2);
    mlp_stage_weights(P, s.w, tid, 64 * MI_WAVES);
    mlp_load_row(ft, norm, m < a.n, s.x[w][lane]);
    __syncthreads();
    if (!wave_need) {
        if (m < a.n) {
            MXF(a, MX_GETA, m) = 0.0;
            MXF(a, MX_GV, m) = 0.0;
        }
        return;
    }
    float y0, y1;
    mlp_infer_wave(s.w, s.x[w], s.h[w], y0, y1);
    if (m < a.n) {
        const double r0 = y0 * norm[2 * MLP_IN + 2] + norm[2 * MLP_IN];
        const double r1 = y1 * norm[2 * MLP_IN + 3] + norm[2 * MLP_IN + 1];
        MXF(a, MX_GETA, m) = ft.duct ? ft.eta0 * pow(10.0, r0) : 0.0;
        MXF(a, MX_GV, m) = ft.rsf ? ft.vhi * pow(10.0, r1) : 0.0;
    }
}

// Training-set extraction.  Before a plain local update: features and the
// log10 default starts into X[n][MLP_IN + 2] = f0..f8, log10 eta0, log10 vhi,
// with the masks of the two outputs in Mask[n][2].  After it: k_mlp_labels
// turns the last two columns into log10(root / default).
constexpr int MLP_ROW = MLP_IN + 2;
__global__ void k_mlp_features(KArgs a, double dt, float *X, float *Mask)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n) return;
    MlpFeat ft;
    mlp_features(a, m, dt, ft);
    for (int k = 0; k < MLP_IN; k++) X[m * MLP_ROW + k] = ft.f[k];
    X[m * MLP_ROW + MLP_IN] = ft.duct ? (float)log10(ft.eta0) : 0.f;
    X[m * MLP_ROW + MLP_IN + 1] = ft.rsf ? (float)log10(ft.vhi) : 0.f;
    Mask[m * 2 + 0] = ft.duct ? 1.f : 0.f;
    Mask[m * 2 + 1] = ft.rsf ? 1.f : 0.f;
}

__global__ void k_mlp_labels(KArgs a, float *X, float *Mask)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n) return;
    const double re = MXF(a, MX_ROOT_ETA, m), rv = MXF(a, MX_ROOT_V, m);
    float *r = X + m * MLP_ROW + MLP_IN;
    if (Mask[m * 2 + 0] != 0.f && re > 0 && isfinite(re)) r[0] = (float)(log10(re) - (double)r[0]);
    else { Mask[m * 2 + 0] = 0.f; r[0] = 0.f; }
    if (Mask[m * 2 + 1] != 0.f && rv > 0 && isfinite(rv)) r[1] = (float)(log10(rv) - (double)r[1]);
    else { Mask[m * 2 + 1] = 0.f; r[1] = 0.f; }
}

// One training step, one wavefront per 64 samples: forward, masked MSE,
// backward; writes this wave's gradient (sum over its rows) to
// G[blockIdx.x][MLP_NPARAM] and its loss sum to L[blockIdx.x].
// Xs: standardised inputs [nb][6]; Ts: standardised targets [nb][2];
// Ms: masks [nb][2]; nb = rows of this batch; inv_cnt = 1 / (#masked terms).
__global__ __launch_bounds__(64) void k_mlp_train(const float *__restrict__ P, const float *__restrict__ Xs,
                                                  const float *__restrict__ Ts, const float *__restrict__ Ms,
                                                  const uint32_t *__restrict__ idx, long nb, float inv_cnt,
                                                  float *G, float *L)
{
    __shared__ MlpLds s;
    __shared__ float msk[64][2];
    const int lane = threadIdx.x;
    const long rr = (long)blockIdx.x * 64 + lane;
    const long r = rr < nb ? (long)idx[rr] : 0;
    const bool live = rr < nb;
    for (int k = 0; k < 16; k++) s.x[lane][k] = (k < MLP_IN && live) ? f2bf(Xs[r * MLP_IN + k]) : (short)0;
    __syncthreads();
    mlp_forward(P, s);
    // dY = 2 m (y - t) / cnt
    float loss = 0.f;
    for (int o = 0; o < MLP_OUT; o++) {
        const float mk = live ? Ms[r * 2 + o] : 0.f;
        const float t = live ? Ts[r * 2 + o] : 0.f;
        const float e = s.y[lane][o] - t;
        loss += mk * e * e;
        msk[lane][o] = mk;
        s.y[lane][o] = 2.f * mk * e * inv_cnt;
    }
    __syncthreads();
    float *g = G + (long)blockIdx.x * MLP_NPARAM;
    const int n = lane & 15, ib = 4 * (lane >> 4);
    // dW3[o][k] = sum_r dY[r][o] h2[r][k]  (M = o padded to 16, K = rows)
    for (int nt = 0; nt < 4; nt++) {
        f32x4 d = mfma_tile(4, [&](int i, int k) { return i < MLP_OUT ? f2bf(s.y[k][i]) : (short)0; },
                            [&](int k, int c) { return f2bf(s.h2[k][nt * 16 + c]); });
        for (int j = 0; j < 4; j++)
            if (ib + j < MLP_OUT) g[P_W3 + (ib + j) * MLP_H + nt * 16 + n] = d[j];
    }
    if (lane < MLP_OUT) {
        float b = 0.f;
        for (int q = 0; q < 64; q++) b += s.y[q][lane];
        g[P_B3 + lane] = b;
    }
    // dZ2 = (dY W3) * (1 - h2^2)   (K = outputs padded to 16)
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++) {
            f32x4 d = mfma_tile(1, [&](int i, int k) { return k < MLP_OUT ? f2bf(s.y[mt * 16 + i][k]) : (short)0; },
                                [&](int k, int c) { return k < MLP_OUT ? f2bf(P[P_W3 + k * MLP_H + nt * 16 + c]) : (short)0; });
            for (int j = 0; j < 4; j++) {
                const float h = s.h2[mt * 16 + ib + j][nt * 16 + n];
                s.g[mt * 16 + ib + j][nt * 16 + n] = d[j] * (1.f - h * h);
            }
        }
    __syncthreads();
    // dW2[o][k] = sum_r dZ2[r][o] h1[r][k]
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++) {
            f32x4 d = mfma_tile(4, [&](int i, int k) { return f2bf(s.g[k][mt * 16 + i]); },
                                [&](int k, int c) { return f2bf(s.h1[k][nt * 16 + c]); });
            for (int j = 0; j < 4; j++) g[P_W2 + (mt * 16 + ib + j) * MLP_H + nt * 16 + n] = d[j];
        }
    {
        float b = 0.f;
        for (int q = 0; q < 64; q++) b += s.g[q][lane];
        g[P_B2 + lane] = b;
    }
    // dZ1 = (dZ2 W2) * (1 - h1^2); reuse h2 as the staging buffer
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++) {
            f32x4 d = mfma_tile(4, [&](int i, int k) { return f2bf(s.g[mt * 16 + i][k]); },
                                [&](int k, int c) { return f2bf(P[P_W2 + k * MLP_H + nt * 16 + c]); });
            for (int j = 0; j < 4; j++) {
                const float h = s.h1[mt * 16 + ib + j][nt * 16 + n];
                s.h2[mt * 16 + ib + j][nt * 16 + n] = d[j] * (1.f - h * h);
            }
        }
    __syncthreads();
    // dW1[o][k] = sum_r dZ1[r][o] x[r][k]   (N = inputs padded to 16)
    for (int mt = 0; mt < 4; mt++) {
        f32x4 d = mfma_tile(4, [&](int i, int k) { return f2bf(s.h2[k][mt * 16 + i]); },
                            [&](int k, int c) { return s.x[k][c]; });
        if (n < MLP_IN)
            for (int j = 0; j < 4; j++) g[P_W1 + (mt * 16 + ib + j) * MLP_IN + n] = d[j];
    }
    {
        float b = 0.f;
        for (int q = 0; q < 64; q++) b += s.h2[q][lane];
        g[P_B1 + lane] = b;
    }
    // loss sum of the wave, fixed order
    __syncthreads();
    s.g[lane][0] = loss;
    __syncthreads();
    if (lane == 0) {
        float t = 0.f;
        for (int q = 0; q < 64; q++) t += s.g[q][0];
        L[blockIdx.x] = t;
    }
}

// Sum the per-wave gradients in wave order and take an Adam step on the FP32
// master weights.  One thread per parameter.
__global__ void k_mlp_adam(float *P, float *Mom, float *Var, const float *G, int nwaves, float lr, float b1,
                           float b2, float eps, float bc1, float bc2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    float g = 0.f;
    for (int w = 0; w < nwaves; w++) g += G[(long)w * MLP_NPARAM + p];
    const float m = b1 * Mom[p] + (1.f - b1) * g;
    const float v = b2 * Var[p] + (1.f - b2) * g * g;
    Mom[p] = m;
    Var[p] = v;
    P[p] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

// Gradient of one batch without an update (for the gradient check).
__global__ void k_mlp_gradsum(const float *G, int nwaves, float *out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    float g = 0.f;
    for (int w = 0; w < nwaves; w++) g += G[(long)w * MLP_NPARAM + p];
    out[p] = g;
}

// Forward on standardised inputs Xs[nb][6] -> Ys[nb][2] (for tests/evaluation).
__global__ __launch_bounds__(64) void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ Xs,
                                                    long nb, float *Ys)
{
    __shared__ MlpLds s;
    const int lane = threadIdx.x;
    const long r = (long)blockIdx.x * 64 + lane;
    for (int k = 0; k < 16; k++) s.x[lane][k] = (k < MLP_IN && r < nb) ? f2bf(Xs[r * MLP_IN + k]) : (short)0;
    __syncthreads();
    mlp_forward(P, s);
    if (r < nb) {
        Ys[r * 2 + 0] = s.y[lane][0];
        Ys[r * 2 + 1] = s.y[lane][1];
    }
}

}  // namespace v2c
