// Accuracy of quot(), log_pos() and exp_red() (velo2cycles_amd/csrc/v2c_vmath.hpp)
// against long-double references, on the host (IEEE reciprocal) or on the
// GPU (hardware reciprocal estimate refined), over seeded arguments:
// log-uniform over the whole supported range, the neighbourhood of 1, the
// reduction boundary sqrt(1/2) and its neighbours, powers of two, and the
// viscosity-like magnitudes the solves use.
//
//   vmath_check host|device [n]
// prints "log_max_ulp <u> log_mean_ulp <u> quot_max_ulp <u> quot_exact <frac>
// exp_max_ulp <u> exp_mean_ulp <u> range_violations <k> n <n>"
// (ulp of the long-double result rounded to double; range_violations counts
// out-of-range arguments that did not give NaN); exit 0 on success.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <cmath>
#include <vector>

#include "../../velo2cycles_amd/csrc/v2c_vmath.hpp"

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static uint64_t next_u64()
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return rng_state;
}
static double uniform01() { return (double)(next_u64() >> 11) * 0x1p-53; }

// Arguments outside the fast ranges, which must give NaN (first in the arrays).
static const double BAD_X[] = {0.0, -1.0, INFINITY, NAN, 1e-310, 0x1p1021, 0x1p-1021};
static const double BAD_U[] = {708.0, -708.0, 1000.0, -INFINITY, INFINITY, NAN};
constexpr size_t NBAD_X = sizeof(BAD_X) / sizeof(BAD_X[0]), NBAD_U = sizeof(BAD_U) / sizeof(BAD_U[0]);

static std::vector<double> log_args(size_t n)
{
    std::vector<double> x(BAD_X, BAD_X + NBAD_X);
    x.reserve(n);
    const double r = sqrt(0.5);
    for (int k = -40; k <= 40; k++) {              // k ulps around the reduction boundary and 1
        double v = r;
        for (int j = 0; j < abs(k); j++) v = k < 0 ? nextafter(v, 0.0) : nextafter(v, 2.0);
        x.push_back(v);
        double w = 1.0;
        for (int j = 0; j < abs(k); j++) w = k < 0 ? nextafter(w, 0.0) : nextafter(w, 2.0);
        x.push_back(w);
    }
    for (int e = -1019; e <= 1019; e++) x.push_back(ldexp(1.0, e));
    while (x.size() < n) {
        const int kind = (int)(next_u64() % 4);
        double v;
        if (kind == 0) v = exp2(-1019.0 + 2038.0 * uniform01());       // whole range
        else if (kind == 1) v = 1.0 + (uniform01() - 0.5) * 1e-3;      // near 1
        else if (kind == 2) v = pow(10.0, 14.0 + 18.0 * uniform01());  // viscosity-like
        else v = 0.5 + uniform01();                                    // [0.5, 1.5)
        x.push_back(v);
    }
    x.resize(n);
    return x;
}

__global__ void k_vmath(const double *x, const double *a, const double *b, const double *u, double *lg, double *q,
                        double *ex, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        lg[i] = v2c::log_pos(x[i]);
        q[i] = v2c::quot(a[i], b[i]);
        ex[i] = v2c::exp_red(u[i]);
    }
}

static double ulp_err(double got, long double ref)
{
    const double r = (double)ref;
    if (got == r) return 0.0;
    const double u = fabs(nextafter(r, INFINITY) - r);
    return (double)(fabsl((long double)got - ref) / u);
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: vmath_check host|device [n]\n");
        return 2;
    }
    const bool dev = strcmp(argv[1], "device") == 0;
    const size_t n = argc > 2 ? (size_t)atol(argv[2]) : 2000000;
    std::vector<double> x = log_args(n), a(n), b(n), u(n), lg(n), q(n), ex(n);
    for (size_t i = 0; i < n; i++) {
        const int kind = (int)(next_u64() % 3);
        if (kind == 0) {
            a[i] = pow(10.0, -10.0 + 40.0 * uniform01());
            b[i] = pow(10.0, -10.0 + 40.0 * uniform01());
        } else if (kind == 1) {     // the iteration's shapes: 1/eta, eta/eta, residual/slope
            a[i] = 1.0;
            b[i] = pow(10.0, 16.0 + 14.0 * uniform01());
        } else {
            a[i] = (uniform01() - 0.5) * pow(10.0, 18.0 + 6.0 * uniform01());
            b[i] = 0.5 + uniform01();
        }
        // exponents: the whole fast range, the reduction's half-integer
        // multiples of ln 2, and small arguments
        const int ek = (int)(next_u64() % 3);
        if (ek == 0) u[i] = (uniform01() - 0.5) * 1415.0;
        else if (ek == 1) u[i] = ((double)(int)(next_u64() % 2000) - 1000.5) * 0.6931471805599453 + (uniform01() - 0.5) * 1e-9;
        else u[i] = (uniform01() - 0.5) * 1e-3;
    }
    for (size_t i = 0; i < NBAD_U; i++) u[i] = BAD_U[i];
    if (dev) {
        double *d[7];
        for (int k = 0; k < 7; k++)
            if (hipMalloc(&d[k], n * sizeof(double)) != hipSuccess) { fprintf(stderr, "hipMalloc failed\n"); return 3; }
        if (hipMemcpy(d[0], x.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d[1], a.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d[2], b.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d[3], u.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess) return 3;
        hipLaunchKernelGGL(k_vmath, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, d[0], d[1], d[2], d[3], d[4],
                           d[5], d[6], (long)n);
        if (hipDeviceSynchronize() != hipSuccess) { fprintf(stderr, "kernel failed\n"); return 3; }
        if (hipMemcpy(lg.data(), d[4], n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(q.data(), d[5], n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(ex.data(), d[6], n * 8, hipMemcpyDeviceToHost) != hipSuccess) return 3;
        for (int k = 0; k < 7; k++) (void)hipFree(d[k]);
    } else {
        for (size_t i = 0; i < n; i++) {
            lg[i] = v2c::log_pos(x[i]);
            q[i] = v2c::quot(a[i], b[i]);
            ex[i] = v2c::exp_red(u[i]);
        }
    }
    int bad = 0;
    for (size_t i = 0; i < NBAD_X; i++)
        if (!std::isnan(lg[i])) bad++;
    for (size_t i = 0; i < NBAD_U; i++)
        if (!std::isnan(ex[i])) bad++;
    double lmax = 0, lsum = 0, qmax = 0, xmax = 0, xsum = 0;
    size_t qexact = 0;
    for (size_t i = 0; i < n; i++) {
        const double el = i < NBAD_X ? 0.0 : ulp_err(lg[i], logl((long double)x[i]));
        const double eq = ulp_err(q[i], (long double)a[i] / (long double)b[i]);
        if (el > lmax) lmax = el;
        lsum += el;
        if (eq > qmax) qmax = eq;
        if (q[i] == a[i] / b[i]) qexact++;
        const double exr = i < NBAD_U ? 0.0 : ulp_err(ex[i], expl((long double)u[i]));
        if (exr > xmax) xmax = exr;
        xsum += exr;
    }
    printf("log_max_ulp %.4f log_mean_ulp %.4f quot_max_ulp %.4f quot_exact %.8f exp_max_ulp %.4f exp_mean_ulp %.4f "
           "range_violations %d n %zu\n", lmax, lsum / n, qmax, (double)qexact / n, xmax, xsum / n, bad, n);
    return 0;
}
