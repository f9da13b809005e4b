/*
 * v2c_oracle.c -- FP64 CPU restatement of the VELO2CYCLEs per-marker hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Each routine names the reference lines it restates.  Operation order inside
 * every expression follows the reference so that a single-threaded reference
 * run and this file agree bit for bit when both are built without FMA
 * contraction.  Pinned by: tests/test_oracle_golden.py (golden vectors written
 * by the in-place reference build), tests/test_oracle_math.py (closed forms,
 * invariants and brute force).
 */
#include "v2c_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NODAL(s, f) ((s)->nod[V2C_N_##f])

/* Ternary MIN of headevps.c:3: returns b when a is NaN. */
static inline double ref_minv(double a, double b) { return a <= b ? a : b; }
static inline double ref_abs(double a) { return a >= 0 ? a : -a; }

/* ---- cell search (interpolation.c:1536-1573) -------------------------------
 * Bisection for the gridline left of x; clamped to the last cell.  The loop is
 * kept as a bisection (not a formula) because its NaN behaviour (NaN lands in
 * the last cell) is part of the reference result. */
static long bisect_left(const double *g, long n, double x)
{
    long lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        long mid = (lo + hi) / 2;
        if (g[mid] > x) hi = mid; else lo = mid;
    }
    return lo > n - 2 ? n - 2 : lo;
}
long or_nxsearch(const v2c_grid *g, double x) { return bisect_left(g->gx, g->xnumx, x); }
long or_nysearch(const v2c_grid *g, double y) { return bisect_left(g->gy, g->ynumy, y); }
/* nxsearch / nysearch of n points (for array checks of cell indices). */
void or_cells(const v2c_grid *g, long n, const double *x, const double *y, long *nx, long *ny)
{
    for (long i = 0; i < n; i++) {
        nx[i] = or_nxsearch(g, x[i]);
        ny[i] = or_nysearch(g, y[i]);
    }
}

static inline double clampd(double v, double lo, double hi)
{
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

/* Bilinear weights of a point inside a pressure-node (cell-centre) stencil.
 * Shared by allinterpomp (interpolation.c:1942-1948), the centre part of
 * interpolate_mechanical (:1212-1224) and allinteri (:1723-1730). */
static void centre_stencil(const v2c_grid *g, long nx, long ny, double x, double y,
                           long *m3, double *e, double *n)
{
    const double *gx = g->gx, *gy = g->gy;
    long m10 = nx, m20 = ny;
    if (x > (gx[m10] + gx[m10 + 1]) / 2.0) m10++;
    if (y > (gy[m20] + gy[m20 + 1]) / 2.0) m20++;
    if (m10 < 1) m10 = 1;
    if (m10 > g->xnumx - 2) m10 = g->xnumx - 2;
    if (m20 < 1) m20 = 1;
    if (m20 > g->ynumy - 2) m20 = g->ynumy - 2;
    *e = (x - (gx[m10 - 1] + gx[m10]) / 2.0) / ((gx[m10 + 1] - gx[m10 - 1]) / 2.0);
    *n = (y - (gy[m20 - 1] + gy[m20]) / 2.0) / ((gy[m20 + 1] - gy[m20 - 1]) / 2.0);
    *m3 = m10 * g->ynumy + m20;
}

/* Basic-node stencil with ghost exclusion (interpolation.c:1069-1083, 1711-1717). */
static void basic_stencil(const v2c_grid *g, long nx, long ny, double x, double y,
                          long *m3, double *e, double *n)
{
    long m10 = nx, m20 = ny;
    if (m10 < 1) m10 = 1;
    if (m10 > g->xnumx - 3) m10 = g->xnumx - 3;
    if (m20 < 1) m20 = 1;
    if (m20 > g->ynumy - 3) m20 = g->ynumy - 3;
    *e = (x - g->gx[m10]) / (g->gx[m10 + 1] - g->gx[m10]);
    *n = (y - g->gy[m20]) / (g->gy[m20 + 1] - g->gy[m20]);
    *m3 = m10 * g->ynumy + m20;
}

/* vx-node stencil (interpolation.c:1377-1395 / 1738-1745); lowx clamps m10
 * to [0, xnumx-2] as allinteri does, interpolate_mechanical leaves it. */
static void vx_stencil(const v2c_grid *g, long nx, long ny, double x, double y,
                       int clampx, long *m3, double *e, double *n)
{
    const double *gx = g->gx, *gy = g->gy;
    long m10 = nx, m20 = ny;
    if (y < (gy[m20] + gy[m20 + 1]) / 2.0) m20 -= 1;
    if (clampx) {
        if (m10 < 0) m10 = 0;
        if (m10 > g->xnumx - 2) m10 = g->xnumx - 2;
    }
    if (m20 < 0) m20 = 0;
    if (m20 > g->ynumy - 3) m20 = g->ynumy - 3;
    *e = (x - gx[m10]) / (gx[m10 + 1] - gx[m10]);
    *n = (y - (gy[m20] + gy[m20 + 1]) / 2.0) / ((gy[m20 + 2] - gy[m20]) / 2.0);
    *m3 = m10 * g->ynumy + m20;
}

/* vy-node stencil (interpolation.c:1434-1452 / 1749-1756). */
static void vy_stencil(const v2c_grid *g, long nx, long ny, double x, double y,
                       int clampy, long *m3, double *e, double *n)
{
    const double *gx = g->gx, *gy = g->gy;
    long m10 = nx, m20 = ny;
    if (x < (gx[m10] + gx[m10 + 1]) / 2.0) m10 -= 1;
    if (m10 < 0) m10 = 0;
    if (m10 > g->xnumx - 3) m10 = g->xnumx - 3;
    if (clampy) {
        if (m20 < 0) m20 = 0;
        if (m20 > g->ynumy - 2) m20 = g->ynumy - 2;
    }
    *e = (x - (gx[m10] + gx[m10 + 1]) / 2.0) / ((gx[m10 + 2] - gx[m10]) / 2.0);
    *n = (y - gy[m20]) / (gy[m20 + 1] - gy[m20]);
    *m3 = m10 * g->ynumy + m20;
}

/* 4-point bilinear sum in the reference's term order. */
static inline double bil(const double *f, long m3, long ny, double e, double n)
{
    double d0 = (1.0 - e) * (1.0 - n), d1 = (1.0 - e) * n, d2 = e * (1.0 - n), d3 = e * n;
    return d0 * f[m3] + d1 * f[m3 + 1] + d2 * f[m3 + ny] + d3 * f[m3 + ny + 1];
}
/* allinteri writes its products as ((1-e)*(1-n))*f, identical to bil(). */

/* ---- allinterpomp (interpolation.c:1923-1958): P at a point. ------------- */
double or_allinterpomp(const or_state *s, double x, double y)
{
    const v2c_grid *g = &s->grid;
    long nx = or_nxsearch(g, x), ny = or_nysearch(g, y);
    x = clampd(x, 0, g->xsize);
    y = clampd(y, 0, g->ysize);
    long m3; double e, n;
    centre_stencil(g, nx, ny, x, y, &m3, &e, &n);
    return bil(NODAL(s, pr), m3, g->ynumy, e, n);
}

/* ---- dencalcomp (density_ductilevisc_calc.c:2-27) ------------------------- */
static void dencalc(const or_state *s, double mtk, double mpb, int t,
                    double *mro, double *mbb, double *maa)
{
    const v2c_materials *M = s->mat;
    *mbb = M->bb[t] / (1.0 - M->bb[t] * (mtk - 298.15));
    *maa = 1.e-8 * M->aa[t] / (1.0 + M->aa[t] * (mpb - 1.0) * 1e-3);
    if (s->prm.densimod == 0) *mro = M->ro[t];
    else *mro = M->ro[t] * (1.0 - M->bb[t] * (mtk - 298.15)) * (1.0 + M->aa[t] * (mpb - 1.0) * 1e-3);
}

/* ---- conductivity (recalc_marker_properties.c:51-54) ---------------------- */
static double conductivity(const v2c_materials *M, int t, double mtk, double mpb)
{
    double k = (M->kt[t] + M->kf[t] / (mtk + 77.0)) * exp(M->kp[t] * mpb);
    if (k < 0) k = -k / (1.0 + M->kf[t] * (mtk - M->kp[t]) / M->kp[t]);
    return k;
}

/* ---- ductilevisc (density_ductilevisc_calc.c:29-223) ----------------------
 * Three branches: sticky air (type 1), Newtonian, non-linear diffusion +
 * dislocation creep solved by the reference's quasi-Newton iteration
 * (:157-203; its derivative is not the residual's, and it stops on
 * |log10(mu_k/mu_{k-1})| < 1e-6 after more than 3 iterations).  The
 * uninitialised-return branch (:55-65, types with dh=dv=0, ss!=0, n!=1) yields
 * NaN here. */
static int power_law_type(const v2c_materials *M, int t)
{
    return M->nu[t] > 0 && (M->dh[t] != 0 || M->dv[t] != 0) && M->ss[t] != 0 && M->mm[t] != 1.0;
}

static double clamp_visc(const or_state *s, int t, double v)
{
    const v2c_materials *M = s->mat;
    if (v < s->prm.nubeg) v = s->prm.nubeg;
    if (v > s->prm.nuend) v = s->prm.nuend;
    if (v < M->n0[t]) v = M->n0[t];
    if (v > M->n1[t]) v = M->n1[t];
    return v;
}

double or_ductilevisc(const or_state *s, int t, double sxy, double sxx,
                      double exy, double exx, double mtk, double mpb, int *iters)
{
    const v2c_materials *M = s->mat;
    (void)exy; (void)exx;   /* strain rates are accepted but unused (:114) */
    if (iters) *iters = 0;
    if (t == 1) {
        double v = M->nu[1];
        if (s->prm.stickymod == 1 && s->timesum > 4e7) {
            double cap = M->gg[2] * s->timestep / 100;
            if (v > cap) v = cap;
        }
        return v;
    }
    if (M->dh[t] == 0 && M->dv[t] == 0 && (M->ss[t] == 0 || M->mm[t] == 1.0))
        return M->nu[t];
    if (!power_law_type(M, t))
        return NAN;

    double n = M->mm[t];
    double sigin = pow(sxx * sxx + sxy * sxy, 0.5);
    double e1 = (M->dh[t] + M->dv[t] * mpb) / (8.314 * mtk);
    if (e1 > 150.0) e1 = 150.0;
    e1 = exp(e1);
    double k1 = M->nu[t] / pow(M->ss[t], n - 1.0);
    double inv_newt = 1.0 / (0.5 * k1 * e1);
    double inv_powl = 0;
    if (sigin > 0) inv_powl = 1.0 / (0.5 * sigin * M->nu[t] * e1 / pow(sigin, n));
    if (sigin > 0) {
        double eta_disl = 0.5 * sigin * M->nu[t] * e1 / pow(sigin, n);
        double eta_new = eta_disl;
        double eta_diff = 0.5 * k1 * e1;
        double coeff = pow(eta_disl, 1.0 / n);
        double gdt = M->gg[t] * s->timestep;
        double mu_prev = 0, mu = 0;
        int k = 0;
        for (;;) {
            if (k >= 50) break;
            k++;
            double eta_old = eta_new;
            mu = 1.0 / (1.0 / eta_diff + 1.0 / eta_new);
            double pref = (eta_diff + eta_old) / eta_diff;
            double f1 = pow(pref * (gdt + mu), (n - 1) / n);
            double f2 = eta_old - coeff * f1;
            double df = 1 - (n - 1) / (n * eta_diff) * coeff * pow(pref * (gdt + mu), -1 / n)
                        * ((gdt + mu) + pow(eta_diff, 2 / (eta_diff + eta_old)));
            eta_new = eta_old - f2 / df;
            if (k > 3 && fabs(log10(mu / mu_prev)) < 1.0e-6) break;
            mu_prev = mu;
        }
        if (iters) *iters = k;
        inv_powl = 1 / eta_new;
    }
    return clamp_visc(s, t, 1.0 / (inv_newt + inv_powl));
}

/* Build-defined NEWTON mode (not reference behaviour): the residual of the
 * reference iteration, f(eta) = eta - coeff*X^((n-1)/n) with
 * X = (eta_d+eta)/eta_d*G dt + eta (density_ductilevisc_calc.c:188-189),
 * solved with its exact derivative
 * f'(eta) = 1 - (n-1)/n*coeff*X^(-1/n)*(G dt/eta_d + 1) to |f|/eta <= tol.
 * Start eta_disl0 (:159); with a guess > 0 the residual is evaluated at both
 * starts and Newton starts from the smaller; a guessed start whose residual
 * grows after its first step falls back to eta_disl0.  *iters counts Newton
 * updates, *evals residual evaluations, *fallback the fall-back; *root is the
 * converged unknown.  Returns the clamped effective viscosity. */
static double resid_duct(double eta, double eta_diff, double gdt, double coeff, double p1,
                         double *f, double *X, double *Xp)
{
    *X = (eta_diff + eta) / eta_diff * gdt + eta;
    *Xp = pow(*X, p1);
    *f = eta - coeff * *Xp;
    return fabs(*f) / eta;
}

double or_ductile_newton_ex(const or_state *s, int t, double sxy, double sxx, double mtk, double mpb,
                            double guess, double tol, int *iters, int *evals, int *fallback, double *root)
{
    const v2c_materials *M = s->mat;
    *iters = 0; *evals = 0; *fallback = 0; *root = 0;
    if (!power_law_type(M, t) || t == 1)
        return or_ductilevisc(s, t, sxy, sxx, 0, 0, mtk, mpb, NULL);
    double n = M->mm[t];
    double sigin = pow(sxx * sxx + sxy * sxy, 0.5);
    double e1 = (M->dh[t] + M->dv[t] * mpb) / (8.314 * mtk);
    if (e1 > 150.0) e1 = 150.0;
    e1 = exp(e1);
    double k1 = M->nu[t] / pow(M->ss[t], n - 1.0);
    double eta_diff = 0.5 * k1 * e1;
    if (!(sigin > 0)) return clamp_visc(s, t, eta_diff);
    double eta_disl0 = 0.5 * sigin * M->nu[t] * e1 / pow(sigin, n);
    double coeff = pow(eta_disl0, 1.0 / n);
    double gdt = M->gg[t] * s->timestep;
    double p1 = (n - 1) / n, slope = gdt / eta_diff + 1;
    double eta = eta_disl0, f, X, Xp;
    double r = resid_duct(eta, eta_diff, gdt, coeff, p1, &f, &X, &Xp);
    int ev = 1, k = 0, from_guess = 0;
    if (r > tol && guess > 0) {
        double fg, Xg, Xpg;
        double rg = resid_duct(guess, eta_diff, gdt, coeff, p1, &fg, &Xg, &Xpg);
        ev++;
        if (rg < r) { eta = guess; f = fg; X = Xg; Xp = Xpg; r = rg; from_guess = 1; }
    }
    double rprev = r;
    while (r > tol && k < 200) {
        /* X^(-1/n) evaluated as X^((n-1)/n) / X */
        double df = 1 - p1 * coeff * (Xp / X) * slope;
        double next;
        if (f < 0 && !(df > 0)) next = 2.0 * eta;
        else next = eta - f / df;
        if (!(next > 0)) next = 0.5 * eta;
        eta = next;
        k++;
        r = resid_duct(eta, eta_diff, gdt, coeff, p1, &f, &X, &Xp);
        ev++;
        if (from_guess && k == 1 && r > rprev) {
            from_guess = 0;
            *fallback = 1;
            eta = eta_disl0;
            r = resid_duct(eta, eta_diff, gdt, coeff, p1, &f, &X, &Xp);
            ev++;
        }
        rprev = r;
    }
    *iters = k; *evals = ev; *root = eta;
    return clamp_visc(s, t, 1.0 / (1.0 / eta_diff + 1.0 / eta));
}

double or_ductilevisc_newton(const or_state *s, int t, double sxy, double sxx, double mtk, double mpb,
                             double guess, double tol, int *iters, double *resid)
{
    int ev, fb;
    double root;
    double v = or_ductile_newton_ex(s, t, sxy, sxx, mtk, mpb, guess, tol, iters, &ev, &fb, &root);
    if (resid) {
        /* residual at the returned root */
        *resid = 0;
        const v2c_materials *M = s->mat;
        if (root > 0 && power_law_type(M, t)) {
            double n = M->mm[t], sigin = pow(sxx * sxx + sxy * sxy, 0.5);
            double e1 = (M->dh[t] + M->dv[t] * mpb) / (8.314 * mtk);
            if (e1 > 150.0) e1 = 150.0;
            e1 = exp(e1);
            double eta_diff = 0.5 * (M->nu[t] / pow(M->ss[t], n - 1.0)) * e1;
            double coeff = pow(0.5 * sigin * M->nu[t] * e1 / pow(sigin, n), 1.0 / n);
            double f, X, Xp;
            *resid = resid_duct(root, eta_diff, M->gg[t] * s->timestep, coeff, (n - 1) / n, &f, &X, &Xp);
        }
    }
    return v;
}

/* Build-defined per-marker RSF update (not reference behaviour).  With the
 * state advanced by the aging law over dt at the previous slip rate
 * (calc_RSF.c:96-103), solve the regularised RSF yield condition with
 * radiation damping for the slip rate,
 *   g(s) = nP a asinh(e^s K) + eta_rad e^s - sigII = 0,  K = exp((mu0+b phi')/a)/(2 V0),
 * by Newton in s = ln V from s_hi = ln min(closed form of calc_RSF.c:192,
 * sigII/eta_rad) (g is increasing and convex in s, so Newton from s_hi
 * converges monotonically); guesses as in or_ductile_newton_ex.  Without
 * damping the root is the closed form.  The result is capped at EII*W and
 * floored at 1e-40 like calc_RSF.c:193-196.  Returns the Newton update count. */
static double gfun_rsf(double s, double K, double nP, double a, double eta_rad, double sig, double *g, double *dg)
{
    double e = exp(s), z = e * K;
    *g = nP * a * asinh(z) + eta_rad * e - sig;
    *dg = nP * a * z / sqrt(1 + z * z) + eta_rad * e;
    return fabs(*g) / sig;
}

int or_rsf_marker(double sig, double nP, double a, double b, double L, double V0, double mu0,
                  double phi, double Vprev, double dt, double eta_rad, double EII, double W,
                  double guessV, double tol, double *V_out, double *phi_out, int *fallback)
{
    int evals;
    double root;
    return or_rsf_marker_ex(sig, nP, a, b, L, V0, mu0, phi, Vprev, dt, eta_rad, EII, W, guessV, tol,
                            V_out, phi_out, fallback, &root, &evals);
}

int or_rsf_marker_ex(double sig, double nP, double a, double b, double L, double V0, double mu0,
                     double phi, double Vprev, double dt, double eta_rad, double EII, double W,
                     double guessV, double tol, double *V_out, double *phi_out, int *fallback,
                     double *root, int *evals)
{
    double phin;
    if (Vprev * dt / L <= 1e-6) phin = log(exp(phi) * (1 - Vprev * dt / L) + V0 * dt / L);
    else phin = log(V0 / Vprev + (exp(phi) - V0 / Vprev) * exp(-Vprev * dt / L));
    *phi_out = phin;
    *fallback = 0;
    *root = 0;
    *evals = 0;
    if (!(sig > 0)) { *V_out = 0; return 0; }
    double K = exp((mu0 + b * phin) / a) / (2 * V0);
    double Vu = 2 * V0 * sinh(sig / (a * nP)) * exp(-(mu0 + b * phin) / a);
    double Vhi = Vu;
    if (eta_rad > 0 && sig / eta_rad < Vhi) Vhi = sig / eta_rad;
    double V;
    int k = 0;
    if (eta_rad == 0 || !(Vhi > 0) || !isfinite(Vhi)) {
        V = Vhi;
    } else {
        double shi = log(Vhi), s = shi, g, dg;
        double r = gfun_rsf(s, K, nP, a, eta_rad, sig, &g, &dg);
        int ev = 1, from_guess = 0;
        if (r > tol && guessV > 0) {
            double sg = log(guessV);
            if (sg > shi) sg = shi;
            double gg, dgg;
            double rg = gfun_rsf(sg, K, nP, a, eta_rad, sig, &gg, &dgg);
            ev++;
            if (rg < r) { s = sg; g = gg; dg = dgg; r = rg; from_guess = 1; }
        }
        double rprev = r;
        while (r > tol && k < 200) {
            double sn = s - g / dg;
            if (!(sn <= shi)) sn = shi;
            s = sn;
            k++;
            r = gfun_rsf(s, K, nP, a, eta_rad, sig, &g, &dg);
            ev++;
            if (from_guess && k == 1 && r > rprev) {
                from_guess = 0;
                *fallback = 1;
                s = shi;
                r = gfun_rsf(s, K, nP, a, eta_rad, sig, &g, &dg);
                ev++;
            }
            rprev = r;
        }
        *evals = ev;
        V = exp(s);
    }
    *root = V;
    V = ref_minv(V, EII * W);
    if (V < 1e-40) V = 1e-40;
    *V_out = V;
    return k;
}

/* ---- calc_frictionpara (calc_RSF.c:224-295): per-marker RSF parameters from
 * the hard-wired fault-band geometry.  Writes marker fields in place. */
static void friction_marker(or_state *s, long i, int t)
{
    const v2c_params *p = &s->prm;
    v2c_markers *m = &s->m;
    double x = m->x[i], y = m->y[i];
    double radius = ref_abs(y - p->band_y);
    double hr2 = p->half_range;
    if (t > 0) {
        if (radius <= p->band_halfwidth) {
            double a, b, L, V0;
            if (x > p->end_updip && x < p->start_downdip) {
                a = p->a_vw; b = p->b_vw; L = p->L_vw; V0 = p->V0_vw;
            } else if (x > p->start_updip && x < p->end_updip && p->half_range > 0) {
                double f = (x - (p->start_updip - p->xstpx / 2)) / (2 * hr2);
                a = p->a_vs + f * (p->a_vw - p->a_vs);
                b = p->b_vs + f * (p->b_vw - p->b_vs);
                L = p->L_vs + f * (p->L_vw - p->L_vs);
                V0 = p->V0_vs + f * (p->V0_vw - p->V0_vs);
            } else if (x >= p->start_downdip && x < p->end_downdip && p->half_range > 0) {
                double f = (x - (p->start_downdip - p->xstpx / 2)) / (2 * hr2);
                a = p->a_vw - f * (p->a_vw - p->a_vs);
                b = p->b_vw - f * (p->b_vw - p->b_vs);
                L = p->L_vw - f * (p->L_vw - p->L_vs);
                V0 = p->V0_vw - f * (p->V0_vw - p->V0_vs);
            } else {
                a = p->a_vs; b = p->b_vs; L = p->L_vs; V0 = p->V0_vs;
            }
            m->fric_static[i] = 0.2;
            m->a[i] = a; m->b[i] = b; m->L0[i] = L; m->L[i] = L; m->V0[i] = V0;
            m->C0[i] = 0.0;
            if (s->timesum == s->timestep) m->phi[i] = 0.0;
        } else {
            m->phi[i] = 50.0;
            m->fric_static[i] = 0.6;
            m->a[i] = p->a_hr; m->b[i] = p->b_hr; m->L0[i] = p->L_hr; m->L[i] = p->L_hr;
            m->V0[i] = p->V0_hr;
        }
    } else {
        m->fric_static[i] = 0.6;
        m->a[i] = 1.0; m->b[i] = 0.01; m->L0[i] = 0.2; m->V0[i] = p->V0_hr;
    }
}

/* ---- calc_frictionpara_nodes (calc_RSF.c:298-350) ------------------------- */
static void friction_node(or_state *s, double x, double y, long m3)
{
    const v2c_params *p = &s->prm;
    double radius = ref_abs(y - p->band_y);
    double hr2 = p->half_range;
    if (radius <= p->band_halfwidth) {
        double a, b, L, V0;
        if (x > p->end_updip && x < p->start_downdip) {
            a = p->a_vw; b = p->b_vw; L = p->L_vw; V0 = p->V0_vw;
        } else if (x >= p->start_updip && x <= p->end_updip && p->half_range > 0) {
            double f = (x - (p->start_updip)) / (2 * hr2);
            a = p->a_vs + f * (p->a_vw - p->a_vs);
            b = p->b_vs + f * (p->b_vw - p->b_vs);
            L = p->L_vs + f * (p->L_vw - p->L_vs);
            V0 = p->V0_vs + f * (p->V0_vw - p->V0_vs);
        } else if (x >= p->start_downdip && x < p->end_downdip && p->half_range > 0) {
            double f = (x - (p->start_downdip)) / (2 * hr2);
            a = p->a_vw - f * (p->a_vw - p->a_vs);
            b = p->b_vw - f * (p->b_vw - p->b_vs);
            L = p->L_vw - f * (p->L_vw - p->L_vs);
            V0 = p->V0_vw - f * (p->V0_vw - p->V0_vs);
        } else {
            a = p->a_vs; b = p->b_vs; L = p->L_vs; V0 = p->V0_vs;
        }
        NODAL(s, mus_s)[m3] = 0.2;
        NODAL(s, as)[m3] = a; NODAL(s, bs)[m3] = b; NODAL(s, Ls)[m3] = L; NODAL(s, V0s)[m3] = V0;
        NODAL(s, cohesion)[m3] = 0.0;
    } else {
        NODAL(s, mus_s)[m3] = 0.6;
        NODAL(s, as)[m3] = p->a_hr; NODAL(s, bs)[m3] = p->b_hr; NODAL(s, Ls)[m3] = p->L_hr;
        NODAL(s, V0s)[m3] = p->V0_hr;
    }
}

/* ---- marker2nodes (interpolation.c:55-824) -------------------------------- */
enum { B_NU, B_GG, B_PE, B_SXY, B_EXY, B_COH, B_PHI, B_RO, B_ET, B_HT, B_CP, B_KT, B_TK, B_W, NB };
enum { C_BM, C_ND, C_GD, C_SXX, C_EXX, C_P, C_W, NC };
enum { V_MV, V_MR, V_W, NV };

void or_marker2nodes(or_state *s)
{
    const v2c_grid *g = &s->grid;
    const v2c_params *p = &s->prm;
    const v2c_materials *M = s->mat;
    v2c_markers *m = &s->m;
    const long ynumy = g->ynumy, nodenum = g->xnumx * g->ynumy;

    /* K1 (interpolation.c:58-153): zero outputs and accumulators. */
    static const int zeroed[] = {
        V2C_N_bm, V2C_N_ro, V2C_N_et, V2C_N_nu, V2C_N_nd, V2C_N_gg, V2C_N_gd, V2C_N_sppe,
        V2C_N_phis, V2C_N_phin, V2C_N_sxyen, V2C_N_sxxen, V2C_N_exyen, V2C_N_exxen,
        V2C_N_sppen, V2C_N_mus_s, V2C_N_as, V2C_N_bs, V2C_N_Ls, V2C_N_V0s, V2C_N_sxxe,
        V2C_N_sxye, V2C_N_exxe, V2C_N_exye, V2C_N_cohesion, V2C_N_dro, V2C_N_drp,
        V2C_N_cp, V2C_N_kt, V2C_N_ht, V2C_N_tk, V2C_N_mrx, V2C_N_mry, V2C_N_mvx, V2C_N_mvy,
        V2C_N_wtphisum, V2C_N_phisum, V2C_N_sxysum, V2C_N_exysum, V2C_N_esum,
        V2C_N_wtsxxsum, V2C_N_psum, V2C_N_sxxsum, V2C_N_exxsum, V2C_N_wtvxsum,
        V2C_N_vxsum, V2C_N_wtvysum, V2C_N_vysum, V2C_N_mrx0, V2C_N_mry0 };
    for (size_t z = 0; z < sizeof zeroed / sizeof zeroed[0]; z++)
        memset(s->nod[zeroed[z]], 0, sizeof(double) * nodenum);

    double *B = calloc((size_t)nodenum * NB, sizeof(double));
    double *C = calloc((size_t)nodenum * NC, sizeof(double));
    double *VX = calloc((size_t)nodenum * NV, sizeof(double));
    double *VY = calloc((size_t)nodenum * NV, sizeof(double));

    /* K2 (interpolation.c:215-629), interpolationmod == 2 weights. */
    for (long i = 0; i < s->n; i++) {
        double x = m->x[i], y = m->y[i];
        if (!(x > 0 && y > 0 && x < g->xsize && y < g->ysize)) continue;
        long nx = or_nxsearch(g, x), ny = or_nysearch(g, y);
        int t = (int)m->t[i];
        double mbm = p->bulk_modulus;
        double mgg = M->gg[t];
        double mtk = m->k[i];
        double mpb = 1e-5 * or_allinterpomp(s, x, y);
        /* recalc_marker_properties (recalc_marker_properties.c:43-77) */
        double mro, mbb, maa;
        dencalc(s, mtk, mpb, t, &mro, &mbb, &maa);
        double mkt = conductivity(M, t, mtk, mpb);
        double mnu = or_ductilevisc(s, t, m->xy[i], m->xx[i], m->exy[i], m->exx[i], mtk, mpb, NULL);
        friction_marker(s, i, t);
        double me = m->e[i];

        double swt1 = 1.0 / (g->gx[nx + 1] - g->gx[nx]) / (g->gy[ny + 1] - g->gy[ny]);
        double cx = (x - g->gx[nx]) / (g->gx[nx + 1] - g->gx[nx]);
        double cy = (y - g->gy[ny]) / (g->gy[ny + 1] - g->gy[ny]);

        /* basic node: the quadrant's corner (A.2 table) */
        long bn = -1; double bw = 0;
        if (cx < 0.5 && cy < 0.5) { bn = nx * ynumy + ny; bw = swt1 * (1.0 - 2.0 * cx) * (1.0 - 2.0 * cy); }
        else if (cx < 0.5 && cy > 0.5) { bn = nx * ynumy + ny + 1; bw = swt1 * (1.0 - 2.0 * cx) * (2.0 * cy - 1.0); }
        else if (cx > 0.5 && cy < 0.5) { bn = (nx + 1) * ynumy + ny; bw = swt1 * (2.0 * cx - 1.0) * (1.0 - 2.0 * cy); }
        else if (cx > 0.5 && cy > 0.5) { bn = (nx + 1) * ynumy + ny + 1; bw = swt1 * (2.0 * cx - 1.0) * (2.0 * cy - 1.0); }
        if (bn >= 0) {
            double *b = B + bn * NB;
            b[B_NU] += mnu * bw;
            b[B_GG] += mnu / mgg * bw;
            b[B_PE] += me * bw;
            b[B_SXY] += m->xy[i] * bw;
            b[B_EXY] += m->exy[i] * bw;
            b[B_COH] += m->C[i] * bw;
            b[B_PHI] += m->phi[i] * bw;
            b[B_RO] += mro * bw;
            b[B_ET] += mbb * bw;
            b[B_HT] += M->ht[t] * bw;
            b[B_CP] += M->cp[t] * bw;
            b[B_KT] += mkt * bw;
            b[B_TK] += mtk * bw;
            b[B_W] += bw;
        }
        /* vx node: left or right edge midpoint */
        long vn = -1; double vw = 0;
        if (cx < 0.5) { vn = nx * ynumy + ny; vw = swt1 * (1.0 - cx) * (1.0 - ref_abs(cy - 0.5)); }
        else if (cx > 0.5) { vn = (nx + 1) * ynumy + ny; vw = swt1 * cx * (1.0 - ref_abs(cy - 0.5)); }
        if (vn >= 0) {
            double *v = VX + vn * NV;
            v[V_MV] += m->vx[i] * mro * vw;
            v[V_MR] += mro * vw;
            v[V_W] += vw;
        }
        /* vy node: top or bottom edge midpoint */
        vn = -1;
        if (cy < 0.5) { vn = nx * ynumy + ny; vw = swt1 * (1.0 - ref_abs(cx - 0.5)) * (1.0 - cy); }
        else if (cy > 0.5) { vn = nx * ynumy + ny + 1; vw = swt1 * (1.0 - ref_abs(cx - 0.5)) * cy; }
        if (vn >= 0) {
            double *v = VY + vn * NV;
            v[V_MV] += m->vy[i] * mro * vw;
            v[V_MR] += mro * vw;
            v[V_W] += vw;
        }
        /* centre (pressure) node of the cell */
        {
            long cn = (nx + 1) * ynumy + ny + 1;
            double cw = swt1 * (1.0 - 2.0 * ref_abs(cx - 0.5)) * (1.0 - 2.0 * ref_abs(cy - 0.5));
            double *c = C + cn * NC;
            c[C_BM] += mbm * cw;
            c[C_ND] += mnu * cw;
            c[C_GD] += mnu / mgg * cw;
            c[C_SXX] += m->xx[i] * cw;
            c[C_EXX] += m->exx[i] * cw;
            c[C_P] += m->p[i] * cw;
            c[C_W] += cw;
        }
    }

    /* K3 (interpolation.c:714-824): normalisation per node. */
    const int anti = s->timesum > 0 && p->antidiffinter == 1;
    for (long m3 = 0; m3 < nodenum; m3++) {
        long m22 = m3 % ynumy, m11 = (m3 - m22) / ynumy;
        const double *b = B + m3 * NB, *c = C + m3 * NC, *vx = VX + m3 * NV, *vy = VY + m3 * NV;
        if (b[B_W] != 0 && m11 >= 1 && m22 >= 1 && m11 < g->xnumx - 1 && m22 < ynumy - 1) {
            double phin = b[B_PHI] / b[B_W], sxyen = b[B_SXY] / b[B_W];
            double exyen = b[B_EXY] / b[B_W], pen = b[B_PE] / b[B_W];
            NODAL(s, phin)[m3] = phin; NODAL(s, sxyen)[m3] = sxyen;
            NODAL(s, exyen)[m3] = exyen; NODAL(s, pen)[m3] = pen;
            double pe;
            if (anti) {
                NODAL(s, phis)[m3] = NODAL(s, phis_it)[m3] + (phin - NODAL(s, phie)[m3]);
                NODAL(s, sxye)[m3] = NODAL(s, sxy)[m3] + (sxyen - NODAL(s, sxyee)[m3]);
                NODAL(s, exye)[m3] = NODAL(s, exy)[m3] + (exyen - NODAL(s, exyee)[m3]);
                pe = NODAL(s, pes)[m3] + (pen - NODAL(s, pee)[m3]);
            } else {
                NODAL(s, phis)[m3] = phin; NODAL(s, sxye)[m3] = sxyen;
                NODAL(s, exye)[m3] = exyen; pe = pen;
            }
            if (pe < 0) pe = 0;
            NODAL(s, pe)[m3] = pe;
            friction_node(s, g->gx[m11], g->gy[m22], m3);
            NODAL(s, cohesion)[m3] = b[B_COH] / b[B_W];
            double nu_m2c = b[B_NU] / b[B_W];
            NODAL(s, nu_m2c)[m3] = nu_m2c;
            NODAL(s, gg)[m3] = nu_m2c / (b[B_GG] / b[B_W]);
        }
        if (c[C_W] != 0 && m11 >= 1 && m22 >= 1) {
            double sxxen = c[C_SXX] / c[C_W], sppen = c[C_P] / c[C_W], exxen = c[C_EXX] / c[C_W];
            NODAL(s, sxxen)[m3] = sxxen; NODAL(s, sppen)[m3] = sppen; NODAL(s, exxen)[m3] = exxen;
            double sppe;
            if (anti) {
                NODAL(s, sxxe)[m3] = NODAL(s, sxx)[m3] + (sxxen - NODAL(s, sxxee)[m3]);
                sppe = NODAL(s, pr)[m3] + (sppen - NODAL(s, sppee)[m3]);
                NODAL(s, exxe)[m3] = NODAL(s, exx)[m3] + (exxen - NODAL(s, exxee)[m3]);
            } else {
                NODAL(s, sxxe)[m3] = sxxen; sppe = sppen; NODAL(s, exxe)[m3] = exxen;
            }
            if (s->timesum == 0 && p->reloadfromDP == 0 && p->compressiblemode == 1) sppe = p->nstress;
            NODAL(s, sppe)[m3] = sppe;
            double nd = c[C_ND] / c[C_W];
            NODAL(s, nd_m2c)[m3] = nd;
            NODAL(s, gd)[m3] = nd / (c[C_GD] / c[C_W]);
            NODAL(s, bm)[m3] = c[C_BM] / c[C_W];
            NODAL(s, dro)[m3] = 0 / c[C_W];
            NODAL(s, drp)[m3] = 0 / c[C_W];
        }
        if (vx[V_W] != 0) {
            double mvxn = vx[V_MV] / vx[V_MR];
            NODAL(s, mvxn)[m3] = mvxn;
            NODAL(s, mvx)[m3] = anti ? NODAL(s, vx)[m3] + (mvxn - NODAL(s, mvxe)[m3]) : mvxn;
            NODAL(s, mrx)[m3] = vx[V_MR] / vx[V_W];
        }
        if (vy[V_W] != 0) {
            double mvyn = vy[V_MV] / vy[V_MR];
            NODAL(s, mvyn)[m3] = mvyn;
            NODAL(s, mvy)[m3] = anti ? NODAL(s, vy)[m3] + (mvyn - NODAL(s, mvye)[m3]) : mvyn;
            NODAL(s, mry)[m3] = vy[V_MR] / vy[V_W];
        }
        if (b[B_W] != 0) {
            NODAL(s, ro)[m3] = b[B_RO] / b[B_W];
            NODAL(s, et)[m3] = b[B_ET] / b[B_W];
            NODAL(s, ht)[m3] = b[B_HT] / b[B_W];
            NODAL(s, kt)[m3] = b[B_KT] / b[B_W];
            NODAL(s, cp)[m3] = b[B_CP] / b[B_W];
            NODAL(s, tk)[m3] = b[B_TK] / b[B_W];
        }
    }
    /* State-variable boundary copy when diffusion is on (interpolation.c:827-859). */
    if (p->Dphis > 0) {
        double *phis = NODAL(s, phis);
        for (long m3 = 0; m3 < nodenum; m3++) {
            long m22 = m3 % ynumy, m11 = (m3 - m22) / ynumy;
            if (m11 == 0 && m22 >= 1 && m22 <= ynumy - 2) phis[m3] = phis[m3 + ynumy];
            else if (m11 == g->xnumx - 1 && m22 >= 1 && m22 <= ynumy - 2) phis[m3] = phis[m3 - ynumy];
            else if (m22 == 0 && m11 >= 1 && m11 <= g->xnumx - 2) phis[m3] = phis[m3 + 1];
            else if (m22 == ynumy - 1 && m11 > 0 && m11 <= g->xnumx - 2) phis[m3] = phis[m3 - 1];
            else if ((m11 == 0 && m22 == 0) || (m11 == 0 && m22 == ynumy - 1) ||
                     (m11 == g->xnumx - 1 && m22 == 0) || (m11 == g->xnumx - 1 && m22 == ynumy - 1))
                phis[m3] = 0;
        }
    }
    free(B); free(C); free(VX); free(VY);
}

/* ---- interpolate_mechanical (interpolation.c:1050-1490), mode 2 ----------- */
typedef struct mech_vals {
    double EXY, EXYE, SXY, SXYE, PHIS, PHISE, SXX, SXXE, EXX, EXXE, PR, SPPE;
    double VX, MVX, VY, MVY, ESP, ES, ESE;
} mech_vals;

static void interpolate_mechanical(or_state *s, double x, double y, double mro, mech_vals *o)
{
    const v2c_grid *g = &s->grid;
    const long ynumy = g->ynumy;
    const int anti = s->prm.antidiffinter == 1;
    x = clampd(x, 0, g->xsize);
    y = clampd(y, 0, g->ysize);
    long nx = or_nxsearch(g, x), ny = or_nysearch(g, y);
    long m3; double e, n;

    basic_stencil(g, nx, ny, x, y, &m3, &e, &n);
    o->EXY = bil(NODAL(s, exy), m3, ynumy, e, n);
    o->EXYE = bil(NODAL(s, exye), m3, ynumy, e, n);
    o->SXY = bil(NODAL(s, sxy), m3, ynumy, e, n);
    o->SXYE = bil(NODAL(s, sxye), m3, ynumy, e, n);
    o->ES = bil(NODAL(s, pes), m3, ynumy, e, n);
    o->ESE = bil(NODAL(s, pe), m3, ynumy, e, n);
    o->ESP = bil(NODAL(s, esp), m3, ynumy, e, n);
    o->PHISE = bil(NODAL(s, phis), m3, ynumy, e, n);
    o->PHIS = bil(NODAL(s, phis_it), m3, ynumy, e, n);

    double swt1 = 1.0 / (g->gx[nx + 1] - g->gx[nx]) / (g->gy[ny + 1] - g->gy[ny]);
    double cx = (x - g->gx[nx]) / (g->gx[nx + 1] - g->gx[nx]);
    double cy = (y - g->gy[ny]) / (g->gy[ny + 1] - g->gy[ny]);
    if (anti) {
        long bn = -1; double bw = 0;
        if (cx < 0.5 && cy < 0.5) { bn = nx * ynumy + ny; bw = swt1 * (1.0 - 2 * cx) * (1.0 - 2 * cy); }
        else if (cx < 0.5 && cy > 0.5) { bn = nx * ynumy + ny + 1; bw = swt1 * (1.0 - 2.0 * cx) * (2.0 * cy - 1.0); }
        else if (cx > 0.5 && cy < 0.5) { bn = (nx + 1) * ynumy + ny; bw = swt1 * (2.0 * cx - 1.0) * (1.0 - 2.0 * cy); }
        else if (cx > 0.5 && cy > 0.5) { bn = (nx + 1) * ynumy + ny + 1; bw = swt1 * (2.0 * cx - 1.0) * (2.0 * cy - 1.0); }
        if (bn >= 0) {
            NODAL(s, phisum)[bn] += o->PHIS * bw;
            NODAL(s, sxysum)[bn] += o->SXY * bw;
            NODAL(s, exysum)[bn] += o->EXY * bw;
            NODAL(s, esum)[bn] += o->ES * bw;
            NODAL(s, wtphisum)[bn] += bw;
        }
    }

    centre_stencil(g, nx, ny, x, y, &m3, &e, &n);
    o->EXX = bil(NODAL(s, exx), m3, ynumy, e, n);
    o->EXXE = bil(NODAL(s, exxe), m3, ynumy, e, n);
    o->SXX = bil(NODAL(s, sxx), m3, ynumy, e, n);
    o->SXXE = bil(NODAL(s, sxxe), m3, ynumy, e, n);
    o->PR = bil(NODAL(s, pr), m3, ynumy, e, n);
    o->SPPE = bil(NODAL(s, sppe), m3, ynumy, e, n);
    if (anti) {
        long cn = (nx + 1) * ynumy + ny + 1;
        double cw = swt1 * (1.0 - 2 * ref_abs(cx - 0.5)) * (1.0 - 2 * ref_abs(cy - 0.5));
        NODAL(s, psum)[cn] += o->PR * cw;
        NODAL(s, sxxsum)[cn] += o->SXX * cw;
        NODAL(s, exxsum)[cn] += o->EXX * cw;
        NODAL(s, wtsxxsum)[cn] += cw;
    }

    vx_stencil(g, nx, ny, x, y, 0, &m3, &e, &n);
    o->VX = bil(NODAL(s, vx), m3, ynumy, e, n);
    o->MVX = bil(NODAL(s, mvx), m3, ynumy, e, n);
    if (anti) {
        long vn = -1; double vw = 0;
        if (cx < 0.5) { vn = nx * ynumy + ny; vw = swt1 * (1.0 - cx) * (1.0 - ref_abs(cy - 0.5)); }
        else if (cx > 0.5) { vn = (nx + 1) * ynumy + ny; vw = swt1 * cx * (1.0 - ref_abs(cy - 0.5)); }
        if (vn >= 0) {
            NODAL(s, vxsum)[vn] += o->VX * mro * vw;
            NODAL(s, mrx0)[vn] += mro * vw;
            NODAL(s, wtvxsum)[vn] += vw;
        }
    }

    vy_stencil(g, nx, ny, x, y, 0, &m3, &e, &n);
    o->VY = bil(NODAL(s, vy), m3, ynumy, e, n);
    o->MVY = bil(NODAL(s, mvy), m3, ynumy, e, n);
    if (anti) {
        long vn = -1; double vw = 0;
        if (cy < 0.5) { vn = nx * ynumy + ny; vw = swt1 * (1.0 - ref_abs(cx - 0.5)) * (1.0 - cy); }
        else if (cy > 0.5) { vn = nx * ynumy + ny + 1; vw = swt1 * (1.0 - ref_abs(cx - 0.5)) * cy; }
        if (vn >= 0) {
            NODAL(s, vysum)[vn] += o->VY * mro * vw;
            NODAL(s, mry0)[vn] += mro * vw;
            NODAL(s, wtvysum)[vn] += vw;
        }
    }
}

/* Marker field update of nodes2markers (interpolation.c:943-1011). */
static void n2m_update_marker(or_state *s, long i, const mech_vals *v)
{
    const v2c_grid *g = &s->grid;
    const v2c_params *p = &s->prm;
    v2c_markers *m = &s->m;
    double x = m->x[i], y = m->y[i];
    int t = (int)m->t[i];
    if (x > 0 && y > 0 && x < g->xsize && y < g->ysize && t < 50 && m->k[i] <= 0) {
        if (p->antidiffinter == 1) {
            m->xx[i] = v->SXX; m->xy[i] = v->SXY; m->p[i] = v->PR; m->exx[i] = v->EXX;
            m->exy[i] = v->EXY; m->vx[i] = v->VX; m->vy[i] = v->VY; m->phi[i] = v->PHIS;
            m->e[i] = v->ES;
        } else {
            m->xx[i] = v->SXXE; m->xy[i] = v->SXYE; m->p[i] = v->SPPE; m->exx[i] = v->EXXE;
            m->exy[i] = v->EXYE; m->vx[i] = v->MVX; m->vy[i] = v->MVY; m->phi[i] = v->PHISE;
            m->e[i] = v->ESE;
        }
        if (!p->tempmod) m->k[i] = p->incoming_T;
    }
    if (x >= 0 && y >= 0 && x <= g->xsize && y <= g->ysize && t < 50 && m->k[i] > 0) {
        if (p->antidiffinter == 1) {
            m->phi[i] = v->PHIS; m->exx[i] = v->EXX; m->exy[i] = v->EXY; m->xx[i] = v->SXX;
            m->xy[i] = v->SXY; m->p[i] = v->PR; m->vx[i] = v->VX; m->vy[i] = v->VY;
            m->e[i] = v->ES;
        } else {
            m->phi[i] += (v->PHIS - v->PHISE); m->exx[i] += (v->EXX - v->EXXE);
            m->exy[i] += (v->EXY - v->EXYE); m->xx[i] += (v->SXX - v->SXXE);
            m->xy[i] += (v->SXY - v->SXYE); m->p[i] += (v->PR - v->SPPE);
            m->vx[i] += (v->VX - v->MVX); m->vy[i] += (v->VY - v->MVY);
            m->e[i] += (v->ES - v->ESE);
        }
    }
}

/* ---- nodes2markers (interpolation.c:929-1048) ----------------------------- */
void or_nodes2markers(or_state *s)
{
    const v2c_grid *g = &s->grid;
    const long ynumy = g->ynumy, nodenum = g->xnumx * g->ynumy;
    mech_vals v;
    for (long i = 0; i < s->n; i++) {
        interpolate_mechanical(s, s->m.x[i], s->m.y[i], s->mat->ro[(int)s->m.t[i]], &v);
        n2m_update_marker(s, i, &v);
    }
    /* The extra iteration of the reference's `m3<=marknum` loop: a zero marker
     * of type 0 at (0,0).  Its own fields are never written (both update
     * conditions fail at markk == 0, x == 0). */
    if (s->prm.n2m_phantom)
        interpolate_mechanical(s, 0.0, 0.0, s->mat->ro[0], &v);

    if (s->prm.antidiffinter == 1) {
        for (long m3 = 0; m3 < nodenum; m3++) {
            long m22 = m3 % ynumy, m11 = (m3 - m22) / ynumy;
            double w = NODAL(s, wtphisum)[m3];
            if (w != 0 && m11 >= 1 && m22 >= 1 && m11 < g->xnumx - 1 && m22 < ynumy - 1) {
                NODAL(s, phie)[m3] = NODAL(s, phisum)[m3] / w;
                NODAL(s, sxyee)[m3] = NODAL(s, sxysum)[m3] / w;
                NODAL(s, exyee)[m3] = NODAL(s, exysum)[m3] / w;
                NODAL(s, pee)[m3] = NODAL(s, esum)[m3] / w;
            }
            w = NODAL(s, wtsxxsum)[m3];
            if (w != 0 && m11 >= 1 && m22 >= 1) {
                NODAL(s, sppee)[m3] = NODAL(s, psum)[m3] / w;
                NODAL(s, sxxee)[m3] = NODAL(s, sxxsum)[m3] / w;
                NODAL(s, exxee)[m3] = NODAL(s, exxsum)[m3] / w;
            }
            if (NODAL(s, wtvxsum)[m3] != 0) NODAL(s, mvxe)[m3] = NODAL(s, vxsum)[m3] / NODAL(s, mrx0)[m3];
            if (NODAL(s, wtvysum)[m3] != 0) NODAL(s, mvye)[m3] = NODAL(s, vysum)[m3] / NODAL(s, mry0)[m3];
        }
    }
}

/* ---- allinteri (interpolation.c:1688-1758) -------------------------------- */
void or_allinteri(const or_state *s, double x, double y, double *VX, double *VY,
                  double *ESP, double *EE)
{
    const v2c_grid *g = &s->grid;
    const double *gx = g->gx, *gy = g->gy;
    const long ynumy = g->ynumy;
    x = clampd(x, 0, g->xsize);
    y = clampd(y, 0, g->ysize);
    double xrat = 2.0 / 3.0;
    if (x < (gx[0] + gx[1]) / 2.0) xrat = 1.0;
    if (x > (gx[g->xnumx - 2] + gx[g->xnumx - 1]) / 2.0) xrat = 1.0;
    if (y < (gy[0] + gy[1]) / 2.0) xrat = 1.0;
    if (y > (gy[ynumy - 2] + gy[ynumy - 1]) / 2.0) xrat = 1.0;
    long nx = or_nxsearch(g, x), ny = or_nysearch(g, y);
    long m3; double e, n;

    basic_stencil(g, nx, ny, x, y, &m3, &e, &n);
    double EXY = bil(NODAL(s, exy), m3, ynumy, e, n);
    *ESP = bil(NODAL(s, esp), m3, ynumy, e, n);

    centre_stencil(g, nx, ny, x, y, &m3, &e, &n);
    double EXX = bil(NODAL(s, exx), m3, ynumy, e, n);
    const double *vx = NODAL(s, vx), *vy = NODAL(s, vy);
    double d0 = (1.0 - e) * (1.0 - n), d1 = (1.0 - e) * n, d2 = e * (1.0 - n), d3 = e * n;
    double vxa = (d0 * (vx[m3 - 1] + vx[m3 - ynumy - 1]) + d1 * (vx[m3] + vx[m3 - ynumy])
                  + d2 * (vx[m3 + ynumy - 1] + vx[m3 - 1]) + d3 * (vx[m3 + ynumy] + vx[m3])) * 0.5 * (1.0 - xrat);
    double vya = (d0 * (vy[m3 - ynumy] + vy[m3 - ynumy - 1]) + d1 * (vy[m3 - ynumy + 1] + vy[m3 - ynumy])
                  + d2 * (vy[m3] + vy[m3 - 1]) + d3 * (vy[m3 + 1] + vy[m3])) * 0.5 * (1.0 - xrat);
    *EE = pow(EXX * EXX + EXY * EXY, 0.5);

    vx_stencil(g, nx, ny, x, y, 1, &m3, &e, &n);
    *VX = vxa + bil(vx, m3, ynumy, e, n) * xrat;
    vy_stencil(g, nx, ny, x, y, 1, &m3, &e, &n);
    *VY = vya + bil(vy, m3, ynumy, e, n) * xrat;
}

/* ---- moveandrotatemarkers (moveandrotatemarkers.c:1-138), types < 100 ----- */
long or_moveandrotatemarkers(or_state *s)
{
    const v2c_grid *g = &s->grid;
    const v2c_params *p = &s->prm;
    v2c_markers *m = &s->m;
    const double dt = s->timestep;
    long unsupported = 0;
    for (long i = 0; i < s->n; i++) {
        int t = (int)m->t[i];
        double x = m->x[i], y = m->y[i];
        int inside = x >= 0 && y >= 0 && x <= g->xsize && y <= g->ysize;
        if (!((inside || p->outgrid != 1) && !s->mat->im[t])) continue;
        double vx0, vy0, sp0, ee0;
        if (p->markmod == 1) {
            or_allinteri(s, x, y, &vx0, &vy0, &sp0, &ee0);
        } else {
            double vx1, vy1, sp1, ee1, vx2, vy2, sp2, ee2, vx3, vy3, sp3, ee3, vx4, vy4, sp4, ee4;
            or_allinteri(s, x, y, &vx1, &vy1, &sp1, &ee1);
            or_allinteri(s, x + vx1 * dt / 2.0, y + vy1 * dt / 2.0, &vx2, &vy2, &sp2, &ee2);
            or_allinteri(s, x + vx2 * dt / 2.0, y + vy2 * dt / 2.0, &vx3, &vy3, &sp3, &ee3);
            or_allinteri(s, x + vx3 * dt, y + vy3 * dt, &vx4, &vy4, &sp4, &ee4);
            vx0 = (vx1 + 2.0 * vx2 + 2.0 * vx3 + vx4) / 6.0;
            vy0 = (vy1 + 2.0 * vy2 + 2.0 * vy3 + vy4) / 6.0;
            if (p->markmod == 2) {
                sp0 = (sp1 + 2.0 * sp2 + 2.0 * sp3 + sp4) / 6.0;
                ee0 = (ee1 + 2.0 * ee2 + 2.0 * ee3 + ee4) / 6.0;
            } else {
                sp0 = sp1; ee0 = ee1;
            }
        }
        (void)ee0;
        if (p->outgrid == 2) {
            if (x < 0 || x > g->xsize) vy0 = 0;
            if (y < 0 || y > g->ysize) vx0 = 0;
        }
        if (t >= 100) { unsupported++; continue; }
        x += dt * vx0;
        y += dt * vy0;
        m->x[i] = x; m->y[i] = y;
        double th = sp0 * dt;
        double sxx = m->xx[i], sxy = m->xy[i];
        if (p->stoksmod == 1) {
            double c = cos(th), sn = sin(th);
            double nxx = sxx * c * c - sxx * sn * sn + sxy * sin(2.0 * th);
            double nxy = 0.5 * (-sxx - sxx) * sin(2.0 * th) + sxy * cos(2.0 * th);
            m->xx[i] = nxx; m->xy[i] = nxy;
        } else if (p->stoksmod == 2) {
            double nxx = sxx + sxy * 2.0 * th;
            double nxy = sxy + 0.5 * (-sxx - sxx) * 2.0 * th;
            m->xx[i] = nxx; m->xy[i] = nxy;
        }
        if (y > p->zdeep && vy0 < 0 && m->k[i] < p->tdeep) m->k[i] = p->tdeep;
        if (x < 0 || y < 0 || x > g->xsize || y > g->ysize) {
            m->k[i] = 0; m->d[i] = -1.0; m->w[i] = -1.0; m->e[i] = 0;
        }
    }
    return unsupported;
}

/* ---- calc_vpvisc (calc_RSF.c:33-169) -------------------------------------- */
static double vpvisc_node(or_state *s, int rk, double dt, long m3, double nP, double W, int *err)
{
    const v2c_params *p = &s->prm;
    double V = 0;
    if (rk == 0) {
        if (s->timesum > 0) V = NODAL(s, RK_sliprs_final)[m3];
        else if (p->RSFmode == 1) V = NODAL(s, RK_sliprs_final)[m3];
        else if (p->RSFmode == 0) V = 0;
    }
    if ((p->rkmod == 0 && rk > 0) || (rk > 0 && rk < 3)) {
        if (p->rkmod == 1) {
            V = NODAL(s, RK_sliprs_solved)[m3];
        } else {
            if (p->rkmod == 3) {
                double *rs = rk - 1 == 0 ? NODAL(s, RK_sliprs0) : NODAL(s, RK_sliprs1);
                double *rp = rk - 1 == 0 ? NODAL(s, RK_phidts0) : NODAL(s, RK_phidts1);
                rs[m3] = NODAL(s, RK_sliprs_solved)[m3];
                if (p->RSFmode == 1) rp[m3] = NODAL(s, RK_phidts_solved)[m3];
            }
            V = NODAL(s, RK_sliprs_solved)[m3];
        }
    }
    if (rk == 3 && p->rkmod == 3) V = NODAL(s, RK_sliprs_solved)[m3];

    NODAL(s, pes)[m3] = NODAL(s, pe)[m3] + V / W * dt;
    double newstate = 0, syield, vp = 0;
    const double as = NODAL(s, as)[m3], bs = NODAL(s, bs)[m3], Ls = NODAL(s, Ls)[m3];
    const double V0s = NODAL(s, V0s)[m3], mus = NODAL(s, mus_s)[m3];
    const double nu_m2c = NODAL(s, nu_m2c)[m3];
    if (p->RSFmode == 1) {
        double phis = NODAL(s, phis)[m3];
        if (V * dt / Ls <= 1e-6) newstate = log(exp(phis) * (1 - V * dt / Ls) + V0s * dt / Ls);
        else newstate = log(V0s / V + (exp(phis) - V0s / V) * exp(-V * dt / Ls));
        double rad = p->inertyn == 0 ? p->rad_damping * V : 0.0;
        if (p->power_law_RSF > 0) {
            double sy0 = nP * mus;
            syield = sy0 * pow(V / V0s, as / mus) * exp((newstate * bs) / mus) + rad;
        } else {
            syield = nP * as * asinh(V / (2 * V0s) * exp((mus + newstate * bs) / as)) + rad;
        }
    } else {
        syield = nP * mus + NODAL(s, cohesion)[m3];
    }
    if (p->RSFmode == 1) {
        if (syield > 0) vp = nu_m2c * syield / (2 * nu_m2c * V / W + syield);
        else vp = nu_m2c;
    } else if (p->RSFmode == 0) {
        vp = nu_m2c * syield / (2 * nu_m2c * V / W + syield);
    }
    NODAL(s, phis_it)[m3] = newstate;
    NODAL(s, syields)[m3] = syield;
    NODAL(s, nuvp_it)[m3] = vp;
    NODAL(s, sliprs)[m3] = V;
    if ((s->timesum > 0 && (isnan(newstate) || isnan(syield) || isnan(vp) || vp == 0 || isnan(V)))
        || V < 0 || syield < 0)
        *err = 1;
    return vp;
}

/* ---- compute_before_solve (plastic.c:2-56) -------------------------------- */
long or_compute_before_solve(or_state *s, double dt, int rk)
{
    const v2c_grid *g = &s->grid;
    const v2c_params *p = &s->prm;
    const long ynumy = g->ynumy, nodenum = g->xnumx * g->ynumy;
    long nerr = 0;
    for (long m3 = 0; m3 < nodenum; m3++) {
        long m2 = m3 % ynumy, m1 = (m3 - m2) / ynumy;
        double W = p->faultwidth;
        double nps = (p->pmode == 0 || s->timesum == 0) ? p->nstress : NODAL(s, sppes)[m3];
        if (m1 >= 1 && m2 >= 1 && m1 < g->xnumx - 1 && m2 < ynumy - 1) {
            double sxxs = NODAL(s, sxxs)[m3], sxy = NODAL(s, sxy)[m3];
            double sigin = pow(sxxs * sxxs + sxy * sxy, 0.5);
            int err = 0;
            double nu;
            if (sigin > 0 || p->RSFmode == 0) nu = vpvisc_node(s, rk, dt, m3, nps, W, &err);
            else nu = NODAL(s, nu_m2c)[m3];
            NODAL(s, nu)[m3] = nu;
            if (isnan(nu) || nu <= 0) err = 1;
            nerr += err;
        }
    }
    return nerr;
}

/* ---- calc_sliprate_after_solve (calc_RSF.c:172-221) ----------------------- */
static void sliprate_node(or_state *s, long m3, double dt, double sigin, double G,
                          double EXY, double EPXX, double nP, double NUVP, double NU, double W)
{
    const v2c_params *p = &s->prm;
    double *Vs = NODAL(s, RK_sliprs_solved);
    const double as = NODAL(s, as)[m3], bs = NODAL(s, bs)[m3], Ls = NODAL(s, Ls)[m3];
    const double V0s = NODAL(s, V0s)[m3], mus = NODAL(s, mus_s)[m3];
    const double phi = NODAL(s, phis_it)[m3];
    if (p->RSFmode == 1) {
        if (sigin > 0) {
            double V;
            if (p->power_law_RSF > 0) {
                double sy0 = nP * mus;
                V = V0s * pow(sigin / sy0, (mus / as)) * exp((-phi * bs) / as);
            } else {
                V = 2 * V0s * sinh(sigin / (as * nP)) * exp(-(mus + bs * phi) / as);
            }
            double EII = pow(EPXX * EPXX + EXY * EXY, 0.5);
            V = ref_minv(V, EII * W);
            if (V < 1e-40) V = 1e-40;
            Vs[m3] = V;
        } else {
            Vs[m3] = 0;
        }
        NODAL(s, RK_phidts_solved)[m3] = V0s / Ls * exp(-phi) - Vs[m3] / Ls;
    } else {
        double E2EL = sigin / (NUVP / (G * dt + NUVP));
        double A = NODAL(s, syields)[m3] / E2EL;
        if (A < 1) {
            double nup = G * dt * A / (1.0 - A);
            Vs[m3] = W * (sigin / 2.0 * (1.0 / (nup) - 1.0 / NU));
        } else {
            Vs[m3] = 0;
        }
    }
}

/* ---- compute_after_solve (plastic.c:61-129) ------------------------------- */
double or_compute_after_solve(or_state *s, double dt, int rk, long *nerr)
{
    const v2c_grid *g = &s->grid;
    const v2c_params *p = &s->prm;
    const long ynumy = g->ynumy, nodenum = g->xnumx * g->ynumy;
    (void)rk;
    double deltas = 0.0;
    long count = 0;
    for (long m3 = 0; m3 < nodenum; m3++) {
        long m2 = m3 % ynumy, m1 = (m3 - m2) / ynumy;
        double W = p->faultwidth;
        double nps = (p->pmode == 0 || s->timesum == 0) ? p->nstress : NODAL(s, sppes)[m3];
        if (m1 >= 1 && m2 >= 1 && m1 < g->xnumx - 1 && m2 < ynumy - 1) {
            double sxxs = NODAL(s, sxxs)[m3], sxy = NODAL(s, sxy)[m3];
            double sigin = pow(sxxs * sxxs + sxy * sxy, 0.5);
            sliprate_node(s, m3, dt, sigin, NODAL(s, gg)[m3], NODAL(s, exy)[m3],
                          NODAL(s, exxpl)[m3], nps, NODAL(s, nu)[m3], NODAL(s, nu_m2c)[m3], W);
            if (NODAL(s, RK_sliprs_solved)[m3] > 0) {
                double d = NODAL(s, syields)[m3] - sigin;
                NODAL(s, RK_deltas)[m3] = d;
                deltas += pow(d, 2);
                count = count + 1;
            } else {
                NODAL(s, RK_deltas)[m3] = 0;
            }
        }
    }
    deltas = count > 0 ? pow(deltas / count, 0.5) : 0;
    if (nerr) *nerr = isnan(deltas) ? 1 : 0;
    return deltas;
}

/* Per-marker local update over all markers (build-defined).  mx holds 14
 * fields x n in the device layout: V, phi, a, b, L, V0, mu0, eta, eta_vp, sy,
 * guess_eta, guess_V, root_eta, root_V.  stats: [n, ductile, ductile iters,
 * fault, rsf iters, max ductile, max rsf, fallbacks, nonfinite, ductile evals,
 * rsf evals]. */
void or_local_update(or_state *s, double dt, const unsigned char *fault, double *mx, int warm,
                     double tol, long *stats, int *ditr, int *ritr)
{
    const long n = s->n;
    double *V = mx, *PHI = mx + n, *A = mx + 2 * n, *B = mx + 3 * n, *L = mx + 4 * n, *V0 = mx + 5 * n;
    double *MU0 = mx + 6 * n, *ETA = mx + 7 * n, *ETAVP = mx + 8 * n, *SY = mx + 9 * n;
    double *GETA = mx + 10 * n, *GV = mx + 11 * n, *RETA = mx + 12 * n, *RV = mx + 13 * n;
    for (int k = 0; k < 11; k++) stats[k] = 0;
    for (long i = 0; i < n; i++) {
        int t = (int)s->m.t[i];
        double sxx = s->m.xx[i], sxy = s->m.xy[i], tk = s->m.k[i], p = s->m.p[i];
        int it = 0, ev = 0, fb = 0;
        double root;
        double eta = or_ductile_newton_ex(s, t, sxy, sxx, tk, 1e-5 * p, warm ? GETA[i] : 0.0, tol, &it, &ev,
                                          &fb, &root);
        ETA[i] = eta;
        RETA[i] = root;
        stats[0]++;
        int bad = !isfinite(eta);
        if (ditr) ditr[i] = it;
        if (power_law_type(s->mat, t) && t != 1) {
            stats[1]++; stats[2] += it; stats[9] += ev; if (it > stats[5]) stats[5] = it;
        }
        if (ritr) ritr[i] = 0;
        if (fault[i]) {
            double exx = s->m.exx[i], exy = s->m.exy[i];
            double sig = pow(sxx * sxx + sxy * sxy, 0.5);
            double nP = s->prm.pmode == 0 ? s->prm.nstress : p;
            double EII = pow(exx * exx + exy * exy, 0.5);
            double eta_rad = s->prm.inertyn == 0 ? s->prm.rad_damping : 0.0;
            double W = s->prm.faultwidth, v, phin, rroot;
            int rfb = 0, rev = 0;
            int rit = or_rsf_marker_ex(sig, nP, A[i], B[i], L[i], V0[i], MU0[i], PHI[i], V[i], dt, eta_rad, EII,
                                       W, warm ? GV[i] : 0.0, tol, &v, &phin, &rfb, &rroot, &rev);
            V[i] = v; PHI[i] = phin; SY[i] = sig; RV[i] = rroot;
            ETAVP[i] = sig > 0 ? eta * sig / (2 * eta * v / W + sig) : eta;
            fb |= rfb;
            if (!isfinite(v) || !isfinite(phin)) bad = 1;
            stats[3]++; stats[4] += rit; stats[10] += rev; if (rit > stats[6]) stats[6] = rit;
            if (ritr) ritr[i] = rit;
        }
        stats[7] += fb;
        stats[8] += bad;
    }
}
