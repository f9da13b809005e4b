/*
 * v2c_oracle_post.c -- FP64 CPU oracle of the reference's per-node routines
 * between the Stokes solve and the next marker2nodes.
 *
 * TEST INFRASTRUCTURE ONLY (see v2c_oracle.h).  Single-threaded restatement
 * in the reference's operation and node order:
 *   recompute_stress_strainrate  momentum_continuity_iterate.c:214-263
 *     sxxcalc / sxycalc / nueffcalc momentum_continuity_addmatrix_checkerror.c:
 *     380-393, 446-473, 521-569
 *   interpolate_center2basic_old interpolation.c:1575-1598
 *   interpolate_center2basic     interpolation.c:1600-1637
 *   interpolate_basic2center     interpolation.c:1640-1685
 *   copyarrays                   momentum_continuity_iterate.c:265-283
 *   calc_timestep                calc_RSF.c:352-572 (sticky-air statistics
 *                                separate, printf omitted; chi == 0 returns
 *                                -1 where the reference exits)
 *   sticky-air marks             interpolation.c:121, 861-916 (the end of
 *                                marker2nodes)
 *   sticky-air statistics        calc_RSF.c:364-374, 505-550 (inside
 *                                calc_timestep's node loop)
 *   marker temperatures          temperature_iterate.c:181-419 (the part of
 *                                solve_temperature after the solve)
 */
#include <math.h>
#include <stdlib.h>
#include "v2c_oracle.h"

#define N(f) (s->nod[V2C_N_##f])

static inline double minv(double a, double b) { return a <= b ? a : b; }
static inline double maxv(double a, double b) { return a >= b ? a : b; }
static inline double absv(double a) { return a >= 0 ? a : -a; }

static double nueffcalc(double nvp, double ggeff, double dt, double nubeg, double *Z)
{
    *Z = ggeff * dt / (ggeff * dt + nvp);
    double nueff = *Z * nvp;
    if (nueff < nubeg) nueff = nubeg;
    return nueff;
}

void or_recompute_stress_strainrate(or_state *s, double dt, int rk, double *stats)
{
    (void)rk;
    const long xnumx = s->grid.xnumx, ynumy = s->grid.ynumy, nodenum = xnumx * ynumy;
    const double *gx = s->grid.gx, *gy = s->grid.gy;
    double maxds = 0, minnu = -1e+50, maxnu = 1e+50, maxdphi = 0, maxnvp = 1e+30;
    for (long m3 = 0; m3 < nodenum; m3++) {
        const long m2 = m3 % ynumy, m1 = (m3 - m2) / ynumy;
        if (!(m1 < xnumx && m2 < ynumy && m1 >= 1 && m2 >= 1)) continue;
        const long v0 = (m1 - 1) * ynumy + (m2 - 1), v1 = v0 + 1, v2 = v0 + ynumy, v3 = v2 + 1;
        double Z = 1.0;
        /* sxxcalc */
        double nueff = nueffcalc(N(nd)[v3], N(gd)[v3], dt, s->prm.nubeg, &Z);
        double edotxx = 0.5 * ((N(vx)[v2] - N(vx)[v0]) / (gx[m1] - gx[m1 - 1]) -
                               (N(vy)[v1] - N(vy)[v0]) / (gy[m2] - gy[m2 - 1]));
        N(sxx)[m3] = 2.0 * nueff * edotxx + N(sxxe)[v3] * (1.0 - Z);
        N(exx)[m3] = edotxx;
        maxds = maxv(maxds, absv(N(sxx)[m3] - N(sxxe)[m3]));
        N(deltasxx)[m3] = (N(sxx)[m3] - N(sxxe)[m3]) / s->timestep;
        minnu = minv(minnu, nueff);
        maxnu = maxv(maxnu, nueff);
        if (m1 < xnumx - 1 && m2 < ynumy - 1) {
            /* sxycalc */
            Z = 1.0;
            nueff = nueffcalc(N(nu)[v3], N(gg)[v3], dt, s->prm.nubeg, &Z);
            double exy1 = (N(vx)[v3] - N(vx)[v2]) / (gy[m2 + 1] - gy[m2 - 1]);
            double exy2 = (N(vy)[v3] - N(vy)[v1]) / (gx[m1 + 1] - gx[m1 - 1]);
            double edotxy = exy1 + exy2;
            N(sxy)[m3] = 2.0 * nueff * edotxy + N(sxye)[v3] * (1.0 - Z);
            N(exy)[m3] = edotxy;
            N(esp)[m3] = exy1 - exy2;
            N(deltasxy)[m3] = (N(sxy)[m3] - N(sxye)[m3]) / s->timestep;
            maxnvp = minv(maxnvp, N(nu)[m3] / N(gg)[m3]);
            maxds = maxv(maxds, absv(N(sxy)[m3] - N(sxye)[m3]));
            maxdphi = maxv(maxdphi, absv(N(phis)[m3] - N(phis_it)[m3]));
            minnu = minv(minnu, nueff);
            maxnu = maxv(maxnu, nueff);
        }
    }
    if (stats) {
        stats[0] = maxds; stats[1] = minnu; stats[2] = maxnu; stats[3] = maxdphi; stats[4] = maxnvp;
    }
}

void or_interpolate_center2basic_old(or_state *s)
{
    const long xnumx = s->grid.xnumx, ynumy = s->grid.ynumy, nodenum = xnumx * ynumy;
    for (long m3 = 0; m3 < nodenum; m3++) {
        const long m2 = m3 % ynumy, m1 = (m3 - m2) / ynumy;
        if (m1 >= 1 && m2 >= 1 && m1 < xnumx - 1 && m2 < ynumy - 1) {
            const long a = m3, b = m3 + 1, c = m3 + ynumy, d = m3 + ynumy + 1;
            N(exxes)[m3] = 0.25 * (N(exxe)[a] + N(exxe)[b] + N(exxe)[c] + N(exxe)[d]);
            N(sxxes)[m3] = 0.25 * (N(sxxe)[a] + N(sxxe)[b] + N(sxxe)[c] + N(sxxe)[d]);
            N(sppes)[m3] = 0.25 * (N(sppe)[a] + N(sppe)[b] + N(sppe)[c] + N(sppe)[d]);
        }
    }
}

static double epl_at(const or_state *s, long i, double dt)
{
    return N(exx)[i] - (N(sxx)[i] - N(sxxe)[i]) / (2 * N(gd)[i] * dt) - N(sxx)[i] / (2 * N(nd_m2c)[i]);
}

static double exxel_at(const or_state *s, long i, double dt)
{
    return 2 * N(gd)[i] * dt * N(exx)[i] + N(sxxe)[i];
}

void or_interpolate_center2basic(or_state *s, double dt)
{
    const long xnumx = s->grid.xnumx, ynumy = s->grid.ynumy, nodenum = xnumx * ynumy;
    for (long m3 = 0; m3 < nodenum; m3++) {
        const long m2 = m3 % ynumy, m1 = (m3 - m2) / ynumy;
        if (!(m1 >= 1 && m2 >= 1 && m1 < xnumx - 1 && m2 < ynumy - 1)) continue;
        const long a = m3, b = m3 + 1, c = m3 + ynumy, d = m3 + ynumy + 1;
        N(exxs)[m3] = 0.25 * (N(exx)[a] + N(exx)[b] + N(exx)[c] + N(exx)[d]);
        const double e1 = epl_at(s, a, dt), e2 = epl_at(s, b, dt), e3 = epl_at(s, c, dt), e4 = epl_at(s, d, dt);
        N(exxpl)[m3] = 0.25 * (e1 + e2 + e3 + e4);
        N(sxxs)[m3] = 0.25 * (N(sxx)[a] + N(sxx)[b] + N(sxx)[c] + N(sxx)[d]);
        N(prs)[m3] = 0.25 * (N(pr)[a] + N(pr)[b] + N(pr)[c] + N(pr)[d]);
        const double l1 = exxel_at(s, a, dt), l2 = exxel_at(s, b, dt), l3 = exxel_at(s, c, dt),
                     l4 = exxel_at(s, d, dt);
        N(exxel)[m3] = 0.25 * (l1 + l2 + l3 + l4);
    }
}

void or_interpolate_basic2center(or_state *s)
{
    const v2c_grid *g = &s->grid;
    const long xnumx = g->xnumx, ynumy = g->ynumy, nodenum = xnumx * ynumy;
    for (long m3 = 0; m3 < nodenum; m3++) {
        const long m2 = m3 % ynumy, m1 = (m3 - m2) / ynumy;
        if (!(m1 >= 1 && m2 >= 1)) continue;
        const double x = (g->gx[m1] + g->gx[m1 - 1]) / 2, y = (g->gy[m2] + g->gy[m2 - 1]) / 2;
        long m10 = or_nxsearch(g, x), m20 = or_nysearch(g, y);
        if (m10 < 1) m10 = 1;
        if (m10 > xnumx - 3) m10 = xnumx - 3;
        if (m20 < 1) m20 = 1;
        if (m20 > ynumy - 3) m20 = ynumy - 3;
        double e = (x - g->gx[m10]) / (g->gx[m10 + 1] - g->gx[m10]);
        double n = (y - g->gy[m20]) / (g->gy[m20 + 1] - g->gy[m20]);
        if (e < 0) e = 0.0;
        if (n < 0) n = 0.0;
        if (e > 1) e = 1.0;
        if (n > 1) n = 1.0;
        const double dm0 = (1.0 - e) * (1.0 - n), dm1 = (1.0 - e) * n, dm2 = e * (1.0 - n), dm3 = e * n;
        const long m4 = m10 * ynumy + m20;
        N(nd)[m3] = 1.0 / (dm0 * 1.0 / N(nu)[m4] + dm1 * 1.0 / N(nu)[m4 + 1] + dm2 * 1.0 / N(nu)[m4 + ynumy] +
                           dm3 * 1.0 / N(nu)[m4 + ynumy + 1]);
    }
}

void or_copyarrays(or_state *s, int rk)
{
    (void)rk;
    const long xnumx = s->grid.xnumx, ynumy = s->grid.ynumy, nodenum = xnumx * ynumy;
    for (long m3 = 0; m3 < nodenum; m3++) {
        const long m2 = m3 % ynumy, m1 = (m3 - m2) / ynumy;
        if (m1 >= 1 && m2 >= 1 && m1 < xnumx - 1 && m2 < ynumy - 1) {
            N(deltaVs)[m3] = (log(N(RK_sliprs_solved)[m3]) - log(N(RK_sliprs_final)[m3])) / s->timestep;
            N(sliprs)[m3] = N(RK_sliprs_final)[m3] = N(RK_sliprs_solved)[m3];
            N(RK_phidts_final)[m3] = N(RK_phidts_solved)[m3];
        }
    }
}

int or_calc_timestep(or_state *s, int oldornew, const v2c_timestep_params *p, v2c_timestep_result *r)
{
    const v2c_grid *g = &s->grid;
    const long xnumx = g->xnumx, ynumy = g->ynumy, nodenum = xnumx * ynumy;
    double maxvel = 0.0, maxVx = 0.0;
    double dt_f = p->maxtmstep;
    double dtnvp = 1e+30, dtd = 1e+30, dtphi = 1e+30, dtL = 1e+30;
    double lambda = p->lambda0;          /* the reference leaves it uninitialised */
    for (long m3 = 0; m3 < nodenum; m3++) {
        const long m2 = m3 % ynumy, m1 = (m3 - m2) / ynumy;
        if (!(m1 >= 1 && m2 >= 1 && m1 < xnumx - 1 && m2 < ynumy - 1)) continue;
        const double dx = g->gx[m1 + 1] - g->gx[m1], dy = g->gy[m2 + 1] - g->gy[m2];
        const double faultwidth = 2 * 0.5 * (dx + dy);
        double NUVP;
        if (oldornew == 0) {
            if (N(nu_m2c)[m3] > 1.00e+26) {
                const double sigin = pow(N(sxxes)[m3] * N(sxxes)[m3] + N(sxye)[m3] * N(sxye)[m3], 0.5);
                if (sigin > 0 && N(RK_sliprs_final)[m3] > 0)
                    NUVP = N(nu_m2c)[m3] * sigin /
                           (2 * N(nu_m2c)[m3] * N(RK_sliprs_final)[m3] / faultwidth + sigin);
                else
                    NUVP = N(nu_m2c)[m3];
            } else {
                NUVP = 5.00e+26;
            }
            dtnvp = minv(dtnvp, p->tcnvp * NUVP / N(gg)[m3]);
        } else {
            NUVP = N(nu_m2c)[m3] > 1.00e+26 ? N(nu)[m3] : 5.00e+26;
            dtnvp = minv(dtnvp, p->tcnvp * NUVP / N(gg)[m3]);
        }
        if (oldornew == 1 && N(vx)[m3] && N(vy)[m3]) {
            maxVx = maxv(maxVx, maxv(absv(N(vx)[m3]), absv(N(vy)[m3])));
            const double maxdr = maxv(absv(N(vx)[m3] / dx), absv(N(vy)[m3] / dy));
            dtd = minv(dtd, p->tcd * 1.0 / maxdr);
        } else if (oldornew == 0 && N(mvx)[m3] && N(mvy)[m3]) {
            maxVx = maxv(maxVx, maxv(absv(N(mvx)[m3]), absv(N(mvy)[m3])));
            const double maxdr = maxv(absv(N(mvx)[m3] / dx), absv(N(mvy)[m3] / dy));
            dtd = minv(dtd, p->tcd * 1.0 / maxdr);
        }
        if (s->prm.RSFmode == 1) {
            double nps;
            if (s->prm.pmode == 0) nps = s->prm.nstress;
            else if (s->prm.pmode == 1) nps = oldornew == 0 ? N(sppes)[m3] : N(prs)[m3];
            else return -2;
            const double bm = N(bm)[m3], gg = N(gg)[m3], Ls = N(Ls)[m3], as = N(as)[m3], bs = N(bs)[m3];
            const double v = (3 * bm - 2 * gg) / (2 * (3 * bm + gg));
            const double k = 2.0 / 3.14 * (gg / (1 - v)) / s->prm.xstpx;
            const double chi = 1.0 / 4.0 * pow(k * Ls / (as * nps) - (bs - as) / as, 2) - k * Ls / (as * nps);
            if (chi > 0) lambda = as * nps / (k * Ls - (bs - as) * nps);
            else if (chi < 0) lambda = 1 - (bs - as) * nps / (k * Ls);
            else if (chi == 0) return -1;
            lambda = minv(lambda, 1.0 / 5.0);
            if (bs > 0) {
                if (oldornew == 0 && N(RK_sliprs_final)[m3] > 0) {
                    const double phisc = N(V0s)[m3] / Ls * exp(-N(phis)[m3]);
                    dtphi = minv(dtphi, p->tcphih * 1.0 / phisc);
                    dtL = minv(dtL, lambda * Ls / (N(RK_sliprs_final)[m3]));
                } else if (oldornew == 1 && N(RK_sliprs_solved)[m3] > 0) {
                    const double phisc = N(V0s)[m3] / Ls * exp(-N(phis_it)[m3]);
                    dtphi = minv(dtphi, p->tcphih * 1.0 / phisc);
                    dtL = minv(dtL, lambda * Ls / (N(RK_sliprs_solved)[m3]));
                }
            } else {
                dtphi = p->maxtmstep;
                dtL = p->maxtmstep;
            }
        }
        if (oldornew == 0 && N(RK_sliprs_final)[m3] > 0) maxvel = maxv(maxvel, N(RK_sliprs_final)[m3]);
        if (oldornew == 1 && N(RK_sliprs_solved)[m3] > 0) maxvel = maxv(maxvel, N(RK_sliprs_solved)[m3]);
    }
    dt_f = p->dtfactor * minv(dt_f, minv(dtphi, minv(minv(dtL, dtnvp), dtd)));
    r->dt = dt_f;
    r->dtnvp = dtnvp; r->dtd = dtd; r->dtphi = dtphi; r->dtL = dtL;
    r->maxvel = maxvel; r->maxVx = maxVx;
    return 0;
}

/* interpolate_temperature (interpolation.c:1492-1533) */
static void interp_temperature(const or_state *s, double x, double y, double *TK, double *TK2)
{
    const v2c_grid *g = &s->grid;
    if (x < 0) x = 0; else if (x > g->xsize) x = g->xsize;
    if (y < 0) y = 0; else if (y > g->ysize) y = g->ysize;
    long m10 = or_nxsearch(g, x), m20 = or_nysearch(g, y);
    if (m10 < 0) m10 = 0;
    if (m10 > g->xnumx - 2) m10 = g->xnumx - 2;
    if (m20 < 0) m20 = 0;
    if (m20 > g->ynumy - 2) m20 = g->ynumy - 2;
    const double e = (x - g->gx[m10]) / (g->gx[m10 + 1] - g->gx[m10]);
    const double n = (y - g->gy[m20]) / (g->gy[m20 + 1] - g->gy[m20]);
    const double dm0 = (1.0 - e) * (1.0 - n), dm1 = (1.0 - e) * n, dm2 = e * (1.0 - n), dm3 = e * n;
    const long m3 = m10 * g->ynumy + m20, ny = g->ynumy;
    const double *tk = N(tk), *tk2 = N(tk2);
    *TK = dm0 * tk[m3] + dm1 * tk[m3 + 1] + dm2 * tk[m3 + ny] + dm3 * tk[m3 + ny + 1];
    *TK2 = dm0 * tk2[m3] + dm1 * tk2[m3 + 1] + dm2 * tk2[m3 + ny] + dm3 * tk2[m3 + ny + 1];
}

static int inside(const v2c_grid *g, double x, double y)
{
    return x > 0 && y > 0 && x < g->xsize && y < g->ysize;
}

/* The marker part of solve_temperature (temperature_iterate.c:181-419),
 * single-threaded in marker order.  The reference's extra iteration of its
 * m3<=marknum loops visits a zero marker at x = 0, which the inside test
 * skips, so the loops run over the n markers. */
void or_temperature_markers(or_state *s, double heatdif)
{
    const v2c_grid *g = &s->grid;
    const long ny = g->ynumy, nodenum = g->xnumx * ny;
    double *x = s->m.x, *y = s->m.y, *mk = s->m.k;
    const char *t = s->m.t;
    double *tol0 = N(tol0), *tol1 = N(tol1);
    for (long m1 = 0; m1 < nodenum; m1++) tol0[m1] = tol1[m1] = 0;
    double TK, TK2;
    for (long m3 = 0; m3 < s->n; m3++)
        if (inside(g, x[m3], y[m3]) && t[m3] >= 50 && t[m3] < 100) {
            interp_temperature(s, x[m3], y[m3], &TK, &TK2);
            mk[m3] = TK;
        }
    for (long m3 = 0; m3 < s->n; m3++) {
        if (!(inside(g, x[m3], y[m3]) && t[m3] < 50)) continue;
        const long m10 = or_nxsearch(g, x[m3]), m20 = or_nysearch(g, y[m3]);
        interp_temperature(s, x[m3], y[m3], &TK, &TK2);
        if (mk[m3] <= 0) {
            mk[m3] = TK2;
            if (mk[m3] < s->prm.tdeep) mk[m3] = s->prm.tdeep;
        }
        if (heatdif) {
            const double dtk = TK2 - mk[m3];
            double celdx = g->gx[m10 + 1] - g->gx[m10], celdy = g->gy[m20 + 1] - g->gy[m20];
            const double swt1 = 1.0 / celdx / celdy;
            celdx = (x[m3] - g->gx[m10]) / (g->gx[m10 + 1] - g->gx[m10]);
            celdy = (y[m3] - g->gy[m20]) / (g->gy[m20 + 1] - g->gy[m20]);
            double mkt = 0, mrocp = 0;
            for (long m1 = m10; m1 <= m10 + 1; m1++)
                for (long m2 = m20; m2 <= m20 + 1; m2++) {
                    const long m4 = m1 * ny + m2;
                    double ival = 0;
                    if (m1 == m10 && m2 == m20) ival = (1.0 - celdx) * (1.0 - celdy);
                    else if (m1 == m10 + 1 && m2 == m20) ival = (celdx) * (1.0 - celdy);
                    else if (m1 == m10 && m2 == m20 + 1) ival = (1.0 - celdx) * (celdy);
                    else ival = (celdx) * (celdy);
                    mkt += N(kt)[m4] * ival;
                    mrocp += N(cp)[m4] * N(ro)[m4] * ival;
                }
            celdx = g->gx[m10 + 1] - g->gx[m10];
            celdy = g->gy[m20 + 1] - g->gy[m20];
            double dtk1 = -heatdif * mkt / mrocp * 2.0 * (1.0 / celdx / celdx + 1.0 / celdy / celdy) * s->timestep;
            if (dtk1 < -150.0) dtk1 = -150.0;
            dtk1 = dtk * (1.0 - exp(dtk1));
            mk[m3] += dtk1;
            celdx = (x[m3] - g->gx[m10]) / (g->gx[m10 + 1] - g->gx[m10]);
            celdy = (y[m3] - g->gy[m20]) / (g->gy[m20 + 1] - g->gy[m20]);
            const long m4 = m10 * ny + m20;
            const double ival = (1.0 - celdx) * (1.0 - celdy) * swt1;
            tol0[m4] += ival;
            tol1[m4] += dtk1 * ival;
        }
        mk[m3] += (TK - TK2);
    }
    if (heatdif) {
        for (long m1 = 0; m1 < nodenum; m1++)
            if (tol0[m1]) tol1[m1] /= tol0[m1];
        for (long m3 = 0; m3 < s->n; m3++) {
            if (!(inside(g, x[m3], y[m3]) && t[m3] < 50)) continue;
            const long m10 = or_nxsearch(g, x[m3]), m20 = or_nysearch(g, y[m3]);
            const double celdx = (x[m3] - g->gx[m10]) / (g->gx[m10 + 1] - g->gx[m10]);
            const double celdy = (y[m3] - g->gy[m20]) / (g->gy[m20 + 1] - g->gy[m20]);
            const long m4 = m10 * ny + m20;
            if (tol1[m4]) mk[m3] -= (tol1[m4] * ((1.0 - celdx) * (1.0 - celdy)));
        }
    }
}

/* ---- boundary-condition tail ------------------------------------------------
 * Restatement of the reference's row-by-row band elimination
 * (gausmat3, gausevps.c:216-385): each added row is reduced by the earlier
 * rows it refers to (fill-in included) and stored normalised; the solve is a
 * back substitution from the last row.  Rows are added in the routines' node
 * order; an unknown without a row solves to 0.  Returns 0, -1 where a row
 * refers to an absent earlier row, -2 for a zero diagonal (where the
 * reference exits). */
typedef struct {
    long n;
    double *bufv, *fre, *sol, *val;
    long *lin, *pos, *num;
    long cur, cap;
} band_t;

static int band_init(band_t *B, long n)
{
    B->n = n;
    B->bufv = calloc(n, sizeof(double));
    B->fre = calloc(n, sizeof(double));
    B->sol = calloc(n, sizeof(double));
    B->pos = calloc(n, sizeof(long));
    B->num = calloc(n, sizeof(long));
    B->cap = 8 * n + 16;
    B->val = malloc(B->cap * sizeof(double));
    B->lin = malloc(B->cap * sizeof(long));
    B->cur = 0;
    return B->bufv && B->fre && B->sol && B->pos && B->num && B->val && B->lin ? 0 : -3;
}

static void band_free(band_t *B)
{
    free(B->bufv); free(B->fre); free(B->sol); free(B->pos); free(B->num); free(B->val); free(B->lin);
}

/* row r: r's own coefficient 1 first, then nw couplings (w[i], -k[i]) */
static int band_add(band_t *B, long r, double rhs, int nw, const long *w, const double *k)
{
    long lo = r, hi = r;
    if (nw > 0) hi = w[nw - 1];        /* the reference starts from the last listed parameter */
    if (hi < r) hi = r;                /* ... and widens the range by every nonzero entry, r included */
    B->fre[r] = rhs;
    B->bufv[r] += 1.0;
    for (int i = 0; i < nw; i++) {
        B->bufv[w[i]] += -k[i];
        if (-k[i]) {
            if (lo > w[i]) lo = w[i];
            if (hi < w[i]) hi = w[i];
        }
    }
    for (long m2 = lo; m2 < r; m2++) {
        if (!B->bufv[m2]) continue;
        const double iv = B->bufv[m2];
        B->bufv[m2] = 0;
        if (!B->num[m2]) return -1;
        for (long m3 = 1; m3 < B->num[m2]; m3++) B->bufv[B->lin[B->pos[m2] + m3]] -= iv * B->val[B->pos[m2] + m3];
        B->fre[r] -= B->fre[m2] * iv;
        const long last = B->lin[B->pos[m2] + B->num[m2] - 1];
        if (hi < last) hi = last;
    }
    if (!B->bufv[r]) return -2;
    B->pos[r] = B->cur;
    B->num[r] = 0;
    const double iv = B->bufv[r];
    for (long m2 = r; m2 <= hi; m2++) {
        if (B->bufv[m2]) {
            if (B->cur >= B->cap) {
                B->cap *= 2;
                B->val = realloc(B->val, B->cap * sizeof(double));
                B->lin = realloc(B->lin, B->cap * sizeof(long));
            }
            B->lin[B->cur] = m2;
            B->val[B->cur] = B->bufv[m2] / iv;
            B->cur++;
            B->num[r]++;
        }
        B->bufv[m2] = 0;
    }
    B->fre[r] /= iv;
    return 0;
}

static void band_solve(band_t *B)
{
    for (long m1 = B->n - 1; m1 >= 0; m1--) {
        if (B->num[m1]) {
            double iv = B->fre[m1];
            for (long m2 = 0; m2 < B->num[m1]; m2++) iv -= B->val[B->pos[m1] + m2] * B->sol[B->lin[B->pos[m1] + m2]];
            B->sol[m1] = iv;
        } else {
            B->sol[m1] = 0;
        }
    }
}

/* A bounded row: its constant and the listed couplings, each parameter
 * located by map(q) from bondn = q. */
static int bc_listed(const long *bondn, const double *bondv, long k, long (*map)(long, long, long), long nn,
                     long u, long *w, double *kv)
{
    int nw = 0;
    for (int i = 0; i < 3; i++) {
        const long q = bondn[3 * k + i];
        if (!q) continue;
        w[nw] = map(q, nn, u);
        kv[nw] = bondv[4 * k + i + 1];
        nw++;
    }
    return nw;
}

static long map_vx(long q, long nn, long u) { (void)nn; (void)u; return ((q - 1) - 1) / 3; }
static long map_vy(long q, long nn, long u) { (void)u; return ((q - 1) - 2) / 3 + nn; }
static long map_p(long q, long nn, long u) { (void)nn; (void)u; return (q - 1) / 3; }
static long map_t(long q, long nn, long u) { (void)u; return q - 1 - 3 * nn; }

int or_vxvy_boundarycorrection(or_state *s, const long *bondm, const long *bondn, const double *bondv)
{
    const long nx = s->grid.xnumx, ny = s->grid.ynumy, nn = nx * ny;
    band_t B;
    if (band_init(&B, 2 * nn)) return -3;
    double *mvx = N(mvx), *mvy = N(mvy);
    int rc = 0;
    long w[3];
    double kv[3];
    for (long m1 = 0; m1 < nx && !rc; m1++)
        for (long m2 = 0; m2 < ny && !rc; m2++) {
            const long m3 = m1 * ny + m2;
            if (m2 < ny - 1) {
                const long k = bondm[3 * m3 + 1];
                if (!k) rc = band_add(&B, m3, mvx[m3], 0, w, kv);
                else rc = band_add(&B, m3, bondv[4 * k], bc_listed(bondn, bondv, k, map_vx, nn, m3, w, kv), w, kv);
            }
            if (!rc && m1 < nx - 1) {
                const long k = bondm[3 * m3 + 2];
                if (!k) rc = band_add(&B, m3 + nn, mvy[m3], 0, w, kv);
                else rc = band_add(&B, m3 + nn, bondv[4 * k], bc_listed(bondn, bondv, k, map_vy, nn, m3, w, kv), w, kv);
            }
        }
    if (!rc) {
        band_solve(&B);
        for (long m3 = 0; m3 < nn; m3++) {
            const long m2 = m3 % ny, m1 = m3 / ny;
            mvx[m3] = m2 < ny - 1 ? B.sol[m3] : 0.0;
            mvy[m3] = m1 < nx - 1 ? B.sol[m3 + nn] : 0.0;
        }
    }
    band_free(&B);
    return rc;
}

int or_P_boundarycorrection(or_state *s, const long *bondm, const long *bondn, const double *bondv)
{
    const long nx = s->grid.xnumx, ny = s->grid.ynumy, nn = nx * ny;
    band_t B;
    if (band_init(&B, nn)) return -3;
    double *sppe = N(sppe);
    int rc = 0;
    long w[3];
    double kv[3];
    for (long m1 = 0; m1 < nx && !rc; m1++)
        for (long m2 = 0; m2 < ny && !rc; m2++) {
            const long m3 = m1 * ny + m2;
            if (!(m1 && m2)) continue;
            const long k = bondm[3 * m3];
            if (!k || s->prm.compressiblemode == 1) rc = band_add(&B, m3, sppe[m3], 0, w, kv);
            else rc = band_add(&B, m3, bondv[4 * k], bc_listed(bondn, bondv, k, map_p, nn, m3, w, kv), w, kv);
        }
    if (!rc) {
        band_solve(&B);
        for (long m3 = 0; m3 < nn; m3++) {
            const long m2 = m3 % ny, m1 = m3 / ny;
            sppe[m3] = (m1 && m2) ? B.sol[m3] : 0.0;
        }
    }
    band_free(&B);
    return rc;
}

int or_tkrecalc(or_state *s, const long *bondm, const long *bondn, const double *bondv)
{
    const long nx = s->grid.xnumx, ny = s->grid.ynumy, nn = nx * ny;
    band_t B;
    if (band_init(&B, nn)) return -3;
    double *tk = N(tk);
    int rc = 0;
    long w[3];
    double kv[3];
    for (long m1 = 0; m1 < nx && !rc; m1++)
        for (long m2 = 0; m2 < ny && !rc; m2++) {
            const long m3 = m1 * ny + m2;
            const long k = bondm[3 * nn + m3];
            if (!k) rc = band_add(&B, m3, tk[m3], 0, w, kv);
            else rc = band_add(&B, m3, bondv[4 * k], bc_listed(bondn, bondv, k, map_t, nn, m3, w, kv), w, kv);
        }
    if (!rc) {
        band_solve(&B);
        for (long m3 = 0; m3 < nn; m3++) tk[m3] = B.sol[m3];
    }
    band_free(&B);
    return rc;
}

/* interpolation.c:121 clears the field over every node; :866-916 then marks
 * in marker order, each write replacing the last. */
void or_sticky_marks(or_state *s)
{
    const long ny = s->grid.ynumy, nodenum = (long)s->grid.xnumx * ny;
    const double *gx = s->grid.gx, *gy = s->grid.gy;
    double *st = N(sticky01);
    for (long m3 = 0; m3 < nodenum; m3++) st[m3] = 0;
    for (long mm1 = 0; mm1 < s->n; mm1++) {
        const int t = s->m.t[mm1];
        if (t != 1 && t != 2) continue;
        const long nx = or_nxsearch(&s->grid, s->m.x[mm1]);
        const long nyc = or_nysearch(&s->grid, s->m.y[mm1]);
        const long m3 = nx * ny + nyc + 1;
        const double celdx = (s->m.x[mm1] - gx[nx]) / (gx[nx + 1] - gx[nx]);
        const double celdy = (s->m.y[mm1] - gy[nyc]) / (gy[nyc + 1] - gy[nyc]);
        long k = -1;
        if (celdx < 0.5 && celdy < 0.5) k = m3;
        else if (celdx < 0.5 && celdy > 0.5) k = m3 + 1;
        else if (celdx > 0.5 && celdy < 0.5) k = m3 + ny;
        else if (celdx > 0.5 && celdy > 0.5) k = m3 + ny + 1;
        /* indices at or past nodenum land beyond every array the reference
         * reads (its arrays are MAXNOD long) */
        if (k >= 0 && k < nodenum) st[k] = t;
    }
}

/* calc_RSF.c:364-374 (initial values) and :505-550 (per interior node, in
 * node order); the std members are computed by host code elsewhere. */
void or_calc_sticky_stats(or_state *s, v2c_sticky_stats *o)
{
    const long ny = s->grid.ynumy, nx = s->grid.xnumx, nodenum = nx * ny;
    const double *sxxs = N(sxxs), *sxy = N(sxy), *nu = N(nu), *st = N(sticky01);
    o->nr_1 = o->nr_0 = 0;
    o->max_sigin_1 = o->max_sigin_0 = 0.0;
    o->min_sigin_1 = o->min_sigin_0 = 1e30;
    o->sum_sigin_1 = o->sum2_sigin_1 = 0.0;
    o->sum_sigin_0 = o->sum2_sigin_0 = 0.0;
    o->max_visc_1 = o->max_visc_0 = 0.0;
    o->min_visc_1 = o->min_visc_0 = 1e30;
    o->sum_visc_1 = o->sum2_visc_1 = 0.0;
    o->sum_visc_0 = o->sum2_visc_0 = 0.0;
    o->sigin_P1 = o->sigin_P2 = 0.0;
    o->visco_P1 = o->visco_P2 = 0.0;
    for (long m3 = 0; m3 < nodenum; m3++) {
        const long m2 = m3 % ny, m1 = (m3 - m2) / ny;
        if (!(m1 >= 1 && m2 >= 1 && m1 < nx - 1 && m2 < ny - 1)) continue;
        const double sigin = pow(sxxs[m3] * sxxs[m3] + sxy[m3] * sxy[m3], 0.5);
        if (st[m3] == 1 && m1 == 50 && m2 == 50) { o->sigin_P1 = sigin; o->visco_P1 = nu[m3]; }
        if (st[m3] == 1 && m1 == 500 && m2 == 50) { o->sigin_P2 = sigin; o->visco_P2 = nu[m3]; }
        if (st[m3] == 2) {
            o->nr_1 = o->nr_1 + 1;
            o->max_sigin_1 = maxv(o->max_sigin_1, sigin);
            o->min_sigin_1 = minv(o->min_sigin_1, sigin);
            o->sum_sigin_1 += sigin;
            o->sum2_sigin_1 += pow(sigin, 2);
            o->max_visc_1 = maxv(o->max_visc_1, nu[m3]);
            o->min_visc_1 = minv(o->min_visc_1, nu[m3]);
            o->sum_visc_1 += log10(nu[m3]);
            o->sum2_visc_1 += pow(log10(nu[m3]), 2);
        }
        if (st[m3] == 1) {
            o->nr_0 = o->nr_0 + 1;
            o->max_sigin_0 = maxv(o->max_sigin_0, sigin);
            o->min_sigin_0 = minv(o->min_sigin_0, sigin);
            o->sum_sigin_0 += sigin;
            o->sum2_sigin_0 += pow(sigin, 2);
            o->max_visc_0 = maxv(o->max_visc_0, nu[m3]);
            o->min_visc_0 = minv(o->min_visc_0, nu[m3]);
            o->sum_visc_0 += log10(nu[m3]);
            o->sum2_visc_0 += pow(log10(nu[m3]), 2);
        }
    }
}
