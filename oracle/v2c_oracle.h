/*
 * v2c_oracle.h -- FP64 CPU oracle of the per-marker hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is compared
 * against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; the product library (libv2c) never links or calls it.
 *
 * It restates the reference algorithm single-threaded, in the reference's own
 * operation order, so that it reproduces a single-threaded run of the
 * reference routines bit for bit (the reference is built with gcc, no FMA
 * contraction; this file is built with -ffp-contract=off).
 */
#ifndef V2C_ORACLE_H
#define V2C_ORACLE_H
#include "../include/v2c.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_state {
    v2c_grid grid;
    v2c_params prm;
    const v2c_materials *mat;
    double timesum, timestep;
    v2c_markers m;          /* caller-owned marker arrays                    */
    long n;                 /* marker count                                  */
    double *nod[V2C_NF];    /* caller-owned nodal arrays, length nodenum     */
} or_state;

long or_nxsearch(const v2c_grid *g, double x);
long or_nysearch(const v2c_grid *g, double y);
void or_cells(const v2c_grid *g, long n, const double *x, const double *y, long *nx, long *ny);
double or_allinterpomp(const or_state *s, double x, double y);
double or_ductilevisc(const or_state *s, int type, double sxy, double sxx,
                      double exy, double exx, double tk, double pb, int *iters);
/* Build-defined NEWTON mode of the ductile solve: same residual, exact
 * derivative, stop at |f|/eta <= tol.  `guess` <= 0 selects the reference's
 * default start.  Returns the clamped effective viscosity. */
double or_ductilevisc_newton(const or_state *s, int type, double sxy,
                             double sxx, double tk, double pb, double guess,
                             double tol, int *iters, double *resid);

double or_ductile_newton_ex(const or_state *s, int t, double sxy, double sxx, double mtk, double mpb,
                            double guess, double tol, int *iters, int *evals, int *fallback, double *root);
int or_rsf_marker_ex(double sig, double nP, double a, double b, double L, double V0, double mu0,
                     double phi, double Vprev, double dt, double eta_rad, double EII, double W,
                     double guessV, double tol, double *V_out, double *phi_out, int *fallback,
                     double *root, int *evals);
int or_rsf_marker(double sig, double nP, double a, double b, double L, double V0, double mu0,
                  double phi, double Vprev, double dt, double eta_rad, double EII, double W,
                  double guessV, double tol, double *V_out, double *phi_out, int *fallback);

void or_local_update(or_state *s, double dt, const unsigned char *fault, double *mx, int warm,
                     double tol, long *stats, int *ditr, int *ritr);

void or_marker2nodes(or_state *s);
void or_nodes2markers(or_state *s);
long or_moveandrotatemarkers(or_state *s);   /* returns #markers of type >= 100 */
long or_compute_before_solve(or_state *s, double dt, int rk);   /* returns #err */
double or_compute_after_solve(or_state *s, double dt, int rk, long *nerr);
void or_allinteri(const or_state *s, double x, double y, double *vx,
                  double *vy, double *esp, double *ee);

/* post-solve nodal routines (v2c_oracle_post.c) */
void or_recompute_stress_strainrate(or_state *s, double dt, int rk, double *stats /* 5, may be NULL */);
void or_interpolate_center2basic(or_state *s, double dt);
void or_interpolate_center2basic_old(or_state *s);
void or_interpolate_basic2center(or_state *s);
void or_copyarrays(or_state *s, int rk);
/* 0, or -1 where chi == 0 (the reference exits), -2 for an unknown pmode */
int or_calc_timestep(or_state *s, int oldornew, const v2c_timestep_params *p, v2c_timestep_result *r);
/* sticky-air marks of marker2nodes and statistics of calc_timestep */
void or_sticky_marks(or_state *s);
void or_calc_sticky_stats(or_state *s, v2c_sticky_stats *o);
/* marker temperatures after the heat solve */
void or_temperature_markers(or_state *s, double heatdif);
/* boundary-condition tail: 0, or -1 / -2 where the reference exits */
int or_vxvy_boundarycorrection(or_state *s, const long *bondm, const long *bondn, const double *bondv);
int or_P_boundarycorrection(or_state *s, const long *bondm, const long *bondn, const double *bondv);
int or_tkrecalc(or_state *s, const long *bondm, const long *bondn, const double *bondv);

#ifdef __cplusplus
}
#endif
#endif
