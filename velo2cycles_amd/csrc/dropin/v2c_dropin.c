/*
 * v2c_dropin.c -- reference-side binding of the VELO2CYCLEs hot path to
 * libv2c.  A maintainer compiles this file into the reference translation unit
 * in place of the hot-path routines, so every call site stays as it is:
 *
 *   marker2nodes()                  i2evps.c:89, in2evps.c:1345
 *   nodes2markers()                 i2evps.c:108
 *   moveandrotatemarkers()          i2evps.c:113
 *   compute_before_solve()          momentum_continuity_iterate.c:69
 *   compute_after_solve()           momentum_continuity_iterate.c:89
 *   recompute_stress_strainrate()   momentum_continuity_iterate.c:85
 *   interpolate_center2basic()      momentum_continuity_iterate.c:87
 *   interpolate_center2basic_old()  i2evps.c:90
 *   interpolate_basic2center()      momentum_continuity_iterate.c:71
 *   copyarrays()                    momentum_continuity_iterate.c:128
 *   calc_timestep()                 momentum_continuity_iterate.c:44, 92
 *                                   (with its sticky-air statistics)
 *   vxvy_boundarycorrection(), P_boundarycorrection(), tkrecalc()
 *                                   interpolation.c:919-924, temperature_iterate.c
 *
 * The reference globals (headevps.c:33-250, input_RSF.c:3-11) are declared
 * extern here with the reference's own types; the file owns no state beyond
 * the device context.  The device owns the marker cloud between
 * v2c_dropin_attach() (after load_model/extra_init) and
 * v2c_dropin_markers_to_host() (before save_model and solve_temperature,
 * which read markers on the host).  The boundary-condition tail of
 * marker2nodes (interpolation.c:918-925) runs on the device inside
 * marker2nodes().
 *
 * Nodal arrays are resident on both sides, and the binding tracks which side
 * holds the newer copy of each.  A routine sends, in one staged transfer,
 * only the arrays it reads that host code changed since they last crossed,
 * and leaves its results on the device.  Host code that reads nodal arrays
 * asks for them first, and host code that writes them says so afterwards:
 *
 *   v2c_dropin_before_stokes()   before preparematrix
 *                                (momentum_continuity_iterate.c:76): brings
 *                                back what the Stokes assembly, solve, reload
 *                                and checkerror read
 *   v2c_dropin_after_stokes()    after reloadVxVyP (:83): pr, vx, vy changed
 *   v2c_dropin_before_output()   before the sticky-air diagnostics (:129-170)
 *                                and postproc (i2evps.c:131): the arrays they
 *                                read each step
 *   v2c_dropin_host_reads(f, n) / v2c_dropin_host_wrote(f, n)
 *                                any other host code, by V2C_N_* field ids
 *   v2c_dropin_sync_to_host()    every array the device changed (before the
 *                                heat solve, hdf5 output, save_model)
 *   v2c_dropin_host_wrote_all()  after host code that may change any array
 *
 * A step of the reference loop then moves pr, vx, vy to the device and the
 * 18 arrays the host reads back, instead of every array each routine
 * touches.
 *
 * solve_temperature (temperature_iterate.c:2-419) keeps its heat solve on
 * the host; its marker part (:181-419) is replaced by a call to
 * v2c_dropin_temperature_markers(heatdif), which updates the device-resident
 * markers, so the markers need not come to the host for the heat step.
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../../include/v2c.h"

/* ---- reference globals (declared in headevps.c / input_RSF.c) ------------ */
extern long int xnumx, ynumy, marknum, nodenum;
extern double xsize, ysize, xstpx, timesum, timestep, errcheck, nstress, Dphis;
extern double gx[], gy[];
extern char markt[];
extern double markx[], marky[], markk[], markp[], markxx[], markxy[], markexx[], markexy[];
extern double marke[], markphi[], markvx[], markvy[], markC[], markd[], markw[];
extern double mark_a[], mark_b[], mark_L[], mark_L0[], mark_V0[], fric_static[], markC0[];
extern double marknu[], markdh[], markdv[], markss[], markmm[], markgg[], markn0[], markn1[];
extern double markro[], markbb[], markaa[], markcp[], markkt[], markkf[], markkp[], markht[];
extern int markim[];
extern int markmod, stoksmod, outgrid, densimod, interpolationmod, antidiffinter, rkmod, RSFmode;
extern int power_law_RSF, reloadfromDP, compressiblemode, stickymod, inertyn, pmode, tempmod;
extern double nubeg, nuend, tdeep, zdeep, eps[];
extern double half_range, start_updip, end_updip, start_downdip, end_downdip;
extern double a_vw, b_vw, L_vw, V0_vw, a_vs, b_vs, L_vs, V0_vs, a_hr, b_hr, L_hr, V0_hr;
extern double RK_sliprs[][2], RK_phidts[][2];
extern double pr[], vx[], vy[], exx[], exy[], esp[], sxx[], sxy[], sxxs[], sppes[], exxpl[], exxel[];
extern double bm[], ro[], et[], nu[], nd[], gg[], gd[], sppe[], phis[], phin[], sxyen[], sxxen[];
extern double exyen[], exxen[], sppen[], mus_s[], as[], bs[], Ls[], V0s[], sxxe[], sxye[], exxe[];
extern double exye[], cohesion[], dro[], drp[], cp[], kt[], ht[], tk[], mrx[], mry[], mvx[], mvy[];
extern double pe[], pen[], nu_m2c[], nd_m2c[], mvxn[], mvyn[];
extern double phie[], sxyee[], exyee[], pee[], sppee[], sxxee[], exxee[], mvxe[], mvye[];
extern double wtphisum[], phisum[], sxysum[], exysum[], esum[], wtsxxsum[], psum[], sxxsum[];
extern double exxsum[], wtvxsum[], vxsum[], wtvysum[], vysum[], mrx0[], mry0[];
extern double pes[], phis_it[], syields[], nuvp_it[], sliprs[], RK_sliprs_final[], RK_phidts_final[];
extern double RK_sliprs_solved[], RK_phidts_solved[], RK_deltas[];
extern double deltasxx[], deltasxy[], deltaVs[], exxs[], exxes[], sxxes[], prs[], tk2[], tol0[], tol1[];
extern double sticky01[];
extern int nr_1, nr_0;
extern double max_sigin_1, max_sigin_0, min_sigin_1, min_sigin_0, sum_sigin_1, sum2_sigin_1, sum_sigin_0;
extern double sum2_sigin_0, max_visc_1, max_visc_0, min_visc_1, min_visc_0, sum_visc_1, sum2_visc_1;
extern double sum_visc_0, sum2_visc_0, sigin_P1, sigin_P2, visco_P1, visco_P2;
extern long int bondm[], bondn[][3], bondnum;
extern double bondv[][4];
extern double maxvel, maxdphi, maxnvp, dtL, dtphi, dtd, dtnvp, maxtmstep, tcnvp, tcd, tcphih, dtfactor;

static v2c_ctx *g_ctx;

static double *nodal_ptr(int f)
{
    double *const tab[] = {
        pr, vx, vy, exx, exy, esp, sxx, sxy, sxxs, sppes, exxpl, exxel,
        bm, ro, et, nu, nd, gg, gd, sppe, phis, phin, sxyen, sxxen, exyen, exxen, sppen, mus_s, as, bs,
        Ls, V0s, sxxe, sxye, exxe, exye, cohesion, dro, drp, cp, kt, ht, tk, mrx, mry, mvx, mvy, pe, pen,
        nu_m2c, nd_m2c, mvxn, mvyn,
        phie, sxyee, exyee, pee, sppee, sxxee, exxee, mvxe, mvye,
        wtphisum, phisum, sxysum, exysum, esum, wtsxxsum, psum, sxxsum, exxsum, wtvxsum, vxsum,
        wtvysum, vysum, mrx0, mry0,
        pes, phis_it, syields, nuvp_it, sliprs, RK_sliprs_final, RK_phidts_final, RK_sliprs_solved,
        RK_phidts_solved, RK_deltas, NULL, NULL, NULL, NULL,
        deltasxx, deltasxy, deltaVs, exxs, exxes, sxxes, prs, tk2, tol0, tol1, sticky01};
    typedef char one_pointer_per_field[sizeof tab / sizeof tab[0] == V2C_NF ? 1 : -1];
    (void)sizeof(one_pointer_per_field);
    return f >= 0 && f < V2C_NF ? tab[f] : NULL;
}

static void die_on(int rc, const char *what)
{
    if (rc == 0) return;
    /* the reference's own failure convention: flag errcheck and report */
    fprintf(stderr, "libv2c %s failed (%d): %s\n", what, rc, g_ctx ? v2c_last_error(g_ctx) : "");
    errcheck = 1;
}

/* RK_sliprs/RK_phidts are [MAXNOD][2] in the reference; the ABI carries the
 * two columns as separate arrays. */
static void rk_columns(int to_device)
{
    double *tmp = malloc(sizeof(double) * nodenum);
    const int fid[4] = {V2C_N_RK_sliprs0, V2C_N_RK_sliprs1, V2C_N_RK_phidts0, V2C_N_RK_phidts1};
    for (int k = 0; k < 4; k++) {
        double (*src)[2] = k < 2 ? RK_sliprs : RK_phidts;
        int col = k & 1;
        if (to_device) {
            for (long i = 0; i < nodenum; i++) tmp[i] = src[i][col];
            die_on(v2c_nodal_set(g_ctx, fid[k], tmp), "nodal_set");
        } else {
            die_on(v2c_nodal_get(g_ctx, fid[k], tmp), "nodal_get");
            for (long i = 0; i < nodenum; i++) src[i][col] = tmp[i];
        }
    }
    free(tmp);
}

#define NOF(a) ((int)(sizeof(a) / sizeof((a)[0])))

/* ---- residency of the nodal arrays ---------------------------------------
 * g_dev_newer[f]: a device routine wrote array f after its host copy was
 * last brought up to date; g_host_newer[f]: host code wrote it (or it has
 * never been sent) after the device copy.  The RK_sliprs / RK_phidts
 * columns (compute_before_solve with rkmod 3) have one pair of flags. */
static unsigned char g_dev_newer[V2C_NF], g_host_newer[V2C_NF];
static int g_rk_dev_newer, g_rk_host_newer;

/* Send the host-newer arrays among f[0..n) in one staged transfer. */
static void send_newer(const int *f, int n)
{
    int ids[V2C_NF];
    const double *src[V2C_NF];
    int k = 0;
    for (int i = 0; i < n; i++)
        if (g_host_newer[f[i]]) {
            g_host_newer[f[i]] = 0;      /* also drops repeats within f */
            ids[k] = f[i];
            src[k++] = nodal_ptr(f[i]);
        }
    if (k) die_on(v2c_nodal_set_many(g_ctx, ids, src, k), "nodal_set_many");
}

/* Bring the device-newer arrays among f[0..n) to the host arrays. */
static void fetch_newer(const int *f, int n)
{
    int ids[V2C_NF];
    double *dst[V2C_NF];
    int k = 0;
    for (int i = 0; i < n; i++)
        if (g_dev_newer[f[i]]) {
            g_dev_newer[f[i]] = 0;
            ids[k] = f[i];
            dst[k++] = nodal_ptr(f[i]);
        }
    if (k) die_on(v2c_nodal_get_many(g_ctx, ids, dst, k), "nodal_get_many");
}

/* A routine's outputs: the device copy is now the newest, also of arrays
 * the host never sent (a routine that writes an array only in part lists it
 * among its inputs too, so the host's values went first). */
static void device_wrote(const int *f, int n)
{
    for (int i = 0; i < n; i++) {
        g_dev_newer[f[i]] = 1;
        g_host_newer[f[i]] = 0;
    }
}

void v2c_dropin_host_reads(const int *fields, int n)
{
    for (int i = 0; i < n; i++)
        if (fields[i] < 0 || fields[i] >= V2C_NF || !nodal_ptr(fields[i])) {
            die_on(V2C_EINVAL, "host_reads");
            return;
        }
    fetch_newer(fields, n);
}

void v2c_dropin_host_wrote(const int *fields, int n)
{
    for (int i = 0; i < n; i++) {
        if (fields[i] < 0 || fields[i] >= V2C_NF || !nodal_ptr(fields[i])) {
            die_on(V2C_EINVAL, "host_wrote");
            return;
        }
        /* a partial host write over a device-newer array would lose the
         * device's values: such host code must call host_reads first */
        if (g_dev_newer[fields[i]]) die_on(V2C_EINVAL, "host_wrote: array not brought back first");
        g_host_newer[fields[i]] = 1;
    }
}

void v2c_dropin_sync_to_host(void)
{
    int all[V2C_NF], k = 0;
    for (int f = 0; f < V2C_NF; f++)
        if (nodal_ptr(f)) all[k++] = f;
    fetch_newer(all, k);
    if (g_rk_dev_newer) {
        rk_columns(0);
        g_rk_dev_newer = 0;
    }
}

void v2c_dropin_host_wrote_all(void)
{
    v2c_dropin_sync_to_host();
    for (int f = 0; f < V2C_NF; f++)
        if (nodal_ptr(f)) g_host_newer[f] = 1;
    g_rk_host_newer = 1;
}

/* What the host Stokes step reads: preparematrix and its stencils
 * (momentum_continuity_addmatrix_checkerror.c:22-213, 380-1000: ro, nu, nd,
 * gg, gd, sxxe, sxye, sppe, mvx, mvy, mrx, mry, bm, pr, vx, vy, sxx, sxy),
 * reloadVxVyP (momentum_continuity_iterate.c:177-210) and checkerror
 * (:214-378: pr, vx, vy); reloadVxVyP writes pr, vx, vy. */
static const int STOKES_READS[] = {V2C_N_ro, V2C_N_nu, V2C_N_nd, V2C_N_gg, V2C_N_gd, V2C_N_sxxe,
                                   V2C_N_sxye, V2C_N_sppe, V2C_N_mvx, V2C_N_mvy, V2C_N_mrx, V2C_N_mry,
                                   V2C_N_bm, V2C_N_pr, V2C_N_vx, V2C_N_vy, V2C_N_sxx, V2C_N_sxy};
static const int STOKES_WRITES[] = {V2C_N_pr, V2C_N_vx, V2C_N_vy};
void v2c_dropin_before_stokes(void) { fetch_newer(STOKES_READS, NOF(STOKES_READS)); }
void v2c_dropin_after_stokes(void) { v2c_dropin_host_wrote(STOKES_WRITES, NOF(STOKES_WRITES)); }

/* What the per-step diagnostics after copyarrays
 * (momentum_continuity_iterate.c:129-170: sxxs, sxy, nu, sticky01) and
 * postproc (output_RSF.c:2-72: nu, sliprs, phis_it, syields) read. */
static const int OUTPUT_READS[] = {V2C_N_sxxs, V2C_N_sxy, V2C_N_nu, V2C_N_sliprs, V2C_N_phis_it,
                                   V2C_N_syields, V2C_N_sticky01};
void v2c_dropin_before_output(void) { fetch_newer(OUTPUT_READS, NOF(OUTPUT_READS)); }


static v2c_markers marker_view(void)
{
    v2c_markers m = {markx, marky, markk, markp, markxx, markxy, markexx, markexy, marke, markphi,
                     markvx, markvy, markC, markd, markw, mark_a, mark_b, mark_L, mark_L0, mark_V0,
                     fric_static, markC0, markt};
    return m;
}

/* Create the device context from the loaded model and hand it the markers.
 * Call once after load_model()/extra_init() (i2evps.c:48-50). */
int v2c_dropin_attach(long capacity, int device)
{
    v2c_grid g = {xnumx, ynumy, gx, gy, xsize, ysize};
    v2c_params p;
    v2c_params_default(&p);
    p.markmod = markmod; p.stoksmod = stoksmod; p.outgrid = outgrid; p.densimod = densimod;
    p.interpolationmod = interpolationmod; p.antidiffinter = antidiffinter; p.rkmod = rkmod;
    p.RSFmode = RSFmode; p.power_law_RSF = power_law_RSF; p.reloadfromDP = reloadfromDP;
    p.compressiblemode = compressiblemode; p.stickymod = stickymod; p.inertyn = inertyn;
    p.pmode = pmode; p.tempmod = tempmod; p.n2m_phantom = 1;
    p.nubeg = nubeg; p.nuend = nuend; p.tdeep = tdeep; p.zdeep = zdeep; p.incoming_T = eps[2];
    p.nstress = nstress; p.Dphis = Dphis; p.half_range = half_range;
    p.start_updip = start_updip; p.end_updip = end_updip;
    p.start_downdip = start_downdip; p.end_downdip = end_downdip;
    p.a_vw = a_vw; p.b_vw = b_vw; p.L_vw = L_vw; p.V0_vw = V0_vw;
    p.a_vs = a_vs; p.b_vs = b_vs; p.L_vs = L_vs; p.V0_vs = V0_vs;
    p.a_hr = a_hr; p.b_hr = b_hr; p.L_hr = L_hr; p.V0_hr = V0_hr; p.xstpx = xstpx;
    static v2c_materials mat;
    mat.ntypes = 200;
    for (int t = 0; t < 200; t++) {
        mat.nu[t] = marknu[t]; mat.dh[t] = markdh[t]; mat.dv[t] = markdv[t]; mat.ss[t] = markss[t];
        mat.mm[t] = markmm[t]; mat.gg[t] = markgg[t]; mat.n0[t] = markn0[t]; mat.n1[t] = markn1[t];
        mat.ro[t] = markro[t]; mat.bb[t] = markbb[t]; mat.aa[t] = markaa[t]; mat.cp[t] = markcp[t];
        mat.kt[t] = markkt[t]; mat.kf[t] = markkf[t]; mat.kp[t] = markkp[t]; mat.ht[t] = markht[t];
        mat.im[t] = markim[t];
    }
    int rc = v2c_create(&g_ctx, &g, &p, &mat, capacity, device);
    if (rc) return rc;
    /* the device holds no nodal array yet */
    for (int f = 0; f < V2C_NF; f++) {
        g_dev_newer[f] = 0;
        g_host_newer[f] = nodal_ptr(f) != NULL;
    }
    g_rk_dev_newer = 0;
    g_rk_host_newer = 1;
    /* boundary equations as load_model left them (load_save.c:279-304) */
    if (bondnum > 1) {
        rc = v2c_set_boundary(g_ctx, bondm, &bondn[0][0], &bondv[0][0], bondnum);
        if (rc) return rc;
    }
    v2c_markers m = marker_view();
    return v2c_markers_upload(g_ctx, &m, marknum);
}

void v2c_dropin_markers_to_host(void)
{
    v2c_markers m = marker_view();
    die_on(v2c_markers_download(g_ctx, &m, marknum), "markers_download");
}

void v2c_dropin_markers_to_device(void)
{
    v2c_markers m = marker_view();
    die_on(v2c_markers_upload(g_ctx, &m, marknum), "markers_upload");
}

/* Brings every device-newer nodal array back first. */
void v2c_dropin_detach(void)
{
    if (g_ctx) v2c_dropin_sync_to_host();
    v2c_destroy(g_ctx);
    g_ctx = NULL;
}

/* ---- the reference entry points ---------------------------------------- */

void marker2nodes(void)
{
    static const int in[] = {V2C_N_pr, V2C_N_phis_it, V2C_N_sxy, V2C_N_exy, V2C_N_pes, V2C_N_sxx,
                             V2C_N_exx, V2C_N_vx, V2C_N_vy, V2C_N_phie, V2C_N_sxyee, V2C_N_exyee,
                             V2C_N_pee, V2C_N_sxxee, V2C_N_sppee, V2C_N_exxee, V2C_N_mvxe, V2C_N_mvye,
                             /* written only where markers land: keep the host's values elsewhere */
                             V2C_N_pe, V2C_N_pen, V2C_N_nu_m2c, V2C_N_nd_m2c, V2C_N_mvxn, V2C_N_mvyn};
    static const int out[] = {V2C_N_bm, V2C_N_ro, V2C_N_et, V2C_N_nu, V2C_N_nd, V2C_N_gg, V2C_N_gd,
                              V2C_N_sppe, V2C_N_phis, V2C_N_phin, V2C_N_sxyen, V2C_N_sxxen, V2C_N_exyen,
                              V2C_N_exxen, V2C_N_sppen, V2C_N_mus_s, V2C_N_as, V2C_N_bs, V2C_N_Ls,
                              V2C_N_V0s, V2C_N_sxxe, V2C_N_sxye, V2C_N_exxe, V2C_N_exye,
                              V2C_N_cohesion, V2C_N_dro, V2C_N_drp, V2C_N_cp, V2C_N_kt, V2C_N_ht,
                              V2C_N_tk, V2C_N_mrx, V2C_N_mry, V2C_N_mvx, V2C_N_mvy, V2C_N_pe,
                              V2C_N_pen, V2C_N_nu_m2c, V2C_N_nd_m2c, V2C_N_mvxn, V2C_N_mvyn,
                              V2C_N_mrx0, V2C_N_mry0, V2C_N_sticky01};
    v2c_set_time(g_ctx, timesum, timestep);
    send_newer(in, NOF(in));
    die_on(v2c_marker2nodes(g_ctx), "marker2nodes");
    /* the routine's sticky-air marks (interpolation.c:861-916) */
    die_on(v2c_sticky_marks(g_ctx), "sticky_marks");
    /* the routine's boundary-condition tail (interpolation.c:918-925) */
    if (bondnum > 1) {
        die_on(v2c_vxvy_boundarycorrection(g_ctx), "vxvy_boundarycorrection");
        die_on(v2c_P_boundarycorrection(g_ctx), "P_boundarycorrection");
        if (tempmod) die_on(v2c_tkrecalc(g_ctx), "tkrecalc");
    }
    device_wrote(out, NOF(out));
}

double nodes2markers(void)
{
    static const int in[] = {V2C_N_exy, V2C_N_exye, V2C_N_sxy, V2C_N_sxye, V2C_N_pes, V2C_N_pe,
                             V2C_N_esp, V2C_N_phis, V2C_N_phis_it, V2C_N_exx, V2C_N_exxe, V2C_N_sxx,
                             V2C_N_sxxe, V2C_N_pr, V2C_N_sppe, V2C_N_vx, V2C_N_mvx, V2C_N_vy, V2C_N_mvy,
                             V2C_N_phie, V2C_N_sxyee, V2C_N_exyee, V2C_N_pee, V2C_N_sppee, V2C_N_sxxee,
                             V2C_N_exxee, V2C_N_mvxe, V2C_N_mvye};
    static const int out[] = {V2C_N_phie, V2C_N_sxyee, V2C_N_exyee, V2C_N_pee, V2C_N_sppee,
                              V2C_N_sxxee, V2C_N_exxee, V2C_N_mvxe, V2C_N_mvye};
    v2c_set_time(g_ctx, timesum, timestep);
    send_newer(in, NOF(in));
    die_on(v2c_nodes2markers(g_ctx), "nodes2markers");
    device_wrote(out, NOF(out));
    return 0;
}

void moveandrotatemarkers(void)
{
    static const int in[] = {V2C_N_vx, V2C_N_vy, V2C_N_exx, V2C_N_exy, V2C_N_esp};
    v2c_set_time(g_ctx, timesum, timestep);
    send_newer(in, NOF(in));
    die_on(v2c_moveandrotatemarkers(g_ctx), "moveandrotatemarkers");
}

double compute_before_solve(double dt, int rk)
{
    static const int in[] = {V2C_N_sxxs, V2C_N_sxy, V2C_N_sppes, V2C_N_nu_m2c, V2C_N_RK_sliprs_final,
                             V2C_N_RK_phidts_final, V2C_N_RK_sliprs_solved, V2C_N_RK_phidts_solved,
                             V2C_N_pe, V2C_N_phis, V2C_N_as, V2C_N_bs, V2C_N_Ls, V2C_N_V0s,
                             V2C_N_mus_s, V2C_N_cohesion, V2C_N_nu, V2C_N_pes, V2C_N_phis_it,
                             V2C_N_syields, V2C_N_nuvp_it, V2C_N_sliprs};
    static const int out[] = {V2C_N_nu, V2C_N_pes, V2C_N_phis_it, V2C_N_syields, V2C_N_nuvp_it,
                              V2C_N_sliprs};
    long nerr = 0;
    v2c_set_time(g_ctx, timesum, timestep);
    send_newer(in, NOF(in));
    if (rkmod == 3 && g_rk_host_newer) {
        rk_columns(1);
        g_rk_host_newer = 0;
    }
    die_on(v2c_compute_before_solve(g_ctx, dt, rk, &nerr), "compute_before_solve");
    device_wrote(out, NOF(out));
    if (rkmod == 3) g_rk_dev_newer = 1;
    if (nerr) errcheck = 1;
    return 0;
}

double compute_after_solve(double dt, int rk)
{
    static const int in[] = {V2C_N_sxxs, V2C_N_sxy, V2C_N_sxye, V2C_N_gg, V2C_N_exy, V2C_N_exxpl,
                             V2C_N_sppes, V2C_N_nu, V2C_N_nu_m2c, V2C_N_syields, V2C_N_phis_it,
                             V2C_N_as, V2C_N_bs, V2C_N_Ls, V2C_N_V0s, V2C_N_mus_s,
                             V2C_N_RK_sliprs_solved, V2C_N_RK_phidts_solved, V2C_N_RK_deltas};
    static const int out[] = {V2C_N_RK_sliprs_solved, V2C_N_RK_phidts_solved, V2C_N_RK_deltas};
    double deltas = 0;
    long nerr = 0;
    v2c_set_time(g_ctx, timesum, timestep);
    send_newer(in, NOF(in));
    die_on(v2c_compute_after_solve(g_ctx, dt, rk, &deltas, &nerr), "compute_after_solve");
    device_wrote(out, NOF(out));
    if (nerr) errcheck = 1;
    return deltas;
}

/* ---- post-solve nodal routines ------------------------------------------ */

int recompute_stress_strainrate(double dt, int rk)
{
    static const int in[] = {V2C_N_vx, V2C_N_vy, V2C_N_nd, V2C_N_gd, V2C_N_sxxe, V2C_N_nu, V2C_N_gg,
                             V2C_N_sxye, V2C_N_phis, V2C_N_phis_it,
                             /* written on interior nodes only */
                             V2C_N_sxx, V2C_N_exx, V2C_N_deltasxx, V2C_N_sxy, V2C_N_exy, V2C_N_esp,
                             V2C_N_deltasxy};
    static const int out[] = {V2C_N_sxx, V2C_N_exx, V2C_N_deltasxx, V2C_N_sxy, V2C_N_exy, V2C_N_esp,
                              V2C_N_deltasxy};
    v2c_stress_stats st;
    v2c_set_time(g_ctx, timesum, timestep);
    send_newer(in, NOF(in));
    die_on(v2c_recompute_stress_strainrate(g_ctx, dt, rk, &st), "recompute_stress_strainrate");
    device_wrote(out, NOF(out));
    maxdphi = st.maxdphi;
    maxnvp = st.maxnvp;
    printf("VISKOS: min = %e max = %e \n", st.minnu, st.maxnu);
    printf("STRESSS CHANGE:    max = %e \n", st.maxds);
    printf("MAX DPHI:    max = %e \n", st.maxdphi);
    return 0;
}

double interpolate_center2basic(double dt)
{
    static const int in[] = {V2C_N_exx, V2C_N_sxx, V2C_N_sxxe, V2C_N_gd, V2C_N_nd_m2c, V2C_N_pr,
                             V2C_N_exxs, V2C_N_exxpl, V2C_N_sxxs, V2C_N_prs, V2C_N_exxel};
    static const int out[] = {V2C_N_exxs, V2C_N_exxpl, V2C_N_sxxs, V2C_N_prs, V2C_N_exxel};
    send_newer(in, NOF(in));
    die_on(v2c_interpolate_center2basic(g_ctx, dt), "interpolate_center2basic");
    device_wrote(out, NOF(out));
    return 0;
}

double interpolate_center2basic_old(double dt)
{
    static const int in[] = {V2C_N_exxe, V2C_N_sxxe, V2C_N_sppe, V2C_N_exxes, V2C_N_sxxes, V2C_N_sppes};
    static const int out[] = {V2C_N_exxes, V2C_N_sxxes, V2C_N_sppes};
    (void)dt;   /* unused by the reference routine too */
    send_newer(in, NOF(in));
    die_on(v2c_interpolate_center2basic_old(g_ctx), "interpolate_center2basic_old");
    device_wrote(out, NOF(out));
    return 0;
}

double interpolate_basic2center(void)
{
    static const int in[] = {V2C_N_nu, V2C_N_nd};
    static const int out[] = {V2C_N_nd};
    send_newer(in, NOF(in));
    die_on(v2c_interpolate_basic2center(g_ctx), "interpolate_basic2center");
    device_wrote(out, NOF(out));
    return 0;
}

void copyarrays(int rk)
{
    static const int in[] = {V2C_N_RK_sliprs_solved, V2C_N_RK_phidts_solved, V2C_N_RK_sliprs_final,
                             V2C_N_RK_phidts_final, V2C_N_deltaVs, V2C_N_sliprs};
    static const int out[] = {V2C_N_deltaVs, V2C_N_sliprs, V2C_N_RK_sliprs_final, V2C_N_RK_phidts_final};
    v2c_set_time(g_ctx, timesum, timestep);
    send_newer(in, NOF(in));
    die_on(v2c_copyarrays(g_ctx, rk), "copyarrays");
    device_wrote(out, NOF(out));
}

/* The time step and, into the reference's globals, the sticky-air
 * statistics (calc_RSF.c:364-374, 505-550) that host code after copyarrays
 * prints (momentum_continuity_iterate.c:129-170). */
double calc_timestep(int oldornew)
{
    static const int in[] = {V2C_N_sppes, V2C_N_prs, V2C_N_bm, V2C_N_gg, V2C_N_Ls, V2C_N_as, V2C_N_bs,
                             V2C_N_V0s, V2C_N_RK_sliprs_final, V2C_N_RK_sliprs_solved, V2C_N_nu_m2c,
                             V2C_N_sxxes, V2C_N_sxye, V2C_N_nu, V2C_N_vx, V2C_N_mvx, V2C_N_vy, V2C_N_mvy,
                             V2C_N_phis, V2C_N_phis_it, V2C_N_sxxs, V2C_N_sxy, V2C_N_sticky01};
    v2c_sticky_stats st;
    v2c_timestep_params p;
    v2c_timestep_result r;
    v2c_timestep_params_default(&p);
    p.maxtmstep = maxtmstep; p.tcnvp = tcnvp; p.tcd = tcd; p.tcphih = tcphih; p.dtfactor = dtfactor;
    v2c_set_time(g_ctx, timesum, timestep);
    send_newer(in, NOF(in));
    const int rc = v2c_calc_timestep(g_ctx, oldornew, &p, &r);
    if (rc) {
        /* the reference exits on chi == 0 (calc_RSF.c:478) */
        die_on(rc, "calc_timestep");
        exit(0);
    }
    maxvel = r.maxvel; dtnvp = r.dtnvp; dtd = r.dtd; dtphi = r.dtphi; dtL = r.dtL;
    die_on(v2c_calc_sticky_stats(g_ctx, &st), "calc_sticky_stats");
    nr_1 = st.nr_1; nr_0 = st.nr_0;
    max_sigin_1 = st.max_sigin_1; max_sigin_0 = st.max_sigin_0;
    min_sigin_1 = st.min_sigin_1; min_sigin_0 = st.min_sigin_0;
    sum_sigin_1 = st.sum_sigin_1; sum2_sigin_1 = st.sum2_sigin_1;
    sum_sigin_0 = st.sum_sigin_0; sum2_sigin_0 = st.sum2_sigin_0;
    max_visc_1 = st.max_visc_1; max_visc_0 = st.max_visc_0;
    min_visc_1 = st.min_visc_1; min_visc_0 = st.min_visc_0;
    sum_visc_1 = st.sum_visc_1; sum2_visc_1 = st.sum2_visc_1;
    sum_visc_0 = st.sum_visc_0; sum2_visc_0 = st.sum2_visc_0;
    sigin_P1 = st.sigin_P1; sigin_P2 = st.sigin_P2; visco_P1 = st.visco_P1; visco_P2 = st.visco_P2;
    printf("===================================\n");
    printf(">>> dt_f =%e s, Max Slip-Vel=%e m/s, Max Vx=%e m/s", r.dt, r.maxvel, r.maxVx);
    printf("\n>>> dtnvp=%e s, dtd =%e s, dtphi =%e s, dtL=%e s \n", r.dtnvp, r.dtd, r.dtphi, r.dtL);
    printf("===================================\n");
    return r.dt;
}

/* ---- boundary-condition tail --------------------------------------------- */

void vxvy_boundarycorrection(void)
{
    static const int io[] = {V2C_N_mvx, V2C_N_mvy};
    send_newer(io, NOF(io));
    die_on(v2c_vxvy_boundarycorrection(g_ctx), "vxvy_boundarycorrection");
    device_wrote(io, NOF(io));
}

void P_boundarycorrection(void)
{
    static const int io[] = {V2C_N_sppe};
    send_newer(io, NOF(io));
    die_on(v2c_P_boundarycorrection(g_ctx), "P_boundarycorrection");
    device_wrote(io, NOF(io));
}

void tkrecalc(void)
{
    static const int io[] = {V2C_N_tk};
    send_newer(io, NOF(io));
    die_on(v2c_tkrecalc(g_ctx), "tkrecalc");
    device_wrote(io, NOF(io));
}

/* ---- the marker part of solve_temperature -------------------------------- */

/* temperature_iterate.c:181-419 on the device-resident markers; call it where
 * that block stood, after the thermal cycle restored timestep (:178). */
void v2c_dropin_temperature_markers(double heatdif)
{
    static const int in[] = {V2C_N_tk, V2C_N_tk2, V2C_N_kt, V2C_N_cp, V2C_N_ro};
    static const int out[] = {V2C_N_tol0, V2C_N_tol1};
    v2c_set_time(g_ctx, timesum, timestep);
    send_newer(in, NOF(in));
    die_on(v2c_temperature_markers(g_ctx, heatdif), "temperature_markers");
    /* the rest of the heat step reads them on the host */
    device_wrote(out, NOF(out));
    fetch_newer(out, NOF(out));
}
