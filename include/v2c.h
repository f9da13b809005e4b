/*
 * v2c.h -- C-ABI of the MI355X (gfx950) per-marker hot path of VELO2CYCLEs.
 *
 * The reference has no plugin layer: its hot path is a set of `void f(void)`
 * routines that read and write static global arrays (headevps.c:33-250,
 * input_RSF.c:3-11).  This header is the drop-in boundary that replaces them.
 * A C host keeps its call sites and forwards each one to the entry point named
 * in the comment above it; the reference globals are handed over as plain
 * pointers and sizes.  No torch or HIP type appears in any signature.
 *
 * Conventions (identical to the reference):
 *   - nodal arrays are FP64, length nodenum = xnumx*ynumy, indexed
 *     m3 = m1*ynumy + m2 (x-major, y fastest; interpolation.c:267);
 *   - marker arrays are FP64 structure-of-arrays, marker type is `char`
 *     (headevps.c:210);
 *   - times are in seconds, pressures in Pa, temperatures in K.
 *
 * Ownership: after v2c_markers_upload the device owns the marker state until
 * v2c_markers_download.  Nodal arrays live in a device mirror; the host moves
 * them with v2c_nodal_set / v2c_nodal_get.
 *
 * Every entry point returns 0 on success or a negative V2C_E* code; it never
 * calls exit().  A context is not thread-safe: one host thread per context.
 * Entry points are synchronous on return unless the context was switched to
 * asynchronous mode with v2c_set_async (then only the calls that return a
 * value to the host synchronise).  A context owns two HIP streams of the
 * default priority: work that shares no data runs side by side on them
 * within one call, and every call joins them before it returns or hands the
 * next call's work to the first.
 */
#ifndef V2C_H
#define V2C_H

#ifdef __cplusplus
extern "C" {
#endif

#define V2C_ABI_VERSION 1
#define V2C_MAX_TYPES 256

/* ---- error codes ------------------------------------------------------- */
#define V2C_OK 0
#define V2C_EINVAL (-1)      /* bad argument, size or flag value            */
#define V2C_ENOMEM (-2)      /* device or host allocation failed            */
#define V2C_EHIP (-3)        /* HIP runtime error (message in last_error)   */
#define V2C_ECAPACITY (-4)   /* marker count above the context capacity     */
#define V2C_EUNSUPPORTED (-5)/* mode the library does not implement         */
#define V2C_ENOGPU (-6)      /* no usable gfx950 device                      */

/* ---- grid (headevps.c:36-40,91) ---------------------------------------- */
typedef struct v2c_grid {
    long xnumx, ynumy;      /* number of gridlines in x / y                 */
    const double *gx;       /* x gridline coordinates [xnumx], increasing   */
    const double *gy;       /* y gridline coordinates [ynumy], increasing   */
    double xsize, ysize;    /* model size (m)                               */
} v2c_grid;

/* ---- run flags and hard-wired constants ----------------------------------
 * Flags are the reference's mode_*.t3c switches (load_save.c:84-154).
 * Physical constants that the reference hard-codes are exposed as fields,
 * each defaulting (v2c_params_default) to the reference value. */
typedef struct v2c_params {
    int markmod;            /* 1 Euler, else RK4; 2 = RK4-averaged spin     */
    int stoksmod;           /* 1 exact rotation, 2 Jaumann                  */
    int outgrid;            /* 0 move out, 1 stay, 2 orthogonal only        */
    int densimod;           /* 0 constant density, else P-T dependent       */
    int interpolationmod;   /* only 2 is implemented                        */
    int antidiffinter;      /* 1 = antidiffusion correction                  */
    int rkmod;              /* Picard (0) / RK variants (1..3)               */
    int RSFmode;            /* 1 rate-and-state, 0 Drucker-Prager           */
    int power_law_RSF;      /* >0 power-law RSF form                        */
    int reloadfromDP;
    int compressiblemode;
    int stickymod;
    int inertyn;            /* 0 = radiation damping on                     */
    int pmode;              /* 0 = constant normal stress nstress           */
    int tempmod;            /* heat solve active (n2m incoming-marker rule) */
    int n2m_phantom;        /* 1 = replicate the extra marker that the
                               reference's `m3<=marknum` loop visits
                               (interpolation.c:937): a zero marker of type 0
                               at (0,0) that scatters into the antidiffusion
                               sums                                          */
    double nubeg, nuend;    /* global viscosity limits (load_save.c:117-118) */
    double tdeep, zdeep;    /* deep-marker temperature reset (load_save.c:145-146) */
    double incoming_T;      /* the reference's eps[2], used as T of newly
                               incoming markers when tempmod==0
                               (interpolation.c:979)                         */
    /* RSF set-up (input_RSF.c:26-59) */
    double nstress, Dphis;
    double half_range, start_updip, end_updip, start_downdip, end_downdip;
    double a_vw, b_vw, L_vw, V0_vw;
    double a_vs, b_vs, L_vs, V0_vs;
    double a_hr, b_hr, L_hr, V0_hr;
    double xstpx;           /* grid step used by calc_frictionpara (calc_RSF.c:243) */
    /* constants hard-coded in the reference */
    double faultwidth;      /* 400 m (plastic.c:22,80)                      */
    double rad_damping;     /* 4.3303e6 Pa s/m (calc_RSF.c:111)             */
    double bulk_modulus;    /* 50e9 Pa (interpolation.c:226)                */
    double band_y, band_halfwidth; /* fault band |y-10 km| <= 100 m (calc_RSF.c:226-230) */
} v2c_params;

/* ---- rock table (headevps.c:231-243) ----------------------------------- */
typedef struct v2c_materials {
    int ntypes;                                   /* types 0..ntypes-1     */
    double nu[V2C_MAX_TYPES], dh[V2C_MAX_TYPES], dv[V2C_MAX_TYPES];
    double ss[V2C_MAX_TYPES], mm[V2C_MAX_TYPES], gg[V2C_MAX_TYPES];
    double n0[V2C_MAX_TYPES], n1[V2C_MAX_TYPES];
    double ro[V2C_MAX_TYPES], bb[V2C_MAX_TYPES], aa[V2C_MAX_TYPES];
    double cp[V2C_MAX_TYPES], kt[V2C_MAX_TYPES], kf[V2C_MAX_TYPES];
    double kp[V2C_MAX_TYPES], ht[V2C_MAX_TYPES];
    int im[V2C_MAX_TYPES];                        /* immobile flag         */
} v2c_materials;

/* ---- marker SoA (headevps.c:210-246, input_RSF.c:11) --------------------
 * Field names follow the reference globals with the `mark` prefix dropped.
 * Every pointer must be non-NULL for upload/download. */
typedef struct v2c_markers {
    double *x, *y;          /* markx, marky                                  */
    double *k;              /* markk   temperature (K)                       */
    double *p;              /* markp   pressure (Pa)                         */
    double *xx, *xy;        /* markxx, markxy deviatoric stress              */
    double *exx, *exy;      /* markexx, markexy strain rate                  */
    double *e;              /* marke   plastic strain                        */
    double *phi;            /* markphi RSF state                             */
    double *vx, *vy;        /* markvx, markvy                                */
    double *C;              /* markC   cohesion                              */
    double *d, *w;          /* markd, markw (reset to -1 out of the grid)    */
    double *a, *b, *L, *L0, *V0, *fric_static, *C0; /* calc_frictionpara outputs */
    char *t;                /* markt   rock type                             */
} v2c_markers;

/* ---- nodal fields (reference global names) ------------------------------ */
enum v2c_nodal_field {
    /* solution / solver-side fields (inputs of the hot path) */
    V2C_N_pr = 0, V2C_N_vx, V2C_N_vy, V2C_N_exx, V2C_N_exy, V2C_N_esp,
    V2C_N_sxx, V2C_N_sxy, V2C_N_sxxs, V2C_N_sppes, V2C_N_exxpl, V2C_N_exxel,
    /* marker2nodes outputs */
    V2C_N_bm, V2C_N_ro, V2C_N_et, V2C_N_nu, V2C_N_nd, V2C_N_gg, V2C_N_gd,
    V2C_N_sppe, V2C_N_phis, V2C_N_phin, V2C_N_sxyen, V2C_N_sxxen, V2C_N_exyen,
    V2C_N_exxen, V2C_N_sppen, V2C_N_mus_s, V2C_N_as, V2C_N_bs, V2C_N_Ls,
    V2C_N_V0s, V2C_N_sxxe, V2C_N_sxye, V2C_N_exxe, V2C_N_exye, V2C_N_cohesion,
    V2C_N_dro, V2C_N_drp, V2C_N_cp, V2C_N_kt, V2C_N_ht, V2C_N_tk, V2C_N_mrx,
    V2C_N_mry, V2C_N_mvx, V2C_N_mvy, V2C_N_pe, V2C_N_pen, V2C_N_nu_m2c,
    V2C_N_nd_m2c, V2C_N_mvxn, V2C_N_mvyn,
    /* nodes2markers antidiffusion outputs and accumulators */
    V2C_N_phie, V2C_N_sxyee, V2C_N_exyee, V2C_N_pee, V2C_N_sppee, V2C_N_sxxee,
    V2C_N_exxee, V2C_N_mvxe, V2C_N_mvye,
    V2C_N_wtphisum, V2C_N_phisum, V2C_N_sxysum, V2C_N_exysum, V2C_N_esum,
    V2C_N_wtsxxsum, V2C_N_psum, V2C_N_sxxsum, V2C_N_exxsum, V2C_N_wtvxsum,
    V2C_N_vxsum, V2C_N_wtvysum, V2C_N_vysum, V2C_N_mrx0, V2C_N_mry0,
    /* plastic / RSF nodal state (plastic.c, calc_RSF.c) */
    V2C_N_pes, V2C_N_phis_it, V2C_N_syields, V2C_N_nuvp_it, V2C_N_sliprs,
    V2C_N_RK_sliprs_final, V2C_N_RK_phidts_final, V2C_N_RK_sliprs_solved,
    V2C_N_RK_phidts_solved, V2C_N_RK_deltas,
    V2C_N_RK_sliprs0, V2C_N_RK_sliprs1, V2C_N_RK_phidts0, V2C_N_RK_phidts1,
    /* post-solve nodal fields (momentum_continuity_iterate.c:214-283,
     * interpolation.c:1575-1637) */
    V2C_N_deltasxx, V2C_N_deltasxy, V2C_N_deltaVs, V2C_N_exxs, V2C_N_exxes,
    V2C_N_sxxes, V2C_N_prs,
    /* heat solve: temperature before it (temperature_iterate.c:58-62) and
     * the marker-temperature averages (temperature_iterate.c:183-184,353-364) */
    V2C_N_tk2, V2C_N_tol0, V2C_N_tol1,
    /* sticky-air marks of marker2nodes (interpolation.c:121, 861-916) */
    V2C_N_sticky01,
    V2C_NF
};

typedef struct v2c_ctx v2c_ctx;

/* Fill *p with the reference defaults (input_RSF.c:13-59 and the shipped
 * mode_RSF.t3c flags, except markmod=2 which selects RK4). */
void v2c_params_default(v2c_params *p);

/* Name of a nodal field ("pr", "sxye", ...), NULL if out of range. */
const char *v2c_nodal_name(int field);

/* Create a context on HIP device `device`.  `capacity` is the largest marker
 * count the context will hold (device memory is sized for it).
 * Replaces the static allocation of headevps.c:8-21. */
int v2c_create(v2c_ctx **ctx, const v2c_grid *grid, const v2c_params *prm,
               const v2c_materials *mat, long capacity, int device);
void v2c_destroy(v2c_ctx *ctx);
const char *v2c_last_error(const v2c_ctx *ctx);
int v2c_set_async(v2c_ctx *ctx, int async_mode);
/* 1 (default): marker2nodes writes the per-marker friction parameters that
 * calc_frictionpara produces (calc_RSF.c:264-293: fric_static, mark_a/_b/_L/
 * _L0/_V0, markC0).  The hot path never reads them back (their only reader is
 * the checkpoint writer, input_RSF.c:106-116); 0 skips these writes. */
int v2c_set_friction_fields(v2c_ctx *ctx, int on);
int v2c_synchronize(v2c_ctx *ctx);

/* Mirror of the reference globals timesum / timestep (headevps.c:343-348). */
int v2c_set_time(v2c_ctx *ctx, double timesum, double timestep);

/* Marker state upload/download (load_model / save_model use it,
 * load_save.c:164-347,353-527).  Upload also builds the cell ordering. */
int v2c_markers_upload(v2c_ctx *ctx, const v2c_markers *m, long n);
int v2c_markers_download(v2c_ctx *ctx, v2c_markers *m, long n);
long v2c_marker_count(const v2c_ctx *ctx);

/* Statistics of the cell ordering: markers whose sub-cell changed at the last
 * re-ordering, number of full radix sorts and of incremental merges so far. */
int v2c_order_stats(v2c_ctx *ctx, long *last_movers, long *full_sorts, long *incremental);

/* Brings the cell ordering up to date (as the next marker2nodes would) and
 * copies it out: order[i] is the marker id at sorted position i and skey[i]
 * its sub-cell key (cell*4 + quadrant), i < n <= marker count.  No reference
 * counterpart: the reference keeps no ordering; this exposes the library's own
 * state for inspection and tests. */
int v2c_order_get(v2c_ctx *ctx, unsigned int *order, unsigned int *skey, long n);

/* Nodal mirror transfer, one field at a time (length nodenum). */
int v2c_nodal_set(v2c_ctx *ctx, int field, const double *host);
int v2c_nodal_get(v2c_ctx *ctx, int field, double *host);

/* Several nodal fields per transfer: fields[i] (V2C_N_*) from / to host[i]
 * (nodenum doubles each).  The library packs them into one pinned staging
 * buffer and moves them in one copy per direction (24 fields per copy; the
 * buffers, pinned host and device, are allocated at the first call).  No
 * reference counterpart: the reference keeps its nodal arrays in host
 * memory; the drop-in binding uses these for its residency hand-overs. */
int v2c_nodal_set_many(v2c_ctx *ctx, const int *fields, const double *const *host, int n);
int v2c_nodal_get_many(v2c_ctx *ctx, const int *fields, double *const *host, int n);

/* ---- hot-path entry points ---------------------------------------------- */

/* marker2nodes() (interpolation.c:55-859, called from i2evps.c:89):
 * marker->node scatter with interpolationmod=2 weights, normalisation,
 * antidiffusion and nodal friction parameters.  The routine's sticky-air
 * marks (:861-916) are v2c_sticky_marks and its boundary-condition tail
 * (:918-925) the three boundary routines below; the reference-side binding
 * calls both after this one. */
int v2c_marker2nodes(v2c_ctx *ctx);

/* The sticky-air marks at the end of marker2nodes (interpolation.c:861-916;
 * the field is cleared at :121), in the nodal field V2C_N_sticky01: every
 * node is 0 except those a marker of type 1 or 2 marks with its type.  A
 * marker's cell is nx = nxsearch(x), ny = nysearch(y); its quadrant, from
 * celdx and celdy against 0.5, picks node m3, m3+1, m3+ynumy or m3+ynumy+1
 * with the reference's m3 = nx*ynumy+ny+1 (one node below the cell's
 * upper-left node); a marker on a cell midline marks nothing.  Indices past
 * the end of a column wrap into the next column, as in the reference; those
 * at or beyond nodenum, which no reference loop reads, are dropped.  Where
 * several markers mark one node, the one with the highest index wins, as in
 * the reference's serial loop. */
int v2c_sticky_marks(v2c_ctx *ctx);

/* nodes2markers() (interpolation.c:929-1048, called from i2evps.c:108). */
int v2c_nodes2markers(v2c_ctx *ctx);

/* moveandrotatemarkers() (moveandrotatemarkers.c:1-178, called from
 * i2evps.c:113).  Marker types >= 100 return V2C_EUNSUPPORTED. */
int v2c_moveandrotatemarkers(v2c_ctx *ctx);

/* nodes2markers() immediately followed by moveandrotatemarkers() in one pass
 * over the markers; valid when nothing runs between the two calls (the
 * reference time loop with tempmod==0 or timesum>timestep, i2evps.c:98-113). */
int v2c_nodes2markers_advect(v2c_ctx *ctx);

/* compute_before_solve(dt, rk) (plastic.c:2-56, called from
 * momentum_continuity_iterate.c:69).  *nerr receives the number of nodes that
 * set the reference's errcheck flag. */
int v2c_compute_before_solve(v2c_ctx *ctx, double dt, int rk, long *nerr);

/* compute_after_solve(dt, rk) (plastic.c:61-129, called from
 * momentum_continuity_iterate.c:89).  Returns the RMS yield misfit in
 * *deltas. */
int v2c_compute_after_solve(v2c_ctx *ctx, double dt, int rk, double *deltas,
                            long *nerr);

/* ductilevisc() (density_ductilevisc_calc.c:29-223) evaluated for n
 * independent inputs; a pure helper that the host may call directly
 * (headevps.c:556).  Arrays are host arrays of length n; `iters` may be NULL. */
int v2c_ductilevisc_batch(v2c_ctx *ctx, long n, const char *type,
                          const double *sxy, const double *sxx,
                          const double *exy, const double *exx,
                          const double *tk, const double *pb,
                          double *nu_out, int *iters);

/* v2c_ductilevisc_batch without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  The
 * inputs and results are staged in a static device array of the library
 * (45 bytes per input, 11.25 MiB per device, held until the process ends),
 * so n is at most V2C_DUCTILEVISC_NOALLOC_MAX (V2C_EINVAL with a message
 * above it).  Every call on one device shares that device's array: calls
 * that overlap on one device (from several threads or contexts) race.
 * Results equal v2c_ductilevisc_batch's. */
#define V2C_DUCTILEVISC_NOALLOC_MAX (1L << 18)
int v2c_ductilevisc_batch_noalloc(v2c_ctx *ctx, long n, const char *type,
                                  const double *sxy, const double *sxx,
                                  const double *exy, const double *exx,
                                  const double *tk, const double *pb,
                                  double *nu_out, int *iters);

/* ---- post-solve nodal stencils (device-resident nodal fields) ------------
 * The reference's per-node routines between the Stokes solve and the next
 * marker2nodes, on the context's nodal fields.  Each writes exactly the
 * fields the reference routine writes. */

/* recompute_stress_strainrate(dt, rk) (momentum_continuity_iterate.c:214-263,
 * sxxcalc / sxycalc / nueffcalc momentum_continuity_addmatrix_checkerror.c:
 * 380-393, 446-473, 521-569): sxx, exx, deltasxx on every node with m1, m2
 * >= 1; sxy, exy, esp, deltasxy on the interior ones.  deltas divide by the
 * context's timestep.  stats (may be NULL) receives the routine's extrema as
 * the sequential MINV/MAXV folds of its node loop (the reference reduces them
 * with OpenMP, which leaves NaN handling to the compiler; without NaN the
 * two agree); minnu / maxnu start from the reference's -1e50 / 1e50 and so
 * stay there, as in the reference. */
typedef struct {
    double maxds;      /* max |sxx - sxxe|, |sxy - sxye|                      */
    double minnu;      /* min(-1e50, effective viscosities)                  */
    double maxnu;      /* max(1e50, effective viscosities)                   */
    double maxdphi;    /* max |phis - phis_it| over interior nodes (global)  */
    double maxnvp;     /* min(1e30, nu / gg) over interior nodes (global)    */
} v2c_stress_stats;
int v2c_recompute_stress_strainrate(v2c_ctx *ctx, double dt, int rk, v2c_stress_stats *stats);

/* interpolate_center2basic(dt) (interpolation.c:1600-1637): exxs, exxpl,
 * sxxs, prs, exxel on interior basic nodes from the four surrounding
 * pressure nodes. */
int v2c_interpolate_center2basic(v2c_ctx *ctx, double dt);
/* interpolate_center2basic_old(dt) (interpolation.c:1575-1598): exxes,
 * sxxes, sppes. */
int v2c_interpolate_center2basic_old(v2c_ctx *ctx);
/* interpolate_basic2center() (interpolation.c:1640-1685): nd on every
 * pressure node with m1, m2 >= 1, harmonic average of the basic-node nu. */
int v2c_interpolate_basic2center(v2c_ctx *ctx);
/* copyarrays(rk) (momentum_continuity_iterate.c:265-283): deltaVs, sliprs,
 * RK_sliprs_final, RK_phidts_final on interior nodes (rk is unused, as in
 * the reference). */
int v2c_copyarrays(v2c_ctx *ctx, int rk);

/* calc_timestep(oldornew) (calc_RSF.c:352-572) without its sticky-air
 * statistics (v2c_calc_sticky_stats): the adaptive time step from the
 * visco-plastic, displacement and rate-and-state limits over interior
 * nodes.  Every running minimum and
 * maximum follows the reference's sequential MINV/MAXV semantics exactly
 * (including NaN and the reset of dtphi/dtL to maxtmstep at nodes with
 * bs <= 0), whatever the reduction order on the device.  Where the reference
 * leaves its state-step factor lambda unset at a node (chi NaN) it keeps the
 * previous node's value; before any node sets it, lambda0 is used (the
 * reference's value there is uninitialised).  chi == 0 at a node, where the
 * reference exits, returns V2C_EUNSUPPORTED; pmode other than 0 or 1 with
 * RSFmode == 1 returns V2C_EUNSUPPORTED. */
typedef struct {
    double maxtmstep;  /* upper limit, the mode file's time step (i2evps.c:78) */
    double tcnvp, tcd, tcphih, dtfactor;      /* input_RSF.c:17-20           */
    double lambda0;    /* lambda before the first node that sets it        */
} v2c_timestep_params;
typedef struct {
    double dt;         /* dtfactor * min(maxtmstep, dtphi, dtL, dtnvp, dtd)   */
    double dtnvp, dtd, dtphi, dtL;
    double maxvel, maxVx;
} v2c_timestep_result;
/* reference defaults (input_RSF.c:17-20), maxtmstep = 1 yr, lambda0 = NaN */
void v2c_timestep_params_default(v2c_timestep_params *p);
int v2c_calc_timestep(v2c_ctx *ctx, int oldornew, const v2c_timestep_params *p, v2c_timestep_result *out);

/* The sticky-air statistics of calc_timestep (calc_RSF.c:364-374, 505-550;
 * the reference globals of headevps.c:361-370 without the std_* members,
 * which host code after copyarrays computes, momentum_continuity_iterate.c:
 * 129-158), over the interior nodes 1 <= m1 < xnumx-1, 1 <= m2 < ynumy-1,
 * from sxxs, sxy, nu and sticky01, with sigin = (sxxs^2 + sxy^2)^0.5.  As in
 * the reference, the "_1" members gather the nodes with sticky01 == 2 and the
 * "_0" members those with sticky01 == 1; P1 is node (50, 50) and P2 node
 * (500, 50), set where that node has sticky01 == 1.  Counts, extrema and the
 * P values follow the reference's serial loop exactly (MAXV/MINV, NaN
 * included, from 0 and 1e30); the sums are added in a fixed order
 * (deterministic, within 1e-12 relative of the serial sums). */
typedef struct {
    int nr_1, nr_0;
    double max_sigin_1, max_sigin_0, min_sigin_1, min_sigin_0;
    double sum_sigin_1, sum2_sigin_1, sum_sigin_0, sum2_sigin_0;
    double max_visc_1, max_visc_0, min_visc_1, min_visc_0;
    double sum_visc_1, sum2_visc_1, sum_visc_0, sum2_visc_0;
    double sigin_P1, sigin_P2, visco_P1, visco_P2;
} v2c_sticky_stats;
int v2c_calc_sticky_stats(v2c_ctx *ctx, v2c_sticky_stats *out);

/* ---- boundary-condition tail of marker2nodes and the heat solve -----------
 * Boundary equations in the reference's layout (headevps.c:197-202, read by
 * load_save.c:279-304): parameter p of node m3 is 3*m3 (P), 3*m3+1 (vx),
 * 3*m3+2 (vy) or 3*nodenum+m3 (T); bondm[p] = 0 for none, else the row k
 * (1 <= k < nbond) of bondn / bondv, which reads
 *     PAR = bondv[k][0] + sum_i bondv[k][i+1] * PAR(bondn[k][i] - 1)
 * over the i with bondn[k][i] != 0.  The arrays are copied to the device. */
int v2c_set_boundary(v2c_ctx *ctx, const long *bondm /* 4*nodenum */, const long *bondn /* nbond x 3 */,
                     const double *bondv /* nbond x 4 */, long nbond);
/* vxvy_boundarycorrection (momentum_continuity_addmatrix_checkerror.c:
 * 1000-1108), P_boundarycorrection (:1110-1184) and tkrecalc
 * (temperature_iterate.c:828-891): mvx / mvy, sppe and tk replaced by the
 * solution of the identity-plus-boundary system (free rows keep their value,
 * bounded rows take their equation; the coupled parameter of a vx, vy or P
 * row is located with the reference's integer formulas).  The system is
 * solved on the device by substitution to a fixed point, which chains of
 * equations referring to bounded parameters reach after one sweep per link;
 * equations that refer to each other in a cycle are solved if the
 * substitution settles within 64 sweeps and return V2C_EUNSUPPORTED
 * otherwise.  A row referring to an absent earlier row or to its own
 * parameter returns V2C_EUNSUPPORTED (the reference exits on the first and
 * on a zero diagonal). */
int v2c_vxvy_boundarycorrection(v2c_ctx *ctx);
int v2c_P_boundarycorrection(v2c_ctx *ctx);
int v2c_tkrecalc(v2c_ctx *ctx);

/* ---- marker temperatures after the heat solve ------------------------------
 * The marker part of solve_temperature (temperature_iterate.c:181-419), on
 * the device-resident markers: fluid markers (types 50..99) take the new
 * nodal temperature; rock markers (types < 50) get the solved increment
 * tk - tk2, with (heatdif != 0) the numerical subgrid diffusion toward the
 * old nodal temperature and its antidiffusion correction through the
 * per-node averages tol1 = sum(dT w) / sum(w) (tol0 = sum(w)), each sum over
 * the markers of the cell whose upper-left node it is.  Inputs: nodal tk
 * (after the solve), tk2 (before it), kt, cp, ro; the context's timestep
 * and tdeep.  Outputs: marker temperatures, nodal tol0 and tol1.  Markers
 * outside the open domain are untouched, as in the reference. */
int v2c_temperature_markers(v2c_ctx *ctx, double heatdif);

/* ---- .prn checkpoint files (host only) -----------------------------------
 * The reference's binary model file: written by save_model
 * (load_save.c:353-527 with extra_save_grid / extra_save_marker,
 * input_RSF.c:106-132), read by load_model (load_save.c:164-347).  Layout:
 * four size bytes (int, long, float, double), the header, the rock table
 * (markim as int, then 27 FP64 columns), one record per node (19 FP64, the
 * four bondm entries as long, 2 FP64), the gridlines, boundary rows
 * 1..bondnum-1 (4 FP64 + 3 long), timesave[5], and one record per marker
 * (14 FP64, the type byte, 5 FP64).  Arrays are column-wise in file order;
 * v2c_prn_read allocates them (v2c_prn_free releases them), v2c_prn_write
 * takes them from the caller.  Files whose size bytes differ from this
 * host's int/long/float/double are rejected, as the reference rejects
 * them. */
enum v2c_prn_rock {    /* rock[] columns after markim */
    V2C_PRN_R_n0 = 0, V2C_PRN_R_n1, V2C_PRN_R_s0, V2C_PRN_R_s1, V2C_PRN_R_nu, V2C_PRN_R_dh,
    V2C_PRN_R_dv, V2C_PRN_R_ss, V2C_PRN_R_mm, V2C_PRN_R_gg, V2C_PRN_R_ll, V2C_PRN_R_a0,
    V2C_PRN_R_a1, V2C_PRN_R_b0, V2C_PRN_R_b1, V2C_PRN_R_e0, V2C_PRN_R_e1, V2C_PRN_R_f0,
    V2C_PRN_R_f1, V2C_PRN_R_ro, V2C_PRN_R_bb, V2C_PRN_R_aa, V2C_PRN_R_cp, V2C_PRN_R_kt,
    V2C_PRN_R_kf, V2C_PRN_R_kp, V2C_PRN_R_ht, V2C_PRN_NROCK
};
enum v2c_prn_node {    /* node[] columns; the file stores RK_sliprs_final twice */
    V2C_PRN_N_sxxee = 0, V2C_PRN_N_sxx, V2C_PRN_N_exxee, V2C_PRN_N_exx, V2C_PRN_N_sxyee,
    V2C_PRN_N_sxy, V2C_PRN_N_exyee, V2C_PRN_N_exy, V2C_PRN_N_sppee, V2C_PRN_N_pr,
    V2C_PRN_N_phie, V2C_PRN_N_phis_it, V2C_PRN_N_mvxe, V2C_PRN_N_vx, V2C_PRN_N_mvye,
    V2C_PRN_N_vy, V2C_PRN_N_pes, V2C_PRN_N_pee, V2C_PRN_N_RK_sliprs_final_a,
    V2C_PRN_N_RK_sliprs_final, V2C_PRN_N_RK_phidts_final, V2C_PRN_NNODE
};
enum v2c_prn_mark {    /* mark[] columns */
    V2C_PRN_M_x = 0, V2C_PRN_M_y, V2C_PRN_M_k, V2C_PRN_M_e, V2C_PRN_M_xx, V2C_PRN_M_v,
    V2C_PRN_M_xy, V2C_PRN_M_p, V2C_PRN_M_exx, V2C_PRN_M_exy, V2C_PRN_M_d, V2C_PRN_M_w,
    V2C_PRN_M_vx, V2C_PRN_M_vy, V2C_PRN_M_phi, V2C_PRN_M_L, V2C_PRN_M_L0, V2C_PRN_M_C,
    V2C_PRN_M_C0, V2C_PRN_NMARK
};
typedef struct v2c_prn {
    long xnumx, ynumy, mnumx, mnumy, marknum;
    double xsize, ysize, pinit, pkf[4], timestep, timesum, GXKOEF, GYKOEF;
    int rocknum;
    long bondnum;            /* boundary rows incl. the unused row 0          */
    int cycle;               /* save_model's circle number (n0)               */
    int *markim;             /* rocknum                                       */
    double *rock[V2C_PRN_NROCK];     /* rocknum each                          */
    double *node[V2C_PRN_NNODE];     /* xnumx*ynumy each                      */
    long *bondm;             /* 4*nodenum: P, vx, vy of node m3 at 3*m3+0..2, T at 3*nodenum+m3 */
    double *gx, *gy;
    double *bondv;           /* 4*bondnum, row 0 unused                       */
    long *bondn;             /* 3*bondnum, row 0 unused                       */
    double timesave[5];
    double *mark[V2C_PRN_NMARK];     /* marknum each                          */
    char *t;                 /* marknum                                       */
} v2c_prn;
int v2c_prn_read(const char *path, v2c_prn *prn);
int v2c_prn_write(const char *path, const v2c_prn *prn);
void v2c_prn_free(v2c_prn *prn);
/* message of the last failed v2c_prn_read / v2c_prn_write on this thread */
const char *v2c_prn_last_error(void);

/* ---- build-defined per-marker local update (no reference counterpart) ----
 * The reference evaluates rate-and-state friction per node (plastic.c,
 * calc_RSF.c:33-221) and the ductile viscosity with a quasi-Newton that stops
 * on |log10(mu_k/mu_{k-1})| < 1e-6 (density_ductilevisc_calc.c:170-200).  The
 * local update evaluates both per marker to a stated residual instead:
 *   ductile: Newton on f(eta) = eta - coeff*X^((n-1)/n) (the reference's own
 *            residual, density_ductilevisc_calc.c:188-193) to |f|/eta <= tol;
 *   RSF (fault markers): state aged over dt at the previous slip rate
 *            (calc_RSF.c:96-103), then the slip rate V solving
 *            sigII = nP a asinh(V/(2V0) exp((mu0+b phi')/a)) + eta_rad V
 *            (calc_RSF.c:111,129) by Newton in ln V to |g|/sigII <= tol,
 *            capped at EII*W and floored at 1e-40 (calc_RSF.c:193-196).
 * Warm start: the context's MLP proposes eta and V; a start whose residual
 * grows after one step falls back to the default start. */
typedef struct v2c_lu_stats {
    long markers, ductile_markers, ductile_iters, fault_markers, rsf_iters;
    long max_ductile_iters, max_rsf_iters, fallbacks, nonfinite;
    long ductile_evals, rsf_evals;           /* residual evaluations (the cost measure) */
    long hist_ductile[32], hist_rsf[32];     /* iteration histograms; last bin = overflow */
} v2c_lu_stats;

/* Per-marker RSF inputs: fault flag, slip rate V, state phi and the marker's
 * own a, b, L, V0, mu0 (all length n, host arrays).  The fault markers' ids
 * are kept as a list: the slip-rate solve runs over them alone. */
int v2c_rsf_upload(v2c_ctx *ctx, const unsigned char *fault, const double *V, const double *phi,
                   const double *a, const double *b, const double *L, const double *V0,
                   const double *mu0, long n);
/* warm: 0 default starts; 1 the stored guesses (v2c_mlp_predict or
 * v2c_local_set_guess) for both solves; 2 the slip-rate guess evaluated by
 * the network inside the slip-rate solve (v2c_mlp_set must have run; the
 * ductile solve starts from its default, where it needs about one Newton
 * step).  The ductile solve runs over all markers, the slip-rate solve over
 * the fault markers on the context's second stream beside it; their kernel
 * times are V2C_K_LOCAL and V2C_K_LOCAL_RSF, the whole call's device time
 * (fork to join) V2C_K_LOCAL_WALL. */
int v2c_local_update(v2c_ctx *ctx, double dt, int warm, double tol, v2c_lu_stats *stats);
/* Outputs of the last local update (any pointer may be NULL). */
int v2c_local_download(v2c_ctx *ctx, double *eta, double *V, double *phi, double *eta_vp,
                       double *syield, long n);
/* Set the warm-start guesses directly (host arrays; NULL keeps a field). */
int v2c_local_set_guess(v2c_ctx *ctx, const double *eta, const double *V, long n);
/* Read the current warm-start guesses (set directly or proposed by
 * v2c_mlp_predict; 0 where a solve does not apply). */
int v2c_local_get_guess(v2c_ctx *ctx, double *eta, double *V, long n);

/* ---- learned warm start (no reference counterpart) ------------------------
 * 9 -> 64 -> 64 -> 2 tanh MLP, BF16 MFMA with FP32 accumulation.  It predicts
 * log10(root / default start) of the ductile unknown and of the fault slip
 * rate from standardised (log10 strain-rate invariant, T, P, log10 stress
 * invariant, log10 theta, a-b, and for fault markers log10 of the slip-rate
 * default start, log10 L, log10 of the previous slip rate).  params: 4930
 * FP32 values, weights [out][in] in the order W1 b1 W2 b2 W3 b3; norm:
 * in_mean[9] in_std[9] out_mean[2] out_std[2]. */
int v2c_mlp_set(v2c_ctx *ctx, const float *params, const float *norm);
int v2c_mlp_get(v2c_ctx *ctx, float *params, float *norm);
/* Write warm-start guesses for every resident marker (inputs of dt). */
int v2c_mlp_predict(v2c_ctx *ctx, double dt);
/* The same for a subset of the outputs: bit 0 the ductile guess, bit 1 the
 * slip-rate guess (the other is set to 0, the default start).  Workgroups of
 * 256 markers none of which needs a selected guess skip the network, so with
 * outputs = 2 the cost follows the fault markers. */
int v2c_mlp_predict_outputs(v2c_ctx *ctx, double dt, int outputs);
/* Training samples from one plain local update of the resident markers:
 * X[n][11] = 9 raw features + 2 labels, mask[n][2] (1 where a label exists).
 * Runs the plain local update (state advances as in v2c_local_update). */
int v2c_mlp_samples(v2c_ctx *ctx, double dt, double tol, float *X, float *mask);
/* Train on standardised Xs[ns][9], Ts[ns][2] with masks Ms[ns][2]; batch is a
 * multiple of 64; loss_out[epochs] receives each epoch's mean masked MSE. */
int v2c_mlp_train(v2c_ctx *ctx, const float *Xs, const float *Ts, const float *Ms, long ns, int epochs,
                  int batch, float lr, unsigned seed, float *loss_out);
/* Gradient of the masked mean squared error over one batch (no update). */
int v2c_mlp_grad(v2c_ctx *ctx, const float *Xs, const float *Ts, const float *Ms, long ns, float *grad,
                 float *loss);
/* Network outputs on standardised inputs. */
int v2c_mlp_forward(v2c_ctx *ctx, const float *Xs, long ns, float *Ys);

/* Timing of the last entry point's kernels (HIP events on the context
 * stream), in milliseconds; `kernel` selects one of V2C_K_* . */
enum v2c_kernel_id {
    V2C_K_M2N_CELLS = 0, V2C_K_M2N_NODES, V2C_K_N2M_CELLS, V2C_K_N2M_NODES,
    V2C_K_ADVECT, V2C_K_SORT, V2C_K_BEFORE, V2C_K_AFTER, V2C_K_LOCAL, V2C_K_MLP,
    V2C_K_M2N_PROPS, V2C_K_N2M_UPDATE, V2C_K_POST, V2C_K_LOCAL_RSF, V2C_K_LOCAL_WALL,
    V2C_K_COUNT
};
int v2c_enable_timing(v2c_ctx *ctx, int on);
/* Measurement aid (no reference counterpart): while timing is on, record
 * events only around the kernels whose bit (1u << V2C_K_*) is set (default:
 * all).  An event pair costs a few microseconds of stream time, so a timed
 * run that wants the whole step's duration undisturbed times one kernel. */
int v2c_timing_select(v2c_ctx *ctx, unsigned int kernel_mask);
/* Measurement aid (no reference counterpart): time a plain device copy of
 * `bytes` (reads and writes `bytes` each), used to calibrate HBM counters. */
int v2c_copy_calibration(v2c_ctx *ctx, long bytes, double *ms);
int v2c_kernel_time(v2c_ctx *ctx, int kernel, double *ms, long *launches);
int v2c_reset_timing(v2c_ctx *ctx);

/* ---- batched ensembles (no reference counterpart) -----------------------
 * Independent models, each an ordinary context (own grid, markers and nodal
 * state, same device), advanced together.  Each stage of the step runs as one
 * launch over a work list of (model, range) entries: per-marker and per-node
 * stages as one 256-thread block per entry of 256 items, the scatter stages
 * as a persistent grid of wavefronts striding over cell groups.  The stage
 * sequence is captured once into a HIP graph and replayed.
 * A step is, per model: marker2nodes, compute_before_solve(timestep, 0),
 * compute_after_solve(timestep, 0), [the local update], nodes2markers +
 * advection, [copyarrays and calc_timestep], then timesum += timestep (and,
 * with adaptive steps, the next timestep).  Results equal the per-model entry points bit for bit
 * whatever the shapes.  Models are grouped in size classes by marker count
 * (< 0.5 M, < 2 M, larger); shapes apply per class. */
typedef struct v2c_ens v2c_ens;
#define V2C_ENS_CLASSES 3
typedef struct {
    int cells_per_group;    /* cells per scatter wavefront group, 1..1024 */
    int markers_per_tile;   /* markers of one model kept on one XCD in a row
                               (a run of entries), a multiple of 256 */
    int nodes_per_tile;     /* the same for nodes, a multiple of 256 */
} v2c_tile_shape;
int v2c_ens_create(v2c_ens **ens, v2c_ctx *const *models, int nmodels);
void v2c_ens_destroy(v2c_ens *ens);
const char *v2c_ens_last_error(const v2c_ens *ens);
int v2c_ens_size_class(const v2c_ens *ens, int model);
int v2c_ens_set_shape(v2c_ens *ens, int size_class, const v2c_tile_shape *shape);
int v2c_ens_get_shape(const v2c_ens *ens, int size_class, v2c_tile_shape *shape);
/* persistent scatter grid = 4 x blocks_per_cu x compute units wavefronts */
int v2c_ens_set_blocks_per_cu(v2c_ens *ens, int blocks);
/* Run the step as `branches` (1..4) concurrent branches of the graph, each a
 * contiguous group of models of about equal marker count with its own stage
 * sequence on its own stream, so that stages of different groups overlap.
 * Results are the same whatever the count.  v2c_ens_get_branches returns the
 * count in use (at most the number of models). */
int v2c_ens_set_branches(v2c_ens *ens, int branches);
int v2c_ens_get_branches(const v2c_ens *ens);
/* one batched step; use_graph: replay the captured stage graph (every model's
 * property pass runs beside the batched re-ordering, before the graph) */
int v2c_ens_step(v2c_ens *ens, int use_graph);
/* bring every model's cell ordering up to date in one batched pass (the
 * movers of all models in one radix sort); a step does this itself */
int v2c_ens_order(v2c_ens *ens);
/* the same step through the per-model entry points (one stream per model) */
int v2c_ens_step_looped(v2c_ens *ens);
/* Adds the per-marker local update (v2c_local_update with this warm mode and
 * tolerance, on each model's RSF inputs and network) to every batched step,
 * after compute_after_solve; warm < 0 removes it.  v2c_ens_step_looped runs
 * the same sequence through the per-model entry points. */
int v2c_ens_set_local(v2c_ens *ens, int warm, double tol);
/* local-update statistics of one model's last step */
int v2c_ens_local_stats(v2c_ens *ens, int model, v2c_lu_stats *stats);
/* Adaptive time steps: every batched step ends with copyarrays(0)
 * (momentum_continuity_iterate.c:265-283) and calc_timestep(oldornew)
 * (calc_RSF.c:352-572) of every model on the device; the model's next step is
 * that result clamped to [dt_min, dt_max].  p == NULL turns it off. */
int v2c_ens_set_adaptive(v2c_ens *ens, const v2c_timestep_params *p, int oldornew, double dt_min, double dt_max);
/* every model's (timesum, timestep) after the last step: 2 * nmodels values */
int v2c_ens_times(v2c_ens *ens, double *timesum_timestep);
/* compute_after_solve misfit of every model's last batched step:
 * rms_count[2*k] = RMS, rms_count[2*k+1] = node count */
int v2c_ens_rms(v2c_ens *ens, double *rms_count);
/* time tile shapes per size class, the grid size and the branch count on
 * the ensemble itself (coordinate descent, reps timed steps each, the
 * graph's time); keeps the fastest.  Advances the models by the steps it
 * runs. */
int v2c_ens_autotune(v2c_ens *ens, int reps, double *best_ms);

#ifdef __cplusplus
}
#endif
#endif /* V2C_H */
