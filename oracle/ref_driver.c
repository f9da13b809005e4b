/*
 * ref_driver.c -- drives the reference's own hot-path routines, compiled in
 * place from /root/reference/src (the include path is given by
 * oracle/Makefile; no reference source is copied into this repository).
 *
 * TEST INFRASTRUCTURE ONLY: it produces golden vectors (tests/golden/) and
 * checks the oracle.  It never travels as part of the product.
 *
 * Protocol: `ref_driver IN OUT op...`.  IN and OUT are sequences of records
 *   char name[32]; char kind ('d' double, 'l' long, 'i' int, 'c' char);
 *   int64 count; payload
 * IN sets reference globals by name; the ops run in order; OUT receives every
 * global of the table after the last op.  Ops:
 *   m2n | n2m | move | before:<dt>:<rk> | after:<dt>:<rk> | ductile | allinteri
 *   hostsolve | hostout (tests/native/host_stand_in.h; hostout's sum in drv_ret[3])
 *   stress:<dt>:<rk> | c2b:<dt> | c2bold:<dt> | b2c | copy:<rk> | tstep:<oldornew>
 *   bcvxvy | bcp | bctk | save:<cycle>:<path> | load:<path>
 * (tstep leaves the returned time step in drv_ret[2]; save and load are the
 * reference's .prn writer and reader, save also writes file.t3c and
 * file_hdf5.t3c into the working directory).
 *
 * Built with -DV2C_DROPIN (oracle/_ref/dropin_driver) the reference's
 * hot-path routines are compiled under ref_* names and the reference-side
 * binding velo2cycles_amd/csrc/dropin/v2c_dropin.c supplies the live ones,
 * so every op goes through libv2c on the GPU exactly as the reference's own
 * call sites would once the binding is linked in.  The device context is
 * attached after IN is read and the markers come back before OUT is written.
 * Extra op there: heat:<heatdif> (v2c_dropin_temperature_markers).
 */
#ifndef _GNU_SOURCE
#define _GNU_SOURCE 1
#endif
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <stdbool.h>
#include <time.h>

#ifdef V2C_DROPIN
#define marker2nodes ref_marker2nodes
#define nodes2markers ref_nodes2markers
#define moveandrotatemarkers ref_moveandrotatemarkers
#define compute_before_solve ref_compute_before_solve
#define compute_after_solve ref_compute_after_solve
#define recompute_stress_strainrate ref_recompute_stress_strainrate
#define interpolate_center2basic ref_interpolate_center2basic
#define interpolate_center2basic_old ref_interpolate_center2basic_old
#define interpolate_basic2center ref_interpolate_basic2center
#define copyarrays ref_copyarrays
#define calc_timestep ref_calc_timestep
#define vxvy_boundarycorrection ref_vxvy_boundarycorrection
#define P_boundarycorrection ref_P_boundarycorrection
#define tkrecalc ref_tkrecalc
#endif

/* The reference translation unit (i2evps.c:12-30 order, hot-path files only). */
#include "headevps.c"
#include "input_RSF.c"
#include "calc_RSF.c"
#include "load_save.c"
#include "gausevps.c"
#include "interpolation.c"
#include "density_ductilevisc_calc.c"
#include "moveandrotatemarkers.c"
#include "temperature_iterate.c"
#include "momentum_continuity_addmatrix_checkerror.c"
#include "momentum_continuity_iterate.c"
#include "plastic.c"
#include "recalc_marker_properties.c"

#ifdef V2C_DROPIN
#undef marker2nodes
#undef nodes2markers
#undef moveandrotatemarkers
#undef compute_before_solve
#undef compute_after_solve
#undef recompute_stress_strainrate
#undef interpolate_center2basic
#undef interpolate_center2basic_old
#undef interpolate_basic2center
#undef copyarrays
#undef calc_timestep
#undef vxvy_boundarycorrection
#undef P_boundarycorrection
#undef tkrecalc
#include "../velo2cycles_amd/csrc/dropin/v2c_dropin.c"
#endif

/* host stand-ins shared with the binding's test host */
#include "../tests/native/host_stand_in.h"

/* driver-local arrays for the pure-function ops */
#define DRV_MAX 2000000
static double dv_sxy[DRV_MAX], dv_sxx[DRV_MAX], dv_tk[DRV_MAX], dv_pb[DRV_MAX], dv_out[DRV_MAX];
static char dv_type[DRV_MAX];
static double ai_x[DRV_MAX], ai_y[DRV_MAX], ai_vx[DRV_MAX], ai_vy[DRV_MAX], ai_esp[DRV_MAX], ai_ee[DRV_MAX];
static long drv_n;
static double drv_ret[8];
static double drv_time;   /* wall seconds of the op sequence */
/* wall seconds per routine: marker2nodes, compute_before_solve,
 * compute_after_solve, nodes2markers, moveandrotatemarkers */
static double drv_optime[5];
static double RK_sliprs0[MAXNOD], RK_sliprs1[MAXNOD], RK_phidts0[MAXNOD], RK_phidts1[MAXNOD];

typedef struct { const char *name; char kind; void *ptr; long cap; } entry;
#define D(a) {#a, 'd', (void *)(a), (long)(sizeof(a) / sizeof(double))}
#define S(a) {#a, 'd', (void *)&(a), 1}
#define L(a) {#a, 'l', (void *)&(a), 1}
#define LA(a) {#a, 'l', (void *)(a), (long)(sizeof(a) / sizeof(long))}
#define I(a) {#a, 'i', (void *)&(a), 1}
static entry table[] = {
    /* grid and scalars */
    D(gx), D(gy), L(xnumx), L(ynumy), L(marknum), S(xsize), S(ysize), S(timesum), S(timestep),
    I(markmod), I(stoksmod), I(outgrid), I(densimod), I(interpolationmod), I(antidiffinter),
    I(rkmod), I(RSFmode), I(power_law_RSF), I(reloadfromDP), I(compressiblemode), I(stickymod),
    I(inertyn), I(pmode), I(tempmod), S(nubeg), S(nuend), S(tdeep), S(zdeep), S(nstress), S(Dphis),
    S(half_range), S(start_updip), S(end_updip), S(start_downdip), S(end_downdip),
    S(a_vw), S(b_vw), S(L_vw), S(V0_vw), S(a_vs), S(b_vs), S(L_vs), S(V0_vs),
    S(a_hr), S(b_hr), S(L_hr), S(V0_hr), S(errcheck), D(eps),
    /* rock table */
    D(marknu), D(markdh), D(markdv), D(markss), D(markmm), D(markgg), D(markn0), D(markn1),
    D(markro), D(markbb), D(markaa), D(markcp), D(markkt), D(markkf), D(markkp), D(markht),
    {"markim", 'i', (void *)markim, MAXTMR},
    /* markers */
    D(markx), D(marky), D(markk), D(markp), D(markxx), D(markxy), D(markexx), D(markexy),
    D(marke), D(markphi), D(markvx), D(markvy), D(markC), D(markd), D(markw), D(mark_a),
    D(mark_b), D(mark_L), D(mark_L0), D(mark_V0), D(fric_static), D(markC0),
    {"markt", 'c', (void *)markt, MAXMRK},
    /* nodal */
    D(pr), D(vx), D(vy), D(exx), D(exy), D(esp), D(sxx), D(sxy), D(sxxs), D(sppes), D(exxpl), D(exxel),
    D(bm), D(ro), D(et), D(nu), D(nd), D(gg), D(gd), D(sppe), D(phis), D(phin), D(sxyen), D(sxxen),
    D(exyen), D(exxen), D(sppen), D(mus_s), D(as), D(bs), D(Ls), D(V0s), D(sxxe), D(sxye), D(exxe),
    D(exye), D(cohesion), D(dro), D(drp), D(cp), D(kt), D(ht), D(tk), D(mrx), D(mry), D(mvx), D(mvy),
    D(pe), D(pen), D(nu_m2c), D(nd_m2c), D(mvxn), D(mvyn),
    D(phie), D(sxyee), D(exyee), D(pee), D(sppee), D(sxxee), D(exxee), D(mvxe), D(mvye),
    D(wtphisum), D(phisum), D(sxysum), D(exysum), D(esum), D(wtsxxsum), D(psum), D(sxxsum),
    D(exxsum), D(wtvxsum), D(vxsum), D(wtvysum), D(vysum), D(mrx0), D(mry0),
    D(pes), D(phis_it), D(syields), D(nuvp_it), D(sliprs), D(RK_sliprs_final), D(RK_phidts_final),
    D(RK_sliprs_solved), D(RK_phidts_solved), D(RK_deltas),
    D(RK_sliprs0), D(RK_sliprs1), D(RK_phidts0), D(RK_phidts1),
    D(deltasxx), D(deltasxy), D(deltaVs), D(exxs), D(exxes), D(sxxes), D(prs), D(tk2), D(tol0), D(tol1),
    /* sticky-air marks and statistics (headevps.c:34, 361-370) */
    D(sticky01), I(nr_1), I(nr_0), S(max_sigin_1), S(max_sigin_0), S(min_sigin_1), S(min_sigin_0),
    S(sum_sigin_1), S(sum2_sigin_1), S(sum_sigin_0), S(sum2_sigin_0), S(max_visc_1), S(max_visc_0),
    S(min_visc_1), S(min_visc_0), S(sum_visc_1), S(sum2_visc_1), S(sum_visc_0), S(sum2_visc_0),
    S(sigin_P1), S(sigin_P2), S(visco_P1), S(visco_P2),
    /* post-solve scalars (headevps.c:120-122, 181, 339, 345-346) */
    S(maxvel), S(maxdphi), S(maxnvp), S(dtL), S(dtphi), S(dtd), S(dtnvp),
    S(maxtmstep), S(tcnvp), S(tcd), S(tcphih), S(dtfactor),
    /* boundary equations (headevps.c:197-202) */
    LA(bondm), LA(bondn), D(bondv), L(bondnum),
    /* the rest of the .prn header and rock table (headevps.c:37-58, 180, 205,
     * 230-241) */
    L(mnumx), L(mnumy), S(pinit), D(pkf), S(GXKOEF), S(GYKOEF), I(rocknum), D(timesave),
    D(marks0), D(marks1), D(markll), D(marka0), D(marka1), D(markb0), D(markb1),
    D(marke0), D(marke1), D(markf0), D(markf1), D(markv),
    /* pure-function op arrays */
    D(dv_sxy), D(dv_sxx), D(dv_tk), D(dv_pb), D(dv_out), {"dv_type", 'c', (void *)dv_type, DRV_MAX},
    D(ai_x), D(ai_y), D(ai_vx), D(ai_vy), D(ai_esp), D(ai_ee), L(drv_n), D(drv_ret), S(drv_time), D(drv_optime),
};
#define NTAB ((long)(sizeof table / sizeof table[0]))

static size_t ksize(char k) { return k == 'd' ? 8 : k == 'l' ? 8 : k == 'i' ? 4 : 1; }

static int read_input(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) { perror(path); return -1; }
    for (;;) {
        char name[32]; char kind; long long cnt;
        if (fread(name, 1, 32, f) != 32) break;
        if (fread(&kind, 1, 1, f) != 1 || fread(&cnt, 8, 1, f) != 1) { fclose(f); return -1; }
        name[31] = 0;
        long i;
        for (i = 0; i < NTAB; i++) if (!strcmp(table[i].name, name)) break;
        if (i == NTAB || table[i].kind != kind || cnt > table[i].cap) {
            fprintf(stderr, "ref_driver: bad record %s kind %c count %lld\n", name, kind, cnt);
            fclose(f); return -1;
        }
        if (fread(table[i].ptr, ksize(kind), (size_t)cnt, f) != (size_t)cnt) { fclose(f); return -1; }
    }
    fclose(f);
    return 0;
}

static long used_count(const entry *e)
{
    long nodenum_ = xnumx * ynumy;
    if (e->cap == 1) return 1;
    if (!strcmp(e->name, "bondm")) return 4 * nodenum_;
    if (!strcmp(e->name, "bondn")) return 3 * bondnum;
    if (!strcmp(e->name, "bondv")) return 4 * bondnum;
    if (!strcmp(e->name, "gx")) return xnumx;
    if (!strcmp(e->name, "gy")) return ynumy;
    if (!strncmp(e->name, "mark", 4) && e->cap == MAXMRK) return marknum + 1;
    if (e->cap == MAXMRK) return marknum + 1;     /* fric_static */
    if (!strncmp(e->name, "mark", 4) || !strcmp(e->name, "markim")) return MAXTMR;
    if (!strncmp(e->name, "dv_", 3) || !strncmp(e->name, "ai_", 3)) return drv_n;
    if (!strcmp(e->name, "eps") || !strcmp(e->name, "drv_ret") || !strcmp(e->name, "pkf") ||
        !strcmp(e->name, "timesave"))
        return e->cap;
    return nodenum_ + 1;
}

static int write_output(const char *path)
{
    FILE *f = fopen(path, "wb");
    if (!f) { perror(path); return -1; }
    for (long i = 0; i < NTAB; i++) {
        char name[32] = {0};
        strncpy(name, table[i].name, 31);
        long long cnt = used_count(&table[i]);
        fwrite(name, 1, 32, f);
        fwrite(&table[i].kind, 1, 1, f);
        fwrite(&cnt, 8, 1, f);
        fwrite(table[i].ptr, ksize(table[i].kind), (size_t)cnt, f);
    }
    fclose(f);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3) { fprintf(stderr, "usage: ref_driver IN OUT op...\n"); return 2; }
    load_extrainput();   /* reference defaults (input_RSF.c:13-63); IN may override */
    if (read_input(argv[1])) return 3;
    /* derived grid globals (load_save.c:599-641) */
    nodenum = xnumx * ynumy; nodenum2 = nodenum * 2; nodenum3 = nodenum * 3;
    cellnum = (xnumx - 1) * (ynumy - 1);
    xstpx = xsize / (double)(xnumx - 1); ystpy = ysize / (double)(ynumy - 1);
    for (long i = 0; i < nodenum; i++) {
        RK_sliprs[i][0] = RK_sliprs0[i]; RK_sliprs[i][1] = RK_sliprs1[i];
        RK_phidts[i][0] = RK_phidts0[i]; RK_phidts[i][1] = RK_phidts1[i];
    }
#ifdef V2C_DROPIN
    {
        const int rc = v2c_dropin_attach(marknum > 0 ? marknum : 1, 0);
        if (rc) { fprintf(stderr, "ref_driver: v2c_dropin_attach failed (%d)\n", rc); return 6; }
    }
#endif
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int a = 3; a < argc; a++) {
        const char *op = argv[a];
        double dt; int rk;
        struct timespec o0, o1;
        int slot = -1;
        clock_gettime(CLOCK_MONOTONIC, &o0);
        if (!strcmp(op, "m2n")) { marker2nodes(); slot = 0; }
        else if (!strcmp(op, "n2m")) { nodes2markers(); slot = 3; }
        else if (!strcmp(op, "move")) { moveandrotatemarkers(); slot = 4; }
        else if (sscanf(op, "before:%lf:%d", &dt, &rk) == 2) { drv_ret[0] = compute_before_solve(dt, rk); slot = 1; }
        else if (sscanf(op, "after:%lf:%d", &dt, &rk) == 2) { drv_ret[1] = compute_after_solve(dt, rk); slot = 2; }
        else if (!strcmp(op, "hostsolve")) stand_in_stokes();
        else if (!strcmp(op, "hostout")) drv_ret[3] = stand_in_output();
        else if (!strcmp(op, "ductile")) {
            for (long i = 0; i < drv_n; i++)
                dv_out[i] = ductilevisc((int)dv_type[i], dv_sxy[i], dv_sxx[i], 0, 0, dv_tk[i], dv_pb[i]);
        } else if (!strcmp(op, "allinteri")) {
            for (long i = 0; i < drv_n; i++)
                allinteri(ai_x[i], ai_y[i], &ai_vx[i], &ai_vy[i], &ai_esp[i], &ai_ee[i]);
        } else if (sscanf(op, "stress:%lf:%d", &dt, &rk) == 2) recompute_stress_strainrate(dt, rk);
        else if (sscanf(op, "c2bold:%lf", &dt) == 1) interpolate_center2basic_old(dt);
        else if (sscanf(op, "c2b:%lf", &dt) == 1) interpolate_center2basic(dt);
        else if (!strcmp(op, "b2c")) interpolate_basic2center();
        else if (sscanf(op, "copy:%d", &rk) == 1) copyarrays(rk);
        else if (sscanf(op, "tstep:%d", &rk) == 1) drv_ret[2] = calc_timestep(rk);
        else if (!strcmp(op, "bcvxvy")) vxvy_boundarycorrection();
        else if (!strcmp(op, "bcp")) P_boundarycorrection();
        else if (!strcmp(op, "bctk")) tkrecalc();
        else if (sscanf(op, "save:%d:", &rk) == 1 && strlen(strchr(op + 5, ':') + 1) < sizeof fl1out) {
            strcpy(fl1out, strchr(op + 5, ':') + 1);
            save_model(0, rk, 0);
        } else if (!strncmp(op, "load:", 5) && strlen(op + 5) < sizeof fl1in) {
            strcpy(fl1in, op + 5);
            load_model();
            nodenum = xnumx * ynumy; nodenum3 = nodenum * 3;
        }
#ifdef V2C_DROPIN
        else if (sscanf(op, "heat:%lf", &dt) == 1) v2c_dropin_temperature_markers(dt);
#endif
        else { fprintf(stderr, "ref_driver: unknown op %s\n", op); return 4; }
        if (slot >= 0) {
            clock_gettime(CLOCK_MONOTONIC, &o1);
            drv_optime[slot] += (double)(o1.tv_sec - o0.tv_sec) + 1e-9 * (double)(o1.tv_nsec - o0.tv_nsec);
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    drv_time = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
#ifdef V2C_DROPIN
    v2c_dropin_markers_to_host();
    v2c_dropin_detach();
    if (errcheck) fprintf(stderr, "ref_driver: errcheck set by the binding\n");
#endif
    for (long i = 0; i < nodenum; i++) {
        RK_sliprs0[i] = RK_sliprs[i][0]; RK_sliprs1[i] = RK_sliprs[i][1];
        RK_phidts0[i] = RK_phidts[i][0]; RK_phidts1[i] = RK_phidts[i][1];
    }
    return write_output(argv[2]) ? 5 : 0;
}
