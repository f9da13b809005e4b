/*
 * host_stand_in.h -- host code that stands in for the reference's host-side
 * steps in the drop-in tests, compiled into both the reference driver
 * (oracle/ref_driver.c) and the binding's test host
 * (tests/native/dropin_harness.c), so the two run the same host arithmetic
 * on the same globals.
 *
 * stand_in_stokes: in place of preparematrix / gausmat4_solve / reloadVxVyP
 * (momentum_continuity_iterate.c:76-83): reads arrays that marker2nodes,
 * compute_before_solve and interpolate_basic2center wrote (ro, bm, sppe,
 * sxxe, sxye, mvx, mvy, nd, nu) and writes pr, vx, vy on the nodes
 * reloadVxVyP writes (:190-206).  The velocities take only marker2nodes'
 * outputs, which the device reproduces bit for bit, so the marker positions
 * after the step can be compared bit for bit.
 *
 * stand_in_output: in place of the per-step diagnostics and postproc
 * (momentum_continuity_iterate.c:129-170, output_RSF.c:2-72): a sum over the
 * arrays they read, returned for comparison.
 *
 * Needs the reference globals (xnumx, ynumy, nodenum and the nodal arrays)
 * declared before inclusion.
 */
#ifndef HOST_STAND_IN_H
#define HOST_STAND_IN_H
#include <math.h>

static void stand_in_stokes(void)
{
    for (long m3 = 0; m3 < nodenum; m3++) {
        const long m2 = m3 % ynumy, m1 = m3 / ynumy;
        if (m1 && m2) pr[m3] = sppe[m3] + 1e-3 * ro[m3] + 1e-12 * bm[m3] + 1e-30 * (nd[m3] + nu[m3]);
        if (m2 < ynumy - 1) vx[m3] = mvx[m3] * (1.0 + 1e-3 * tanh(1e-7 * sxxe[m3]));
        if (m1 < xnumx - 1) vy[m3] = mvy[m3] * (1.0 + 1e-3 * tanh(1e-7 * sxye[m3]));
    }
}

static double stand_in_output(void)
{
    double s = 0.0;
    for (long m3 = 0; m3 < nodenum; m3++)
        s += fabs(sxxs[m3]) + fabs(sxy[m3]) + 1e-20 * fabs(nu[m3]) + fabs(sliprs[m3]) + fabs(phis_it[m3]) +
             fabs(syields[m3]);
    return s;
}
#endif
