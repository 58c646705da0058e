/* A C caller of the library, as a maintainer would write one from
 * include/ocfd_hip.h: a periodic 32^3 box of uniform flow, hybrid
 * characteristic scheme with LLF splitting and CD6 viscous terms, 20 whole
 * steps, then the primitives back.  A uniform state has zero right-hand
 * side, so the state must come back unchanged up to the rounding of the RK3
 * stage combinations and of the primitive recovery (relative 1e-13).  Exit
 * status: 0 = unchanged, 1 = the library refused (message printed),
 * 2 = changed. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ocfd_hip.h"

int main(void) {
    const int n = 32, lap = 4, nh = n + 2 * lap;
    const size_t np = (size_t)nh * nh * nh;
    double *q[5], *r[5];
    const double init[5] = {1.0, 0.3, -0.2, 0.1, 1.0}; /* d, u, v, w, T */
    for (int m = 0; m < 5; m++) {
        q[m] = malloc(np * sizeof(double));
        r[m] = malloc(np * sizeof(double));
        if (!q[m] || !r[m]) return 1;
        for (size_t p = 0; p < np; p++) q[m][p] = init[m];
    }
    ocfd_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.nx = cfg.ny = cfg.nz = n;
    cfg.lap = lap;
    cfg.iperiodic[0] = cfg.iperiodic[1] = cfg.iperiodic[2] = 1;
    cfg.scheme_invis = OCFD_SCHEME_HYBRID;
    cfg.scheme_vis = OCFD_SCHEME_CD6;
    cfg.flux_splitting = OCFD_SPLIT_LLF;
    cfg.if_scheme_character = 1;
    cfg.if_viscous = 1;
    cfg.ma = 0.5; cfg.re = 500.0; cfg.gamma = 1.4; cfg.pr = 0.7; cfg.ref_amu_t0 = 288.15; cfg.dt = 1e-3;
    cfg.ud7l_diss = 1.0;
    cfg.hybrid_para[0] = 1; cfg.hybrid_para[1] = 0.01; cfg.hybrid_para[2] = 0.05;
    cfg.istep_show = 1000000;
    cfg.metric_mode = OCFD_METRIC_UNIFORM;

    ocfd_handle *h = NULL;
    if (ocfd_create(&cfg, 0, &h)) {
        printf("ocfd_create: %s\n", ocfd_last_error());
        return 1;
    }
    /* uniform metrics of a 2 pi box on the computational grid h = 1/(n-1) */
    const double pi = 3.14159265358979323846, s = 2.0 * pi / n * (n - 1), jac = 1.0 / (s * s * s);
    const double m10[10] = {jac * s * s, 0, 0, 0, jac * s * s, 0, 0, 0, jac * s * s, jac};
    int rc = ocfd_set_uniform_metrics(h, m10);
    rc = rc ? rc : ocfd_upload_primitives(h, q[0], q[1], q[2], q[3], q[4], 0, 0.0);
    rc = rc ? rc : ocfd_bc(h);
    rc = rc ? rc : ocfd_step(h, 20);
    rc = rc ? rc : ocfd_download_primitives(h, r[0], r[1], r[2], r[3], r[4]);
    if (rc) {
        printf("ocfd: %s\n", ocfd_last_error());
        ocfd_destroy(h);
        return 1;
    }
    int32_t istep = 0;
    double tt = 0.0;
    ocfd_get_time(h, &istep, &tt);
    ocfd_destroy(h);
    size_t changed = 0;
    double worst = 0.0;
    for (int m = 0; m < 5; m++)
        for (size_t p = 0; p < np; p++) {
            const double e = fabs(r[m][p] - q[m][p]) / fabs(q[m][p]);
            worst = e > worst ? e : worst;
            changed += !(e <= 1e-13);
        }
    printf("steps %d, t %.6f, largest relative change %.3e, values changed %zu\n", istep, tt, worst, changed);
    return changed ? 2 : 0;
}
