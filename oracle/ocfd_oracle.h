/*
 * ocfd_oracle.h -- CPU restatement of the OpenCFD-SC right-hand side and RK3
 * step, used ONLY as a test oracle and CPU baseline (never linked into or
 * called by the product library).  See ocfd_oracle.c for the per-function
 * citations of the reference and the list of checks that pin each function.
 *
 * The API mirrors the product C ABI (include/ocfd_hip.h) with an oracle_
 * prefix so that tests can drive both through the same call sequence.
 */
#ifndef OCFD_ORACLE_H
#define OCFD_ORACLE_H

#include <stdint.h>
#include "../include/ocfd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_block oracle_block;

int oracle_create(const ocfd_config *cfg, oracle_block **out);
int oracle_destroy(oracle_block *b);

int oracle_set_grid(oracle_block *b, const double *x, const double *y, const double *z,
                    const double spans[9]);
int oracle_set_grid_and_jacobian(oracle_block *b, const double *x, const double *y,
                                 const double *z, const double spans[9],
                                 const double *const m10[10]);
int oracle_upload_metrics(oracle_block *b, const double *const metrics[19]);
int oracle_download_metrics(oracle_block *b, double *const metrics[19]);

int oracle_upload_primitives(oracle_block *b, const double *d, const double *u,
                             const double *v, const double *w, const double *T,
                             int32_t istep, double tt);
int oracle_upload_f(oracle_block *b, const double *f);
int oracle_download_f(oracle_block *b, double *f);
int oracle_download_fn(oracle_block *b, double *fn);
int oracle_download_du(oracle_block *b, double *du);
int oracle_upload_du(oracle_block *b, const double *du);
int oracle_download_primitives(oracle_block *b, double *d, double *u, double *v,
                               double *w, double *T);
int oracle_download_rhybrid(oracle_block *b, double *r);
int oracle_get_time(oracle_block *b, int32_t *istep, double *tt);

int oracle_exchange_primitives(oracle_block *b);
int oracle_comput_rhybrid(oracle_block *b, int64_t counts[4]);
int oracle_du_inviscous(oracle_block *b);
int oracle_du_inviscous_character(oracle_block *b);
int oracle_du_viscous(oracle_block *b);
int oracle_time_adv_rk3(oracle_block *b, int32_t krk);
int oracle_save_fn(oracle_block *b);
int oracle_comput_duvwt(oracle_block *b);
int oracle_bc(oracle_block *b);
int oracle_filtering(oracle_block *b);
int oracle_advance_clock(oracle_block *b);
int oracle_step(oracle_block *b, int32_t nsteps);
int oracle_show_flow_msg(oracle_block *b, double out[3]);
int oracle_ana_box(oracle_block *b, double out[20]);

/* Pointwise / pencil building blocks, exported for the unit tests. */
double oracle_hh_weno5p(const double *v0);
double oracle_hh_weno5m(const double *v0);
double oracle_hh_weno7p(const double *v0);
double oracle_hh_weno7m(const double *v0);
double oracle_hh_omp6p(const double *v0);
double oracle_hh_omp6m(const double *v0);
double oracle_hh_weno5p_boundary(const double *v0, int del);
double oracle_hh_weno5m_boundary(const double *v0, int del);
void oracle_split(int which, int n, int lap, double gamma, const double *c, const double *d,
                  const double *u, const double *v, const double *w, const double *a1,
                  const double *a2, const double *a3, double *fp, double *fm);
void oracle_char_matrices(double gamma, const double st[7], double S[25], double S1[25]);
/* CD6/CD8 first derivative along axis (0,1,2) of a halo'd array -> interior. */
int oracle_dxyz0(oracle_block *b, int axis, const double *fh, double *out);

#ifdef __cplusplus
}
#endif
#endif
