/*
 * ocfd_oracle.c -- TEST ORACLE.  A scalar C restatement of the OpenCFD-SC
 * right-hand side and RK3 time step for one block on one rank.
 *
 * Role: this file is test infrastructure.  Only the tests, the smoke check and
 * the CPU-baseline leg of the benchmark load it, and only as the checker or the
 * CPU baseline.  The product (opencfd-sc_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Expressions keep
 * the reference's operand order and association, and the file is compiled with
 * -ffp-contract=off, so on IEEE-754 hardware it reproduces the reference bit
 * for bit (pinned against the compiled reference in tests/golden/).
 *
 * Pinning (tests/test_oracle_*.py):
 *  - golden vectors written by the compiled reference (oracle/ref_probe.f90)
 *    for every routine of a stage and for whole steps, on several schemes,
 *    splittings, boundary kinds and grids: bitwise or <= 1e-14 relative;
 *  - WENO5/WENO7/UD7L: linear data -> closed-form linear reconstructions;
 *    P/M mirror symmetry; constant data -> zero flux difference;
 *  - SW/LLF: fp + fm equals the projected Euler flux; SW fm = 0 when all
 *    eigenvalues are positive;
 *  - S*S1 = I for the characteristic matrices;
 *  - CD6/CD8: exact on polynomials of degree <= 6 / 8;
 *  - RK3: amplification factor 1 + z + z^2/2 + z^3/6 on du = lambda*f;
 *  - Fo9p: D(0) = 0, D(pi) = 1;
 *  - periodic box conserves sum f/J.
 */
#include "ocfd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* block state (flow_data, OCFD_parameters.f90:99-108, plus work arrays)   */
/* ---------------------------------------------------------------------- */
enum { DEL_NONE = 0, DEL_LIFT = 1, DEL_RIGHT = 2 };
enum { M_AKX, M_AKY, M_AKZ, M_AIX, M_AIY, M_AIZ, M_ASX, M_ASY, M_ASZ, M_AJAC,
       M_AKX1, M_AKY1, M_AKZ1, M_AIX1, M_AIY1, M_AIZ1, M_ASX1, M_ASY1, M_ASZ1, M_N };

struct oracle_block {
    ocfd_config c;
    int nx, ny, nz, lap, sx, sy, sz;
    size_t nh, ni;
    double hx, hy, hz, cp, cv;
    int bidx[3][2];            /* Scheme%Bound_index(1:2, 1:3) */
    int ka1, kb1, ka2, kb2;    /* Scheme%Ka1 ... */
    int kh[3][4];              /* [H1,H2,H3][Ka1,Kb1,Ka2,Kb2] */
    double ud7l[8];
    int istep;
    double tt;
    double *d, *u, *v, *w, *T; /* halo'd */
    double *f, *fn, *du;       /* (nx,ny,nz,5) */
    double *amu;               /* (nx,ny,nz) */
    double *axx, *ayy, *azz;   /* halo'd coordinates */
    double *met[M_N];          /* halo'd metrics */
    double *rhyb;              /* halo'd */
    double *cc;                /* halo'd sound speed */
    double *ev[3][4];          /* halo'd viscous fluxes Ev1..Ev3 (1:4) */
    double *grad[12];          /* uk,ui,us, vk,vi,vs, wk,wi,ws, Tk,Ti,Ts */
    double *f0, *pf;           /* filter scratch (halo'd) */
    int nmax;                  /* max(nx,ny,nz) */
    double *pen[8], *fp, *fm, *hh1, *hh2, *hhc; /* pencil work */
    int *scm;
};

#define HI(b, i, j, k) ((size_t)((i) + (b)->lap - 1) + (size_t)(b)->sx * ((size_t)((j) + (b)->lap - 1) + (size_t)(b)->sy * (size_t)((k) + (b)->lap - 1)))
#define II(b, i, j, k) ((size_t)((i) - 1) + (size_t)(b)->nx * ((size_t)((j) - 1) + (size_t)(b)->ny * (size_t)((k) - 1)))
#define FI(b, i, j, k, m) (II(b, i, j, k) + (b)->ni * (size_t)((m) - 1))

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static void fill(double *a, size_t n, double v) { for (size_t i = 0; i < n; i++) a[i] = v; }

/* Fortran sign(1.d0, x): +1 for x >= +0, -1 for x < 0 or x = -0. */
static double sgn1(double x) { return copysign(1.0, x); }
static double dmin2(double a, double b) { return a < b ? a : b; }
static double dmax2(double a, double b) { return a > b ? a : b; }
static double dmin3(double a, double b, double c) { return dmin2(dmin2(a, b), c); }
static double dmax3(double a, double b, double c) { return dmax2(dmax2(a, b), c); }

/* ---------------------------------------------------------------------- */
/* set_parameters (OCFD_readpara.f90:261-341)                              */
/* ---------------------------------------------------------------------- */
static int set_parameters(oracle_block *b) {
    const ocfd_config *c = &b->c;
    b->hx = 1.0 / ((double)c->nx - 1.0);
    b->hy = 1.0 / ((double)c->ny - 1.0);
    b->hz = 1.0 / ((double)c->nz - 1.0);
    b->cv = 1.0 / (c->gamma * (c->gamma - 1.0) * c->ma * c->ma);
    b->cp = b->cv * c->gamma;
    for (int a = 0; a < 3; a++) {
        b->bidx[a][0] = c->iperiodic[a] == 0 ? 1 : 0;
        b->bidx[a][1] = c->iperiodic[a] == 0 ? 1 : 0;
    }
    switch (c->scheme_invis) {
    case OCFD_SCHEME_WENO5: b->ka1 = -2; b->kb1 = 2; b->ka2 = -1; b->kb2 = 3; break;
    case OCFD_SCHEME_WENO7: b->ka1 = -3; b->kb1 = 3; b->ka2 = -2; b->kb2 = 4; break;
    case OCFD_SCHEME_OMP6:
    case OCFD_SCHEME_UD7L:
    case OCFD_SCHEME_HYBRID: b->ka1 = -3; b->kb1 = 4; b->ka2 = -3; b->kb2 = 4; break;
    default: return -1;
    }
    int h[3][4] = {{-3, 4, -3, 4}, {-3, 3, -2, 4}, {-2, 2, -1, 3}};
    memcpy(b->kh, h, sizeof h);
    /* set_scheme_para (OCFD_readpara.f90:334-341) */
    const double ud7[8] = {-3.0, 25.0, -101.0, 319.0, 214.0, -38.0, 4.0, 0.0};
    const double cd8[8] = {-3.0, 29.0, -139.0, 533.0, 533.0, -139.0, 29.0, -3.0};
    for (int i = 0; i < 8; i++)
        b->ud7l[i] = c->ud7l_diss * ud7[i] / 420.0 + (1.0 - c->ud7l_diss) * cd8[i] / 840.0;
    if (c->scheme_vis != OCFD_SCHEME_CD6 && c->scheme_vis != OCFD_SCHEME_CD8) return -1;
    return 0;
}

int oracle_create(const ocfd_config *cfg, oracle_block **out) {
    if (!cfg || !out || cfg->lap != 4 || cfg->nx < 8 || cfg->ny < 8 || cfg->nz < 8) return -1;
    oracle_block *b = (oracle_block *)calloc(1, sizeof *b);
    b->c = *cfg;
    b->nx = cfg->nx; b->ny = cfg->ny; b->nz = cfg->nz; b->lap = cfg->lap;
    b->sx = b->nx + 2 * b->lap; b->sy = b->ny + 2 * b->lap; b->sz = b->nz + 2 * b->lap;
    b->nh = (size_t)b->sx * b->sy * b->sz;
    b->ni = (size_t)b->nx * b->ny * b->nz;
    if (set_parameters(b)) { free(b); return -1; }
    /* allocate_flow_data (OCFD_NS_Solver.f90:88-151): primitives and metrics
     * start at 1.0, du and Amu at 0 */
    double **h1[] = {&b->d, &b->u, &b->v, &b->w, &b->T, &b->axx, &b->ayy, &b->azz};
    for (size_t q = 0; q < sizeof h1 / sizeof h1[0]; q++) { *h1[q] = zalloc(b->nh); fill(*h1[q], b->nh, 1.0); }
    for (int m = 0; m < M_N; m++) { b->met[m] = zalloc(b->nh); fill(b->met[m], b->nh, 1.0); }
    b->rhyb = zalloc(b->nh);
    b->cc = zalloc(b->nh);
    b->f0 = zalloc(b->nh); b->pf = zalloc(b->nh);
    for (int a = 0; a < 3; a++) for (int m = 0; m < 4; m++) b->ev[a][m] = zalloc(b->nh);
    b->f = zalloc(5 * b->ni); b->fn = zalloc(5 * b->ni); b->du = zalloc(5 * b->ni);
    b->amu = zalloc(b->ni);
    for (int q = 0; q < 12; q++) b->grad[q] = zalloc(b->ni);
    b->nmax = b->nx > b->ny ? b->nx : b->ny;
    if (b->nz > b->nmax) b->nmax = b->nz;
    int np = b->nmax + 2 * b->lap;
    for (int q = 0; q < 8; q++) b->pen[q] = zalloc(np);
    b->fp = zalloc(5 * np); b->fm = zalloc(5 * np);
    b->hh1 = zalloc(b->nmax + 1); b->hh2 = zalloc(b->nmax + 1);
    b->hhc = zalloc(5 * (b->nmax + 1));
    b->scm = (int *)calloc(b->nmax + 1, sizeof(int));
    *out = b;
    return 0;
}

int oracle_destroy(oracle_block *b) {
    if (!b) return 0;
    free(b->d); free(b->u); free(b->v); free(b->w); free(b->T);
    free(b->axx); free(b->ayy); free(b->azz);
    for (int m = 0; m < M_N; m++) free(b->met[m]);
    free(b->rhyb); free(b->cc); free(b->f0); free(b->pf);
    for (int a = 0; a < 3; a++) for (int m = 0; m < 4; m++) free(b->ev[a][m]);
    free(b->f); free(b->fn); free(b->du); free(b->amu);
    for (int q = 0; q < 12; q++) free(b->grad[q]);
    for (int q = 0; q < 8; q++) free(b->pen[q]);
    free(b->fp); free(b->fm); free(b->hh1); free(b->hh2); free(b->hhc); free(b->scm);
    free(b);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* exchange_boundary_x/y/z on a single periodic rank (OCFD_mpi.f90:198-233):
 * the two Sendrecv's with ID_*M1 = ID_*P1 = self copy interior planes 1..LAP
 * to n+1..n+LAP, then n+1-LAP..n to 1-LAP..0, over the interior range of the
 * other two indices only (TYPE_LAP*2, OCFD_mpi.f90:114-122).  Non-periodic:
 * MPI_PROC_NULL, nothing happens.                                           */
/* ---------------------------------------------------------------------- */
static void exchange_x(oracle_block *b, double *a) {
    if (b->c.iperiodic[0] != 1) return;
    int n = b->nx, L = b->lap;
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) {
        for (int l = 1; l <= L; l++) a[HI(b, n + l, j, k)] = a[HI(b, l, j, k)];
        for (int l = 1; l <= L; l++) a[HI(b, l - L, j, k)] = a[HI(b, n - L + l, j, k)];
    }
}
static void exchange_y(oracle_block *b, double *a) {
    if (b->c.iperiodic[1] != 1) return;
    int n = b->ny, L = b->lap;
    for (int k = 1; k <= b->nz; k++) {
        for (int l = 1; l <= L; l++) for (int i = 1; i <= b->nx; i++) a[HI(b, i, n + l, k)] = a[HI(b, i, l, k)];
        for (int l = 1; l <= L; l++) for (int i = 1; i <= b->nx; i++) a[HI(b, i, l - L, k)] = a[HI(b, i, n - L + l, k)];
    }
}
static void exchange_z(oracle_block *b, double *a) {
    if (b->c.iperiodic[2] != 1) return;
    int n = b->nz, L = b->lap;
    for (int l = 1; l <= L; l++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++)
        a[HI(b, i, j, n + l)] = a[HI(b, i, j, l)];
    for (int l = 1; l <= L; l++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++)
        a[HI(b, i, j, l - L)] = a[HI(b, i, j, n - L + l)];
}
static void exchange_xyz(oracle_block *b, double *a) { exchange_x(b, a); exchange_y(b, a); exchange_z(b, a); }

/* ---------------------------------------------------------------------- */
/* flux-vector splitting (OCFD_flux_split.f90)                              */
/* arrays are pencils starting at index 1-LAP; fp/fm are [5][n+2LAP]        */
/* ---------------------------------------------------------------------- */
/* split_Stager_Warming (OCFD_flux_split.f90:5-70) */
static void split_sw(int n, int lap, double gamma, const double *c1, const double *d1,
                     const double *u1, const double *v1, const double *w1, const double *A1,
                     const double *A2, const double *A3, double *fp, double *fm) {
    int np = n + 2 * lap;
    double tmp1 = 2.0 * (gamma - 1.0);
    double tmp3 = (3.0 - gamma) / (2.0 * (gamma - 1.0));
    for (int i = 0; i < np; i++) {
        double ss = sqrt(A1[i] * A1[i] + A2[i] * A2[i] + A3[i] * A3[i]);
        double ss1 = 1.0 / ss;
        double ak1 = A1[i] * ss1, ak2 = A2[i] * ss1, ak3 = A3[i] * ss1;
        double vs = A1[i] * u1[i] + A2[i] * v1[i] + A3[i] * w1[i];
        double E1 = vs, E2 = vs - c1[i] * ss, E3 = vs + c1[i] * ss;
        double E1P = (E1 + fabs(E1)) * 0.5, E2P = (E2 + fabs(E2)) * 0.5, E3P = (E3 + fabs(E3)) * 0.5;
        double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
        double tmp0 = d1[i] / (2.0 * gamma);
        double uc1 = u1[i] - c1[i] * ak1, uc2 = u1[i] + c1[i] * ak1;
        double vc1 = v1[i] - c1[i] * ak2, vc2 = v1[i] + c1[i] * ak2;
        double wc1 = w1[i] - c1[i] * ak3, wc2 = w1[i] + c1[i] * ak3;
        double vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * 0.5;
        double vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * 0.5;
        double vv = (gamma - 1.0) * (u1[i] * u1[i] + v1[i] * v1[i] + w1[i] * w1[i]);
        double W2 = tmp3 * c1[i] * c1[i];
        fp[0 * np + i] = tmp0 * (tmp1 * E1P + E2P + E3P);
        fp[1 * np + i] = tmp0 * (tmp1 * E1P * u1[i] + E2P * uc1 + E3P * uc2);
        fp[2 * np + i] = tmp0 * (tmp1 * E1P * v1[i] + E2P * vc1 + E3P * vc2);
        fp[3 * np + i] = tmp0 * (tmp1 * E1P * w1[i] + E2P * wc1 + E3P * wc2);
        fp[4 * np + i] = tmp0 * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
        fm[0 * np + i] = tmp0 * (tmp1 * E1M + E2M + E3M);
        fm[1 * np + i] = tmp0 * (tmp1 * E1M * u1[i] + E2M * uc1 + E3M * uc2);
        fm[2 * np + i] = tmp0 * (tmp1 * E1M * v1[i] + E2M * vc1 + E3M * vc2);
        fm[3 * np + i] = tmp0 * (tmp1 * E1M * w1[i] + E2M * wc1 + E3M * wc2);
        fm[4 * np + i] = tmp0 * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
    }
}

/* split_Local_LaxFriedrichs (OCFD_flux_split.f90:73-104) */
static void split_llf(int n, int lap, double gamma, const double *c1, const double *d1,
                      const double *u1, const double *v1, const double *w1, const double *A1,
                      const double *A2, const double *A3, double *fp, double *fm) {
    int np = n + 2 * lap;
    double tmp0 = 1.0 / gamma, tmp1 = 1.0 / (gamma - 1.0);
    for (int i = 0; i < np; i++) {
        double ss = sqrt(A1[i] * A1[i] + A2[i] * A2[i] + A3[i] * A3[i]);
        double un = A1[i] * u1[i] + A2[i] * v1[i] + A3[i] * w1[i];
        double lamda0 = fabs(un) + ss * c1[i];
        double p1 = tmp0 * d1[i] * c1[i] * c1[i];
        double E0 = tmp1 * p1 + 0.5 * d1[i] * (u1[i] * u1[i] + v1[i] * v1[i] + w1[i] * w1[i]);
        double h0 = E0 + p1;
        fp[0 * np + i] = 0.5 * (d1[i] * un + lamda0 * d1[i]);
        fp[1 * np + i] = 0.5 * (d1[i] * u1[i] * un + A1[i] * p1 + lamda0 * d1[i] * u1[i]);
        fp[2 * np + i] = 0.5 * (d1[i] * v1[i] * un + A2[i] * p1 + lamda0 * d1[i] * v1[i]);
        fp[3 * np + i] = 0.5 * (d1[i] * w1[i] * un + A3[i] * p1 + lamda0 * d1[i] * w1[i]);
        fp[4 * np + i] = 0.5 * (h0 * un + lamda0 * E0);
        fm[0 * np + i] = 0.5 * (d1[i] * un - lamda0 * d1[i]);
        fm[1 * np + i] = 0.5 * (d1[i] * u1[i] * un + A1[i] * p1 - lamda0 * d1[i] * u1[i]);
        fm[2 * np + i] = 0.5 * (d1[i] * v1[i] * un + A2[i] * p1 - lamda0 * d1[i] * v1[i]);
        fm[3 * np + i] = 0.5 * (d1[i] * w1[i] * un + A3[i] * p1 - lamda0 * d1[i] * w1[i]);
        fm[4 * np + i] = 0.5 * (h0 * un - lamda0 * E0);
    }
}

void oracle_split(int which, int n, int lap, double gamma, const double *c, const double *d,
                  const double *u, const double *v, const double *w, const double *a1,
                  const double *a2, const double *a3, double *fp, double *fm) {
    if (which == OCFD_SPLIT_SW) split_sw(n, lap, gamma, c, d, u, v, w, a1, a2, a3, fp, fm);
    else split_llf(n, lap, gamma, c, d, u, v, w, a1, a2, a3, fp, fm);
}

/* ---------------------------------------------------------------------- */
/* Point-wise reconstructions, v0 points at v(0) = value at the face's     */
/* left cell (OCFD_Schemes_1.f90).  The pencil versions of OCFD_Schemes.f90 */
/* evaluate the same expressions in the same order (OCFD_Schemes.f90:6-358) */
/* ---------------------------------------------------------------------- */
static const double W5_EP = 1.0e-6, W5_C0 = 3.0 / 10.0, W5_C1 = 3.0 / 5.0, W5_C2 = 1.0 / 10.0;

/* hh_weno5P (OCFD_Schemes_1.f90:9-29) == OCFD_weno5P (OCFD_Schemes.f90:6-28) */
double oracle_hh_weno5p(const double *v) {
    double t, S0, S1, S2;
    t = v[0] - 2.0 * v[1] + v[2]; S0 = 13.0 / 12.0 * (t * t);
    t = 3.0 * v[0] - 4.0 * v[1] + v[2]; S0 = S0 + 1.0 / 4.0 * (t * t);
    t = v[-1] - 2.0 * v[0] + v[1]; S1 = 13.0 / 12.0 * (t * t);
    t = v[-1] - v[1]; S1 = S1 + 1.0 / 4.0 * (t * t);
    t = v[-2] - 2.0 * v[-1] + v[0]; S2 = 13.0 / 12.0 * (t * t);
    t = v[-2] - 4.0 * v[-1] + 3.0 * v[0]; S2 = S2 + 1.0 / 4.0 * (t * t);
    double a0 = W5_C0 / ((W5_EP + S0) * (W5_EP + S0));
    double a1 = W5_C1 / ((W5_EP + S1) * (W5_EP + S1));
    double a2 = W5_C2 / ((W5_EP + S2) * (W5_EP + S2));
    double am = a0 + a1 + a2;
    double q03 = 1.0 / 3.0 * v[0] + 5.0 / 6.0 * v[1] - 1.0 / 6.0 * v[2];
    double q13 = -(1.0 / 6.0 * v[-1]) + 5.0 / 6.0 * v[0] + 1.0 / 3.0 * v[1];
    double q23 = 1.0 / 3.0 * v[-2] - 7.0 / 6.0 * v[-1] + 11.0 / 6.0 * v[0];
    return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

/* hh_weno5M (OCFD_Schemes_1.f90:31-52) == OCFD_weno5M (OCFD_Schemes.f90:30-53) */
double oracle_hh_weno5m(const double *v) {
    double t, S0, S1, S2;
    t = v[1] - 2.0 * v[0] + v[-1]; S0 = 13.0 / 12.0 * (t * t);
    t = 3.0 * v[1] - 4.0 * v[0] + v[-1]; S0 = S0 + 1.0 / 4.0 * (t * t);
    t = v[2] - 2.0 * v[1] + v[0]; S1 = 13.0 / 12.0 * (t * t);
    t = v[2] - v[0]; S1 = S1 + 1.0 / 4.0 * (t * t);
    t = v[3] - 2.0 * v[2] + v[1]; S2 = 13.0 / 12.0 * (t * t);
    t = v[3] - 4.0 * v[2] + 3.0 * v[1]; S2 = S2 + 1.0 / 4.0 * (t * t);
    double a0 = W5_C0 / ((W5_EP + S0) * (W5_EP + S0));
    double a1 = W5_C1 / ((W5_EP + S1) * (W5_EP + S1));
    double a2 = W5_C2 / ((W5_EP + S2) * (W5_EP + S2));
    double am = a0 + a1 + a2;
    double q03 = 1.0 / 3.0 * v[1] + 5.0 / 6.0 * v[0] - 1.0 / 6.0 * v[-1];
    double q13 = -(1.0 / 6.0 * v[2]) + 5.0 / 6.0 * v[1] + 1.0 / 3.0 * v[0];
    double q23 = 1.0 / 3.0 * v[3] - 7.0 / 6.0 * v[2] + 11.0 / 6.0 * v[1];
    return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

/* hh_weno5P_boundary (OCFD_Schemes_1.f90:266-295): WENO5 with the
 * out-of-domain sub-stencil removed */
double oracle_hh_weno5p_boundary(const double *v, int del) {
    double t, S0, S1, S2, a0 = 0.0, q03 = 0.0, a2 = 0.0, q23 = 0.0;
    if (del != DEL_RIGHT) {
        t = v[0] - 2.0 * v[1] + v[2]; S0 = 13.0 / 12.0 * (t * t);
        t = 3.0 * v[0] - 4.0 * v[1] + v[2]; S0 = S0 + 1.0 / 4.0 * (t * t);
        a0 = W5_C0 / ((W5_EP + S0) * (W5_EP + S0));
        q03 = 1.0 / 3.0 * v[0] + 5.0 / 6.0 * v[1] - 1.0 / 6.0 * v[2];
    }
    t = v[-1] - 2.0 * v[0] + v[1]; S1 = 13.0 / 12.0 * (t * t);
    t = v[-1] - v[1]; S1 = S1 + 1.0 / 4.0 * (t * t);
    double a1 = W5_C1 / ((W5_EP + S1) * (W5_EP + S1));
    double q13 = -(1.0 / 6.0 * v[-1]) + 5.0 / 6.0 * v[0] + 1.0 / 3.0 * v[1];
    if (del != DEL_LIFT) {
        t = v[-2] - 2.0 * v[-1] + v[0]; S2 = 13.0 / 12.0 * (t * t);
        t = v[-2] - 4.0 * v[-1] + 3.0 * v[0]; S2 = S2 + 1.0 / 4.0 * (t * t);
        a2 = W5_C2 / ((W5_EP + S2) * (W5_EP + S2));
        q23 = 1.0 / 3.0 * v[-2] - 7.0 / 6.0 * v[-1] + 11.0 / 6.0 * v[0];
    }
    double am = a0 + a1 + a2;
    return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

/* hh_weno5M_boundary (OCFD_Schemes_1.f90:297-325) */
double oracle_hh_weno5m_boundary(const double *v, int del) {
    double t, S0, S1, S2, a0 = 0.0, q03 = 0.0, a2 = 0.0, q23 = 0.0;
    if (del != DEL_LIFT) {
        t = v[1] - 2.0 * v[0] + v[-1]; S0 = 13.0 / 12.0 * (t * t);
        t = 3.0 * v[1] - 4.0 * v[0] + v[-1]; S0 = S0 + 1.0 / 4.0 * (t * t);
        a0 = W5_C0 / ((W5_EP + S0) * (W5_EP + S0));
        q03 = 1.0 / 3.0 * v[1] + 5.0 / 6.0 * v[0] - 1.0 / 6.0 * v[-1];
    }
    t = v[2] - 2.0 * v[1] + v[0]; S1 = 13.0 / 12.0 * (t * t);
    t = v[2] - v[0]; S1 = S1 + 1.0 / 4.0 * (t * t);
    double a1 = W5_C1 / ((W5_EP + S1) * (W5_EP + S1));
    double q13 = -(1.0 / 6.0 * v[2]) + 5.0 / 6.0 * v[1] + 1.0 / 3.0 * v[0];
    if (del != DEL_RIGHT) {
        t = v[3] - 2.0 * v[2] + v[1]; S2 = 13.0 / 12.0 * (t * t);
        t = v[3] - 4.0 * v[2] + 3.0 * v[1]; S2 = S2 + 1.0 / 4.0 * (t * t);
        a2 = W5_C2 / ((W5_EP + S2) * (W5_EP + S2));
        q23 = 1.0 / 3.0 * v[3] - 7.0 / 6.0 * v[2] + 11.0 / 6.0 * v[1];
    }
    double am = a0 + a1 + a2;
    return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

/* WENO7-JS constants (OCFD_Schemes.f90:66-80) */
#define W7_C0 (1.0 / 35.0)
#define W7_C1 (12.0 / 35.0)
#define W7_C2 (18.0 / 35.0)
#define W7_C3 (4.0 / 35.0)
#define A11 (-2.0 / 6.0)
#define A12 (9.0 / 6.0)
#define A13 (-18.0 / 6.0)
#define A14 (11.0 / 6.0)
#define A21 (1.0 / 6.0)
#define A23 (3.0 / 6.0)
#define A24 (2.0 / 6.0)
#define A31 (-2.0 / 6.0)
#define A32 (-3.0 / 6.0)
#define A34 (-1.0 / 6.0)
#define A41 (-11.0 / 6.0)
#define A42 (18.0 / 6.0)
#define A43 (-9.0 / 6.0)
#define A44 (2.0 / 6.0)
#define B12 4.0
#define B13 (-5.0)
#define B14 2.0
#define B22 (-2.0)
#define B41 2.0
#define B42 (-5.0)
#define B43 4.0
#define C12 3.0
#define D12 (13.0 / 12.0)
#define D13 (1043.0 / 960.0)
#define D14 (1.0 / 12.0)
#define E11 (-3.0 / 12.0)
#define E12 (13.0 / 12.0)
#define E13 (-23.0 / 12.0)
#define E14 (25.0 / 12.0)
#define E21 (1.0 / 12.0)
#define E22 (-5.0 / 12.0)
#define E23 (13.0 / 12.0)
#define E24 (3.0 / 12.0)
#define E31 (-1.0 / 12.0)
#define E32 (7.0 / 12.0)
#define E33 (7.0 / 12.0)
#define E34 (-1.0 / 12.0)
#define E41 (3.0 / 12.0)
#define E42 (13.0 / 12.0)
#define E43 (-5.0 / 12.0)
#define E44 (1.0 / 12.0)
#define W7_EP 1.0e-8

/* WENO7 core on the 7 values r[0..6] ordered from the upwind end: for the
 * P flux r = v(-3..3), for the M flux r = v(4..-2) (OCFD_Schemes.f90:83-123
 * and 151-191 use the same expressions on the mirrored index). */
static double weno7_core(double m3, double m2, double m1, double c0, double p1, double p2, double p3) {
    double S10 = A11 * m3 + A12 * m2 + A13 * m1 + A14 * c0;
    double S11 = A21 * m2 - m1 + A23 * c0 + A24 * p1;
    double S12 = A31 * m1 + A32 * c0 + p1 + A34 * p2;
    double S13 = A41 * c0 + A42 * p1 + A43 * p2 + A44 * p3;
    double S20 = -m3 + B12 * m2 + B13 * m1 + B14 * c0;
    double S21 = m1 + B22 * c0 + p1;
    double S22 = c0 + B22 * p1 + p2;
    double S23 = B41 * c0 + B42 * p1 + B43 * p2 - p3;
    double S30 = -m3 + C12 * (m2 - m1) + c0;
    double S31 = -m2 + C12 * (m1 - c0) + p1;
    double S32 = -m1 + C12 * (c0 - p1) + p2;
    double S33 = -c0 + C12 * (p1 - p2) + p3;
    double S0 = S10 * S10 + D12 * S20 * S20 + D13 * S30 * S30 + D14 * S10 * S30;
    double S1 = S11 * S11 + D12 * S21 * S21 + D13 * S31 * S31 + D14 * S11 * S31;
    double S2 = S12 * S12 + D12 * S22 * S22 + D13 * S32 * S32 + D14 * S12 * S32;
    double S3 = S13 * S13 + D12 * S23 * S23 + D13 * S33 * S33 + D14 * S13 * S33;
    double a0 = W7_C0 / ((W7_EP + S0) * (W7_EP + S0));
    double a1 = W7_C1 / ((W7_EP + S1) * (W7_EP + S1));
    double a2 = W7_C2 / ((W7_EP + S2) * (W7_EP + S2));
    double a3 = W7_C3 / ((W7_EP + S3) * (W7_EP + S3));
    double am = a0 + a1 + a2 + a3;
    double q0 = E11 * m3 + E12 * m2 + E13 * m1 + E14 * c0;
    double q1 = E21 * m2 + E22 * m1 + E23 * c0 + E24 * p1;
    double q2 = E31 * m1 + E32 * c0 + E33 * p1 + E34 * p2;
    double q3 = E41 * c0 + E42 * p1 + E43 * p2 + E44 * p3;
    return (a0 * q0 + a1 * q1 + a2 * q2 + a3 * q3) / am;
}
/* hh_weno7P (OCFD_Schemes_1.f90:58-121) */
double oracle_hh_weno7p(const double *v) { return weno7_core(v[-3], v[-2], v[-1], v[0], v[1], v[2], v[3]); }
/* hh_weno7M (OCFD_Schemes_1.f90:124-187) */
double oracle_hh_weno7m(const double *v) { return weno7_core(v[4], v[3], v[2], v[1], v[0], v[-1], v[-2]); }

/* minmod2 / minmod4 (OCFD_Schemes.f90:274-289) */
static double minmod2(double x1, double x2) {
    return 0.5 * (sgn1(x1) + sgn1(x2)) * dmin2(fabs(x1), fabs(x2));
}
static double minmod4(double x1, double x2, double x3, double x4) {
    double r = 0.5 * (sgn1(x1) + sgn1(x2));
    r = r * fabs(0.5 * (sgn1(x1) + sgn1(x3)));
    r = r * fabs(0.5 * (sgn1(x1) + sgn1(x4)));
    return r * dmin2(dmin2(dmin2(fabs(x1), fabs(x2)), fabs(x3)), fabs(x4));
}

/* OMP6 constants (OCFD_Schemes.f90:209-215); ep is a REAL(4) literal. */
#define OMP_EP ((double)1.e-10f)
#define OMP_KAPPA 4.0
static double omp_coef(int n) {
    const double m = 0.015, nn = 0.0;
    const double a = 0.5 * (m + nn), bb = 0.5 * (m - nn);
    switch (n) {
    case 1: return a;
    case 2: return 1.0 / 60.0 - bb - 6.0 * a;
    case 3: return -(2.0 / 15.0) + 6.0 * bb + 15.0 * a;
    case 4: return 37.0 / 60.0 - 15.0 * bb - 20.0 * a;
    case 5: return 37.0 / 60.0 + 20.0 * bb + 15.0 * a;
    case 6: return -(2.0 / 15.0) - 6.0 * a - 15.0 * bb;
    case 7: return 1.0 / 60.0 + a + 6.0 * bb;
    default: return -bb;
    }
}
/* MP-limited OMP6 on 8 values ordered from the upwind end:
 * P: (x[-3..4]) = v(-3..4);  M: (x[-3..4]) = v(4..-3)
 * (hh_OMP6P OCFD_Schemes_1.f90:195-226, hh_OMP6M :228-260).
 * The reference's M branch is not the exact mirror of P in one term: its
 * curvature-limited value lc uses v(1)-v(0), i.e. x0-x1 in upwind order
 * (OCFD_Schemes_1.f90:253, OCFD_Schemes.f90:263), where P uses x0-xm1
 * (OCFD_Schemes_1.f90:219).  `mflux` selects the M form. */
static double omp6_core(double xm3, double xm2, double xm1, double x0, double x1, double x2,
                        double x3, double x4, int mflux) {
    double mid = omp_coef(1) * x4 + omp_coef(2) * x3 + omp_coef(3) * x2 + omp_coef(4) * x1 +
                 omp_coef(5) * x0 + omp_coef(6) * xm1 + omp_coef(7) * xm2 + omp_coef(8) * xm3;
    double mp = x0 + minmod2((x1 - x0), OMP_KAPPA * (x0 - xm1));
    if ((mid - x0) * (mid - mp) >= OMP_EP) {
        double d1 = xm2 + x0 - 2.0 * xm1;
        double d2 = xm1 + x1 - 2.0 * x0;
        double d3 = x0 + x2 - 2.0 * x1;
        double ul = x0 + OMP_KAPPA * (x0 - xm1);
        double md = 0.5 * (x0 + x1) - 0.5 * minmod4(4.0 * d2 - d3, 4.0 * d3 - d2, d2, d3);
        double lc = x0 + 0.5 * (x0 - (mflux ? x1 : xm1)) + OMP_KAPPA * minmod4(4.0 * d1 - d2, 4.0 * d2 - d1, d2, d1) / 3.0;
        double fmin = dmax2(dmin3(x0, x1, md), dmin3(x0, ul, lc));
        double fmax = dmin2(dmax3(x0, x1, md), dmax3(x0, ul, lc));
        mid = mid + minmod2(fmax - mid, fmin - mid);
    }
    return mid;
}
double oracle_hh_omp6p(const double *v) { return omp6_core(v[-3], v[-2], v[-1], v[0], v[1], v[2], v[3], v[4], 0); }
double oracle_hh_omp6m(const double *v) { return omp6_core(v[4], v[3], v[2], v[1], v[0], v[-1], v[-2], v[-3], 1); }

/* OCFD_UD7L_P / hh_UD7L_P (OCFD_Schemes.f90:292-303, OCFD_Schemes_1.f90:328-337) */
static double ud7l_p(const double *c, const double *v) {
    return c[0] * v[-3] + c[1] * v[-2] + c[2] * v[-1] + c[3] * v[0] + c[4] * v[1] + c[5] * v[2] +
           c[6] * v[3] + c[7] * v[4];
}
/* OCFD_UD7L_M / hh_UD7L_M (OCFD_Schemes.f90:305-316, OCFD_Schemes_1.f90:339-348) */
static double ud7l_m(const double *c, const double *v) {
    return c[0] * v[4] + c[1] * v[3] + c[2] * v[2] + c[3] * v[1] + c[4] * v[0] + c[5] * v[-1] +
           c[6] * v[-2] + c[7] * v[-3];
}

/* set_hybrid_scheme (OCFD_SetHybrid.f90:2-15) */
static int set_hybrid_scheme(const oracle_block *b, double r0) {
    if (r0 <= b->c.hybrid_para[1]) return 1;
    if (r0 <= b->c.hybrid_para[2]) return 2;
    return 3;
}

/* Inner-scheme flux of one face, +/- (the select-case of OCFD2d_flux1/2,
 * OCFD_inviscous.f90:230-244 and 307-320, with OCFD_HybridP/M of
 * OCFD_Schemes.f90:319-358 for the hybrid case). */
static double inner_p(const oracle_block *b, const double *v, int scm) {
    switch (b->c.scheme_invis) {
    case OCFD_SCHEME_OMP6: return oracle_hh_omp6p(v);
    case OCFD_SCHEME_WENO7: return oracle_hh_weno7p(v);
    case OCFD_SCHEME_WENO5: return oracle_hh_weno5p(v);
    case OCFD_SCHEME_UD7L: return ud7l_p(b->ud7l, v);
    default:
        if (scm == 1) return ud7l_p(b->ud7l, v);
        if (scm == 2) return oracle_hh_weno7p(v);
        return oracle_hh_weno5p(v);
    }
}
static double inner_m(const oracle_block *b, const double *v, int scm) {
    switch (b->c.scheme_invis) {
    case OCFD_SCHEME_OMP6: return oracle_hh_omp6m(v);
    case OCFD_SCHEME_WENO7: return oracle_hh_weno7m(v);
    case OCFD_SCHEME_WENO5: return oracle_hh_weno5m(v);
    case OCFD_SCHEME_UD7L: return ud7l_m(b->ud7l, v);
    default:
        if (scm == 1) return ud7l_m(b->ud7l, v);
        if (scm == 2) return oracle_hh_weno7m(v);
        return oracle_hh_weno5m(v);
    }
}

/* OCFD2d_flux1 (OCFD_inviscous.f90:210-283).  f1d points at f1d(1-LAP);
 * hh is hh(0:n). */
static void flux1(const oracle_block *b, const double *f1d, double *hh, int n,
                  const int bidx[2], const int *scm, const int sb[2]) {
    int L = b->lap, Ka = b->ka1, Kb = b->kb1;
#define F1(i) (f1d + ((i) + L - 1))
    int ib = (bidx[0] == 0 || sb[0] == -1) ? 0 : -Ka + 1;
    int ie = (bidx[1] == 0 || sb[1] == -1) ? n : n - Kb;
    for (int i = ib; i <= ie; i++) hh[i] = inner_p(b, F1(i), scm[i]);
    if (bidx[0] == 1) {
        if (sb[0] == 0) {
            for (int i = 0; i <= -Ka; i++) {
                if (i == 0 || i == 1) hh[i] = (2.0 * *F1(1) + 5.0 * *F1(2) - *F1(3)) / 6.0;
                else if (i == 2) hh[i] = oracle_hh_weno5p_boundary(F1(i), DEL_LIFT);
                else hh[i] = oracle_hh_weno5p(F1(i));
            }
        } else if (sb[0] == 1) {
            for (int i = 0; i <= -Ka; i++) hh[i] = oracle_hh_weno5p(F1(i));
        }
    }
    if (bidx[1] == 1) {
        if (sb[1] == 0) {
            for (int i = n - Kb + 1; i <= n; i++) {
                if (i == n) hh[i] = (2.0 * *F1(n - 2) - 7.0 * *F1(n - 1) + 11.0 * *F1(n)) / 6.0;
                else if (i == n - 1) hh[i] = oracle_hh_weno5p_boundary(F1(i), DEL_RIGHT);
                else hh[i] = oracle_hh_weno5p(F1(i));
            }
        } else if (sb[1] == 1) {
            for (int i = n - Kb + 1; i <= n; i++) hh[i] = oracle_hh_weno5p(F1(i));
        }
    }
#undef F1
}

/* OCFD2d_flux2 (OCFD_inviscous.f90:287-361) */
static void flux2(const oracle_block *b, const double *f1d, double *hh, int n,
                  const int bidx[2], const int *scm, const int sb[2]) {
    int L = b->lap, Ka = b->ka2, Kb = b->kb2;
#define F1(i) (f1d + ((i) + L - 1))
    int ib = (bidx[0] == 0 || sb[0] == -1) ? 0 : -Ka + 1;
    int ie = (bidx[1] == 0 || sb[1] == -1) ? n : n - Kb;
    for (int i = ib; i <= ie; i++) hh[i] = inner_m(b, F1(i), scm[i]);
    if (bidx[0] == 1) {
        if (sb[0] == 0) {
            for (int i = 0; i <= -Ka; i++) {
                if (i == 0) hh[i] = (2.0 * *F1(3) - 7.0 * *F1(2) + 11.0 * *F1(1)) / 6.0;
                else if (i == 1) hh[i] = oracle_hh_weno5m_boundary(F1(i), DEL_LIFT);
                else hh[i] = oracle_hh_weno5m(F1(i));
            }
        } else if (sb[0] == 1) {
            for (int i = 0; i <= -Ka; i++) hh[i] = oracle_hh_weno5m(F1(i));
        }
    }
    if (bidx[1] == 1) {
        if (sb[1] == 0) {
            for (int i = n - Kb + 1; i <= n; i++) {
                if (i == n || i == n - 1) hh[i] = (2.0 * *F1(n) + 5.0 * *F1(n - 1) - *F1(n - 2)) / 6.0;
                else if (i == n - 2) hh[i] = oracle_hh_weno5m_boundary(F1(i), DEL_RIGHT);
                else hh[i] = oracle_hh_weno5m(F1(i));
            }
        } else if (sb[1] == 1) {
            for (int i = n - Kb + 1; i <= n; i++) hh[i] = oracle_hh_weno5m(F1(i));
        }
    }
#undef F1
}

/* ---------------------------------------------------------------------- */
/* pencil gather helpers                                                   */
/* ---------------------------------------------------------------------- */
/* axis 0: pencil along i at (j,k); 1: along j at (i,k); 2: along k at (i,j).
 * Pencil index p = s + LAP - 1 for s = 1-LAP .. n+LAP. */
static size_t pen_index(const oracle_block *b, int axis, int s, int a, int c) {
    if (axis == 0) return HI(b, s, a, c);
    if (axis == 1) return HI(b, a, s, c);
    return HI(b, a, c, s);
}
static int axis_n(const oracle_block *b, int axis) { return axis == 0 ? b->nx : (axis == 1 ? b->ny : b->nz); }

/* Gather c,d,u,v,w and the normalised metrics of the axis (inviscous.f90:83-92,
 * 123-132, 161-170) and split them. */
static void gather_split(oracle_block *b, int axis, int a, int c) {
    int n = axis_n(b, axis), L = b->lap;
    double *A[3];
    A[0] = b->met[M_AKX1 + 3 * axis]; A[1] = b->met[M_AKX1 + 3 * axis + 1]; A[2] = b->met[M_AKX1 + 3 * axis + 2];
    for (int s = 1 - L; s <= n + L; s++) {
        size_t q = pen_index(b, axis, s, a, c);
        int p = s + L - 1;
        b->pen[0][p] = b->cc[q]; b->pen[1][p] = b->d[q]; b->pen[2][p] = b->u[q];
        b->pen[3][p] = b->v[q]; b->pen[4][p] = b->w[q];
        b->pen[5][p] = A[0][q]; b->pen[6][p] = A[1][q]; b->pen[7][p] = A[2][q];
    }
    oracle_split(b->c.flux_splitting, n, L, b->c.gamma, b->pen[0], b->pen[1], b->pen[2], b->pen[3],
                 b->pen[4], b->pen[5], b->pen[6], b->pen[7], b->fp, b->fm);
}

/* Scm_Hb of each face (inviscous.f90:76-81): max of Rhybrid on both sides. */
static void hybrid_index(oracle_block *b, int axis, int a, int c) {
    int n = axis_n(b, axis);
    for (int s = 0; s <= n; s++) {
        double r0 = dmax2(b->rhyb[pen_index(b, axis, s, a, c)], b->rhyb[pen_index(b, axis, s + 1, a, c)]);
        b->scm[s] = set_hybrid_scheme(b, r0);
    }
}

static void sound_speed(oracle_block *b) {
    /* cc = sqrt(T)/Ma on the halo'd block (inviscous.f90:64-70) */
    for (size_t q = 0; q < b->nh; q++) b->cc[q] = sqrt(b->T[q]) / b->c.ma;
}

/* Mass force (OCFD_inviscous.f90:189-201) */
static void mass_force(oracle_block *b) {
    if (b->c.if_mass_force != 1) return;
#define g b->c.mass_force
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        size_t q = HI(b, i, j, k);
        b->du[FI(b, i, j, k, 2)] = b->du[FI(b, i, j, k, 2)] + b->d[q] * g[0];
        b->du[FI(b, i, j, k, 3)] = b->du[FI(b, i, j, k, 3)] + b->d[q] * g[1];
        b->du[FI(b, i, j, k, 4)] = b->du[FI(b, i, j, k, 4)] + b->d[q] * g[2];
        b->du[FI(b, i, j, k, 5)] = b->du[FI(b, i, j, k, 5)] +
            b->d[q] * (b->u[q] * g[0] + b->v[q] * g[1] + b->w[q] * g[2]);
    }
#undef g
}

/* ---------------------------------------------------------------------- */
/* du_inviscous (OCFD_inviscous.f90:53-203)                                */
/* ---------------------------------------------------------------------- */
int oracle_du_inviscous(oracle_block *b) {
    int L = b->lap;
    int hyb = b->c.scheme_invis == OCFD_SCHEME_HYBRID;
    double hinv[3] = {1.0 / b->hx, 1.0 / b->hy, 1.0 / b->hz};
    sound_speed(b);
    for (int axis = 0; axis < 3; axis++) {
        int n = axis_n(b, axis), np = n + 2 * L;
        int na = axis == 0 ? b->ny : b->nx;          /* first index of the pencil family */
        int nc = axis == 2 ? b->ny : b->nz;          /* second index */
        const int *sb = &b->c.scheme_boundary[2 * axis];
        for (int c = 1; c <= nc; c++) for (int a = 1; a <= na; a++) {
            if (hyb) hybrid_index(b, axis, a, c);
            gather_split(b, axis, a, c);
            for (int m = 1; m <= 5; m++) {
                flux1(b, b->fp + (m - 1) * np, b->hh1, n, b->bidx[axis], b->scm, sb);
                flux2(b, b->fm + (m - 1) * np, b->hh2, n, b->bidx[axis], b->scm, sb);
                for (int s = 1; s <= n; s++) {
                    int i = axis == 0 ? s : a;
                    int j = axis == 0 ? a : (axis == 1 ? s : c);
                    int k = axis == 2 ? s : c;
                    double dd = (b->hh1[s] - b->hh1[s - 1] + b->hh2[s] - b->hh2[s - 1]) * hinv[axis];
                    size_t q = FI(b, i, j, k, m);
                    if (axis == 0) b->du[q] = dd;
                    else if (axis == 1) b->du[q] = b->du[q] + dd;
                    else b->du[q] = -(b->du[q] + dd) * b->met[M_AJAC][HI(b, i, j, k)];
                }
            }
        }
    }
    mass_force(b);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* characteristic flux (OCFD_inviscous_character.f90:162-416)              */
/* ---------------------------------------------------------------------- */
/* Left (S) and right (S1) eigenvector matrices from the face-averaged state
 * st = {u1,v1,w1,c1,a1,a2,a3} (OCFD_inviscous_character.f90:216-248).
 * Row-major: S[r*5+c] = S(r+1,c+1). */
void oracle_char_matrices(double gamma, const double st[7], double S[25], double S1[25]) {
    double u1 = st[0], v1 = st[1], w1 = st[2], c1 = st[3], a1 = st[4], a2 = st[5], a3 = st[6];
    double ss = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
    double n1 = a1 / ss, n2 = a2 / ss, n3 = a3 / ss, l1, l2, l3;
    if (fabs(n3) <= fabs(n2)) {
        ss = sqrt(n1 * n1 + n2 * n2);
        l1 = -n2 / ss; l2 = n1 / ss; l3 = 0.0;
    } else {
        ss = sqrt(n1 * n1 + n3 * n3);
        l1 = -n3 / ss; l2 = 0.0; l3 = n1 / ss;
    }
    double m1 = n2 * l3 - n3 * l2, m2 = n3 * l1 - n1 * l3, m3 = n1 * l2 - n2 * l1;
    double un = u1 * n1 + v1 * n2 + w1 * n3;
    double ul = u1 * l1 + v1 * l2 + w1 * l3;
    double um = u1 * m1 + v1 * m2 + w1 * m3;
    double V2 = (u1 * u1 + v1 * v1 + w1 * w1) * 0.5;
    double KK = (gamma - 1.0) / (c1 * c1);
    double KK1 = KK * 0.5;
    double X1 = 1.0 / (2.0 * c1);
    double H = V2 + 1.0 / KK;
    double s[25] = {1.0 - KK * V2, KK * u1, KK * v1, KK * w1, -KK,
                    -ul, l1, l2, l3, 0.0,
                    -um, m1, m2, m3, 0.0,
                    KK1 * V2 + X1 * un, -(X1 * n1) - KK1 * u1, -(X1 * n2) - KK1 * v1, -(X1 * n3) - KK1 * w1, KK1,
                    KK1 * V2 - X1 * un, X1 * n1 - KK1 * u1, X1 * n2 - KK1 * v1, X1 * n3 - KK1 * w1, KK1};
    double s1[25] = {1.0, 0.0, 0.0, 1.0, 1.0,
                     u1, l1, m1, u1 - c1 * n1, u1 + c1 * n1,
                     v1, l2, m2, v1 - c1 * n2, v1 + c1 * n2,
                     w1, l3, m3, w1 - c1 * n3, w1 + c1 * n3,
                     V2, ul, um, H - c1 * un, H + c1 * un};
    memcpy(S, s, sizeof s);
    memcpy(S1, s1, sizeof s1);
}

/* flux_x1 (OCFD_inviscous_character.f90:313-364): ff points at ff(0). */
static double flux_x1(const oracle_block *b, const double *ff, int i, int nn, const int bidx[2],
                      int ihyb, const int sb[2]) {
    int Ka = b->ka1, Kb = b->kb1;
    if ((bidx[0] == 0 || i > -Ka || sb[0] == -1) && (bidx[1] == 0 || i <= nn - Kb || sb[1] == -1))
        return inner_p(b, ff, ihyb);
    if (i <= -Ka && bidx[0] == 1 && sb[0] != -1) {
        if (sb[0] == 0) {
            if (i == 1) return (2.0 * ff[0] + 5.0 * ff[1] - ff[2]) / 6.0;
            if (i == 2) return oracle_hh_weno5p_boundary(ff, DEL_LIFT);
            return oracle_hh_weno5p(ff);
        }
        if (sb[0] == 1) return oracle_hh_weno5p(ff);
    } else if (i > nn - Kb && bidx[1] == 1 && sb[1] != -1) {
        if (sb[1] == 0) {
            if (i == nn - 1) return oracle_hh_weno5p_boundary(ff, DEL_RIGHT);
            return oracle_hh_weno5p(ff);
        }
        if (sb[1] == 1) return oracle_hh_weno5p(ff);
    }
    return NAN; /* unreachable for valid Scheme_boundary values */
}

/* flux_x2 (OCFD_inviscous_character.f90:367-416) */
static double flux_x2(const oracle_block *b, const double *ff, int i, int nn, const int bidx[2],
                      int ihyb, const int sb[2]) {
    int Ka = b->ka2, Kb = b->kb2;
    if ((bidx[0] == 0 || i > -Ka || sb[0] == -1) && (bidx[1] == 0 || i <= nn - Kb || sb[1] == -1))
        return inner_m(b, ff, ihyb);
    if (i <= -Ka && bidx[0] == 1 && sb[0] != -1) {
        if (sb[0] == 0) {
            if (i == 1) return oracle_hh_weno5m_boundary(ff, DEL_LIFT);
            return oracle_hh_weno5m(ff);
        }
        if (sb[0] == 1) return oracle_hh_weno5m(ff);
    } else if (i > nn - Kb && bidx[1] == 1 && sb[1] != -1) {
        if (sb[1] == 0) {
            if (i == nn - 1) return (2.0 * ff[1] + 5.0 * ff[0] - ff[-1]) / 6.0;
            if (i == nn - 2) return oracle_hh_weno5m_boundary(ff, DEL_RIGHT);
            return oracle_hh_weno5m(ff);
        }
        if (sb[1] == 1) return oracle_hh_weno5m(ff);
    }
    return NAN;
}

/* flux_charteric (OCFD_inviscous_character.f90:162-310).  Pencil arrays in
 * b->pen / b->fp / b->fm (index p = s+LAP-1); output hhc[m*(n+1) + s]. */
static void flux_charteric(oracle_block *b, int nn, const int bidx[2], const int sb[2]) {
    int L = b->lap, np = nn + 2 * L;
    int Ka1 = b->ka1, Kb1 = b->kb1, Ka2 = b->ka2, Kb2 = b->kb2;
    double fpc[5][8], fmc[5][8]; /* index mm - Ka (Ka >= -3, Kb <= 4) */
    double S[25], S1[25], h1[5], h2[5], h0[5];
    memset(fpc, 0, sizeof fpc); memset(fmc, 0, sizeof fmc);
    int ia = (bidx[0] == 0 || sb[0] != 0) ? 0 : 1;
    int ib = (bidx[1] == 0 || sb[1] != 0) ? nn : nn - 1;
#define P(s) ((s) + L - 1)
    for (int i = ia; i <= ib; i++) {
        int sc = b->scm[i];
        if (sc == 1) {
            for (int mm = Ka1; mm <= Kb1; mm++) for (int m = 0; m < 5; m++) fpc[m][mm - Ka1] = b->fp[m * np + P(i + mm)];
            for (int mm = Ka2; mm <= Kb2; mm++) for (int m = 0; m < 5; m++) fmc[m][mm - Ka2] = b->fm[m * np + P(i + mm)];
        } else {
            int a1h, b1h, a2h, b2h;
            if (sc == 0) { a1h = Ka1; b1h = Kb1; a2h = Ka2; b2h = Kb2; }
            else if (sc == 2) { a1h = b->kh[1][0]; b1h = b->kh[1][1]; a2h = b->kh[1][2]; b2h = b->kh[1][3]; }
            else { a1h = b->kh[2][0]; b1h = b->kh[2][1]; a2h = b->kh[2][2]; b2h = b->kh[2][3]; }
            double st[7];
            st[0] = (b->pen[2][P(i)] + b->pen[2][P(i + 1)]) * 0.5;
            st[1] = (b->pen[3][P(i)] + b->pen[3][P(i + 1)]) * 0.5;
            st[2] = (b->pen[4][P(i)] + b->pen[4][P(i + 1)]) * 0.5;
            st[3] = (b->pen[0][P(i)] + b->pen[0][P(i + 1)]) * 0.5;
            st[4] = (b->pen[5][P(i)] + b->pen[5][P(i + 1)]) * 0.5;
            st[5] = (b->pen[6][P(i)] + b->pen[6][P(i + 1)]) * 0.5;
            st[6] = (b->pen[7][P(i)] + b->pen[7][P(i + 1)]) * 0.5;
            oracle_char_matrices(b->c.gamma, st, S, S1);
            for (int mm = a1h; mm <= b1h; mm++) {
                int q = P(i + mm);
                for (int r = 0; r < 5; r++)
                    fpc[r][mm - Ka1] = S[r * 5 + 0] * b->fp[0 * np + q] + S[r * 5 + 1] * b->fp[1 * np + q] +
                                       S[r * 5 + 2] * b->fp[2 * np + q] + S[r * 5 + 3] * b->fp[3 * np + q] +
                                       S[r * 5 + 4] * b->fp[4 * np + q];
            }
            for (int mm = a2h; mm <= b2h; mm++) {
                int q = P(i + mm);
                for (int r = 0; r < 5; r++)
                    fmc[r][mm - Ka2] = S[r * 5 + 0] * b->fm[0 * np + q] + S[r * 5 + 1] * b->fm[1 * np + q] +
                                       S[r * 5 + 2] * b->fm[2 * np + q] + S[r * 5 + 3] * b->fm[3 * np + q] +
                                       S[r * 5 + 4] * b->fm[4 * np + q];
            }
        }
        for (int m = 0; m < 5; m++) {
            h1[m] = flux_x1(b, &fpc[m][-Ka1], i, nn, bidx, sc, sb);
            h2[m] = flux_x2(b, &fmc[m][-Ka2], i, nn, bidx, sc, sb);
        }
        for (int m = 0; m < 5; m++) h0[m] = h1[m] + h2[m];
        if (sc == 1) {
            for (int m = 0; m < 5; m++) b->hhc[m * (nn + 1) + i] = h0[m];
        } else {
            for (int r = 0; r < 5; r++)
                b->hhc[r * (nn + 1) + i] = S1[r * 5 + 0] * h0[0] + S1[r * 5 + 1] * h0[1] + S1[r * 5 + 2] * h0[2] +
                                           S1[r * 5 + 3] * h0[3] + S1[r * 5 + 4] * h0[4];
        }
    }
    /* non-reflecting closure (OCFD_inviscous_character.f90:295-309) */
    if (bidx[0] == 1 && sb[0] == 0) {
        for (int m = 0; m < 5; m++) {
            const double *fp = b->fp + m * np, *fm = b->fm + m * np;
            b->hhc[m * (nn + 1) + 0] = (2.0 * fp[P(1)] + 5.0 * fp[P(2)] - fp[P(3)] + 2.0 * fm[P(3)] -
                                        7.0 * fm[P(2)] + 11.0 * fm[P(1)]) / 6.0;
        }
    }
    if (bidx[1] == 1 && sb[1] == 0) {
        for (int m = 0; m < 5; m++) {
            const double *fp = b->fp + m * np, *fm = b->fm + m * np;
            b->hhc[m * (nn + 1) + nn] = (2.0 * fp[P(nn - 2)] - 7.0 * fp[P(nn - 1)] + 11.0 * fp[P(nn)] +
                                         2.0 * fm[P(nn)] + 5.0 * fm[P(nn - 1)] - fm[P(nn - 2)]) / 6.0;
        }
    }
#undef P
}

/* du_inviscous_Character (OCFD_inviscous_character.f90:3-159) */
int oracle_du_inviscous_character(oracle_block *b) {
    int hyb = b->c.scheme_invis == OCFD_SCHEME_HYBRID;
    double hinv[3] = {1.0 / b->hx, 1.0 / b->hy, 1.0 / b->hz};
    sound_speed(b);
    for (int axis = 0; axis < 3; axis++) {
        int n = axis_n(b, axis);
        int na = axis == 0 ? b->ny : b->nx;
        int nc = axis == 2 ? b->ny : b->nz;
        const int *sb = &b->c.scheme_boundary[2 * axis];
        for (int c = 1; c <= nc; c++) for (int a = 1; a <= na; a++) {
            if (hyb) hybrid_index(b, axis, a, c);
            else for (int s = 0; s <= n; s++) b->scm[s] = 0;
            gather_split(b, axis, a, c);
            flux_charteric(b, n, b->bidx[axis], sb);
            for (int m = 1; m <= 5; m++) {
                const double *h = b->hhc + (m - 1) * (n + 1);
                for (int s = 1; s <= n; s++) {
                    int i = axis == 0 ? s : a;
                    int j = axis == 0 ? a : (axis == 1 ? s : c);
                    int k = axis == 2 ? s : c;
                    double dd = (h[s] - h[s - 1]) * hinv[axis];
                    size_t q = FI(b, i, j, k, m);
                    if (axis == 0) b->du[q] = dd;
                    else if (axis == 1) b->du[q] = b->du[q] + dd;
                    else b->du[q] = -(b->du[q] + dd) * b->met[M_AJAC][HI(b, i, j, k)];
                }
            }
        }
    }
    mass_force(b);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* OCFD_dx0 / dy0 / dz0 (OCFD_dxyz_viscous.f90:3-239)                       */
/* ---------------------------------------------------------------------- */
int oracle_dxyz0(oracle_block *b, int axis, const double *f, double *fx) {
    double h = axis == 0 ? b->hx : (axis == 1 ? b->hy : b->hz);
    int n = axis_n(b, axis);
    double a1 = 1.0 / (60.0 * h), a2 = -(3.0 / (20.0 * h)), a3 = 3.0 / (4.0 * h);
    double c1 = 0.8 / h, c2 = -(0.2 / h), c3 = 3.80952380952380952e-2 / h, c4 = -(3.571428571428571428e-3 / h);
    int cd8 = b->c.scheme_vis == OCFD_SCHEME_CD8;
    int di = axis == 0, dj = axis == 1, dk = axis == 2;
#define FF(s) f[HI(b, i + di * (s), j + dj * (s), k + dk * (s))]
#define FA(t) f[HI(b, di ? (t) : i, dj ? (t) : j, dk ? (t) : k)]
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        double r;
        if (!cd8) r = a1 * (FF(3) - FF(-3)) + a2 * (FF(2) - FF(-2)) + a3 * (FF(1) - FF(-1));
        else r = c1 * (FF(1) - FF(-1)) + c2 * (FF(2) - FF(-2)) + c3 * (FF(3) - FF(-3)) + c4 * (FF(4) - FF(-4));
        int s = di ? i : (dj ? j : k);
        if (b->c.iperiodic[axis] == 0) {
            if (s == 1) r = (-(3.0 * FA(1)) + 4.0 * FA(2) - FA(3)) / (2.0 * h);
            else if (s == 2) r = (FA(3) - FA(1)) / (2.0 * h);
            else if (s == 3) r = (8.0 * (FA(4) - FA(2)) - (FA(5) - FA(1))) / (12.0 * h);
            else if (s == 4 && cd8) r = a1 * (FA(7) - FA(1)) + a2 * (FA(6) - FA(2)) + a3 * (FA(5) - FA(3));
            else if (s == n) r = (FA(n - 2) - 4.0 * FA(n - 1) + 3.0 * FA(n)) / (2.0 * h);
            else if (s == n - 1) r = (FA(n) - FA(n - 2)) / (2.0 * h);
            else if (s == n - 2) r = (8.0 * (FA(n - 1) - FA(n - 3)) - (FA(n) - FA(n - 4))) / (12.0 * h);
            else if (s == n - 3 && cd8) r = a1 * (FA(n) - FA(n - 6)) + a2 * (FA(n - 1) - FA(n - 5)) + a3 * (FA(n - 2) - FA(n - 4));
        }
        fx[II(b, i, j, k)] = r;
    }
#undef FF
#undef FA
    return 0;
}

/* ---------------------------------------------------------------------- */
/* du_viscous + comput_Amu (OCFD_viscous.f90:29-149)                       */
/* ---------------------------------------------------------------------- */
static void comput_amu(oracle_block *b) {
    double Tsb = 110.4 / b->c.ref_amu_t0, TmpR = 1.0 / b->c.re;
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        double t = b->T[HI(b, i, j, k)];
        b->amu[II(b, i, j, k)] = TmpR * (1.0 + Tsb) * sqrt(t * t * t) / (Tsb + t);
    }
}

int oracle_du_viscous(oracle_block *b) {
    const double tmp2_3 = 2.0 / 3.0;
    comput_amu(b);
    double *g[12]; memcpy(g, b->grad, sizeof g);
    double *src[4] = {b->u, b->v, b->w, b->T};
    /* g index: field*3 + axis (uk,ui,us, vk,vi,vs, wk,wi,ws, Tk,Ti,Ts) */
    for (int fld = 0; fld < 4; fld++) for (int axis = 0; axis < 3; axis++) oracle_dxyz0(b, axis, src[fld], g[fld * 3 + axis]);
    double **M = b->met;
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        size_t q = II(b, i, j, k), h = HI(b, i, j, k);
        double ux = g[0][q] * M[M_AKX][h] + g[1][q] * M[M_AIX][h] + g[2][q] * M[M_ASX][h];
        double vx = g[3][q] * M[M_AKX][h] + g[4][q] * M[M_AIX][h] + g[5][q] * M[M_ASX][h];
        double wx = g[6][q] * M[M_AKX][h] + g[7][q] * M[M_AIX][h] + g[8][q] * M[M_ASX][h];
        double Tx = g[9][q] * M[M_AKX][h] + g[10][q] * M[M_AIX][h] + g[11][q] * M[M_ASX][h];
        double uy = g[0][q] * M[M_AKY][h] + g[1][q] * M[M_AIY][h] + g[2][q] * M[M_ASY][h];
        double vy = g[3][q] * M[M_AKY][h] + g[4][q] * M[M_AIY][h] + g[5][q] * M[M_ASY][h];
        double wy = g[6][q] * M[M_AKY][h] + g[7][q] * M[M_AIY][h] + g[8][q] * M[M_ASY][h];
        double Ty = g[9][q] * M[M_AKY][h] + g[10][q] * M[M_AIY][h] + g[11][q] * M[M_ASY][h];
        double uz = g[0][q] * M[M_AKZ][h] + g[1][q] * M[M_AIZ][h] + g[2][q] * M[M_ASZ][h];
        double vz = g[3][q] * M[M_AKZ][h] + g[4][q] * M[M_AIZ][h] + g[5][q] * M[M_ASZ][h];
        double wz = g[6][q] * M[M_AKZ][h] + g[7][q] * M[M_AIZ][h] + g[8][q] * M[M_ASZ][h];
        double Tz = g[9][q] * M[M_AKZ][h] + g[10][q] * M[M_AIZ][h] + g[11][q] * M[M_ASZ][h];
        double div = ux + vy + wz;
        double Amu1 = b->amu[q];
        double Amuk = b->cp * b->amu[q] / b->c.pr;
        double s11 = (2.0 * ux - tmp2_3 * div) * Amu1;
        double s22 = (2.0 * vy - tmp2_3 * div) * Amu1;
        double s33 = (2.0 * wz - tmp2_3 * div) * Amu1;
        double s12 = (uy + vx) * Amu1, s13 = (uz + wx) * Amu1, s23 = (vz + wy) * Amu1;
        double uu = b->u[h], vv = b->v[h], ww = b->w[h];
        double E1 = uu * s11 + vv * s12 + ww * s13 + Amuk * Tx;
        double E2 = uu * s12 + vv * s22 + ww * s23 + Amuk * Ty;
        double E3 = uu * s13 + vv * s23 + ww * s33 + Amuk * Tz;
        for (int a = 0; a < 3; a++) {
            double n1 = M[M_AKX1 + 3 * a][h], n2 = M[M_AKX1 + 3 * a + 1][h], n3 = M[M_AKX1 + 3 * a + 2][h];
            b->ev[a][0][h] = n1 * s11 + n2 * s12 + n3 * s13;
            b->ev[a][1][h] = n1 * s12 + n2 * s22 + n3 * s23;
            b->ev[a][2][h] = n1 * s13 + n2 * s23 + n3 * s33;
            b->ev[a][3][h] = n1 * E1 + n2 * E2 + n3 * E3;
        }
    }
    for (int m = 0; m < 4; m++) {
        exchange_x(b, b->ev[0][m]);
        exchange_y(b, b->ev[1][m]);
        exchange_z(b, b->ev[2][m]);
        oracle_dxyz0(b, 0, b->ev[0][m], g[0]);
        oracle_dxyz0(b, 1, b->ev[1][m], g[1]);
        oracle_dxyz0(b, 2, b->ev[2][m], g[2]);
        for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
            size_t q = II(b, i, j, k), fq = FI(b, i, j, k, m + 2);
            b->du[fq] = b->du[fq] + (g[0][q] + g[1][q] + g[2][q]) * M[M_AJAC][HI(b, i, j, k)];
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* time advance, primitives, BC (OCFD_Time_Adv.f90, OCFD_BC.f90)           */
/* ---------------------------------------------------------------------- */
int oracle_save_fn(oracle_block *b) { memcpy(b->fn, b->f, 5 * b->ni * sizeof(double)); return 0; }

/* OCFD_time_adv_RK3 (OCFD_Time_Adv.f90:4-36) */
int oracle_time_adv_rk3(oracle_block *b, int32_t krk) {
    const double ra[3] = {1.0, 3.0 / 4.0, 1.0 / 3.0}, rb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
    double dt = b->c.dt;
    if (krk < 1 || krk > 3) return -1;
    size_t n = 5 * b->ni;
    if (krk == 1) for (size_t q = 0; q < n; q++) b->f[q] = b->fn[q] + dt * b->du[q];
    else for (size_t q = 0; q < n; q++) b->f[q] = ra[krk - 1] * b->fn[q] + rb[krk - 1] * (b->f[q] + dt * b->du[q]);
    return 0;
}

/* comput_duvwT + handle_NegativeT (OCFD_Time_Adv.f90:40-63,
 * OCFD_handle_NegativeT.f90:1-17).  Returns the number of repaired cells
 * (the reference stops the run beyond 10). */
int oracle_comput_duvwt(oracle_block *b) {
    int nneg = 0;
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        size_t h = HI(b, i, j, k);
        double f1 = b->f[FI(b, i, j, k, 1)], f2 = b->f[FI(b, i, j, k, 2)], f3 = b->f[FI(b, i, j, k, 3)];
        double f4 = b->f[FI(b, i, j, k, 4)], f5 = b->f[FI(b, i, j, k, 5)];
        b->d[h] = f1;
        b->u[h] = f2 / f1;
        b->v[h] = f3 / f1;
        b->w[h] = f4 / f1;
        b->T[h] = (f5 - (f2 * b->u[h] + f3 * b->v[h] + f4 * b->w[h]) * 0.5) / (f1 * b->cv);
        if (b->T[h] <= 0) {
            b->T[h] = (b->T[HI(b, i + 1, j, k)] + b->T[HI(b, i - 1, j, k)] + b->T[HI(b, i, j + 1, k)] +
                       b->T[HI(b, i, j - 1, k)] + b->T[HI(b, i, j, k + 1)] + b->T[HI(b, i, j, k - 1)]) / 6.0;
            nneg++;
        }
    }
    return nneg;
}

/* ocfd_bc_blunt2d (OCFD_BC_Def.f90:4-33) */
static void bc_blunt2d(oracle_block *b) {
    for (int k = 1; k <= b->nz; k++) for (int i = 1; i <= b->nx; i++) {
        size_t h1 = HI(b, i, 1, k), h2 = HI(b, i, 2, k);
        b->d[h1] = b->d[h2]; b->u[h1] = 0.0; b->v[h1] = 0.0; b->w[h1] = 0.0; b->T[h1] = b->T[h2];
    }
    for (int k = 1; k <= b->nz; k++) for (int i = 1; i <= b->nx; i++) {
        size_t h = HI(b, i, b->ny, k);
        b->d[h] = 1.0; b->u[h] = 1.0; b->v[h] = 0.0; b->w[h] = 0.0; b->T[h] = 1.0;
    }
}

/* The case hook ocfd_bc_User (OCFD_BC.f90:18-19), as supplied to the
 * compiled reference by oracle/ref_user_hooks.f90: isothermal (BC_para(1) > 0)
 * or adiabatic no-slip wall at j=1 and free stream at j=ny. */
static void bc_user(oracle_block *b) {
    double Tw = b->c.bc_para[0];
    for (int k = 1; k <= b->nz; k++) for (int i = 1; i <= b->nx; i++) {
        size_t h1 = HI(b, i, 1, k), h2 = HI(b, i, 2, k), h3 = HI(b, i, 3, k);
        b->u[h1] = 0.0; b->v[h1] = 0.0; b->w[h1] = 0.0;
        if (Tw > 0.0) {
            b->d[h1] = (4.0 * b->d[h2] * b->T[h2] - b->d[h3] * b->T[h3]) / (3.0 * Tw);
            b->T[h1] = Tw;
        } else {
            b->d[h1] = (4.0 * b->d[h2] - b->d[h3]) / 3.0;
            b->T[h1] = (4.0 * b->T[h2] - b->T[h3]) / 3.0;
        }
    }
    for (int k = 1; k <= b->nz; k++) for (int i = 1; i <= b->nx; i++) {
        size_t h = HI(b, i, b->ny, k);
        b->d[h] = 1.0; b->u[h] = 1.0; b->v[h] = 0.0; b->w[h] = 0.0; b->T[h] = 1.0;
    }
}

static void refresh_f(oracle_block *b, int i, int j, int k) {
    size_t h = HI(b, i, j, k);
    double dd = b->d[h], uu = b->u[h], vv = b->v[h], ww = b->w[h];
    b->f[FI(b, i, j, k, 1)] = dd;
    b->f[FI(b, i, j, k, 2)] = dd * uu;
    b->f[FI(b, i, j, k, 3)] = dd * vv;
    b->f[FI(b, i, j, k, 4)] = dd * ww;
    b->f[FI(b, i, j, k, 5)] = dd * (b->cv * b->T[h] + (uu * uu + vv * vv + ww * ww) * 0.5);
}

/* Flow_Ghost_Extent_2nd (OCFD_GhostCell.f90:164-286): odd reflection about
 * the boundary cell, density and temperature limited to [0.8, 1.2] x wall. */
static void ghost_reflect(oracle_block *b, size_t g, size_t w, size_t mir) {
    const double s1 = 0.8, s2 = 1.2;
    b->u[g] = 2.0 * b->u[w] - b->u[mir];
    b->v[g] = 2.0 * b->v[w] - b->v[mir];
    b->w[g] = 2.0 * b->w[w] - b->w[mir];
    b->d[g] = 2.0 * b->d[w] - b->d[mir];
    b->T[g] = 2.0 * b->T[w] - b->T[mir];
    if (b->d[g] < s1 * b->d[w]) b->d[g] = s1 * b->d[w];
    if (b->d[g] > s2 * b->d[w]) b->d[g] = s2 * b->d[w];
    if (b->T[g] < s1 * b->T[w]) b->T[g] = s1 * b->T[w];
    if (b->T[g] > s2 * b->T[w]) b->T[g] = s2 * b->T[w];
}
static void flow_ghost(oracle_block *b) {
    int L = b->lap, nx = b->nx, ny = b->ny, nz = b->nz;
    const int *gc = b->c.ghost_cell;
#define GON(n) (gc[n] == 1 || gc[n] == 2)
    if (GON(0)) for (int k = 1; k <= nz; k++) for (int j = 1; j <= ny; j++) for (int i = 1 - L; i <= 0; i++)
        ghost_reflect(b, HI(b, i, j, k), HI(b, 1, j, k), HI(b, 2 - i, j, k));
    if (GON(1)) for (int k = 1; k <= nz; k++) for (int j = 1; j <= ny; j++) for (int i = nx + 1; i <= nx + L; i++)
        ghost_reflect(b, HI(b, i, j, k), HI(b, nx, j, k), HI(b, 2 * nx - i, j, k));
    if (GON(2)) for (int k = 1; k <= nz; k++) for (int j = 1 - L; j <= 0; j++) for (int i = 1; i <= nx; i++)
        ghost_reflect(b, HI(b, i, j, k), HI(b, i, 1, k), HI(b, i, 2 - j, k));
    if (GON(3)) for (int k = 1; k <= nz; k++) for (int j = ny + 1; j <= ny + L; j++) for (int i = 1; i <= nx; i++)
        ghost_reflect(b, HI(b, i, j, k), HI(b, i, ny, k), HI(b, i, 2 * ny - j, k));
    if (GON(4)) for (int k = 1 - L; k <= 0; k++) for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++)
        ghost_reflect(b, HI(b, i, j, k), HI(b, i, j, 1), HI(b, i, j, 2 - k));
    if (GON(5)) for (int k = nz + 1; k <= nz + L; k++) for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++)
        ghost_reflect(b, HI(b, i, j, k), HI(b, i, j, nz), HI(b, i, j, 2 * nz - k));
#undef GON
}

/* OCFD_bc (OCFD_BC.f90:2-99) on one rank: BC hook, then f recomputed from
 * the primitives on all six block faces, then the flow ghost cells. */
int oracle_bc(oracle_block *b) {
    if (b->c.ibc == OCFD_BC_BLUNT2D) bc_blunt2d(b);
    else if (b->c.ibc == OCFD_BC_USER_DEF) bc_user(b);
    else if (b->c.ibc != OCFD_BC_NONE) return -1;
    int nx = b->nx, ny = b->ny, nz = b->nz;
    for (int k = 1; k <= nz; k++) for (int j = 1; j <= ny; j++) refresh_f(b, 1, j, k);
    for (int k = 1; k <= nz; k++) for (int j = 1; j <= ny; j++) refresh_f(b, nx, j, k);
    for (int k = 1; k <= nz; k++) for (int i = 1; i <= nx; i++) refresh_f(b, i, 1, k);
    for (int k = 1; k <= nz; k++) for (int i = 1; i <= nx; i++) refresh_f(b, i, ny, k);
    for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++) refresh_f(b, i, j, 1);
    for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++) refresh_f(b, i, j, nz);
    flow_ghost(b);
    return 0;
}

int oracle_exchange_primitives(oracle_block *b) {
    exchange_xyz(b, b->d); exchange_xyz(b, b->u); exchange_xyz(b, b->v);
    exchange_xyz(b, b->w); exchange_xyz(b, b->T);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* comput_Rhybrid (OCFD_SetHybrid.f90:17-152)                              */
/* ---------------------------------------------------------------------- */
int oracle_comput_rhybrid(oracle_block *b, int64_t counts[4]) {
    double p00 = 1.0 / (b->c.gamma * b->c.ma * b->c.ma);
    int sensor = (int)lround(b->c.hybrid_para[0]);
    double *p = b->pf;
    for (size_t q = 0; q < b->nh; q++) p[q] = p00 * b->d[q] * b->T[q];
    if (sensor == 1) {
        for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
            double pc = p[HI(b, i, j, k)];
            double xm = p[HI(b, i - 1, j, k)], xp = p[HI(b, i + 1, j, k)];
            double ym = p[HI(b, i, j - 1, k)], yp = p[HI(b, i, j + 1, k)];
            double zm = p[HI(b, i, j, k - 1)], zp = p[HI(b, i, j, k + 1)];
            b->rhyb[HI(b, i, j, k)] = fabs((xm - 2.0 * pc + xp) / (xm + 2.0 * pc + xp)) +
                                      fabs((ym - 2.0 * pc + yp) / (ym + 2.0 * pc + yp)) +
                                      fabs((zm - 2.0 * pc + zp) / (zm + 2.0 * pc + zp));
        }
    } else {
        double *pk = b->grad[0], *pi = b->grad[1], *ps = b->grad[2];
        oracle_dxyz0(b, 0, p, pk); oracle_dxyz0(b, 1, p, pi); oracle_dxyz0(b, 2, p, ps);
        double dp0 = 0.0;
        double **M = b->met;
        for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
            size_t q = II(b, i, j, k), h = HI(b, i, j, k);
            double px = pk[q] * M[M_AKX][h] + pi[q] * M[M_AIX][h] + ps[q] * M[M_ASX][h];
            double py = pk[q] * M[M_AKY][h] + pi[q] * M[M_AIY][h] + ps[q] * M[M_ASY][h];
            double pz = pk[q] * M[M_AKZ][h] + pi[q] * M[M_AIZ][h] + ps[q] * M[M_ASZ][h];
            b->rhyb[h] = sqrt(px * px + py * py + pz * pz);
            dp0 = dp0 + b->rhyb[h];
        }
        double dp_av = dp0 / (1.0 * b->nx * b->ny * b->nz);
        double dp_av1 = 1.0 / (dp_av + 1.0e-30);
        for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++)
            b->rhyb[HI(b, i, j, k)] = b->rhyb[HI(b, i, j, k)] * dp_av1;
    }
    /* forced-WENO patches (OCFD_SetHybrid.f90:87-110), with the j range read
     * from Hybrid_para(7:8) as the comments there intend */
    int mpatch = (int)lround(b->c.hybrid_para[3]);
#define hp b->c.hybrid_para
    for (int m = 1; m <= mpatch; m++) {
        int ib = (int)lround(hp[4 + (m - 1) * 6]), ie = (int)lround(hp[5 + (m - 1) * 6]);
        int jb = (int)lround(hp[6 + (m - 1) * 6]), je = (int)lround(hp[7 + (m - 1) * 6]);
        int kb = (int)lround(hp[8 + (m - 1) * 6]), ke = (int)lround(hp[9 + (m - 1) * 6]);
        if (ib < 1) ib = 1;
        if (ie > b->nx) ie = b->nx;
        if (jb < 1) jb = 1;
        if (je > b->ny) je = b->ny;
        if (kb < 1) kb = 1;
        if (ke > b->nz) ke = b->nz;
        for (int k = kb; k <= ke; k++) for (int j = jb; j <= je; j++) for (int i = ib; i <= ie; i++)
            b->rhyb[HI(b, i, j, k)] = b->rhyb[HI(b, i, j, k)] + 100.0;
    }
#undef hp
    exchange_xyz(b, b->rhyb);
    if (counts) {
        int64_t ns[3] = {0, 0, 0};
        for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 0; i <= b->nx; i++) {
            double r0 = dmax2(b->rhyb[HI(b, i, j, k)], b->rhyb[HI(b, i + 1, j, k)]);
            ns[set_hybrid_scheme(b, r0) - 1]++;
        }
        counts[0] = ns[0]; counts[1] = ns[1]; counts[2] = ns[2];
        counts[3] = (int64_t)(b->nx + 1) * b->ny * b->nz;
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* filtering (OCFD_filtering.f90:4-363)                                    */
/* ---------------------------------------------------------------------- */
static const double FO_D0 = 0.243527493120, FO_D1 = -0.204788880640, FO_D2 = 0.120007591680,
                    FO_D3 = -0.045211119360, FO_D4 = 0.008228661760;
static const double SH_C1 = -0.210383, SH_C2 = 0.039617;

/* filter_x3d / filter_y3d / filter_z3d (OCFD_filtering.f90:73-183) */
static void filter_fo9p(oracle_block *b, int axis, double s0, int rng[6]) {
    int di = axis == 0, dj = axis == 1, dk = axis == 2;
    int lo[3] = {rng[0], rng[2], rng[4]}, hi[3] = {rng[1], rng[3], rng[5]};
    int n = axis_n(b, axis);
    if (b->c.iperiodic[axis] != 1) {
        if (lo[axis] < 6) lo[axis] = 6;
        if (hi[axis] > n - 5) hi[axis] = n - 5;
    }
    double *f0 = b->f0;
    for (int m = 1; m <= 5; m++) {
        for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++)
            f0[HI(b, i, j, k)] = b->f[FI(b, i, j, k, m)];
        if (axis == 0) exchange_x(b, f0); else if (axis == 1) exchange_y(b, f0); else exchange_z(b, f0);
#define G(s) f0[HI(b, i + di * (s), j + dj * (s), k + dk * (s))]
        for (int k = lo[2]; k <= hi[2]; k++) for (int j = lo[1]; j <= hi[1]; j++) for (int i = lo[0]; i <= hi[0]; i++)
            b->f[FI(b, i, j, k, m)] = G(0) - s0 * (FO_D0 * G(0) + FO_D1 * (G(-1) + G(1)) + FO_D2 * (G(-2) + G(2)) +
                                                  FO_D3 * (G(-3) + G(3)) + FO_D4 * (G(-4) + G(4)));
#undef G
    }
}

/* filter_x3d_shock / y / z (OCFD_filtering.f90:188-363) */
static void filter_shock(oracle_block *b, int axis, double s0, double rth, int rng[6]) {
    int di = axis == 0, dj = axis == 1, dk = axis == 2;
    int lo[3] = {rng[0], rng[2], rng[4]}, hi[3] = {rng[1], rng[3], rng[5]};
    int n = axis_n(b, axis);
    if (b->c.iperiodic[axis] != 1) {
        if (lo[axis] < 4) lo[axis] = 4;
        if (hi[axis] > n - 3) hi[axis] = n - 3;
    }
    double *f0 = b->f0, *p = b->pf;
    double gm1 = b->c.gamma - 1.0;
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        double f1 = b->f[FI(b, i, j, k, 1)], f2 = b->f[FI(b, i, j, k, 2)], f3 = b->f[FI(b, i, j, k, 3)];
        double f4 = b->f[FI(b, i, j, k, 4)], f5 = b->f[FI(b, i, j, k, 5)];
        p[HI(b, i, j, k)] = (f5 - 0.5 * (f2 * f2 + f3 * f3 + f4 * f4) / f1) * gm1;
    }
    if (axis == 0) exchange_x(b, p); else if (axis == 1) exchange_y(b, p); else exchange_z(b, p);
    double *sc0 = (double *)calloc(n + 2, sizeof(double));
    for (int m = 1; m <= 5; m++) {
        for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++)
            f0[HI(b, i, j, k)] = b->f[FI(b, i, j, k, m)];
        if (axis == 0) exchange_x(b, f0); else if (axis == 1) exchange_y(b, f0); else exchange_z(b, f0);
        /* lines of the two transverse indices */
        int ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
        for (int c2 = lo[tb]; c2 <= hi[tb]; c2++) for (int c1 = lo[ta]; c1 <= hi[ta]; c1++) {
            int base[3];
            base[ta] = c1; base[tb] = c2; base[axis] = 0;
#define PP(s) p[HI(b, base[0] + di * (s), base[1] + dj * (s), base[2] + dk * (s))]
#define GG(s) f0[HI(b, base[0] + di * (s), base[1] + dj * (s), base[2] + dk * (s))]
            for (int s = lo[axis] - 1; s <= hi[axis] + 1; s++) {
                double dp0 = 0.25 * (-PP(s + 1) + 2.0 * PP(s) - PP(s - 1));
                double dp1 = 0.25 * (-PP(s + 2) + 2.0 * PP(s + 1) - PP(s));
                double dp2 = 0.25 * (-PP(s) + 2.0 * PP(s - 1) - PP(s - 2));
                double x = dp0 - dp1, y = dp0 - dp2;
                double ri = 0.5 * (x * x + y * y) / (PP(s) * PP(s)) + 1.0e-16;
                sc0[s] = 0.5 * (1.0 - rth / ri + fabs(1.0 - rth / ri));
            }
            for (int s = lo[axis]; s <= hi[axis]; s++) {
                double sc1 = 0.5 * (sc0[s] + sc0[s + 1]);
                double sc2 = 0.5 * (sc0[s] + sc0[s - 1]);
                int idx[3] = {base[0], base[1], base[2]};
                idx[axis] = s;
                b->f[FI(b, idx[0], idx[1], idx[2], m)] =
                    GG(s) - s0 * (sc1 * (SH_C1 * (GG(s + 1) - GG(s)) + SH_C2 * (GG(s + 2) - GG(s - 1))) -
                                  sc2 * (SH_C1 * (GG(s) - GG(s - 1)) + SH_C2 * (GG(s + 1) - GG(s - 2))));
            }
#undef PP
#undef GG
        }
    }
    free(sc0);
}

/* filtering (OCFD_filtering.f90:4-70) at the block's Istep */
int oracle_filtering(oracle_block *b) {
    int nf = b->c.nfiltering;
    int any = 0;
    for (int m = 0; m < nf; m++) {
        int per = (int)lround(b->c.filter_para[m][0]);
        if (per != 0 && b->istep % per == 0) any = 1;
    }
    if (!any) return 0;
#define fp b->c.filter_para[m]
    for (int m = 0; m < nf; m++) {
        int per = (int)lround(fp[0]);
        if (per == 0 || b->istep % per != 0) continue;
        int fx = (int)lround(fp[1]), fy = (int)lround(fp[2]), fz = (int)lround(fp[3]);
        int rng[6];
        for (int q = 0; q < 6; q++) rng[q] = (int)lround(fp[4 + q]);
        if (rng[0] < 1) rng[0] = 1;
        if (rng[1] > b->nx) rng[1] = b->nx;
        if (rng[2] < 1) rng[2] = 1;
        if (rng[3] > b->ny) rng[3] = b->ny;
        if (rng[4] < 1) rng[4] = 1;
        if (rng[5] > b->nz) rng[5] = b->nz;
        int scheme = (int)lround(fp[10]);
        double s0 = fp[11], rth = fp[12];
        int on[3] = {fx, fy, fz};
        for (int axis = 0; axis < 3; axis++) {
            if (on[axis] != 1) continue;
            if (scheme == 1) filter_fo9p(b, axis, s0, rng);
            else if (scheme == 2) filter_shock(b, axis, s0, rth, rng);
        }
    }
#undef fp
    return 0;
}

/* ---------------------------------------------------------------------- */
/* diagnostics                                                              */
/* ---------------------------------------------------------------------- */
/* show_flow_msg sums (OCFD_Time_Adv.f90:65-89) */
int oracle_show_flow_msg(oracle_block *b, double out[3]) {
    double p00 = 1.0 / (b->c.gamma * b->c.ma * b->c.ma), tmp1 = 1.0 - b->c.gamma;
    double E1[3] = {0, 0, 0};
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        size_t h = HI(b, i, j, k);
        double tmp0 = 1.0 / b->met[M_AJAC][h];
        E1[0] = E1[0] + b->f[FI(b, i, j, k, 5)] * tmp0;
        E1[1] = E1[1] + b->d[h] * (b->u[h] * b->u[h] + b->v[h] * b->v[h] + b->w[h] * b->w[h]) * 0.5 * tmp0;
        E1[2] = E1[2] + p00 * pow(b->d[h], tmp1) * b->T[h] * tmp0;
    }
    double n = (double)(b->nx * b->ny * b->nz);
    for (int q = 0; q < 3; q++) out[q] = E1[q] / n;
    return 0;
}

/* ana_box (OCFD_ANA.f90:178-311) */
int oracle_ana_box(oracle_block *b, double out[20]) {
    double *g = NULL;
    double *gr[9];
    for (int q = 0; q < 9; q++) gr[q] = b->grad[q];
    (void)g;
    exchange_xyz(b, b->u);
    oracle_dxyz0(b, 0, b->u, gr[0]); oracle_dxyz0(b, 1, b->u, gr[1]); oracle_dxyz0(b, 2, b->u, gr[2]);
    exchange_xyz(b, b->v);
    oracle_dxyz0(b, 0, b->v, gr[3]); oracle_dxyz0(b, 1, b->v, gr[4]); oracle_dxyz0(b, 2, b->v, gr[5]);
    exchange_xyz(b, b->w);
    oracle_dxyz0(b, 0, b->w, gr[6]); oracle_dxyz0(b, 1, b->w, gr[7]); oracle_dxyz0(b, 2, b->w, gr[8]);
    comput_amu(b);
    double s[15] = {0};
    double **M = b->met;
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        size_t q = II(b, i, j, k), h = HI(b, i, j, k);
        double ux = gr[0][q] * M[M_AKX][h] + gr[1][q] * M[M_AIX][h] + gr[2][q] * M[M_ASX][h];
        double uy = gr[0][q] * M[M_AKY][h] + gr[1][q] * M[M_AIY][h] + gr[2][q] * M[M_ASY][h];
        double uz = gr[0][q] * M[M_AKZ][h] + gr[1][q] * M[M_AIZ][h] + gr[2][q] * M[M_ASZ][h];
        double vx = gr[3][q] * M[M_AKX][h] + gr[4][q] * M[M_AIX][h] + gr[5][q] * M[M_ASX][h];
        double vy = gr[3][q] * M[M_AKY][h] + gr[4][q] * M[M_AIY][h] + gr[5][q] * M[M_ASY][h];
        double vz = gr[3][q] * M[M_AKZ][h] + gr[4][q] * M[M_AIZ][h] + gr[5][q] * M[M_ASZ][h];
        double wx = gr[6][q] * M[M_AKX][h] + gr[7][q] * M[M_AIX][h] + gr[8][q] * M[M_ASX][h];
        double wy = gr[6][q] * M[M_AKY][h] + gr[7][q] * M[M_AIY][h] + gr[8][q] * M[M_ASY][h];
        double wz = gr[6][q] * M[M_AKZ][h] + gr[7][q] * M[M_AIZ][h] + gr[8][q] * M[M_ASZ][h];
        (void)wz; (void)vy;
        double uu = b->u[h], vv = b->v[h], ww = b->w[h];
        s[7] = s[7] + sqrt(b->T[h]) / b->c.ma;             /* c_aver */
        s[6] = s[6] + b->d[h];                             /* d_aver */
        s[8] = s[8] + b->amu[q];                           /* Amu_aver */
        s[10] = s[10] + b->T[h];                           /* T_aver */
        s[0] = s[0] + uu * uu;
        s[1] = s[1] + vv * vv;
        s[2] = s[2] + ww * ww;
        s[3] = s[3] + ux * ux;
        s[4] = s[4] + ux * ux * ux;
        s[5] = s[5] + (ux * ux) * (ux * ux);
        s[11] = s[11] + uu * uu * uu;
        s[12] = s[12] + (uu * uu) * (uu * uu);
        s[9] = s[9] + (uu * uu + vv * vv + ww * ww) * b->d[h] * 0.5;
        double e1 = wy - vz, e2 = uz - wx, e3 = vx - uy;
        s[13] = s[13] + e1 * e1 + e2 * e2 + e3 * e3;
    }
    double n = (double)(b->nx * b->ny * b->nz);
    for (int q = 0; q < 14; q++) s[q] = s[q] / n;
    double u_rms = s[0], v_rms = s[1], w_rms = s[2], Ux_rms = s[3];
    double Ux_sk = s[4], Ux_fl = s[5], d_aver = s[6], c_aver = s[7], Amu_aver = s[8], E_k = s[9];
    double T_aver = s[10], u_sk = s[11], u_fl = s[12], ens = 0.5 * s[13];
    double VV_rms = sqrt((u_rms + v_rms + w_rms) / 3.0);
    double Amt = sqrt(u_rms + v_rms + w_rms) / c_aver;
    u_rms = sqrt(u_rms); v_rms = sqrt(v_rms); w_rms = sqrt(w_rms); Ux_rms = sqrt(Ux_rms);
    double Alamda = VV_rms / Ux_rms, Alamdax = u_rms / Ux_rms;
    double Re_l = VV_rms * Alamda * d_aver / Amu_aver, Re_lx = VV_rms * Alamdax * d_aver / Amu_aver;
    Ux_sk = Ux_sk / (Ux_rms * Ux_rms * Ux_rms);
    Ux_fl = Ux_fl / ((Ux_rms * Ux_rms) * (Ux_rms * Ux_rms));
    u_sk = u_sk / (u_rms * u_rms * u_rms);
    u_fl = u_fl / ((u_rms * u_rms) * (u_rms * u_rms));
    double d_rms = 0.0;
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        double t = b->d[HI(b, i, j, k)] - d_aver;
        d_rms = d_rms + t * t;
    }
    d_rms = sqrt(d_rms / n);
    double o[20] = {b->tt, Re_l, Re_lx, Amt, VV_rms, u_rms, v_rms, w_rms, u_sk, u_fl, Ux_sk, Ux_fl,
                    E_k, ens, d_rms, T_aver, Amu_aver, d_aver, Ux_rms, Alamdax};
    memcpy(out, o, sizeof o);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* grid metrics (OCFD_mesh.f90:1-185, OCFD_GhostCell.f90:1-160, 290-500)   */
/* ---------------------------------------------------------------------- */
/* updata_Axyz_periodic (OCFD_mesh.f90:109-185) */
static void coords_periodic_shift(oracle_block *b, const double sp[9]) {
    int L = b->lap, nx = b->nx, ny = b->ny, nz = b->nz;
    double *X[3] = {b->axx, b->ayy, b->azz};
    if (b->c.iperiodic[0] == 1) {
        for (int c = 0; c < 3; c++) for (int k = 1; k <= nz; k++) for (int j = 1; j <= ny; j++) {
            for (int i = 1 - L; i <= 0; i++) X[c][HI(b, i, j, k)] = X[c][HI(b, i, j, k)] - sp[c];
            for (int i = nx + 1; i <= nx + L; i++) X[c][HI(b, i, j, k)] = X[c][HI(b, i, j, k)] + sp[c];
        }
    }
    if (b->c.iperiodic[1] == 1) {
        for (int c = 0; c < 3; c++) for (int k = 1; k <= nz; k++) for (int i = 1; i <= nx; i++) {
            for (int j = 1 - L; j <= 0; j++) X[c][HI(b, i, j, k)] = X[c][HI(b, i, j, k)] - sp[3 + c];
            for (int j = ny + 1; j <= ny + L; j++) X[c][HI(b, i, j, k)] = X[c][HI(b, i, j, k)] + sp[3 + c];
        }
    }
    if (b->c.iperiodic[2] == 1) {
        for (int c = 0; c < 3; c++) for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++) {
            for (int k = 1 - L; k <= 0; k++) X[c][HI(b, i, j, k)] = X[c][HI(b, i, j, k)] - sp[6 + c];
            for (int k = nz + 1; k <= nz + L; k++) X[c][HI(b, i, j, k)] = X[c][HI(b, i, j, k)] + sp[6 + c];
        }
    }
}

static void jac_point(oracle_block *b, size_t h, double xi1, double xj1, double xk1, double yi1,
                      double yj1, double yk1, double zi1, double zj1, double zk1) {
    double Jac1 = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 -
                         yi1 * xj1 * zk1 - xi1 * zj1 * yk1);
    double **M = b->met;
    M[M_AJAC][h] = Jac1;
    M[M_AKX][h] = Jac1 * (yj1 * zk1 - zj1 * yk1);
    M[M_AKY][h] = Jac1 * (zj1 * xk1 - xj1 * zk1);
    M[M_AKZ][h] = Jac1 * (xj1 * yk1 - yj1 * xk1);
    M[M_AIX][h] = Jac1 * (yk1 * zi1 - zk1 * yi1);
    M[M_AIY][h] = Jac1 * (zk1 * xi1 - xk1 * zi1);
    M[M_AIZ][h] = Jac1 * (xk1 * yi1 - yk1 * xi1);
    M[M_ASX][h] = Jac1 * (yi1 * zj1 - zi1 * yj1);
    M[M_ASY][h] = Jac1 * (zi1 * xj1 - xi1 * zj1);
    M[M_ASZ][h] = Jac1 * (xi1 * yj1 - yi1 * xj1);
}

/* comput_Jacobian3d (OCFD_mesh.f90:61-107) */
static void comput_jacobian3d(oracle_block *b) {
    double *d[9];
    for (int q = 0; q < 9; q++) d[q] = (double *)malloc(b->ni * sizeof(double));
    double *X[3] = {b->axx, b->ayy, b->azz};
    /* d[c*3+axis]: xi,xj,xk, yi,yj,yk, zi,zj,zk */
    for (int c = 0; c < 3; c++) for (int a = 0; a < 3; a++) oracle_dxyz0(b, a, X[c], d[c * 3 + a]);
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        size_t q = II(b, i, j, k);
        jac_point(b, HI(b, i, j, k), d[0][q], d[1][q], d[2][q], d[3][q], d[4][q], d[5][q], d[6][q], d[7][q], d[8][q]);
    }
    for (int q = 0; q < 9; q++) free(d[q]);
}

/* Jac_Ghost_Extent_1st (OCFD_GhostCell.f90:18-160) */
static void jac_ghost_1st(oracle_block *b) {
    int L = b->lap, nx = b->nx, ny = b->ny, nz = b->nz;
    const int *gc = b->c.ghost_cell;
    double *X[3] = {b->axx, b->ayy, b->azz};
    for (int side = 0; side < 6; side++) {
        if (gc[side] != 1) continue;
        int axis = side / 2, hi = side % 2;
        int n = axis_n(b, axis);
        for (int k = 1 - L; k <= nz + L; k++) for (int j = 1 - L; j <= ny + L; j++) for (int i = 1 - L; i <= nx + L; i++) {
            int idx[3] = {i, j, k};
            int s = idx[axis];
            int inside = 1;
            for (int a = 0; a < 3; a++) if (a != axis && (idx[a] < 1 || idx[a] > axis_n(b, a))) inside = 0;
            if (!inside) continue;
            if (!hi && s > 0) continue;
            if (hi && s <= n) continue;
            int w[3] = {i, j, k}, w2[3] = {i, j, k};
            w[axis] = hi ? n : 1;
            w2[axis] = hi ? n - 1 : 2;
            size_t g = HI(b, i, j, k), hw = HI(b, w[0], w[1], w[2]), hw2 = HI(b, w2[0], w2[1], w2[2]);
            for (int m = 0; m <= M_AJAC; m++) b->met[m][g] = b->met[m][hw];
            double fac = hi ? (double)(s - n) : (1.0 - s);
            for (int c = 0; c < 3; c++) X[c][g] = X[c][hw] + fac * (X[c][hw] - X[c][hw2]);
        }
    }
}

/* one-sided / central coordinate derivative used in comput_Jacobian3d_Ghost
 * (OCFD_GhostCell.f90:429-477) */
static double ghost_deriv(const double *X, oracle_block *b, int axis, int i, int j, int k, int lo, int hi, double h) {
    int di = axis == 0, dj = axis == 1, dk = axis == 2;
#define XX(s) X[HI(b, i + di * (s), j + dj * (s), k + dk * (s))]
    int s = axis == 0 ? i : (axis == 1 ? j : k);
    if (s == lo) return (-(3.0 * XX(0)) + 4.0 * XX(1) - XX(2)) / (2.0 * h);
    if (s == hi) return (XX(-2) - 4.0 * XX(-1) + 3.0 * XX(0)) / (2.0 * h);
    if (s == lo + 1 || s == hi - 1) return (XX(1) - XX(-1)) / (2.0 * h);
    return (8.0 * (XX(1) - XX(-1)) - (XX(2) - XX(-2))) / (12.0 * h);
#undef XX
}

/* Jac_Ghost_Extent_2nd + comput_Jacobian3d_Ghost (OCFD_GhostCell.f90:290-500) */
static void jac_ghost_2nd(oracle_block *b) {
    int L = b->lap, nx = b->nx, ny = b->ny, nz = b->nz;
    const int *gc = b->c.ghost_cell;
    double *X[3] = {b->axx, b->ayy, b->azz};
    for (int nb = 1; nb <= 6; nb++) {
        if (gc[nb - 1] != 2) continue;
        int axis = (nb - 1) / 2, hiSide = (nb - 1) % 2;
        int n = axis_n(b, axis);
        /* reflected coordinates */
        int rb[3] = {1, 1, 1}, re[3] = {nx, ny, nz};
        if (hiSide) { rb[axis] = n + 1; re[axis] = n + L; } else { rb[axis] = 1 - L; re[axis] = 0; }
        for (int k = rb[2]; k <= re[2]; k++) for (int j = rb[1]; j <= re[1]; j++) for (int i = rb[0]; i <= re[0]; i++) {
            int w[3] = {i, j, k}, mr[3] = {i, j, k};
            int s = axis == 0 ? i : (axis == 1 ? j : k);
            w[axis] = hiSide ? n : 1;
            mr[axis] = hiSide ? 2 * n - s : 2 - s;
            for (int c = 0; c < 3; c++)
                X[c][HI(b, i, j, k)] = 2.0 * X[c][HI(b, w[0], w[1], w[2])] - X[c][HI(b, mr[0], mr[1], mr[2])];
        }
        /* derivative ranges (coordinate data range [b1,e1] per axis) */
        int b1[3] = {1, 1, 1}, e1[3] = {nx, ny, nz};
        if (hiSide) { b1[axis] = n; e1[axis] = n + L; } else { b1[axis] = 1 - L; e1[axis] = 1; }
        double hh[3] = {b->hx, b->hy, b->hz};
        for (int k = rb[2]; k <= re[2]; k++) for (int j = rb[1]; j <= re[1]; j++) for (int i = rb[0]; i <= re[0]; i++) {
            double dd[3][3]; /* [coord][axis] */
            for (int c = 0; c < 3; c++) for (int a = 0; a < 3; a++)
                dd[c][a] = ghost_deriv(X[c], b, a, i, j, k, b1[a], e1[a], hh[a]);
            jac_point(b, HI(b, i, j, k), dd[0][0], dd[0][1], dd[0][2], dd[1][0], dd[1][1], dd[1][2],
                      dd[2][0], dd[2][1], dd[2][2]);
        }
    }
}

static void normalise_metrics(oracle_block *b) {
    for (size_t q = 0; q < b->nh; q++) {
        double J = b->met[M_AJAC][q];
        for (int m = 0; m < 9; m++) b->met[M_AKX1 + m][q] = b->met[M_AKX + m][q] / J;
    }
}

static void load_coords(oracle_block *b, const double *x, const double *y, const double *z) {
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        size_t q = II(b, i, j, k), h = HI(b, i, j, k);
        b->axx[h] = x[q]; b->ayy[h] = y[q]; b->azz[h] = z[q];
    }
    exchange_xyz(b, b->axx); exchange_xyz(b, b->ayy); exchange_xyz(b, b->azz);
}

/* init_mesh after the coordinates are read (OCFD_mesh.f90:18-55) */
int oracle_set_grid(oracle_block *b, const double *x, const double *y, const double *z, const double sp[9]) {
    load_coords(b, x, y, z);
    coords_periodic_shift(b, sp);
    comput_jacobian3d(b);
    for (int m = 0; m <= M_AJAC; m++) exchange_xyz(b, b->met[m]);
    jac_ghost_1st(b);
    jac_ghost_2nd(b);
    normalise_metrics(b);
    return 0;
}

/* GRID_AND_JACOBIAN3D input (OCFD_mesh.f90:199-210): metrics read as given */
int oracle_set_grid_and_jacobian(oracle_block *b, const double *x, const double *y, const double *z,
                                 const double sp[9], const double *const m10[10]) {
    load_coords(b, x, y, z);
    coords_periodic_shift(b, sp);
    for (int m = 0; m <= M_AJAC; m++) {
        for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++)
            b->met[m][HI(b, i, j, k)] = m10[m][II(b, i, j, k)];
        exchange_xyz(b, b->met[m]);
    }
    jac_ghost_1st(b);
    jac_ghost_2nd(b);
    normalise_metrics(b);
    return 0;
}

int oracle_upload_metrics(oracle_block *b, const double *const metrics[19]) {
    for (int m = 0; m < M_N; m++) memcpy(b->met[m], metrics[m], b->nh * sizeof(double));
    return 0;
}
int oracle_download_metrics(oracle_block *b, double *const metrics[19]) {
    for (int m = 0; m < M_N; m++) if (metrics[m]) memcpy(metrics[m], b->met[m], b->nh * sizeof(double));
    return 0;
}

/* ---------------------------------------------------------------------- */
/* state transfer                                                           */
/* ---------------------------------------------------------------------- */
/* init3d: f from the primitives (OCFD_init.f90:10-20) */
int oracle_upload_primitives(oracle_block *b, const double *d, const double *u, const double *v,
                             const double *w, const double *T, int32_t istep, double tt) {
    memcpy(b->d, d, b->nh * sizeof(double)); memcpy(b->u, u, b->nh * sizeof(double));
    memcpy(b->v, v, b->nh * sizeof(double)); memcpy(b->w, w, b->nh * sizeof(double));
    memcpy(b->T, T, b->nh * sizeof(double));
    b->istep = istep; b->tt = tt;
    for (int k = 1; k <= b->nz; k++) for (int j = 1; j <= b->ny; j++) for (int i = 1; i <= b->nx; i++) {
        size_t h = HI(b, i, j, k);
        double dd = b->d[h], uu = b->u[h], vv = b->v[h], ww = b->w[h];
        b->f[FI(b, i, j, k, 1)] = dd;
        b->f[FI(b, i, j, k, 2)] = dd * uu;
        b->f[FI(b, i, j, k, 3)] = dd * vv;
        b->f[FI(b, i, j, k, 4)] = dd * ww;
        b->f[FI(b, i, j, k, 5)] = dd * ((uu * uu + vv * vv + ww * ww) * 0.5 + b->cv * b->T[h]);
    }
    return 0;
}
int oracle_upload_f(oracle_block *b, const double *f) { memcpy(b->f, f, 5 * b->ni * sizeof(double)); return 0; }
int oracle_download_f(oracle_block *b, double *f) { memcpy(f, b->f, 5 * b->ni * sizeof(double)); return 0; }
int oracle_download_fn(oracle_block *b, double *f) { memcpy(f, b->fn, 5 * b->ni * sizeof(double)); return 0; }
int oracle_download_du(oracle_block *b, double *f) { memcpy(f, b->du, 5 * b->ni * sizeof(double)); return 0; }
int oracle_upload_du(oracle_block *b, const double *f) { memcpy(b->du, f, 5 * b->ni * sizeof(double)); return 0; }
int oracle_download_primitives(oracle_block *b, double *d, double *u, double *v, double *w, double *T) {
    double *src[5] = {b->d, b->u, b->v, b->w, b->T}, *dst[5] = {d, u, v, w, T};
    for (int q = 0; q < 5; q++) if (dst[q]) memcpy(dst[q], src[q], b->nh * sizeof(double));
    return 0;
}
int oracle_download_rhybrid(oracle_block *b, double *r) { memcpy(r, b->rhyb, b->nh * sizeof(double)); return 0; }
int oracle_get_time(oracle_block *b, int32_t *istep, double *tt) { *istep = b->istep; *tt = b->tt; return 0; }
int oracle_advance_clock(oracle_block *b) { b->istep += 1; b->tt = b->tt + b->c.dt; return 0; }

/* ---------------------------------------------------------------------- */
/* one time step of the driver loop (OCFD_NS_Solver.f90:23-67)              */
/* ---------------------------------------------------------------------- */
int oracle_step(oracle_block *b, int32_t nsteps) {
    int neg_total = 0;
    for (int n = 0; n < nsteps; n++) {
        oracle_save_fn(b);
        for (int krk = 1; krk <= 3; krk++) {
            oracle_exchange_primitives(b);
            if (krk == 1 && b->c.scheme_invis == OCFD_SCHEME_HYBRID) oracle_comput_rhybrid(b, NULL);
            if (b->c.if_scheme_character == 0) oracle_du_inviscous(b);
            else oracle_du_inviscous_character(b);
            if (b->c.if_viscous == 1) oracle_du_viscous(b);
            oracle_time_adv_rk3(b, krk);
            int neg = oracle_comput_duvwt(b);
            neg_total += neg;
            if (neg > 10) return neg_total;
            if (oracle_bc(b)) return -1;
        }
        oracle_advance_clock(b);
        oracle_filtering(b);
    }
    return neg_total;
}
