// ocfd_common.h -- data layout, launch parameters and the point-wise numerics
// shared by every gfx950 kernel of the OpenCFD-SC right-hand side.
//
// Layout in HBM (one block, FP64):
//   primitives d,u,v,w,T, Rhybrid, coordinates, metrics: halo'd arrays
//     X(1-L:nx+L, 1-L:ny+L, 1-L:nz+L), i fastest (the reference's own layout,
//     OCFD_NS_Solver.f90:102-135), so host<->device copies are plain memcpy;
//   f, fn, du: five planes of nx*ny*nz, i fastest (OCFD_NS_Solver.f90:99);
//   viscous fluxes Ev1..Ev3 (4 each): interior only, periodic neighbours are
//     addressed by index wrap instead of halo copies.
//
// All floating-point expressions keep the reference's operand order and are
// compiled with -ffp-contract=off, so the FP64 path is bit-identical to the
// reference on the same inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ocfd {

constexpr int LAP = 4;

// scheme codes (OCFD_parameters.f90:23-29)
enum : int { SCH_WENO5 = 1, SCH_WENO7 = 2, SCH_OMP6 = 3, SCH_UD7L = 4, SCH_CD6 = 5, SCH_CD8 = 6,
             SCH_HYBRID = 10 };
enum : int { DEL_NONE = 0, DEL_LIFT = 1, DEL_RIGHT = 2 };
// metric field indices
enum : int { M_AKX, M_AKY, M_AKZ, M_AIX, M_AIY, M_AIZ, M_ASX, M_ASY, M_ASZ, M_AJAC, M_AKX1,
             M_AKY1, M_AKZ1, M_AIX1, M_AIY1, M_AIZ1, M_ASX1, M_ASY1, M_ASZ1, M_N };

struct Geom {
  int nx, ny, nz;
  int sx, sy, sz;      // halo'd extents
  int64_t nh, ni;      // halo'd / interior point counts
  __host__ __device__ inline int64_t hi(int i, int j, int k) const {
    return (int64_t)(i + LAP - 1) + (int64_t)sx * ((int64_t)(j + LAP - 1) + (int64_t)sy * (k + LAP - 1));
  }
  __host__ __device__ inline int64_t ii(int i, int j, int k) const {
    return (int64_t)(i - 1) + (int64_t)nx * ((int64_t)(j - 1) + (int64_t)ny * (k - 1));
  }
  __host__ __device__ inline int n(int axis) const { return axis == 0 ? nx : (axis == 1 ? ny : nz); }
  // stride of one step along `axis` in a halo'd array / interior array
  __host__ __device__ inline int64_t hstride(int axis) const {
    return axis == 0 ? 1 : (axis == 1 ? (int64_t)sx : (int64_t)sx * sy);
  }
  __host__ __device__ inline int64_t istride(int axis) const {
    return axis == 0 ? 1 : (axis == 1 ? (int64_t)nx : (int64_t)nx * ny);
  }
};

// Metrics: per-point fields, or the ten constants of a uniform grid plus the
// nine normalised values formed as A/Ajac (OCFD_mesh.f90:41-55).
struct Metrics {
  const double *p[M_N];
  double u[M_N];
  int uniform;
  __device__ inline double get(int q, int64_t h) const { return uniform ? u[q] : p[q][h]; }
};

// Scheme parameters derived by set_parameters (OCFD_readpara.f90:261-341).
struct SchemeParams {
  int scheme_invis, scheme_vis, flux_split, is_char;
  int ka1, kb1, ka2, kb2;
  int bidx[3][2];    // Scheme%Bound_index
  int sb[6];         // Scheme%Scheme_boundary
  int iperiodic[3];
  double ud7l[8];
  double hyb2, hyb3; // Hybrid_para(2:3) thresholds
};

struct Phys {
  double gamma, ma, re, pr, cv, cp, dt, ref_amu_t0;
  double h[3];       // hx, hy, hz
  double hinv[3];    // 1/hx ...
  double mass_force[3];
  int if_mass_force;
};

// ---------------------------------------------------------------------------
// point-wise reconstructions.  x[k] holds the stencil value v(k-3), k=0..7,
// i.e. x[3] is the left cell of the face (v(0) in OCFD_Schemes_1.f90).
// ---------------------------------------------------------------------------
template <typename R> struct K;  // constants in the working precision

template <> struct K<double> {
  static constexpr double w5ep = 1.0e-6, w5c0 = 3.0 / 10.0, w5c1 = 3.0 / 5.0, w5c2 = 1.0 / 10.0;
  static constexpr double c1312 = 13.0 / 12.0, c14 = 1.0 / 4.0, c13 = 1.0 / 3.0, c56 = 5.0 / 6.0,
                          c16 = 1.0 / 6.0, c76 = 7.0 / 6.0, c116 = 11.0 / 6.0;
  static constexpr double w7ep = 1.0e-8;
  static constexpr double ompep = (double)1.e-10f;  // REAL(4) literal (OCFD_Schemes.f90:209)
};

// WENO5-JS, flux+ (hh_weno5P, OCFD_Schemes_1.f90:9-29; OCFD_weno5P, OCFD_Schemes.f90:6-28)
__device__ inline double weno5p(const double *x) {
  typedef K<double> C;
  const double vm2 = x[1], vm1 = x[2], v0 = x[3], v1 = x[4], v2 = x[5];
  double t, S0, S1, S2;
  t = v0 - 2.0 * v1 + v2; S0 = C::c1312 * (t * t);
  t = 3.0 * v0 - 4.0 * v1 + v2; S0 = S0 + C::c14 * (t * t);
  t = vm1 - 2.0 * v0 + v1; S1 = C::c1312 * (t * t);
  t = vm1 - v1; S1 = S1 + C::c14 * (t * t);
  t = vm2 - 2.0 * vm1 + v0; S2 = C::c1312 * (t * t);
  t = vm2 - 4.0 * vm1 + 3.0 * v0; S2 = S2 + C::c14 * (t * t);
  double a0 = C::w5c0 / ((C::w5ep + S0) * (C::w5ep + S0));
  double a1 = C::w5c1 / ((C::w5ep + S1) * (C::w5ep + S1));
  double a2 = C::w5c2 / ((C::w5ep + S2) * (C::w5ep + S2));
  double am = a0 + a1 + a2;
  double q03 = C::c13 * v0 + C::c56 * v1 - C::c16 * v2;
  double q13 = -(C::c16 * vm1) + C::c56 * v0 + C::c13 * v1;
  double q23 = C::c13 * vm2 - C::c76 * vm1 + C::c116 * v0;
  return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

// WENO5-JS, flux- (hh_weno5M, OCFD_Schemes_1.f90:31-52)
__device__ inline double weno5m(const double *x) {
  typedef K<double> C;
  const double vm1 = x[2], v0 = x[3], v1 = x[4], v2 = x[5], v3 = x[6];
  double t, S0, S1, S2;
  t = v1 - 2.0 * v0 + vm1; S0 = C::c1312 * (t * t);
  t = 3.0 * v1 - 4.0 * v0 + vm1; S0 = S0 + C::c14 * (t * t);
  t = v2 - 2.0 * v1 + v0; S1 = C::c1312 * (t * t);
  t = v2 - v0; S1 = S1 + C::c14 * (t * t);
  t = v3 - 2.0 * v2 + v1; S2 = C::c1312 * (t * t);
  t = v3 - 4.0 * v2 + 3.0 * v1; S2 = S2 + C::c14 * (t * t);
  double a0 = C::w5c0 / ((C::w5ep + S0) * (C::w5ep + S0));
  double a1 = C::w5c1 / ((C::w5ep + S1) * (C::w5ep + S1));
  double a2 = C::w5c2 / ((C::w5ep + S2) * (C::w5ep + S2));
  double am = a0 + a1 + a2;
  double q03 = C::c13 * v1 + C::c56 * v0 - C::c16 * vm1;
  double q13 = -(C::c16 * v2) + C::c56 * v1 + C::c13 * v0;
  double q23 = C::c13 * v3 - C::c76 * v2 + C::c116 * v1;
  return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

// WENO5 with the out-of-domain sub-stencil removed (hh_weno5P_boundary,
// OCFD_Schemes_1.f90:266-295)
__device__ inline double weno5p_boundary(const double *x, int del) {
  typedef K<double> C;
  const double vm2 = x[1], vm1 = x[2], v0 = x[3], v1 = x[4], v2 = x[5];
  double t, S0, S1, S2, a0 = 0.0, q03 = 0.0, a2 = 0.0, q23 = 0.0;
  if (del != DEL_RIGHT) {
    t = v0 - 2.0 * v1 + v2; S0 = C::c1312 * (t * t);
    t = 3.0 * v0 - 4.0 * v1 + v2; S0 = S0 + C::c14 * (t * t);
    a0 = C::w5c0 / ((C::w5ep + S0) * (C::w5ep + S0));
    q03 = C::c13 * v0 + C::c56 * v1 - C::c16 * v2;
  }
  t = vm1 - 2.0 * v0 + v1; S1 = C::c1312 * (t * t);
  t = vm1 - v1; S1 = S1 + C::c14 * (t * t);
  double a1 = C::w5c1 / ((C::w5ep + S1) * (C::w5ep + S1));
  double q13 = -(C::c16 * vm1) + C::c56 * v0 + C::c13 * v1;
  if (del != DEL_LIFT) {
    t = vm2 - 2.0 * vm1 + v0; S2 = C::c1312 * (t * t);
    t = vm2 - 4.0 * vm1 + 3.0 * v0; S2 = S2 + C::c14 * (t * t);
    a2 = C::w5c2 / ((C::w5ep + S2) * (C::w5ep + S2));
    q23 = C::c13 * vm2 - C::c76 * vm1 + C::c116 * v0;
  }
  double am = a0 + a1 + a2;
  return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

// hh_weno5M_boundary (OCFD_Schemes_1.f90:297-325)
__device__ inline double weno5m_boundary(const double *x, int del) {
  typedef K<double> C;
  const double vm1 = x[2], v0 = x[3], v1 = x[4], v2 = x[5], v3 = x[6];
  double t, S0, S1, S2, a0 = 0.0, q03 = 0.0, a2 = 0.0, q23 = 0.0;
  if (del != DEL_LIFT) {
    t = v1 - 2.0 * v0 + vm1; S0 = C::c1312 * (t * t);
    t = 3.0 * v1 - 4.0 * v0 + vm1; S0 = S0 + C::c14 * (t * t);
    a0 = C::w5c0 / ((C::w5ep + S0) * (C::w5ep + S0));
    q03 = C::c13 * v1 + C::c56 * v0 - C::c16 * vm1;
  }
  t = v2 - 2.0 * v1 + v0; S1 = C::c1312 * (t * t);
  t = v2 - v0; S1 = S1 + C::c14 * (t * t);
  double a1 = C::w5c1 / ((C::w5ep + S1) * (C::w5ep + S1));
  double q13 = -(C::c16 * v2) + C::c56 * v1 + C::c13 * v0;
  if (del != DEL_RIGHT) {
    t = v3 - 2.0 * v2 + v1; S2 = C::c1312 * (t * t);
    t = v3 - 4.0 * v2 + 3.0 * v1; S2 = S2 + C::c14 * (t * t);
    a2 = C::w5c2 / ((C::w5ep + S2) * (C::w5ep + S2));
    q23 = C::c13 * v3 - C::c76 * v2 + C::c116 * v1;
  }
  double am = a0 + a1 + a2;
  return (a0 * q03 + a1 * q13 + a2 * q23) / am;
}

// WENO7-JS core on seven values ordered from the upwind end (OCFD_Schemes.f90:59-193:
// the P flux uses v(-3..3), the M flux the mirrored v(4..-2)).
__device__ inline double weno7_core(double m3, double m2, double m1, double c0, double p1, double p2,
                                    double p3) {
  constexpr double C0 = 1.0 / 35.0, C1 = 12.0 / 35.0, C2 = 18.0 / 35.0, C3 = 4.0 / 35.0;
  constexpr double a11 = -2.0 / 6.0, a12 = 9.0 / 6.0, a13 = -18.0 / 6.0, a14 = 11.0 / 6.0;
  constexpr double a21 = 1.0 / 6.0, a23 = 3.0 / 6.0, a24 = 2.0 / 6.0;
  constexpr double a31 = -2.0 / 6.0, a32 = -3.0 / 6.0, a34 = -1.0 / 6.0;
  constexpr double a41 = -11.0 / 6.0, a42 = 18.0 / 6.0, a43 = -9.0 / 6.0, a44 = 2.0 / 6.0;
  constexpr double b12 = 4.0, b13 = -5.0, b14 = 2.0, b22 = -2.0, b41 = 2.0, b42 = -5.0, b43 = 4.0,
                   c12 = 3.0;
  constexpr double d12 = 13.0 / 12.0, d13 = 1043.0 / 960.0, d14 = 1.0 / 12.0;
  constexpr double e11 = -3.0 / 12.0, e12 = 13.0 / 12.0, e13 = -23.0 / 12.0, e14 = 25.0 / 12.0;
  constexpr double e21 = 1.0 / 12.0, e22 = -5.0 / 12.0, e23 = 13.0 / 12.0, e24 = 3.0 / 12.0;
  constexpr double e31 = -1.0 / 12.0, e32 = 7.0 / 12.0, e33 = 7.0 / 12.0, e34 = -1.0 / 12.0;
  constexpr double e41 = 3.0 / 12.0, e42 = 13.0 / 12.0, e43 = -5.0 / 12.0, e44 = 1.0 / 12.0;
  constexpr double ep = 1.0e-8;
  double S10 = a11 * m3 + a12 * m2 + a13 * m1 + a14 * c0;
  double S11 = a21 * m2 - m1 + a23 * c0 + a24 * p1;
  double S12 = a31 * m1 + a32 * c0 + p1 + a34 * p2;
  double S13 = a41 * c0 + a42 * p1 + a43 * p2 + a44 * p3;
  double S20 = -m3 + b12 * m2 + b13 * m1 + b14 * c0;
  double S21 = m1 + b22 * c0 + p1;
  double S22 = c0 + b22 * p1 + p2;
  double S23 = b41 * c0 + b42 * p1 + b43 * p2 - p3;
  double S30 = -m3 + c12 * (m2 - m1) + c0;
  double S31 = -m2 + c12 * (m1 - c0) + p1;
  double S32 = -m1 + c12 * (c0 - p1) + p2;
  double S33 = -c0 + c12 * (p1 - p2) + p3;
  double S0 = S10 * S10 + d12 * S20 * S20 + d13 * S30 * S30 + d14 * S10 * S30;
  double S1 = S11 * S11 + d12 * S21 * S21 + d13 * S31 * S31 + d14 * S11 * S31;
  double S2 = S12 * S12 + d12 * S22 * S22 + d13 * S32 * S32 + d14 * S12 * S32;
  double S3 = S13 * S13 + d12 * S23 * S23 + d13 * S33 * S33 + d14 * S13 * S33;
  double a0 = C0 / ((ep + S0) * (ep + S0));
  double a1 = C1 / ((ep + S1) * (ep + S1));
  double a2 = C2 / ((ep + S2) * (ep + S2));
  double a3 = C3 / ((ep + S3) * (ep + S3));
  double am = a0 + a1 + a2 + a3;
  double q0 = e11 * m3 + e12 * m2 + e13 * m1 + e14 * c0;
  double q1 = e21 * m2 + e22 * m1 + e23 * c0 + e24 * p1;
  double q2 = e31 * m1 + e32 * c0 + e33 * p1 + e34 * p2;
  double q3 = e41 * c0 + e42 * p1 + e43 * p2 + e44 * p3;
  return (a0 * q0 + a1 * q1 + a2 * q2 + a3 * q3) / am;
}
__device__ inline double weno7p(const double *x) { return weno7_core(x[0], x[1], x[2], x[3], x[4], x[5], x[6]); }
__device__ inline double weno7m(const double *x) { return weno7_core(x[7], x[6], x[5], x[4], x[3], x[2], x[1]); }

// Fortran sign(1.d0, x)
__device__ inline double sgn1(double x) { return copysign(1.0, x); }
__device__ inline double dmin2(double a, double b) { return a < b ? a : b; }
__device__ inline double dmax2(double a, double b) { return a > b ? a : b; }

// minmod2 / minmod4 (OCFD_Schemes.f90:274-289)
__device__ inline double minmod2(double x1, double x2) {
  return 0.5 * (sgn1(x1) + sgn1(x2)) * dmin2(fabs(x1), fabs(x2));
}
__device__ inline double minmod4(double x1, double x2, double x3, double x4) {
  double r = 0.5 * (sgn1(x1) + sgn1(x2));
  r = r * fabs(0.5 * (sgn1(x1) + sgn1(x3)));
  r = r * fabs(0.5 * (sgn1(x1) + sgn1(x4)));
  return r * dmin2(dmin2(dmin2(fabs(x1), fabs(x2)), fabs(x3)), fabs(x4));
}

// OMP6 with the MP limiter on eight values ordered from the upwind end
// (hh_OMP6P OCFD_Schemes_1.f90:195-226, hh_OMP6M :228-260).  The reference's M
// branch uses v(1)-v(0) (x0-x1 in upwind order) in lc where P uses x0-xm1.
__device__ inline double omp6_core(double xm3, double xm2, double xm1, double x0, double x1, double x2,
                                   double x3, double x4, bool mflux) {
  constexpr double om = 0.015, on = 0.0;
  constexpr double a = 0.5 * (om + on), b = 0.5 * (om - on);
  constexpr double a1 = a, a2 = 1.0 / 60.0 - b - 6.0 * a, a3 = -(2.0 / 15.0) + 6.0 * b + 15.0 * a;
  constexpr double a4 = 37.0 / 60.0 - 15.0 * b - 20.0 * a, a5 = 37.0 / 60.0 + 20.0 * b + 15.0 * a;
  constexpr double a6 = -(2.0 / 15.0) - 6.0 * a - 15.0 * b, a7 = 1.0 / 60.0 + a + 6.0 * b, a8 = -b;
  constexpr double kappa = 4.0;
  double mid = a1 * x4 + a2 * x3 + a3 * x2 + a4 * x1 + a5 * x0 + a6 * xm1 + a7 * xm2 + a8 * xm3;
  double mp = x0 + minmod2((x1 - x0), kappa * (x0 - xm1));
  if ((mid - x0) * (mid - mp) >= K<double>::ompep) {
    double d1 = xm2 + x0 - 2.0 * xm1;
    double d2 = xm1 + x1 - 2.0 * x0;
    double d3 = x0 + x2 - 2.0 * x1;
    double ul = x0 + kappa * (x0 - xm1);
    double md = 0.5 * (x0 + x1) - 0.5 * minmod4(4.0 * d2 - d3, 4.0 * d3 - d2, d2, d3);
    double lc = x0 + 0.5 * (x0 - (mflux ? x1 : xm1)) + kappa * minmod4(4.0 * d1 - d2, 4.0 * d2 - d1, d2, d1) / 3.0;
    double fmin = dmax2(dmin2(dmin2(x0, x1), md), dmin2(dmin2(x0, ul), lc));
    double fmax = dmin2(dmax2(dmax2(x0, x1), md), dmax2(dmax2(x0, ul), lc));
    mid = mid + minmod2(fmax - mid, fmin - mid);
  }
  return mid;
}
__device__ inline double omp6p(const double *x) {
  return omp6_core(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], false);
}
__device__ inline double omp6m(const double *x) {
  return omp6_core(x[7], x[6], x[5], x[4], x[3], x[2], x[1], x[0], true);
}

// UD7L (OCFD_Schemes.f90:292-316)
__device__ inline double ud7lp(const double *c, const double *x) {
  return c[0] * x[0] + c[1] * x[1] + c[2] * x[2] + c[3] * x[3] + c[4] * x[4] + c[5] * x[5] + c[6] * x[6] +
         c[7] * x[7];
}
__device__ inline double ud7lm(const double *c, const double *x) {
  return c[0] * x[7] + c[1] * x[6] + c[2] * x[5] + c[3] * x[4] + c[4] * x[3] + c[5] * x[2] + c[6] * x[1] +
         c[7] * x[0];
}

// set_hybrid_scheme (OCFD_SetHybrid.f90:207-220)
__device__ inline int hybrid_index(double r0, double t2, double t3) {
  return r0 <= t2 ? 1 : (r0 <= t3 ? 2 : 3);
}

// Inner-scheme flux of one face (the scheme select of OCFD2d_flux1/2 and
// flux_x1/x2; OCFD_HybridP/M for the hybrid scheme).
template <int SCH>
__device__ inline double inner_p(const SchemeParams &S, const double *x, int scm) {
  if (SCH == SCH_WENO5) return weno5p(x);
  if (SCH == SCH_WENO7) return weno7p(x);
  if (SCH == SCH_OMP6) return omp6p(x);
  if (SCH == SCH_UD7L) return ud7lp(S.ud7l, x);
  if (scm == 1) return ud7lp(S.ud7l, x);
  if (scm == 2) return weno7p(x);
  return weno5p(x);
}
template <int SCH>
__device__ inline double inner_m(const SchemeParams &S, const double *x, int scm) {
  if (SCH == SCH_WENO5) return weno5m(x);
  if (SCH == SCH_WENO7) return weno7m(x);
  if (SCH == SCH_OMP6) return omp6m(x);
  if (SCH == SCH_UD7L) return ud7lm(S.ud7l, x);
  if (scm == 1) return ud7lm(S.ud7l, x);
  if (scm == 2) return weno7m(x);
  return weno5m(x);
}

// Flux+ of face s (left cell s, 0 <= s <= n) including the block-edge closures
// of OCFD2d_flux1 (OCFD_inviscous.f90:210-283).  x[k] = f(s+k-3).
template <int SCH>
__device__ inline double face_p_nonchar(const SchemeParams &S, const double *x, int s, int n, int axis,
                                        int scm) {
  const int lo = S.bidx[axis][0], hi = S.bidx[axis][1];
  const int sl = S.sb[2 * axis], sr = S.sb[2 * axis + 1];
  if (lo == 1 && s <= -S.ka1) {
    if (sl == 0) {
      if (s == 0 || s == 1) return (2.0 * x[4 - s] + 5.0 * x[5 - s] - x[6 - s]) / 6.0;
      if (s == 2) return weno5p_boundary(x, DEL_LIFT);
      return weno5p(x);
    }
    if (sl == 1) return weno5p(x);
  }
  if (hi == 1 && s >= n - S.kb1 + 1) {
    if (sr == 0) {
      if (s == n) return (2.0 * x[1] - 7.0 * x[2] + 11.0 * x[3]) / 6.0;
      if (s == n - 1) return weno5p_boundary(x, DEL_RIGHT);
      return weno5p(x);
    }
    if (sr == 1) return weno5p(x);
  }
  return inner_p<SCH>(S, x, scm);
}

// Flux- of face s with the closures of OCFD2d_flux2 (OCFD_inviscous.f90:287-361).
template <int SCH>
__device__ inline double face_m_nonchar(const SchemeParams &S, const double *x, int s, int n, int axis,
                                        int scm) {
  const int lo = S.bidx[axis][0], hi = S.bidx[axis][1];
  const int sl = S.sb[2 * axis], sr = S.sb[2 * axis + 1];
  if (lo == 1 && s <= -S.ka2) {
    if (sl == 0) {
      if (s == 0) return (2.0 * x[6] - 7.0 * x[5] + 11.0 * x[4]) / 6.0;
      if (s == 1) return weno5m_boundary(x, DEL_LIFT);
      return weno5m(x);
    }
    if (sl == 1) return weno5m(x);
  }
  if (hi == 1 && s >= n - S.kb2 + 1) {
    if (sr == 0) {
      if (s == n) return (2.0 * x[3] + 5.0 * x[2] - x[1]) / 6.0;
      if (s == n - 1) return (2.0 * x[4] + 5.0 * x[3] - x[2]) / 6.0;
      if (s == n - 2) return weno5m_boundary(x, DEL_RIGHT);
      return weno5m(x);
    }
    if (sr == 1) return weno5m(x);
  }
  return inner_m<SCH>(S, x, scm);
}

// flux_x1 (OCFD_inviscous_character.f90:313-364) for face s of the
// characteristic path; x[k] = projected value at s+k-3.
template <int SCH>
__device__ inline double face_p_char(const SchemeParams &S, const double *x, int s, int n, int axis, int scm) {
  const int lo = S.bidx[axis][0], hi = S.bidx[axis][1];
  const int sl = S.sb[2 * axis], sr = S.sb[2 * axis + 1];
  const int Ka = S.ka1, Kb = S.kb1;
  if ((lo == 0 || s > -Ka || sl == -1) && (hi == 0 || s <= n - Kb || sr == -1)) return inner_p<SCH>(S, x, scm);
  if (s <= -Ka && lo == 1 && sl != -1) {
    if (sl == 0) {
      if (s == 1) return (2.0 * x[3] + 5.0 * x[4] - x[5]) / 6.0;
      if (s == 2) return weno5p_boundary(x, DEL_LIFT);
      return weno5p(x);
    }
    return weno5p(x);
  }
  if (sr == 0) {
    if (s == n - 1) return weno5p_boundary(x, DEL_RIGHT);
    return weno5p(x);
  }
  return weno5p(x);
}

// flux_x2 (OCFD_inviscous_character.f90:367-416)
template <int SCH>
__device__ inline double face_m_char(const SchemeParams &S, const double *x, int s, int n, int axis, int scm) {
  const int lo = S.bidx[axis][0], hi = S.bidx[axis][1];
  const int sl = S.sb[2 * axis], sr = S.sb[2 * axis + 1];
  const int Ka = S.ka2, Kb = S.kb2;
  if ((lo == 0 || s > -Ka || sl == -1) && (hi == 0 || s <= n - Kb || sr == -1)) return inner_m<SCH>(S, x, scm);
  if (s <= -Ka && lo == 1 && sl != -1) {
    if (sl == 0) {
      if (s == 1) return weno5m_boundary(x, DEL_LIFT);
      return weno5m(x);
    }
    return weno5m(x);
  }
  if (sr == 0) {
    if (s == n - 1) return (2.0 * x[4] + 5.0 * x[3] - x[2]) / 6.0;
    if (s == n - 2) return weno5m_boundary(x, DEL_RIGHT);
    return weno5m(x);
  }
  return weno5m(x);
}

// ---------------------------------------------------------------------------
// flux-vector splitting of one point (OCFD_flux_split.f90)
// ---------------------------------------------------------------------------
// split_Stager_Warming (OCFD_flux_split.f90:5-70)
__device__ inline void split_sw(double gamma, double c1, double d1, double u1, double v1, double w1, double A1,
                                double A2, double A3, double *fp, double *fm) {
  const double tmp1 = 2.0 * (gamma - 1.0);
  const double tmp3 = (3.0 - gamma) / (2.0 * (gamma - 1.0));
  double ss = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
  double ss1 = 1.0 / ss;
  double ak1 = A1 * ss1, ak2 = A2 * ss1, ak3 = A3 * ss1;
  double vs = A1 * u1 + A2 * v1 + A3 * w1;
  double E1 = vs, E2 = vs - c1 * ss, E3 = vs + c1 * ss;
  double E1P = (E1 + fabs(E1)) * 0.5, E2P = (E2 + fabs(E2)) * 0.5, E3P = (E3 + fabs(E3)) * 0.5;
  double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
  double tmp0 = d1 / (2.0 * gamma);
  double uc1 = u1 - c1 * ak1, uc2 = u1 + c1 * ak1;
  double vc1 = v1 - c1 * ak2, vc2 = v1 + c1 * ak2;
  double wc1 = w1 - c1 * ak3, wc2 = w1 + c1 * ak3;
  double vvc1 = (uc1 * uc1 + vc1 * vc1 + wc1 * wc1) * 0.5;
  double vvc2 = (uc2 * uc2 + vc2 * vc2 + wc2 * wc2) * 0.5;
  double vv = (gamma - 1.0) * (u1 * u1 + v1 * v1 + w1 * w1);
  double W2 = tmp3 * c1 * c1;
  fp[0] = tmp0 * (tmp1 * E1P + E2P + E3P);
  fp[1] = tmp0 * (tmp1 * E1P * u1 + E2P * uc1 + E3P * uc2);
  fp[2] = tmp0 * (tmp1 * E1P * v1 + E2P * vc1 + E3P * vc2);
  fp[3] = tmp0 * (tmp1 * E1P * w1 + E2P * wc1 + E3P * wc2);
  fp[4] = tmp0 * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
  fm[0] = tmp0 * (tmp1 * E1M + E2M + E3M);
  fm[1] = tmp0 * (tmp1 * E1M * u1 + E2M * uc1 + E3M * uc2);
  fm[2] = tmp0 * (tmp1 * E1M * v1 + E2M * vc1 + E3M * vc2);
  fm[3] = tmp0 * (tmp1 * E1M * w1 + E2M * wc1 + E3M * wc2);
  fm[4] = tmp0 * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
}

// split_Local_LaxFriedrichs (OCFD_flux_split.f90:73-104)
__device__ inline void split_llf(double gamma, double c1, double d1, double u1, double v1, double w1,
                                 double A1, double A2, double A3, double *fp, double *fm) {
  const double tmp0 = 1.0 / gamma, tmp1 = 1.0 / (gamma - 1.0);
  double ss = sqrt(A1 * A1 + A2 * A2 + A3 * A3);
  double un = A1 * u1 + A2 * v1 + A3 * w1;
  double lamda0 = fabs(un) + ss * c1;
  double p1 = tmp0 * d1 * c1 * c1;
  double E0 = tmp1 * p1 + 0.5 * d1 * (u1 * u1 + v1 * v1 + w1 * w1);
  double h0 = E0 + p1;
  fp[0] = 0.5 * (d1 * un + lamda0 * d1);
  fp[1] = 0.5 * (d1 * u1 * un + A1 * p1 + lamda0 * d1 * u1);
  fp[2] = 0.5 * (d1 * v1 * un + A2 * p1 + lamda0 * d1 * v1);
  fp[3] = 0.5 * (d1 * w1 * un + A3 * p1 + lamda0 * d1 * w1);
  fp[4] = 0.5 * (h0 * un + lamda0 * E0);
  fm[0] = 0.5 * (d1 * un - lamda0 * d1);
  fm[1] = 0.5 * (d1 * u1 * un + A1 * p1 - lamda0 * d1 * u1);
  fm[2] = 0.5 * (d1 * v1 * un + A2 * p1 - lamda0 * d1 * v1);
  fm[3] = 0.5 * (d1 * w1 * un + A3 * p1 - lamda0 * d1 * w1);
  fm[4] = 0.5 * (h0 * un - lamda0 * E0);
}

// Left (S) / right (S1) eigenvector matrices of the face-averaged state
// (OCFD_inviscous_character.f90:205-248).
struct CharBasis {
  double u1, v1, w1, c1, n1, n2, n3, l1, l2, l3, m1, m2, m3, un, ul, um, V2, KK, KK1, X1, H;
};
__device__ inline void char_basis(double gamma, double u1, double v1, double w1, double c1, double a1, double a2,
                                  double a3, CharBasis &B) {
  double ss = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
  B.n1 = a1 / ss; B.n2 = a2 / ss; B.n3 = a3 / ss;
  if (fabs(B.n3) <= fabs(B.n2)) {
    ss = sqrt(B.n1 * B.n1 + B.n2 * B.n2);
    B.l1 = -B.n2 / ss; B.l2 = B.n1 / ss; B.l3 = 0.0;
  } else {
    ss = sqrt(B.n1 * B.n1 + B.n3 * B.n3);
    B.l1 = -B.n3 / ss; B.l2 = 0.0; B.l3 = B.n1 / ss;
  }
  B.m1 = B.n2 * B.l3 - B.n3 * B.l2;
  B.m2 = B.n3 * B.l1 - B.n1 * B.l3;
  B.m3 = B.n1 * B.l2 - B.n2 * B.l1;
  B.un = u1 * B.n1 + v1 * B.n2 + w1 * B.n3;
  B.ul = u1 * B.l1 + v1 * B.l2 + w1 * B.l3;
  B.um = u1 * B.m1 + v1 * B.m2 + w1 * B.m3;
  B.V2 = (u1 * u1 + v1 * v1 + w1 * w1) * 0.5;
  B.KK = (gamma - 1.0) / (c1 * c1);
  B.KK1 = B.KK * 0.5;
  B.X1 = 1.0 / (2.0 * c1);
  B.H = B.V2 + 1.0 / B.KK;
  B.u1 = u1; B.v1 = v1; B.w1 = w1; B.c1 = c1;
}
// row r of S (OCFD_inviscous_character.f90:238-242)
__device__ inline void char_S_row(const CharBasis &B, int r, double *s) {
  switch (r) {
    case 0: s[0] = 1.0 - B.KK * B.V2; s[1] = B.KK * B.u1; s[2] = B.KK * B.v1; s[3] = B.KK * B.w1; s[4] = -B.KK; break;
    case 1: s[0] = -B.ul; s[1] = B.l1; s[2] = B.l2; s[3] = B.l3; s[4] = 0.0; break;
    case 2: s[0] = -B.um; s[1] = B.m1; s[2] = B.m2; s[3] = B.m3; s[4] = 0.0; break;
    case 3:
      s[0] = B.KK1 * B.V2 + B.X1 * B.un; s[1] = -(B.X1 * B.n1) - B.KK1 * B.u1;
      s[2] = -(B.X1 * B.n2) - B.KK1 * B.v1; s[3] = -(B.X1 * B.n3) - B.KK1 * B.w1; s[4] = B.KK1; break;
    default:
      s[0] = B.KK1 * B.V2 - B.X1 * B.un; s[1] = B.X1 * B.n1 - B.KK1 * B.u1;
      s[2] = B.X1 * B.n2 - B.KK1 * B.v1; s[3] = B.X1 * B.n3 - B.KK1 * B.w1; s[4] = B.KK1; break;
  }
}
// hh = S1 * h0 (OCFD_inviscous_character.f90:244-248, 284-288)
__device__ inline void char_back(const CharBasis &B, const double *h0, double *hh) {
  const double c1 = B.c1;
  hh[0] = 1.0 * h0[0] + 0.0 * h0[1] + 0.0 * h0[2] + 1.0 * h0[3] + 1.0 * h0[4];
  hh[1] = B.u1 * h0[0] + B.l1 * h0[1] + B.m1 * h0[2] + (B.u1 - c1 * B.n1) * h0[3] + (B.u1 + c1 * B.n1) * h0[4];
  hh[2] = B.v1 * h0[0] + B.l2 * h0[1] + B.m2 * h0[2] + (B.v1 - c1 * B.n2) * h0[3] + (B.v1 + c1 * B.n2) * h0[4];
  hh[3] = B.w1 * h0[0] + B.l3 * h0[1] + B.m3 * h0[2] + (B.w1 - c1 * B.n3) * h0[3] + (B.w1 + c1 * B.n3) * h0[4];
  hh[4] = B.V2 * h0[0] + B.ul * h0[1] + B.um * h0[2] + (B.H - c1 * B.un) * h0[3] + (B.H + c1 * B.un) * h0[4];
}

// ---------------------------------------------------------------------------
// central first derivative of a field along an axis at interior point s of n
// (OCFD_dx0/dy0/dz0, OCFD_dxyz_viscous.f90:3-239).  `o(d)` returns the field
// at offset d (|d| <= 4) from s; every branch uses constant offsets, so a
// caller can serve o() from registers.  Block-edge closures apply on
// non-periodic axes only.
// ---------------------------------------------------------------------------
template <typename F>
__device__ inline double dxyz0_point(F o, int s, int n, double h, bool cd8, bool periodic) {
  if (!periodic) {
    if (s == 1) return (-(3.0 * o(0)) + 4.0 * o(1) - o(2)) / (2.0 * h);
    if (s == 2) return (o(1) - o(-1)) / (2.0 * h);
    if (s == 3) return (8.0 * (o(1) - o(-1)) - (o(2) - o(-2))) / (12.0 * h);
    if (cd8 && s == 4) {
      const double a1 = 1.0 / (60.0 * h), a2 = -(3.0 / (20.0 * h)), a3 = 3.0 / (4.0 * h);
      return a1 * (o(3) - o(-3)) + a2 * (o(2) - o(-2)) + a3 * (o(1) - o(-1));
    }
    if (s == n) return (o(-2) - 4.0 * o(-1) + 3.0 * o(0)) / (2.0 * h);
    if (s == n - 1) return (o(1) - o(-1)) / (2.0 * h);
    if (s == n - 2) return (8.0 * (o(1) - o(-1)) - (o(2) - o(-2))) / (12.0 * h);
    if (cd8 && s == n - 3) {
      const double a1 = 1.0 / (60.0 * h), a2 = -(3.0 / (20.0 * h)), a3 = 3.0 / (4.0 * h);
      return a1 * (o(3) - o(-3)) + a2 * (o(2) - o(-2)) + a3 * (o(1) - o(-1));
    }
  }
  if (!cd8) {
    const double a1 = 1.0 / (60.0 * h), a2 = -(3.0 / (20.0 * h)), a3 = 3.0 / (4.0 * h);
    return a1 * (o(3) - o(-3)) + a2 * (o(2) - o(-2)) + a3 * (o(1) - o(-1));
  }
  const double c1 = 0.8 / h, c2 = -(0.2 / h), c3 = 3.80952380952380952e-2 / h, c4 = -(3.571428571428571428e-3 / h);
  return c1 * (o(1) - o(-1)) + c2 * (o(2) - o(-2)) + c3 * (o(3) - o(-3)) + c4 * (o(4) - o(-4));
}

}  // namespace ocfd
