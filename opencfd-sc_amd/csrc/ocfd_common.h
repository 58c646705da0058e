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
// Arithmetic of the right-hand side (values of OCFD_PREC_* in the C ABI).
enum : int { OCFD_PREC_FP64_CODE = 0, OCFD_PREC_MIXED_CODE = 1 };

struct SchemeParams {
  int scheme_invis, scheme_vis, flux_split, is_char;
  int ka1, kb1, ka2, kb2;
  int bidx[3][2];    // Scheme%Bound_index
  int sb[6];         // Scheme%Scheme_boundary
  int iperiodic[3];
  double ud7l[8];
  double hyb2, hyb3; // Hybrid_para(2:3) thresholds
};

// Central-difference coefficients of one axis (OCFD_dxyz_viscous.f90:9-16),
// formed once on the host with the reference's expressions.
struct DCoef {
  double h;
  double a1, a2, a3;       // CD6: 1/(60h), -3/(20h), 3/(4h)
  double c1, c2, c3, c4;   // CD8: 0.8/h, -0.2/h, 3.8095...e-2/h, -3.5714...e-3/h
};
__host__ inline DCoef make_dcoef(double h) {
  DCoef c;
  c.h = h;
  c.a1 = 1.0 / (60.0 * h);
  c.a2 = -(3.0 / (20.0 * h));
  c.a3 = 3.0 / (4.0 * h);
  c.c1 = 0.8 / h;
  c.c2 = -(0.2 / h);
  c.c3 = 3.80952380952380952e-2 / h;
  c.c4 = -(3.571428571428571428e-3 / h);
  return c;
}

struct Phys {
  double gamma, ma, re, pr, cv, cp, dt, ref_amu_t0;
  double h[3];       // hx, hy, hz
  DCoef dc[3];       // derivative coefficients per axis
  double hinv[3];    // 1/hx ...
  double mass_force[3];
  int if_mass_force;
};

// ---------------------------------------------------------------------------
// madd(a, b, c) = a*b + c: two roundings in FP64 (the reference's), one fused
// multiply-add in FP32 (the mixed-precision form; explicit, so its rounding
// does not depend on the surrounding code).  A left-to-right sum of products
// p0 + p1 + ... written as madd(.., .., madd(.., .., p0)) rounds exactly as
// the original in FP64, IEEE addition being commutative.
// ---------------------------------------------------------------------------
__device__ inline double madd(double a, double b, double c) { return a * b + c; }
__device__ inline float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// a / b for the nonlinear weights of the WENO schemes: IEEE division in FP64;
// in FP32 a product with the hardware reciprocal (within 1 ulp), which stays
// far inside the mixed form's accuracy, set by its FP32 fluxes
__device__ inline double wdiv(double a, double b) { return a / b; }
__device__ inline float wdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

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

template <> struct K<float> {
  static constexpr float w5ep = 1.0e-6f, w5c0 = 3.0f / 10.0f, w5c1 = 3.0f / 5.0f, w5c2 = 1.0f / 10.0f;
  static constexpr float c1312 = 13.0f / 12.0f, c14 = 1.0f / 4.0f, c13 = 1.0f / 3.0f, c56 = 5.0f / 6.0f,
                         c16 = 1.0f / 6.0f, c76 = 7.0f / 6.0f, c116 = 11.0f / 6.0f;
  static constexpr float w7ep = 1.0e-8f;
  static constexpr float ompep = 1.e-10f;
};

// WENO5-JS, flux+ (hh_weno5P, OCFD_Schemes_1.f90:9-29; OCFD_weno5P, OCFD_Schemes.f90:6-28)
template <typename T>
__device__ inline T weno5p(const T *x) {
  typedef K<T> C;
  const T vm2 = x[1], vm1 = x[2], v0 = x[3], v1 = x[4], v2 = x[5];
  T t, S0, S1, S2;
  t = madd(T(-2.0), v1, v0) + v2; S0 = C::c1312 * (t * t);
  t = madd(T(-4.0), v1, T(3.0) * v0) + v2; S0 = madd(C::c14, t * t, S0);
  t = madd(T(-2.0), v0, vm1) + v1; S1 = C::c1312 * (t * t);
  t = vm1 - v1; S1 = madd(C::c14, t * t, S1);
  t = madd(T(-2.0), vm1, vm2) + v0; S2 = C::c1312 * (t * t);
  t = madd(T(3.0), v0, madd(T(-4.0), vm1, vm2)); S2 = madd(C::c14, t * t, S2);
  T a0 = wdiv(C::w5c0, (C::w5ep + S0) * (C::w5ep + S0));
  T a1 = wdiv(C::w5c1, (C::w5ep + S1) * (C::w5ep + S1));
  T a2 = wdiv(C::w5c2, (C::w5ep + S2) * (C::w5ep + S2));
  T am = a0 + a1 + a2;
  T q03 = madd(-C::c16, v2, madd(C::c56, v1, C::c13 * v0));
  T q13 = madd(C::c13, v1, madd(C::c56, v0, -(C::c16 * vm1)));
  T q23 = madd(C::c116, v0, madd(-C::c76, vm1, C::c13 * vm2));
  return wdiv(madd(a2, q23, madd(a1, q13, a0 * q03)), am);
}

// WENO5-JS, flux- (hh_weno5M, OCFD_Schemes_1.f90:31-52)
template <typename T>
__device__ inline T weno5m(const T *x) {
  typedef K<T> C;
  const T vm1 = x[2], v0 = x[3], v1 = x[4], v2 = x[5], v3 = x[6];
  T t, S0, S1, S2;
  t = madd(T(-2.0), v0, v1) + vm1; S0 = C::c1312 * (t * t);
  t = madd(T(-4.0), v0, T(3.0) * v1) + vm1; S0 = madd(C::c14, t * t, S0);
  t = madd(T(-2.0), v1, v2) + v0; S1 = C::c1312 * (t * t);
  t = v2 - v0; S1 = madd(C::c14, t * t, S1);
  t = madd(T(-2.0), v2, v3) + v1; S2 = C::c1312 * (t * t);
  t = madd(T(3.0), v1, madd(T(-4.0), v2, v3)); S2 = madd(C::c14, t * t, S2);
  T a0 = wdiv(C::w5c0, (C::w5ep + S0) * (C::w5ep + S0));
  T a1 = wdiv(C::w5c1, (C::w5ep + S1) * (C::w5ep + S1));
  T a2 = wdiv(C::w5c2, (C::w5ep + S2) * (C::w5ep + S2));
  T am = a0 + a1 + a2;
  T q03 = madd(-C::c16, vm1, madd(C::c56, v0, C::c13 * v1));
  T q13 = madd(C::c13, v0, madd(C::c56, v1, -(C::c16 * v2)));
  T q23 = madd(C::c116, v1, madd(-C::c76, v2, C::c13 * v3));
  return wdiv(madd(a2, q23, madd(a1, q13, a0 * q03)), am);
}

// WENO5 with the out-of-domain sub-stencil removed (hh_weno5P_boundary,
// OCFD_Schemes_1.f90:266-295)
template <typename T>
__device__ inline T weno5p_boundary(const T *x, int del) {
  typedef K<T> C;
  const T vm2 = x[1], vm1 = x[2], v0 = x[3], v1 = x[4], v2 = x[5];
  T t, S0, S1, S2, a0 = T(0.0), q03 = T(0.0), a2 = T(0.0), q23 = T(0.0);
  if (del != DEL_RIGHT) {
    t = madd(T(-2.0), v1, v0) + v2; S0 = C::c1312 * (t * t);
    t = madd(T(-4.0), v1, T(3.0) * v0) + v2; S0 = madd(C::c14, t * t, S0);
    a0 = wdiv(C::w5c0, (C::w5ep + S0) * (C::w5ep + S0));
    q03 = madd(-C::c16, v2, madd(C::c56, v1, C::c13 * v0));
  }
  t = madd(T(-2.0), v0, vm1) + v1; S1 = C::c1312 * (t * t);
  t = vm1 - v1; S1 = madd(C::c14, t * t, S1);
  T a1 = wdiv(C::w5c1, (C::w5ep + S1) * (C::w5ep + S1));
  T q13 = madd(C::c13, v1, madd(C::c56, v0, -(C::c16 * vm1)));
  if (del != DEL_LIFT) {
    t = madd(T(-2.0), vm1, vm2) + v0; S2 = C::c1312 * (t * t);
    t = madd(T(3.0), v0, madd(T(-4.0), vm1, vm2)); S2 = madd(C::c14, t * t, S2);
    a2 = wdiv(C::w5c2, (C::w5ep + S2) * (C::w5ep + S2));
    q23 = madd(C::c116, v0, madd(-C::c76, vm1, C::c13 * vm2));
  }
  T am = a0 + a1 + a2;
  return wdiv(madd(a2, q23, madd(a1, q13, a0 * q03)), am);
}

// hh_weno5M_boundary (OCFD_Schemes_1.f90:297-325)
template <typename T>
__device__ inline T weno5m_boundary(const T *x, int del) {
  typedef K<T> C;
  const T vm1 = x[2], v0 = x[3], v1 = x[4], v2 = x[5], v3 = x[6];
  T t, S0, S1, S2, a0 = T(0.0), q03 = T(0.0), a2 = T(0.0), q23 = T(0.0);
  if (del != DEL_LIFT) {
    t = madd(T(-2.0), v0, v1) + vm1; S0 = C::c1312 * (t * t);
    t = madd(T(-4.0), v0, T(3.0) * v1) + vm1; S0 = madd(C::c14, t * t, S0);
    a0 = wdiv(C::w5c0, (C::w5ep + S0) * (C::w5ep + S0));
    q03 = madd(-C::c16, vm1, madd(C::c56, v0, C::c13 * v1));
  }
  t = madd(T(-2.0), v1, v2) + v0; S1 = C::c1312 * (t * t);
  t = v2 - v0; S1 = madd(C::c14, t * t, S1);
  T a1 = wdiv(C::w5c1, (C::w5ep + S1) * (C::w5ep + S1));
  T q13 = madd(C::c13, v0, madd(C::c56, v1, -(C::c16 * v2)));
  if (del != DEL_RIGHT) {
    t = madd(T(-2.0), v2, v3) + v1; S2 = C::c1312 * (t * t);
    t = madd(T(3.0), v1, madd(T(-4.0), v2, v3)); S2 = madd(C::c14, t * t, S2);
    a2 = wdiv(C::w5c2, (C::w5ep + S2) * (C::w5ep + S2));
    q23 = madd(C::c116, v1, madd(-C::c76, v2, C::c13 * v3));
  }
  T am = a0 + a1 + a2;
  return wdiv(madd(a2, q23, madd(a1, q13, a0 * q03)), am);
}

// WENO7-JS core on seven values ordered from the upwind end (OCFD_Schemes.f90:59-193:
// the P flux uses v(-3..3), the M flux the mirrored v(4..-2)).
template <typename T>
__device__ inline T weno7_core(T m3, T m2, T m1, T c0, T p1, T p2,
                                    T p3) {
  constexpr T C0 = T(1.0) / T(35.0), C1 = T(12.0) / T(35.0), C2 = T(18.0) / T(35.0), C3 = T(4.0) / T(35.0);
  constexpr T a11 = -T(2.0) / T(6.0), a12 = T(9.0) / T(6.0), a13 = -T(18.0) / T(6.0), a14 = T(11.0) / T(6.0);
  constexpr T a21 = T(1.0) / T(6.0), a23 = T(3.0) / T(6.0), a24 = T(2.0) / T(6.0);
  constexpr T a31 = -T(2.0) / T(6.0), a32 = -T(3.0) / T(6.0), a34 = -T(1.0) / T(6.0);
  constexpr T a41 = -T(11.0) / T(6.0), a42 = T(18.0) / T(6.0), a43 = -T(9.0) / T(6.0), a44 = T(2.0) / T(6.0);
  constexpr T b12 = T(4.0), b13 = -T(5.0), b14 = T(2.0), b22 = -T(2.0), b41 = T(2.0), b42 = -T(5.0), b43 = T(4.0),
                   c12 = T(3.0);
  constexpr T d12 = T(13.0) / T(12.0), d13 = T(1043.0) / T(960.0), d14 = T(1.0) / T(12.0);
  constexpr T e11 = -T(3.0) / T(12.0), e12 = T(13.0) / T(12.0), e13 = -T(23.0) / T(12.0), e14 = T(25.0) / T(12.0);
  constexpr T e21 = T(1.0) / T(12.0), e22 = -T(5.0) / T(12.0), e23 = T(13.0) / T(12.0), e24 = T(3.0) / T(12.0);
  constexpr T e31 = -T(1.0) / T(12.0), e32 = T(7.0) / T(12.0), e33 = T(7.0) / T(12.0), e34 = -T(1.0) / T(12.0);
  constexpr T e41 = T(3.0) / T(12.0), e42 = T(13.0) / T(12.0), e43 = -T(5.0) / T(12.0), e44 = T(1.0) / T(12.0);
  constexpr T ep = T(1.0e-8);
  T S10 = madd(a14, c0, madd(a13, m1, madd(a12, m2, a11 * m3)));
  T S11 = madd(a24, p1, madd(a23, c0, madd(a21, m2, -m1)));
  T S12 = madd(a34, p2, madd(a32, c0, a31 * m1) + p1);
  T S13 = madd(a44, p3, madd(a43, p2, madd(a42, p1, a41 * c0)));
  T S20 = madd(b14, c0, madd(b13, m1, madd(b12, m2, -m3)));
  T S21 = madd(b22, c0, m1) + p1;
  T S22 = madd(b22, p1, c0) + p2;
  T S23 = madd(b43, p2, madd(b42, p1, b41 * c0)) - p3;
  T S30 = madd(c12, m2 - m1, -m3) + c0;
  T S31 = madd(c12, m1 - c0, -m2) + p1;
  T S32 = madd(c12, c0 - p1, -m1) + p2;
  T S33 = madd(c12, p1 - p2, -c0) + p3;
  T S0 = madd(d14 * S10, S30, madd(d13 * S30, S30, madd(d12 * S20, S20, S10 * S10)));
  T S1 = madd(d14 * S11, S31, madd(d13 * S31, S31, madd(d12 * S21, S21, S11 * S11)));
  T S2 = madd(d14 * S12, S32, madd(d13 * S32, S32, madd(d12 * S22, S22, S12 * S12)));
  T S3 = madd(d14 * S13, S33, madd(d13 * S33, S33, madd(d12 * S23, S23, S13 * S13)));
  T a0 = wdiv(C0, (ep + S0) * (ep + S0));
  T a1 = wdiv(C1, (ep + S1) * (ep + S1));
  T a2 = wdiv(C2, (ep + S2) * (ep + S2));
  T a3 = wdiv(C3, (ep + S3) * (ep + S3));
  T am = a0 + a1 + a2 + a3;
  T q0 = madd(e14, c0, madd(e13, m1, madd(e12, m2, e11 * m3)));
  T q1 = madd(e24, p1, madd(e23, c0, madd(e22, m1, e21 * m2)));
  T q2 = madd(e34, p2, madd(e33, p1, madd(e32, c0, e31 * m1)));
  T q3 = madd(e44, p3, madd(e43, p2, madd(e42, p1, e41 * c0)));
  return wdiv(madd(a3, q3, madd(a2, q2, madd(a1, q1, a0 * q0))), am);
}
template <typename T>
__device__ inline T weno7p(const T *x) { return weno7_core(x[0], x[1], x[2], x[3], x[4], x[5], x[6]); }
template <typename T>
__device__ inline T weno7m(const T *x) { return weno7_core(x[7], x[6], x[5], x[4], x[3], x[2], x[1]); }

// Fortran sign(1.d0, x)
template <typename T>
__device__ inline T sgn1(T x) { return copysign(T(1.0), x); }
template <typename T>
__device__ inline T dmin2(T a, T b) { return a < b ? a : b; }
template <typename T>
__device__ inline T dmax2(T a, T b) { return a > b ? a : b; }

// minmod2 / minmod4 (OCFD_Schemes.f90:274-289)
template <typename T>
__device__ inline T minmod2(T x1, T x2) {
  return T(0.5) * (sgn1(x1) + sgn1(x2)) * dmin2(fabs(x1), fabs(x2));
}
template <typename T>
__device__ inline T minmod4(T x1, T x2, T x3, T x4) {
  T r = T(0.5) * (sgn1(x1) + sgn1(x2));
  r = r * fabs(T(0.5) * (sgn1(x1) + sgn1(x3)));
  r = r * fabs(T(0.5) * (sgn1(x1) + sgn1(x4)));
  return r * dmin2(dmin2(dmin2(fabs(x1), fabs(x2)), fabs(x3)), fabs(x4));
}

// OMP6 with the MP limiter on eight values ordered from the upwind end
// (hh_OMP6P OCFD_Schemes_1.f90:195-226, hh_OMP6M :228-260).  The reference's M
// branch uses v(1)-v(0) (x0-x1 in upwind order) in lc where P uses x0-xm1.
template <typename T>
__device__ inline T omp6_core(T xm3, T xm2, T xm1, T x0, T x1, T x2,
                                   T x3, T x4, bool mflux) {
  constexpr T om = T(0.015), on = T(0.0);
  constexpr T a = T(0.5) * (om + on), b = T(0.5) * (om - on);
  constexpr T a1 = a, a2 = T(1.0) / T(60.0) - b - T(6.0) * a, a3 = -(T(2.0) / T(15.0)) + T(6.0) * b + T(15.0) * a;
  constexpr T a4 = T(37.0) / T(60.0) - T(15.0) * b - T(20.0) * a, a5 = T(37.0) / T(60.0) + T(20.0) * b + T(15.0) * a;
  constexpr T a6 = -(T(2.0) / T(15.0)) - T(6.0) * a - T(15.0) * b, a7 = T(1.0) / T(60.0) + a + T(6.0) * b, a8 = -b;
  constexpr T kappa = T(4.0);
  T mid = a1 * x4 + a2 * x3 + a3 * x2 + a4 * x1 + a5 * x0 + a6 * xm1 + a7 * xm2 + a8 * xm3;
  T mp = x0 + minmod2((x1 - x0), kappa * (x0 - xm1));
  if ((mid - x0) * (mid - mp) >= K<T>::ompep) {
    T d1 = xm2 + x0 - T(2.0) * xm1;
    T d2 = xm1 + x1 - T(2.0) * x0;
    T d3 = x0 + x2 - T(2.0) * x1;
    T ul = x0 + kappa * (x0 - xm1);
    T md = T(0.5) * (x0 + x1) - T(0.5) * minmod4(T(4.0) * d2 - d3, T(4.0) * d3 - d2, d2, d3);
    T lc = x0 + T(0.5) * (x0 - (mflux ? x1 : xm1)) + kappa * minmod4(T(4.0) * d1 - d2, T(4.0) * d2 - d1, d2, d1) / T(3.0);
    T fmin = dmax2(dmin2(dmin2(x0, x1), md), dmin2(dmin2(x0, ul), lc));
    T fmax = dmin2(dmax2(dmax2(x0, x1), md), dmax2(dmax2(x0, ul), lc));
    mid = mid + minmod2(fmax - mid, fmin - mid);
  }
  return mid;
}
template <typename T>
__device__ inline T omp6p(const T *x) {
  return omp6_core(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], false);
}
template <typename T>
__device__ inline T omp6m(const T *x) {
  return omp6_core(x[7], x[6], x[5], x[4], x[3], x[2], x[1], x[0], true);
}

// UD7L (OCFD_Schemes.f90:292-316)
template <typename T>
__device__ inline T ud7lp(const double *c, const T *x) {
  T s = T(c[0]) * x[0];
#pragma unroll
  for (int k = 1; k < 8; k++) s = madd(T(c[k]), x[k], s);
  return s;
}
template <typename T>
__device__ inline T ud7lm(const double *c, const T *x) {
  T s = T(c[0]) * x[7];
#pragma unroll
  for (int k = 1; k < 8; k++) s = madd(T(c[k]), x[7 - k], s);
  return s;
}

// set_hybrid_scheme (OCFD_SetHybrid.f90:2-15)
__device__ inline int hybrid_index(double r0, double t2, double t3) {
  return r0 <= t2 ? 1 : (r0 <= t3 ? 2 : 3);
}

// Inner-scheme flux of one face (the scheme select of OCFD2d_flux1/2 and
// flux_x1/x2; OCFD_HybridP/M for the hybrid scheme).
template <int SCH, typename T>
__device__ inline T inner_p(const SchemeParams &S, const T *x, int scm) {
  if (SCH == SCH_WENO5) return weno5p(x);
  if (SCH == SCH_WENO7) return weno7p(x);
  if (SCH == SCH_OMP6) return omp6p(x);
  if (SCH == SCH_UD7L) return ud7lp(S.ud7l, x);
  if (scm == 1) return ud7lp(S.ud7l, x);
  if (scm == 2) return weno7p(x);
  return weno5p(x);
}
template <int SCH, typename T>
__device__ inline T inner_m(const SchemeParams &S, const T *x, int scm) {
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
template <int SCH, typename T>
__device__ inline T face_p_nonchar(const SchemeParams &S, const T *x, int s, int n, int axis,
                                        int scm) {
  const int lo = S.bidx[axis][0], hi = S.bidx[axis][1];
  const int sl = S.sb[2 * axis], sr = S.sb[2 * axis + 1];
  if (lo == 1 && s <= -S.ka1) {
    if (sl == 0) {
      if (s == 0 || s == 1) return (T(2.0) * x[4 - s] + T(5.0) * x[5 - s] - x[6 - s]) / T(6.0);
      if (s == 2) return weno5p_boundary(x, DEL_LIFT);
      return weno5p(x);
    }
    if (sl == 1) return weno5p(x);
  }
  if (hi == 1 && s >= n - S.kb1 + 1) {
    if (sr == 0) {
      if (s == n) return (T(2.0) * x[1] - T(7.0) * x[2] + T(11.0) * x[3]) / T(6.0);
      if (s == n - 1) return weno5p_boundary(x, DEL_RIGHT);
      return weno5p(x);
    }
    if (sr == 1) return weno5p(x);
  }
  return inner_p<SCH>(S, x, scm);
}

// Flux- of face s with the closures of OCFD2d_flux2 (OCFD_inviscous.f90:287-361).
template <int SCH, typename T>
__device__ inline T face_m_nonchar(const SchemeParams &S, const T *x, int s, int n, int axis,
                                        int scm) {
  const int lo = S.bidx[axis][0], hi = S.bidx[axis][1];
  const int sl = S.sb[2 * axis], sr = S.sb[2 * axis + 1];
  if (lo == 1 && s <= -S.ka2) {
    if (sl == 0) {
      if (s == 0) return (T(2.0) * x[6] - T(7.0) * x[5] + T(11.0) * x[4]) / T(6.0);
      if (s == 1) return weno5m_boundary(x, DEL_LIFT);
      return weno5m(x);
    }
    if (sl == 1) return weno5m(x);
  }
  if (hi == 1 && s >= n - S.kb2 + 1) {
    if (sr == 0) {
      if (s == n) return (T(2.0) * x[3] + T(5.0) * x[2] - x[1]) / T(6.0);
      if (s == n - 1) return (T(2.0) * x[4] + T(5.0) * x[3] - x[2]) / T(6.0);
      if (s == n - 2) return weno5m_boundary(x, DEL_RIGHT);
      return weno5m(x);
    }
    if (sr == 1) return weno5m(x);
  }
  return inner_m<SCH>(S, x, scm);
}

// flux_x1 (OCFD_inviscous_character.f90:313-364) for face s of the
// characteristic path; x[k] = projected value at s+k-3.
template <int SCH, typename T>
__device__ inline T face_p_char(const SchemeParams &S, const T *x, int s, int n, int axis, int scm) {
  const int lo = S.bidx[axis][0], hi = S.bidx[axis][1];
  const int sl = S.sb[2 * axis], sr = S.sb[2 * axis + 1];
  const int Ka = S.ka1, Kb = S.kb1;
  if ((lo == 0 || s > -Ka || sl == -1) && (hi == 0 || s <= n - Kb || sr == -1)) return inner_p<SCH>(S, x, scm);
  if (s <= -Ka && lo == 1 && sl != -1) {
    if (sl == 0) {
      if (s == 1) return (T(2.0) * x[3] + T(5.0) * x[4] - x[5]) / T(6.0);
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
template <int SCH, typename T>
__device__ inline T face_m_char(const SchemeParams &S, const T *x, int s, int n, int axis, int scm) {
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
    if (s == n - 1) return (T(2.0) * x[4] + T(5.0) * x[3] - x[2]) / T(6.0);
    if (s == n - 2) return weno5m_boundary(x, DEL_RIGHT);
    return weno5m(x);
  }
  return weno5m(x);
}

// ---------------------------------------------------------------------------
// flux-vector splitting of one point (OCFD_flux_split.f90)
// ---------------------------------------------------------------------------
// Constants of the splittings at a point: functions of gamma and of the
// metric vector A only.  A sweep forms the gamma part once per thread and,
// on a uniform grid, the metric part too; the operations are the routines'
// own, so the fluxes are unchanged bit for bit.
template <typename T>
struct SplitK {
  T gamma, gm1, g2;              // gamma, gamma - 1, 2 gamma
  T sw1, sw3;                    // split_SW: 2 (gamma - 1), (3 - gamma) / (2 (gamma - 1))
  T llf0, llf1;                  // split_LLF: 1 / gamma, 1 / (gamma - 1)
  T A1, A2, A3, ss, ss1, ak1, ak2, ak3;  // A, |A|, and for split_SW 1/|A|, A/|A|
};
template <typename T>
__device__ inline void split_k_gamma(T gamma, SplitK<T> &K) {
  K.gamma = gamma;
  K.gm1 = gamma - T(1.0);
  K.g2 = T(2.0) * gamma;
  K.sw1 = T(2.0) * (gamma - T(1.0));
  K.sw3 = (T(3.0) - gamma) / (T(2.0) * (gamma - T(1.0)));
  K.llf0 = T(1.0) / gamma;
  K.llf1 = T(1.0) / (gamma - T(1.0));
}
template <typename T>
__device__ inline void split_k_metric(T A1, T A2, T A3, bool sw, SplitK<T> &K) {
  K.A1 = A1;
  K.A2 = A2;
  K.A3 = A3;
  K.ss = sqrt(madd(A3, A3, madd(A2, A2, A1 * A1)));
  if (sw) {
    K.ss1 = T(1.0) / K.ss;
    K.ak1 = A1 * K.ss1;
    K.ak2 = A2 * K.ss1;
    K.ak3 = A3 * K.ss1;
  }
}

// split_Stager_Warming (OCFD_flux_split.f90:5-70)
template <typename T>
__device__ inline void split_sw(const SplitK<T> &K, T c1, T d1, T u1, T v1, T w1, T *fp, T *fm) {
  const T tmp1 = K.sw1, tmp3 = K.sw3, ss = K.ss, ak1 = K.ak1, ak2 = K.ak2, ak3 = K.ak3;
  T vs = madd(K.A3, w1, madd(K.A2, v1, K.A1 * u1));
  T E1 = vs, E2 = madd(-c1, ss, vs), E3 = madd(c1, ss, vs);
  T E1P = (E1 + fabs(E1)) * T(0.5), E2P = (E2 + fabs(E2)) * T(0.5), E3P = (E3 + fabs(E3)) * T(0.5);
  T E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
  T tmp0 = d1 / K.g2;
  T uc1 = madd(-c1, ak1, u1), uc2 = madd(c1, ak1, u1);
  T vc1 = madd(-c1, ak2, v1), vc2 = madd(c1, ak2, v1);
  T wc1 = madd(-c1, ak3, w1), wc2 = madd(c1, ak3, w1);
  T vvc1 = madd(wc1, wc1, madd(vc1, vc1, uc1 * uc1)) * T(0.5);
  T vvc2 = madd(wc2, wc2, madd(vc2, vc2, uc2 * uc2)) * T(0.5);
  T vv = K.gm1 * madd(w1, w1, madd(v1, v1, u1 * u1));
  T W2 = tmp3 * c1 * c1;
  fp[0] = tmp0 * (madd(tmp1, E1P, E2P) + E3P);
  fp[1] = tmp0 * madd(E3P, uc2, madd(E2P, uc1, tmp1 * E1P * u1));
  fp[2] = tmp0 * madd(E3P, vc2, madd(E2P, vc1, tmp1 * E1P * v1));
  fp[3] = tmp0 * madd(E3P, wc2, madd(E2P, wc1, tmp1 * E1P * w1));
  fp[4] = tmp0 * madd(W2, E2P + E3P, madd(E3P, vvc2, madd(E2P, vvc1, E1P * vv)));
  fm[0] = tmp0 * (madd(tmp1, E1M, E2M) + E3M);
  fm[1] = tmp0 * madd(E3M, uc2, madd(E2M, uc1, tmp1 * E1M * u1));
  fm[2] = tmp0 * madd(E3M, vc2, madd(E2M, vc1, tmp1 * E1M * v1));
  fm[3] = tmp0 * madd(E3M, wc2, madd(E2M, wc1, tmp1 * E1M * w1));
  fm[4] = tmp0 * madd(W2, E2M + E3M, madd(E3M, vvc2, madd(E2M, vvc1, E1M * vv)));
}

// split_Local_LaxFriedrichs (OCFD_flux_split.f90:73-104)
template <typename T>
__device__ inline void split_llf(const SplitK<T> &K, T c1, T d1, T u1, T v1, T w1, T *fp, T *fm) {
  const T tmp0 = K.llf0, tmp1 = K.llf1, ss = K.ss, A1 = K.A1, A2 = K.A2, A3 = K.A3;
  T un = madd(A3, w1, madd(A2, v1, A1 * u1));
  T lamda0 = madd(ss, c1, fabs(un));
  T p1 = tmp0 * d1 * c1 * c1;
  T E0 = madd(tmp1, p1, T(0.5) * d1 * madd(w1, w1, madd(v1, v1, u1 * u1)));
  T h0 = E0 + p1;
  // (a*b + c*d (+ e*f)) as nested madd; x - y*z as madd(-y, z, x)
  const T ld = lamda0 * d1, nl = -lamda0, nld = -ld;
  const T qu = madd(A1, p1, d1 * u1 * un), qv = madd(A2, p1, d1 * v1 * un), qw = madd(A3, p1, d1 * w1 * un);
  fp[0] = T(0.5) * madd(lamda0, d1, d1 * un);
  fp[1] = T(0.5) * madd(ld, u1, qu);
  fp[2] = T(0.5) * madd(ld, v1, qv);
  fp[3] = T(0.5) * madd(ld, w1, qw);
  fp[4] = T(0.5) * madd(lamda0, E0, h0 * un);
  fm[0] = T(0.5) * madd(nl, d1, d1 * un);
  fm[1] = T(0.5) * madd(nld, u1, qu);
  fm[2] = T(0.5) * madd(nld, v1, qv);
  fm[3] = T(0.5) * madd(nld, w1, qw);
  fm[4] = T(0.5) * madd(nl, E0, h0 * un);
}

// Left (S) / right (S1) eigenvector matrices of the face-averaged state
// (OCFD_inviscous_character.f90:205-248).
template <typename T>
struct CharBasisT {
  T u1, v1, w1, c1, n1, n2, n3, l1, l2, l3, m1, m2, m3, un, ul, um, V2, KK, KK1, X1, H;
};
typedef CharBasisT<double> CharBasis;
template <typename T>
__device__ inline void char_basis(T gamma, T u1, T v1, T w1, T c1, T a1, T a2,
                                  T a3, CharBasisT<T> &B) {
  T ss = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
  B.n1 = a1 / ss; B.n2 = a2 / ss; B.n3 = a3 / ss;
  if (fabs(B.n3) <= fabs(B.n2)) {
    ss = sqrt(B.n1 * B.n1 + B.n2 * B.n2);
    B.l1 = -B.n2 / ss; B.l2 = B.n1 / ss; B.l3 = T(0.0);
  } else {
    ss = sqrt(B.n1 * B.n1 + B.n3 * B.n3);
    B.l1 = -B.n3 / ss; B.l2 = T(0.0); B.l3 = B.n1 / ss;
  }
  B.m1 = B.n2 * B.l3 - B.n3 * B.l2;
  B.m2 = B.n3 * B.l1 - B.n1 * B.l3;
  B.m3 = B.n1 * B.l2 - B.n2 * B.l1;
  B.un = u1 * B.n1 + v1 * B.n2 + w1 * B.n3;
  B.ul = u1 * B.l1 + v1 * B.l2 + w1 * B.l3;
  B.um = u1 * B.m1 + v1 * B.m2 + w1 * B.m3;
  B.V2 = (u1 * u1 + v1 * v1 + w1 * w1) * T(0.5);
  B.KK = (gamma - T(1.0)) / (c1 * c1);
  B.KK1 = B.KK * T(0.5);
  B.X1 = T(1.0) / (T(2.0) * c1);
  B.H = B.V2 + T(1.0) / B.KK;
  B.u1 = u1; B.v1 = v1; B.w1 = w1; B.c1 = c1;
}
// row r of S (OCFD_inviscous_character.f90:238-242)
template <typename T>
__device__ inline void char_S_row(const CharBasisT<T> &B, int r, T *s) {
  switch (r) {
    case 0: s[0] = T(1.0) - B.KK * B.V2; s[1] = B.KK * B.u1; s[2] = B.KK * B.v1; s[3] = B.KK * B.w1; s[4] = -B.KK; break;
    case 1: s[0] = -B.ul; s[1] = B.l1; s[2] = B.l2; s[3] = B.l3; s[4] = T(0.0); break;
    case 2: s[0] = -B.um; s[1] = B.m1; s[2] = B.m2; s[3] = B.m3; s[4] = T(0.0); break;
    case 3:
      s[0] = B.KK1 * B.V2 + B.X1 * B.un; s[1] = -(B.X1 * B.n1) - B.KK1 * B.u1;
      s[2] = -(B.X1 * B.n2) - B.KK1 * B.v1; s[3] = -(B.X1 * B.n3) - B.KK1 * B.w1; s[4] = B.KK1; break;
    default:
      s[0] = B.KK1 * B.V2 - B.X1 * B.un; s[1] = B.X1 * B.n1 - B.KK1 * B.u1;
      s[2] = B.X1 * B.n2 - B.KK1 * B.v1; s[3] = B.X1 * B.n3 - B.KK1 * B.w1; s[4] = B.KK1; break;
  }
}
// hh = S1 * h0 (OCFD_inviscous_character.f90:244-248, 284-288)
template <typename T>
__device__ inline void char_back(const CharBasisT<T> &B, const T *h0, T *hh) {
  const T c1 = B.c1;
  hh[0] = T(1.0) * h0[0] + T(0.0) * h0[1] + T(0.0) * h0[2] + T(1.0) * h0[3] + T(1.0) * h0[4];
  // row . h0 summed left to right: a0 h0 + a1 h1 + a2 h2 + (q - c1 n) h3 + (q + c1 n) h4
  auto row = [&](T a0, T a1, T a2, T q, T nq) {
    return madd(madd(c1, nq, q), h0[4], madd(madd(-c1, nq, q), h0[3], madd(a2, h0[2], madd(a1, h0[1], a0 * h0[0]))));
  };
  hh[1] = row(B.u1, B.l1, B.m1, B.u1, B.n1);
  hh[2] = row(B.v1, B.l2, B.m2, B.v1, B.n2);
  hh[3] = row(B.w1, B.l3, B.m3, B.w1, B.n3);
  hh[4] = row(B.V2, B.ul, B.um, B.H, B.un);
}

// ---------------------------------------------------------------------------
// central first derivative of a field along an axis at interior point s of n
// (OCFD_dx0/dy0/dz0, OCFD_dxyz_viscous.f90:3-239).  `o(d)` returns the field
// at offset d (|d| <= 4) from s; every branch uses constant offsets, so a
// caller can serve o() from registers.  Block-edge closures apply on
// non-periodic axes only.
// ---------------------------------------------------------------------------
template <typename F>
__device__ inline auto dxyz0_point(F o, int s, int n, const DCoef &c, bool cd8, bool periodic) -> decltype(o(0)) {
  typedef decltype(o(0)) T;
  const T h = T(c.h), a1 = T(c.a1), a2 = T(c.a2), a3 = T(c.a3);
  // CD6 weights (CD4-CD6 closure of the CD8 scheme two points from a side)
  auto cd6 = [&]() { return madd(a3, o(1) - o(-1), madd(a2, o(2) - o(-2), a1 * (o(3) - o(-3)))); };
  if (!periodic) {
    if (s == 1) return (-(T(3.0) * o(0)) + T(4.0) * o(1) - o(2)) / (T(2.0) * h);
    if (s == 2) return (o(1) - o(-1)) / (T(2.0) * h);
    if (s == 3) return (T(8.0) * (o(1) - o(-1)) - (o(2) - o(-2))) / (T(12.0) * h);
    if (cd8 && s == 4) return cd6();
    if (s == n) return (o(-2) - T(4.0) * o(-1) + T(3.0) * o(0)) / (T(2.0) * h);
    if (s == n - 1) return (o(1) - o(-1)) / (T(2.0) * h);
    if (s == n - 2) return (T(8.0) * (o(1) - o(-1)) - (o(2) - o(-2))) / (T(12.0) * h);
    if (cd8 && s == n - 3) return cd6();
  }
  if (!cd8) return cd6();
  return madd(T(c.c4), o(4) - o(-4),
              madd(T(c.c3), o(3) - o(-3), madd(T(c.c2), o(2) - o(-2), T(c.c1) * (o(1) - o(-1)))));
}

}  // namespace ocfd
