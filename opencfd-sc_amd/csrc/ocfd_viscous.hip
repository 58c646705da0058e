// ocfd_viscous.hip -- viscous right-hand side (du_viscous, OCFD_viscous.f90:29-149)
// and the per-cell stage update (OCFD_time_adv_RK3, comput_duvwT and the
// boundary-plane refresh of OCFD_bc, OCFD_Time_Adv.f90:4-63, OCFD_BC.f90:132-200).
//
// k_visc_flux   one thread per cell: 12 central derivatives of u,v,w,T
//               (CD6/CD8 with one-sided block-edge closures), Sutherland mu,
//               stresses, heat flux, contravariant fluxes Ev1..Ev3 (12 values)
// k_visc_div    one thread per cell: d/dxi Ev1 + d/deta Ev2 + d/dzeta Ev3,
//               added to du(2:5) times Ajac; periodic neighbours by index wrap
// k_visc_div_update   the same divergence fused with the RK3 stage combination,
//               primitive recovery, the boundary-plane refresh of f and the
//               T <= 0 trap (whole-step path; du is not written back)
#include "ocfd_kernels.h"

namespace ocfd {

namespace {

constexpr int NT = 256;

__device__ inline void cell_of(const Geom &g, int64_t t, int &i, int &j, int &k) {
  i = (int)(t % g.nx) + 1;
  int64_t r = t / g.nx;
  j = (int)(r % g.ny) + 1;
  k = (int)(r / g.ny) + 1;
}

// derivative of a halo'd field along an axis at interior point (i,j,k)
__device__ inline double dh(const Block &B, const double *a, int axis, int i, int j, int k) {
  const Geom &g = B.g;
  const int64_t base = g.hi(i, j, k);
  const int64_t st = g.hstride(axis);
  const int s = axis == 0 ? i : (axis == 1 ? j : k);
  auto at = [&](int t) { return a[base + (int64_t)(t - s) * st]; };
  return dxyz0_point(at, s, g.n(axis), B.P.h[axis], B.S.scheme_vis == SCH_CD8, B.S.iperiodic[axis] == 1);
}

// derivative of an interior (no halo) field; periodic neighbours wrap
__device__ inline double di(const Block &B, const double *a, int axis, int i, int j, int k) {
  const Geom &g = B.g;
  const int n = g.n(axis);
  const int64_t base = g.ii(i, j, k);
  const int64_t st = g.istride(axis);
  const int s = axis == 0 ? i : (axis == 1 ? j : k);
  auto at = [&](int t) {
    int tt = t < 1 ? t + n : (t > n ? t - n : t);
    return a[base + (int64_t)(tt - s) * st];
  };
  return dxyz0_point(at, s, n, B.P.h[axis], B.S.scheme_vis == SCH_CD8, B.S.iperiodic[axis] == 1);
}

__global__ __launch_bounds__(NT) void k_visc_flux(const Block B) {
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int i, j, k;
  cell_of(g, t, i, j, k);
  const int64_t h = g.hi(i, j, k);
  // comput_Amu (OCFD_viscous.f90:135-149)
  const double Tc = B.T[h];
  const double Tsb = 110.4 / B.P.ref_amu_t0, TmpR = 1.0 / B.P.re;
  const double amu = TmpR * (1.0 + Tsb) * sqrt(Tc * Tc * Tc) / (Tsb + Tc);
  // OCFD_viscous.f90:42-53
  const double uk = dh(B, B.u, 0, i, j, k), ui = dh(B, B.u, 1, i, j, k), us = dh(B, B.u, 2, i, j, k);
  const double vk = dh(B, B.v, 0, i, j, k), vi = dh(B, B.v, 1, i, j, k), vs = dh(B, B.v, 2, i, j, k);
  const double wk = dh(B, B.w, 0, i, j, k), wi = dh(B, B.w, 1, i, j, k), ws = dh(B, B.w, 2, i, j, k);
  const double Tk = dh(B, B.T, 0, i, j, k), Ti = dh(B, B.T, 1, i, j, k), Ts = dh(B, B.T, 2, i, j, k);
  const Metrics &M = B.M;
  const double akx = M.get(M_AKX, h), aky = M.get(M_AKY, h), akz = M.get(M_AKZ, h);
  const double aix = M.get(M_AIX, h), aiy = M.get(M_AIY, h), aiz = M.get(M_AIZ, h);
  const double asx = M.get(M_ASX, h), asy = M.get(M_ASY, h), asz = M.get(M_ASZ, h);
  // OCFD_viscous.f90:60-111
  const double ux = uk * akx + ui * aix + us * asx;
  const double vx = vk * akx + vi * aix + vs * asx;
  const double wx = wk * akx + wi * aix + ws * asx;
  const double Tx = Tk * akx + Ti * aix + Ts * asx;
  const double uy = uk * aky + ui * aiy + us * asy;
  const double vy = vk * aky + vi * aiy + vs * asy;
  const double wy = wk * aky + wi * aiy + ws * asy;
  const double Ty = Tk * aky + Ti * aiy + Ts * asy;
  const double uz = uk * akz + ui * aiz + us * asz;
  const double vz = vk * akz + vi * aiz + vs * asz;
  const double wz = wk * akz + wi * aiz + ws * asz;
  const double Tz = Tk * akz + Ti * aiz + Ts * asz;
  const double div = ux + vy + wz;
  const double Amu1 = amu;
  const double Amuk = B.P.cp * amu / B.P.pr;
  const double tmp2_3 = 2.0 / 3.0;
  const double s11 = (2.0 * ux - tmp2_3 * div) * Amu1;
  const double s22 = (2.0 * vy - tmp2_3 * div) * Amu1;
  const double s33 = (2.0 * wz - tmp2_3 * div) * Amu1;
  const double s12 = (uy + vx) * Amu1, s13 = (uz + wx) * Amu1, s23 = (vz + wy) * Amu1;
  const double uu = B.u[h], vv = B.v[h], ww = B.w[h];
  const double E1 = uu * s11 + vv * s12 + ww * s13 + Amuk * Tx;
  const double E2 = uu * s12 + vv * s22 + ww * s23 + Amuk * Ty;
  const double E3 = uu * s13 + vv * s23 + ww * s33 + Amuk * Tz;
  const int64_t ni = g.ni;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double n1 = M.get(M_AKX1 + 3 * a, h), n2 = M.get(M_AKX1 + 3 * a + 1, h), n3 = M.get(M_AKX1 + 3 * a + 2, h);
    double *e = B.ev + (int64_t)(4 * a) * ni + t;
    e[0] = n1 * s11 + n2 * s12 + n3 * s13;
    e[ni] = n1 * s12 + n2 * s22 + n3 * s23;
    e[2 * ni] = n1 * s13 + n2 * s23 + n3 * s33;
    e[3 * ni] = n1 * E1 + n2 * E2 + n3 * E3;
  }
}

// viscous divergence of component m (1..4) at cell t (OCFD_viscous.f90:114-129)
__device__ inline double visc_div_m(const Block &B, int m, int i, int j, int k) {
  const int64_t ni = B.g.ni;
  const double *e1 = B.ev + (int64_t)(m - 1) * ni;
  const double *e2 = B.ev + (int64_t)(4 + m - 1) * ni;
  const double *e3 = B.ev + (int64_t)(8 + m - 1) * ni;
  return di(B, e1, 0, i, j, k) + di(B, e2, 1, i, j, k) + di(B, e3, 2, i, j, k);
}

__global__ __launch_bounds__(NT) void k_visc_div(const Block B) {
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int i, j, k;
  cell_of(g, t, i, j, k);
  const double ajac = B.M.get(M_AJAC, g.hi(i, j, k));
#pragma unroll
  for (int m = 1; m <= 4; m++) {
    const double s = visc_div_m(B, m, i, j, k);
    B.du[t + (int64_t)m * g.ni] = B.du[t + (int64_t)m * g.ni] + s * ajac;
  }
}

// primitives from conserved variables (OCFD_Time_Adv.f90:48-52)
__device__ inline void prims_of(double cv, const double *fq, double &d, double &u, double &v, double &w, double &T) {
  d = fq[0];
  u = fq[1] / fq[0];
  v = fq[2] / fq[0];
  w = fq[3] / fq[0];
  T = (fq[4] - (fq[1] * u + fq[2] * v + fq[3] * w) * 0.5) / (fq[0] * cv);
}
// f from primitives on the boundary planes (OCFD_BC.f90:135-139)
__device__ inline void refresh_of(double cv, double d, double u, double v, double w, double T, double *fq) {
  fq[0] = d;
  fq[1] = d * u;
  fq[2] = d * v;
  fq[3] = d * w;
  fq[4] = d * (cv * T + (u * u + v * v + w * w) * 0.5);
}

// Fused: du(2:5) += div Ev * Ajac; RK3 stage; primitives; plane refresh.
// krk == 1 reads f as fn (fn = f at the step start) and stores fn.
__global__ __launch_bounds__(NT) void k_visc_div_update(const Block B, const int krk, const int refresh,
                                                        const int viscous, int *neg) {
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int i, j, k;
  cell_of(g, t, i, j, k);
  const int64_t h = g.hi(i, j, k);
  const int64_t ni = g.ni;
  const double ajac = B.M.get(M_AJAC, h);
  const double dt = B.P.dt;
  double fq[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    double du = B.du[t + m * ni];
    if (viscous && m > 0) du = du + visc_div_m(B, m, i, j, k) * ajac;
    const double f = B.f[t + m * ni];
    double r;
    if (krk == 1) {
      B.fn[t + m * ni] = f;
      r = f + dt * du;
    } else {
      const double ra = krk == 2 ? 3.0 / 4.0 : 1.0 / 3.0;
      const double rb = krk == 2 ? 1.0 / 4.0 : 2.0 / 3.0;
      r = ra * B.fn[t + m * ni] + rb * (f + dt * du);
    }
    fq[m] = r;
  }
  double d, u, v, w, T;
  prims_of(B.P.cv, fq, d, u, v, w, T);
  B.d[h] = d; B.u[h] = u; B.v[h] = v; B.w[h] = w; B.T[h] = T;
  if (T <= 0) {
    const int slot = atomicAdd(neg, 1);
    if (slot < 64) neg[1 + slot] = (int)t;
  }
  if (refresh && (i == 1 || i == g.nx || j == 1 || j == g.ny || k == 1 || k == g.nz))
    refresh_of(B.P.cv, d, u, v, w, T, fq);
#pragma unroll
  for (int m = 0; m < 5; m++) B.f[t + m * ni] = fq[m];
}

}  // namespace

static inline unsigned nblocks(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

hipError_t launch_visc_flux(const Block &B, hipStream_t st) {
  hipLaunchKernelGGL(k_visc_flux, dim3(nblocks(B.g.ni)), dim3(NT), 0, st, B);
  return hipGetLastError();
}

hipError_t launch_visc_div(const Block &B, hipStream_t st) {
  hipLaunchKernelGGL(k_visc_div, dim3(nblocks(B.g.ni)), dim3(NT), 0, st, B);
  return hipGetLastError();
}

hipError_t launch_visc_div_update(const Block &B, int krk, int refresh_planes, int *neg, hipStream_t st) {
  const int viscous = B.ev != nullptr;
  hipLaunchKernelGGL(k_visc_div_update, dim3(nblocks(B.g.ni)), dim3(NT), 0, st, B, krk, refresh_planes, viscous,
                     neg);
  return hipGetLastError();
}

}  // namespace ocfd
