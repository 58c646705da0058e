// ocfd_viscous.hip -- viscous right-hand side (du_viscous, OCFD_viscous.f90:29-149)
// and the per-cell stage update (OCFD_time_adv_RK3, comput_duvwT and the
// boundary-plane refresh of OCFD_bc, OCFD_Time_Adv.f90:4-63, OCFD_BC.f90:132-200).
//
// k_visc_flux   12 central derivatives of u,v,w,T (CD6/CD8 with one-sided
//               block-edge closures), Sutherland mu, stresses, heat flux and
//               the contravariant fluxes Ev1..Ev3 (12 values per cell)
// k_visc_div    d/dxi Ev1 + d/deta Ev2 + d/dzeta Ev3 added to du(2:5) times
//               Ajac; periodic neighbours by index wrap
// k_visc_div_update   the same divergence fused with the RK3 stage combination,
//               primitive recovery, the boundary-plane refresh of f and the
//               T <= 0 trap (whole-step path; du is not written back)
//
// k_visc_flux and k_visc_div_update march along k: a workgroup owns a 64x4
// tile of (i,j) columns and a segment of kseg planes (up to 32, fewer on small
// blocks so the launch still fills the chip); every thread keeps the
// nine planes k-4..k+4 of the fields it differentiates along k in registers,
// so each element is fetched from HBM once per segment, while the i and j
// neighbours of the current plane come from the L1/L2 caches.
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
  auto at = [&](int d) { return a[base + (int64_t)d * st]; };
  return dxyz0_point(at, s, g.n(axis), B.P.h[axis], B.S.scheme_vis == SCH_CD8, B.S.iperiodic[axis] == 1);
}

// derivative of an interior (no halo) field; periodic neighbours wrap
__device__ inline double di(const Block &B, const double *a, int axis, int i, int j, int k) {
  const Geom &g = B.g;
  const int n = g.n(axis);
  const int64_t base = g.ii(i, j, k);
  const int64_t st = g.istride(axis);
  const int s = axis == 0 ? i : (axis == 1 ? j : k);
  auto at = [&](int d) {
    const int t = s + d;
    const int tt = t < 1 ? t + n : (t > n ? t - n : t);
    return a[base + (int64_t)(tt - s) * st];
  };
  return dxyz0_point(at, s, n, B.P.h[axis], B.S.scheme_vis == SCH_CD8, B.S.iperiodic[axis] == 1);
}

constexpr int TXV = 64, TYV = 4, KSEG_MAX = 32;

struct MarchTile {
  int i, j, kb, ke;
  bool valid;
};
__device__ inline MarchTile march_tile(const Geom &g, int kseg) {
  MarchTile t;
  const int tx = threadIdx.x % TXV, ty = threadIdx.x / TXV;
  const int ntx = (g.nx + TXV - 1) / TXV, nty = (g.ny + TYV - 1) / TYV;
  int b = blockIdx.x;
  const int bx = b % ntx;
  b /= ntx;
  const int by = b % nty;
  const int bz = b / nty;
  t.i = bx * TXV + tx + 1;
  t.j = by * TYV + ty + 1;
  t.kb = bz * kseg + 1;
  t.ke = t.kb + kseg - 1 < g.nz ? t.kb + kseg - 1 : g.nz;
  t.valid = t.i <= g.nx && t.j <= g.ny;
  return t;
}

__global__ __launch_bounds__(NT) void k_visc_flux(const Block B, const int kseg) {
  const Geom &g = B.g;
  const MarchTile mt = march_tile(g, kseg);
  if (!mt.valid) return;
  const int i = mt.i, j = mt.j;
  const int64_t sk = g.hstride(2);
  const bool cd8 = B.S.scheme_vis == SCH_CD8, perz = B.S.iperiodic[2] == 1;
  const double *F[4] = {B.u, B.v, B.w, B.T};
  double q[4][9];  // u, v, w, T at k-4 .. k+4
  {
    const int64_t h0 = g.hi(i, j, mt.kb);
#pragma unroll
    for (int f = 0; f < 4; f++)
#pragma unroll
      for (int d = 0; d < 8; d++) q[f][d] = F[f][h0 + (int64_t)(d - 4) * sk];
  }
  const Metrics &M = B.M;
  const double Tsb = 110.4 / B.P.ref_amu_t0, TmpR = 1.0 / B.P.re;
  const int64_t ni = g.ni;
  for (int k = mt.kb; k <= mt.ke; k++) {
    const int64_t h = g.hi(i, j, k);
#pragma unroll
    for (int f = 0; f < 4; f++) q[f][8] = F[f][h + 4 * sk];
    double dz[4];
#pragma unroll
    for (int f = 0; f < 4; f++) dz[f] = dxyz0_point([&](int d) { return q[f][4 + d]; }, k, g.nz, B.P.h[2], cd8, perz);
    // comput_Amu (OCFD_viscous.f90:135-149)
    const double Tc = q[3][4];
    const double amu = TmpR * (1.0 + Tsb) * sqrt(Tc * Tc * Tc) / (Tsb + Tc);
    // OCFD_viscous.f90:42-53
    const double uk = dh(B, B.u, 0, i, j, k), ui = dh(B, B.u, 1, i, j, k), us = dz[0];
    const double vk = dh(B, B.v, 0, i, j, k), vi = dh(B, B.v, 1, i, j, k), vs = dz[1];
    const double wk = dh(B, B.w, 0, i, j, k), wi = dh(B, B.w, 1, i, j, k), ws = dz[2];
    const double Tk = dh(B, B.T, 0, i, j, k), Ti = dh(B, B.T, 1, i, j, k), Ts = dz[3];
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
    const double uu = q[0][4], vv = q[1][4], ww = q[2][4];
    const double E1 = uu * s11 + vv * s12 + ww * s13 + Amuk * Tx;
    const double E2 = uu * s12 + vv * s22 + ww * s23 + Amuk * Ty;
    const double E3 = uu * s13 + vv * s23 + ww * s33 + Amuk * Tz;
    const int64_t t = g.ii(i, j, k);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double n1 = M.get(M_AKX1 + 3 * a, h), n2 = M.get(M_AKX1 + 3 * a + 1, h), n3 = M.get(M_AKX1 + 3 * a + 2, h);
      double *e = B.ev + (int64_t)(4 * a) * ni + t;
      e[0] = n1 * s11 + n2 * s12 + n3 * s13;
      e[ni] = n1 * s12 + n2 * s22 + n3 * s23;
      e[2 * ni] = n1 * s13 + n2 * s23 + n3 * s33;
      e[3 * ni] = n1 * E1 + n2 * E2 + n3 * E3;
    }
#pragma unroll
    for (int f = 0; f < 4; f++)
#pragma unroll
      for (int d = 0; d < 8; d++) q[f][d] = q[f][d + 1];
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
                                                        const int viscous, const int kseg, int *neg) {
  const Geom &g = B.g;
  const MarchTile mt = march_tile(g, kseg);
  if (!mt.valid) return;
  const int i = mt.i, j = mt.j, nz = g.nz;
  const int64_t ni = g.ni, skz = g.istride(2);
  const bool cd8 = B.S.scheme_vis == SCH_CD8, perz = B.S.iperiodic[2] == 1;
  const double dt = B.P.dt;
  // Ev3(1:4) at k-4 .. k+4; periodic planes wrap, out-of-block planes are
  // clamped (their values are never used by the one-sided closures)
  auto kwrap = [&](int t) { return perz ? (t < 1 ? t + nz : (t > nz ? t - nz : t)) : (t < 1 ? 1 : (t > nz ? nz : t)); };
  const int64_t col = g.ii(i, j, 1);
  const double *e3 = B.ev + (int64_t)8 * ni;
  double q[4][9];
  if (viscous) {
#pragma unroll
    for (int d = 0; d < 8; d++) {
      const int64_t o = col + (int64_t)(kwrap(mt.kb + d - 4) - 1) * skz;
#pragma unroll
      for (int m = 0; m < 4; m++) q[m][d] = e3[(int64_t)m * ni + o];
    }
  }
  for (int k = mt.kb; k <= mt.ke; k++) {
    const int64_t t = col + (int64_t)(k - 1) * skz;
    const int64_t h = g.hi(i, j, k);
    const double ajac = B.M.get(M_AJAC, h);
    double vd[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (viscous) {
      const int64_t o = col + (int64_t)(kwrap(k + 4) - 1) * skz;
#pragma unroll
      for (int m = 0; m < 4; m++) q[m][8] = e3[(int64_t)m * ni + o];
#pragma unroll
      for (int m = 1; m <= 4; m++) {
        const double *e1 = B.ev + (int64_t)(m - 1) * ni;
        const double *e2 = B.ev + (int64_t)(4 + m - 1) * ni;
        const double d3 = dxyz0_point([&](int d) { return q[m - 1][4 + d]; }, k, nz, B.P.h[2], cd8, perz);
        vd[m] = di(B, e1, 0, i, j, k) + di(B, e2, 1, i, j, k) + d3;
      }
    }
    double fq[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      double du = B.du[t + m * ni];
      if (viscous && m > 0) du = du + vd[m] * ajac;
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
    if (viscous) {
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int dd = 0; dd < 8; dd++) q[m][dd] = q[m][dd + 1];
    }
  }
}

}  // namespace

static inline unsigned nblocks(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

// planes per workgroup: the longest segment (<= 32) that still gives about
// 2048 workgroups
static inline int march_kseg(const Geom &g) {
  const int64_t cols = (int64_t)((g.nx + TXV - 1) / TXV) * ((g.ny + TYV - 1) / TYV);
  int64_t k = cols * g.nz / 2048;
  return (int)(k < 1 ? 1 : (k > KSEG_MAX ? KSEG_MAX : k));
}
static inline unsigned march_blocks(const Geom &g, int kseg) {
  return (unsigned)(((g.nx + TXV - 1) / TXV) * ((g.ny + TYV - 1) / TYV) * ((g.nz + kseg - 1) / kseg));
}

hipError_t launch_visc_flux(const Block &B, hipStream_t st) {
  const int kseg = march_kseg(B.g);
  hipLaunchKernelGGL(k_visc_flux, dim3(march_blocks(B.g, kseg)), dim3(NT), 0, st, B, kseg);
  return hipGetLastError();
}

hipError_t launch_visc_div(const Block &B, hipStream_t st) {
  hipLaunchKernelGGL(k_visc_div, dim3(nblocks(B.g.ni)), dim3(NT), 0, st, B);
  return hipGetLastError();
}

hipError_t launch_visc_div_update(const Block &B, int krk, int refresh_planes, int *neg, hipStream_t st) {
  const int viscous = B.ev != nullptr;
  const int kseg = march_kseg(B.g);
  hipLaunchKernelGGL(k_visc_div_update, dim3(march_blocks(B.g, kseg)), dim3(NT), 0, st, B, krk, refresh_planes, viscous,
                     kseg, neg);
  return hipGetLastError();
}

}  // namespace ocfd
