// ocfd_visc.h -- viscous right-hand side (du_viscous, OCFD_viscous.f90:29-149)
// and the per-cell stage update (OCFD_time_adv_RK3, comput_duvwT and the
// boundary-plane refresh of OCFD_bc, OCFD_Time_Adv.f90:4-63, OCFD_BC.f90:26-95).
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
// k_visc_flux and k_visc_div_update march along k over a 64x4 tile of (i,j)
// columns and a segment of up to 64 planes in FP64, 32 in FP32 (fewer on
// small blocks so the launch still fills the chip).  The x and y neighbours of the current plane
// come from a double-buffered LDS tile: while one plane is computed, every
// lane already holds its share of the next plane's tile in registers, so
// there is one barrier per plane and the loads overlap the arithmetic.  The
// z neighbours of a lane's own column are a register queue of the 2R+1
// planes k-R..k+R (R = 3 for CD6, 4 for CD8), refilled one plane ahead.
#pragma once
#include "ocfd_kernels.h"

namespace ocfd {

namespace {

// derivative of a halo'd field along an axis at interior point (i,j,k)
__device__ inline double dh(const Block &B, const double *a, int axis, int i, int j, int k) {
  const Geom &g = B.g;
  const int64_t base = g.hi(i, j, k);
  const int64_t st = g.hstride(axis);
  const int s = axis == 0 ? i : (axis == 1 ? j : k);
  auto at = [&](int d) { return a[base + (int64_t)d * st]; };
  return dxyz0_point(at, s, g.n(axis), B.P.dc[axis], B.S.scheme_vis == SCH_CD8, B.S.iperiodic[axis] == 1);
}

// derivative of an interior (no halo) field; periodic neighbours wrap
template <typename R>
__device__ inline R di(const Block &B, const R *a, int axis, int i, int j, int k) {
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
  return dxyz0_point(at, s, n, B.P.dc[axis], B.S.scheme_vis == SCH_CD8, B.S.iperiodic[axis] == 1);
}

constexpr int TXT = 64, TYT = 4, KSEG_MAX = 64;

// Plane tile geometry: stencil radius R of the central difference, tile of
// LX x LY points around the TXT x TYT columns.
template <bool CD8>
struct PlaneTile {
  static constexpr int R = CD8 ? 4 : 3;
  static constexpr int LX = TXT + 2 * R, LY = TYT + 2 * R;
};

struct TileCols {
  int i0, j0, tx, ty, i, j, kb, ke;
  bool valid;
};
__device__ inline TileCols tile_cols(const Geom &g, int kseg, int b) {
  TileCols t;
  const int ntx = (g.nx + TXT - 1) / TXT, nty = (g.ny + TYT - 1) / TYT;
  const int bx = b % ntx;
  b /= ntx;
  const int by = b % nty;
  const int bz = b / nty;
  t.tx = threadIdx.x % TXT;
  t.ty = threadIdx.x / TXT;
  t.i0 = bx * TXT + 1;
  t.j0 = by * TYT + 1;
  t.i = t.i0 + t.tx;
  t.j = t.j0 + t.ty;
  t.kb = bz * kseg + 1;
  t.ke = t.kb + kseg - 1 < g.nz ? t.kb + kseg - 1 : g.nz;
  t.valid = t.i <= g.nx && t.j <= g.ny;
  return t;
}

// interior index t of a column position that may lie up to R outside
// [1, n]: periodic wrap (positions further out, which only partial tiles
// touch and nothing reads, are clamped first)
__device__ inline int wrap_index(int t, int n, int R) {
  t = t > n + R ? n + R : t;
  return t < 1 ? t + n : (t > n ? t - n : t);
}

// The viscous fluxes Ev1(1:4), Ev2(1:4), Ev3(1:4) of one point
// (OCFD_viscous.f90:42-111, comput_Amu :135-149) from the derivatives of
// u, v, w, T along the three grid lines (dx, dy, dz), the point's u, v, w,
// T and the metrics at halo'd index h (constants when FAST).
template <bool FAST, typename RT>
__device__ __forceinline__ void ev_point(const Block &B, const RT *dx, const RT *dy, const RT *dz, const RT uu,
                                         const RT vv, const RT ww, const RT Tc, const int64_t h, RT *out) {
  const Metrics &M = B.M;
  auto met = [&](int qm) { return RT(FAST ? M.u[qm] : M.get(qm, h)); };
  const RT Tsb = RT(110.4 / B.P.ref_amu_t0), TmpR = RT(1.0 / B.P.re);
  const RT cp = RT(B.P.cp), pr = RT(B.P.pr);
  // comput_Amu (OCFD_viscous.f90:135-149)
  const RT amu = TmpR * (RT(1.0) + Tsb) * sqrt(Tc * Tc * Tc) / (Tsb + Tc);
  // OCFD_viscous.f90:42-53
  const RT uk = dx[0], ui = dy[0], us = dz[0];
  const RT vk = dx[1], vi = dy[1], vs = dz[1];
  const RT wk = dx[2], wi = dy[2], ws = dz[2];
  const RT Tk = dx[3], Ti = dy[3], Ts = dz[3];
  const RT akx = met(M_AKX), aky = met(M_AKY), akz = met(M_AKZ);
  const RT aix = met(M_AIX), aiy = met(M_AIY), aiz = met(M_AIZ);
  const RT asx = met(M_ASX), asy = met(M_ASY), asz = met(M_ASZ);
  // OCFD_viscous.f90:60-111
  const RT ux = uk * akx + ui * aix + us * asx;
  const RT vx = vk * akx + vi * aix + vs * asx;
  const RT wx = wk * akx + wi * aix + ws * asx;
  const RT Tx = Tk * akx + Ti * aix + Ts * asx;
  const RT uy = uk * aky + ui * aiy + us * asy;
  const RT vy = vk * aky + vi * aiy + vs * asy;
  const RT wy = wk * aky + wi * aiy + ws * asy;
  const RT Ty = Tk * aky + Ti * aiy + Ts * asy;
  const RT uz = uk * akz + ui * aiz + us * asz;
  const RT vz = vk * akz + vi * aiz + vs * asz;
  const RT wz = wk * akz + wi * aiz + ws * asz;
  const RT Tz = Tk * akz + Ti * aiz + Ts * asz;
  const RT div = ux + vy + wz;
  const RT Amu1 = amu;
  const RT Amuk = cp * amu / pr;
  const RT tmp2_3 = RT(2.0 / 3.0);
  const RT s11 = (RT(2.0) * ux - tmp2_3 * div) * Amu1;
  const RT s22 = (RT(2.0) * vy - tmp2_3 * div) * Amu1;
  const RT s33 = (RT(2.0) * wz - tmp2_3 * div) * Amu1;
  const RT s12 = (uy + vx) * Amu1, s13 = (uz + wx) * Amu1, s23 = (vz + wy) * Amu1;
  const RT E1 = uu * s11 + vv * s12 + ww * s13 + Amuk * Tx;
  const RT E2 = uu * s12 + vv * s22 + ww * s23 + Amuk * Ty;
  const RT E3 = uu * s13 + vv * s23 + ww * s33 + Amuk * Tz;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const RT n1 = met(M_AKX1 + 3 * a), n2 = met(M_AKX1 + 3 * a + 1), n3 = met(M_AKX1 + 3 * a + 2);
    out[4 * a + 0] = n1 * s11 + n2 * s12 + n3 * s13;
    out[4 * a + 1] = n1 * s12 + n2 * s22 + n3 * s23;
    out[4 * a + 2] = n1 * s13 + n2 * s23 + n3 * s33;
    out[4 * a + 3] = n1 * E1 + n2 * E2 + n3 * E3;
  }
}

// Lanes whose column lies outside the block (partial tiles only) run the
// same code on clamped indices and store into Block::trash, and the tile
// loads and stores clamp their element index, so the plane loop has no
// per-lane branch: with FAST (all axes periodic, uniform metrics) it is one
// basic block, and the loads of plane k+1 issued at its top are waited for
// only after the arithmetic of plane k.
template <bool CD8, bool FAST, typename RT>
__device__ __forceinline__ void visc_flux_tile(const Block &B, const int kseg, const int tile) {
  typedef PlaneTile<CD8> PT;
  constexpr int R = PT::R, LX = PT::LX, LY = PT::LY, NPL = LX * LY, Q = 2 * R + 1;
  constexpr int NRF = (NPL + NT - 1) / NT;
  __shared__ RT sT[2][4 * NPL];  // u, v, w, T of one plane: [4][LY][LX]
  const Geom &g = B.g;
  const TileCols tc = tile_cols(g, kseg, tile);
  // u, v, w, T in the kernel's real type (FP32 fields in the mixed form)
  const RT *F[4] = {B.u.as<RT>(), B.v.as<RT>(), B.w.as<RT>(), B.T.as<RT>()};
  const int64_t sk = g.hstride(2);
  const int ic = tc.i < g.nx ? tc.i : g.nx, jc = tc.j < g.ny ? tc.j : g.ny;

  // tile of one plane, one share per lane and field; positions beyond the
  // halo'd array (partial tiles only) are clamped; in-plane offsets are 32-bit
  int toff[NRF];
#pragma unroll
  for (int r = 0; r < NRF; r++) {
    const int e = threadIdx.x + r * NT < NPL ? threadIdx.x + r * NT : NPL - 1;
    const int ly = e / LX, lx = e - ly * LX;
    const int ii = tc.i0 - R + lx < g.nx + LAP ? tc.i0 - R + lx : g.nx + LAP;
    const int jj = tc.j0 - R + ly < g.ny + LAP ? tc.j0 - R + ly : g.ny + LAP;
    toff[r] = (ii + LAP - 1) + g.sx * (jj + LAP - 1);
  }
  const int64_t plane = (int64_t)g.sx * g.sy;
  RT pre[4][NRF];
  auto fetch_tile = [&](int k) {
    const int64_t base = plane * (k + LAP - 1);
#pragma unroll
    for (int r = 0; r < NRF; r++)
#pragma unroll
      for (int f = 0; f < 4; f++) pre[f][r] = RT(F[f][base + toff[r]]);
  };
  auto store_tile = [&](RT *dst) {
#pragma unroll
    for (int r = 0; r < NRF; r++) {
      // lanes past the tile rewrite its last point with the same value
      const int e = threadIdx.x + r * NT < NPL ? threadIdx.x + r * NT : NPL - 1;
#pragma unroll
      for (int f = 0; f < 4; f++) dst[f * NPL + e] = pre[f][r];
    }
  };

  // own column: u, v, w, T at planes k-R .. k+R
  RT q[4][Q];
  const int64_t hcol = g.hi(ic, jc, 0);
#pragma unroll
  for (int f = 0; f < 4; f++)
#pragma unroll
    for (int d = 0; d < Q; d++) q[f][d] = RT(F[f][hcol + (int64_t)(tc.kb - R + d) * sk]);
  fetch_tile(tc.kb);
  store_tile(sT[0]);
  __syncthreads();

  RT *const ev = reinterpret_cast<RT *>(B.ev);
  RT *const trash = reinterpret_cast<RT *>(B.trash);
  const bool perx = FAST || B.S.iperiodic[0] == 1, pery = FAST || B.S.iperiodic[1] == 1,
             perz = FAST || B.S.iperiodic[2] == 1;
  const int c0 = (tc.ty + R) * LX + tc.tx + R;  // the lane's own point in the tile
  const int64_t es = tc.valid ? g.ni : NT;       // output plane stride (trash for idle lanes)
#pragma unroll 1
  for (int k = tc.kb; k <= tc.ke; k++) {
    // next plane (the last iteration reloads its own; nothing reads it)
    const int kn = k < tc.ke ? k + 1 : k;
    fetch_tile(kn);
    RT qn[4];
#pragma unroll
    for (int f = 0; f < 4; f++) qn[f] = RT(F[f][hcol + (int64_t)(kn + R) * sk]);

    const RT *cur = sT[(k - tc.kb) & 1];
    const int i = ic, j = jc;
    const int64_t h = g.hi(i, j, k);
    RT dx[4], dy[4], dz[4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
      const RT *cf = cur + f * NPL + c0;
      dx[f] = dxyz0_point([&](int d) { return cf[d]; }, i, g.nx, B.P.dc[0], CD8, perx);
      dy[f] = dxyz0_point([&](int d) { return cf[d * LX]; }, j, g.ny, B.P.dc[1], CD8, pery);
      dz[f] = dxyz0_point([&](int d) { return q[f][R + d]; }, k, g.nz, B.P.dc[2], CD8, perz);
      // one field's stencil values in flight at a time
      if (FAST) __builtin_amdgcn_sched_barrier(0);
    }
    RT e12[12];
    ev_point<FAST, RT>(B, dx, dy, dz, q[0][R], q[1][R], q[2][R], q[3][R], h, e12);
    // Ev is stored non-temporally: the L2 keeps the primitive planes that
    // the neighbouring tiles read as their halos
    RT *eb = tc.valid ? ev + g.ii(i, j, k) : trash + threadIdx.x;
#pragma unroll
    for (int c = 0; c < 12; c++) __builtin_nontemporal_store(e12[c], &eb[(int64_t)c * es]);
    // plane k+1 into the other buffer (its last readers, plane k-1, are past
    // the previous barrier); the queue moves up one plane
    store_tile(sT[(k + 1 - tc.kb) & 1]);
#pragma unroll
    for (int f = 0; f < 4; f++) {
#pragma unroll
      for (int d = 0; d < Q - 1; d++) q[f][d] = q[f][d + 1];
      q[f][Q - 1] = qn[f];
    }
    __syncthreads();
  }
}

template <bool CD8, bool FAST, typename RT>
__global__ __launch_bounds__(NT) void k_visc_flux(const Block B, const int kseg) {
  visc_flux_tile<CD8, FAST, RT>(B, kseg, xcd_tile(blockIdx.x, gridDim.x));
}

// viscous divergence of component m (1..4) at cell t (OCFD_viscous.f90:114-129)
template <typename RT>
__device__ inline RT visc_div_m(const Block &B, int m, int i, int j, int k) {
  const int64_t ni = B.g.ni;
  const RT *ev = reinterpret_cast<const RT *>(B.ev);
  const RT *e1 = ev + (int64_t)(m - 1) * ni;
  const RT *e2 = ev + (int64_t)(4 + m - 1) * ni;
  const RT *e3 = ev + (int64_t)(8 + m - 1) * ni;
  return di(B, e1, 0, i, j, k) + di(B, e2, 1, i, j, k) + di(B, e3, 2, i, j, k);
}

template <typename RT>
__global__ __launch_bounds__(NT) void k_visc_div(const Block B) {
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int i, j, k;
  cell_of(g, t, i, j, k);
  const double ajac = B.M.get(M_AJAC, g.hi(i, j, k));
#pragma unroll
  for (int m = 1; m <= 4; m++) {
    const RT s = visc_div_m<RT>(B, m, i, j, k);
    RT *du = reinterpret_cast<RT *>(B.du) + t + (int64_t)m * g.ni;
    du[0] = RT(double(du[0]) + double(s) * ajac);
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
// f from primitives on the boundary planes (OCFD_BC.f90:27-37)
__device__ inline void refresh_of(double cv, double d, double u, double v, double w, double T, double *fq) {
  fq[0] = d;
  fq[1] = d * u;
  fq[2] = d * v;
  fq[3] = d * w;
  fq[4] = d * (cv * T + (u * u + v * v + w * w) * 0.5);
}

// Fused: du(2:5) += div Ev * Ajac; RK3 stage; primitives; plane refresh.
// krk == 1 reads f as fn (fn = f at the step start) and stores fn.
// LDS per plane: Ev1(1:4) with the x neighbours [4][TYT][LX] and Ev2(1:4)
// with the y neighbours [4][LY][TXT]; Ev3(1:4) of the own column is the
// register queue.  Ev arrays have no halo: neighbours wrap by index
// (periodic), out-of-block planes of a non-periodic z are clamped; the
// one-sided closures never read either.  Cells with T <= 0 are collected
// per lane in a bit mask over the segment and listed after the loop.  The
// loop body is branch-free in the same way as k_visc_flux.  The kernel runs
// one wave per SIMD (its register footprint), so the loads of plane k+1 --
// du, f, fn of the own cell and the Ev tiles -- are all issued at the top of
// iteration k to keep enough bytes in flight.
template <bool VISC, bool CD8, bool FAST, typename RT>
__device__ __forceinline__ void visc_div_update_tile(const Block &B, const int krk, const int refresh, const int kseg,
                                                     int *neg, const int tile) {
  typedef PlaneTile<CD8> PT;
  constexpr int R = PT::R, LX = PT::LX, LY = PT::LY, Q = 2 * R + 1;
  constexpr int N1 = LX * TYT, N2 = TXT * LY;       // points per component
  constexpr int NE = VISC ? 4 * (N1 + N2) : 1;
  constexpr int NR1 = VISC ? (N1 + NT - 1) / NT : 1, NR2 = VISC ? (N2 + NT - 1) / NT : 1;
  __shared__ RT sE[VISC ? 2 : 1][NE];
  const RT *const ev = reinterpret_cast<const RT *>(B.ev);
  const RT *const dus = reinterpret_cast<const RT *>(B.du);  // du in the right-hand side's real type
  const Geom &g = B.g;
  const TileCols tc = tile_cols(g, kseg, tile);
  const int nz = g.nz;
  const int64_t ni = g.ni, skz = g.istride(2);
  const bool perx = FAST || B.S.iperiodic[0] == 1, pery = FAST || B.S.iperiodic[1] == 1,
             perz = FAST || B.S.iperiodic[2] == 1;
  const double dt = B.P.dt;
  const int ic = tc.i < g.nx ? tc.i : g.nx, jc = tc.j < g.ny ? tc.j : g.ny;
  auto kwrap = [&](int t) {
    return perz ? (t < 1 ? t + nz : (t > nz ? t - nz : t)) : (t < 1 ? 1 : (t > nz ? nz : t));
  };
  if (halted(B)) {
    // a whole-step run stopped at too many T <= 0 cells: the state stays, and
    // T carries over into the buffer the host swaps in after the stage
    if (tc.valid)
      for (int k = tc.kb; k <= tc.ke; k++) {
        const int64_t h = g.hi(ic, jc, k);
        B.T2.as<RT>()[h] = B.T.as<RT>()[h];
      }
    return;
  }

  // in-plane offsets of this lane's shares of the Ev1 / Ev2 tiles
  int o1[NR1], o2[NR2];
  if (VISC) {
#pragma unroll
    for (int r = 0; r < NR1; r++) {
      const int e = threadIdx.x + r * NT < N1 ? threadIdx.x + r * NT : N1 - 1;
      const int ly = e / LX, lx = e - ly * LX;
      const int ii = wrap_index(tc.i0 - R + lx, g.nx, R);
      const int jj = tc.j0 + ly < g.ny ? tc.j0 + ly : g.ny;
      o1[r] = (ii - 1) + g.nx * (jj - 1);
    }
#pragma unroll
    for (int r = 0; r < NR2; r++) {
      const int e = threadIdx.x + r * NT < N2 ? threadIdx.x + r * NT : N2 - 1;
      const int ly = e / TXT, lx = e - ly * TXT;
      const int ii = tc.i0 + lx < g.nx ? tc.i0 + lx : g.nx;
      const int jj = wrap_index(tc.j0 - R + ly, g.ny, R);
      o2[r] = (ii - 1) + g.nx * (jj - 1);
    }
  }
  RT pre1[4][NR1], pre2[4][NR2];
  auto fetch_tile = [&](int k) {
    const RT *p1 = ev + skz * (k - 1), *p2 = ev + 4 * ni + skz * (k - 1);
#pragma unroll
    for (int r = 0; r < NR1; r++)
#pragma unroll
      for (int f = 0; f < 4; f++) pre1[f][r] = p1[f * ni + o1[r]];
#pragma unroll
    for (int r = 0; r < NR2; r++)
#pragma unroll
      for (int f = 0; f < 4; f++) pre2[f][r] = p2[f * ni + o2[r]];
  };
  auto store_tile = [&](RT *dst) {
#pragma unroll
    for (int r = 0; r < NR1; r++) {
      const int e = threadIdx.x + r * NT < N1 ? threadIdx.x + r * NT : N1 - 1;
#pragma unroll
      for (int f = 0; f < 4; f++) dst[f * N1 + e] = pre1[f][r];
    }
#pragma unroll
    for (int r = 0; r < NR2; r++) {
      const int e = threadIdx.x + r * NT < N2 ? threadIdx.x + r * NT : N2 - 1;
#pragma unroll
      for (int f = 0; f < 4; f++) dst[4 * N1 + f * N2 + e] = pre2[f][r];
    }
  };

  const RT *e3 = ev + (int64_t)8 * ni;
  const int64_t col = g.ii(ic, jc, 1);
  RT q[4][Q];  // Ev3(1:4) at planes k-R .. k+R
  if (VISC) {
#pragma unroll
    for (int d = 0; d < Q; d++) {
      const int64_t o = col + (int64_t)(kwrap(tc.kb + d - R) - 1) * skz;
#pragma unroll
      for (int m = 0; m < 4; m++) q[m][d] = e3[(int64_t)m * ni + o];
    }
    fetch_tile(tc.kb);
    store_tile(sE[0]);
    __syncthreads();
  }

  const int cx = tc.ty * LX + tc.tx + R;              // own point in the Ev1 tile
  const int cy = 4 * N1 + (tc.ty + R) * TXT + tc.tx;  // own point in the Ev2 tile
  const int i = ic, j = jc;
  const int64_t os = tc.valid ? ni : NT;  // output plane stride (trash for idle lanes)
  // du, f, fn of the lane's cell at plane k, fetched one plane ahead (a lane
  // writes f, fn only at its own cell of the current plane).  Data read or
  // written once per stage -- du, f, fn, the primitives -- moves with
  // non-temporal loads and stores, which leaves the L2 to the Ev tiles the
  // neighbouring tiles share.
  struct CellVals {
    double du[5], f[5], fn[5];
  };
  auto fetch_cell = [&](int k) {
    CellVals c;
    const int64_t t = col + (int64_t)(k - 1) * skz;
#pragma unroll
    for (int m = 0; m < 5; m++) {
      c.du[m] = double(__builtin_nontemporal_load(&dus[t + m * ni]));
      c.f[m] = __builtin_nontemporal_load(&B.f[t + m * ni]);
      c.fn[m] = krk != 1 ? __builtin_nontemporal_load(&B.fn[t + m * ni]) : 0.0;
    }
    return c;
  };
  CellVals cv = fetch_cell(tc.kb);
  unsigned long long negmask = 0;  // bit k - kb: T <= 0 at plane k
#pragma unroll 1
  for (int k = tc.kb; k <= tc.ke; k++) {
    const int64_t t = col + (int64_t)(k - 1) * skz;
    const int64_t h = g.hi(i, j, k);
    // the next plane's cell values and Ev tiles (the last iteration reloads
    // its own plane; nothing reads it)
    const int kn = k < tc.ke ? k + 1 : k;
    const CellVals cn = fetch_cell(kn);
    const double *du = cv.du, *f0 = cv.f, *fn0 = cv.fn;
    RT qn[4];
    if (VISC) {
      fetch_tile(kn);
      const int64_t o = col + (int64_t)(kwrap(kn + R) - 1) * skz;
#pragma unroll
      for (int m = 0; m < 4; m++) qn[m] = e3[(int64_t)m * ni + o];
    }
    const double ajac = FAST ? B.M.u[M_AJAC] : B.M.get(M_AJAC, h);
    RT vd[5] = {RT(0), RT(0), RT(0), RT(0), RT(0)};
    if (VISC) {
      const RT *cur = sE[(k - tc.kb) & 1];
#pragma unroll
      for (int m = 1; m <= 4; m++) {
        const RT *c1 = cur + (m - 1) * N1 + cx;
        const RT *c2 = cur + (m - 1) * N2 + cy;
        const RT d1 = dxyz0_point([&](int d) { return c1[d]; }, i, g.nx, B.P.dc[0], CD8, perx);
        const RT d2 = dxyz0_point([&](int d) { return c2[d * TXT]; }, j, g.ny, B.P.dc[1], CD8, pery);
        const RT d3 = dxyz0_point([&](int d) { return q[m - 1][R + d]; }, k, nz, B.P.dc[2], CD8, perz);
        vd[m] = d1 + d2 + d3;
        if (FAST) __builtin_amdgcn_sched_barrier(0);  // one component's stencil at a time
      }
    }
    double *fo = tc.valid ? B.f + t : B.trash + threadIdx.x;
    double *fno = tc.valid ? B.fn + t : B.trash + 5 * NT + threadIdx.x;
    double fq[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      double dum = du[m];
      if (VISC && m > 0) dum = dum + double(vd[m]) * ajac;
      const double f = f0[m];
      double r;
      if (krk == 1) {
        __builtin_nontemporal_store(f, &fno[m * os]);
        r = f + dt * dum;
      } else {
        const double ra = krk == 2 ? 3.0 / 4.0 : 1.0 / 3.0;
        const double rb = krk == 2 ? 1.0 / 4.0 : 2.0 / 3.0;
        r = ra * fn0[m] + rb * (f + dt * dum);
      }
      fq[m] = r;
    }
    double d, u, v, w, T;
    prims_of(B.P.cv, fq, d, u, v, w, T);
    {
      // the primitives are stored in the kernel's real type
      const int64_t ph = tc.valid ? h : 0;
      RT *pt = reinterpret_cast<RT *>(B.trash + 10 * NT) + threadIdx.x;
      __builtin_nontemporal_store(RT(d), tc.valid ? B.d.as<RT>() + ph : pt);
      __builtin_nontemporal_store(RT(u), tc.valid ? B.u.as<RT>() + ph : pt + NT);
      __builtin_nontemporal_store(RT(v), tc.valid ? B.v.as<RT>() + ph : pt + 2 * NT);
      __builtin_nontemporal_store(RT(w), tc.valid ? B.w.as<RT>() + ph : pt + 3 * NT);
      // T into T2: the T <= 0 repair reads this stage's and the previous
      // stage's values (Block::T2)
      __builtin_nontemporal_store(RT(T), tc.valid ? B.T2.as<RT>() + ph : pt + 4 * NT);
    }
    negmask |= (tc.valid && T <= 0) ? (1ull << (k - tc.kb)) : 0ull;
    if (refresh && (i == 1 || i == g.nx || j == 1 || j == g.ny || k == 1 || k == g.nz))
      refresh_of(B.P.cv, d, u, v, w, T, fq);
#pragma unroll
    for (int m = 0; m < 5; m++) __builtin_nontemporal_store(fq[m], &fo[m * os]);
    cv = cn;
    if (VISC) {
      store_tile(sE[(k + 1 - tc.kb) & 1]);
#pragma unroll
      for (int m = 0; m < 4; m++) {
#pragma unroll
        for (int dd = 0; dd < Q - 1; dd++) q[m][dd] = q[m][dd + 1];
        q[m][Q - 1] = qn[m];
      }
      __syncthreads();
    }
  }
  // T <= 0 cells of this column, in k order
  while (negmask) {
    const int kk = __builtin_ctzll(negmask);
    negmask &= negmask - 1;
    const int slot = atomicAdd(neg, 1);
    if (slot < 64) neg[1 + slot] = (int)(col + (int64_t)(tc.kb + kk - 1) * skz);
  }
}

template <bool VISC, bool CD8, bool FAST, typename RT>
__global__ __launch_bounds__(NT) void k_visc_div_update(const Block B, const int krk, const int refresh,
                                                        const int kseg, int *neg) {
  visc_div_update_tile<VISC, CD8, FAST, RT>(B, krk, refresh, kseg, neg, xcd_tile(blockIdx.x, gridDim.x));
}

// Batches of periodic boxes with uniform metrics (ocfd_batch): the tiles of
// all boxes of the launch in XCD order; box b marches segments of B.kseg
// planes (at most nz) and counts its T <= 0 cells in neg + NEG_STRIDE * b.
__device__ inline int batch_kseg_dev(const Block &B) { return B.g.nz < B.kseg ? B.g.nz : B.kseg; }

template <bool CD8, typename RT>
__global__ __launch_bounds__(NT) void k_visc_flux_batch(const BoxSet S) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int j = S.entry(t);
  // copied before any store, so its fields load once into scalar registers
  // (read in place after the loop's stores they would take vector registers)
  const Block B = S.blk[S.ids[j]];
  visc_flux_tile<CD8, true, RT>(B, batch_kseg_dev(B), t - S.off[j]);
}

template <bool VISC, bool CD8, typename RT>
__global__ __launch_bounds__(NT) void k_visc_div_update_batch(const BoxSet S, const int krk, int *neg) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int j = S.entry(t);
  const int box = S.ids[j];
  const Block B = S.blk[box];  // copied before any store, as above
  visc_div_update_tile<VISC, CD8, true, RT>(B, krk, 1, batch_kseg_dev(B), neg + NEG_STRIDE * box, t - S.off[j]);
}

// ---------------------------------------------------------------------------
// launch shapes and per-precision launchers (RT = double, or float for the
// mixed-precision form; each precision is instantiated in its own
// translation unit)
// ---------------------------------------------------------------------------
inline unsigned nblocks(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

// planes per workgroup: the longest segment that still gives about 2048
// workgroups, at most 32 planes, for the FP32 kernels; about 1024 workgroups
// of at most 64 planes for the FP64 ones, which run fewer waves per SIMD and
// gain from amortising the prologue of their plane queue over a longer march
template <typename RT>
inline int march_kseg(const Geom &g) {
  const int64_t cols = (int64_t)((g.nx + TXT - 1) / TXT) * ((g.ny + TYT - 1) / TYT);
  const int64_t k = cols * g.nz / (sizeof(RT) == 8 ? 1024 : 2048);
  const int kmax = sizeof(RT) == 8 ? KSEG_MAX : 32;
  return (int)(k < 1 ? 1 : (k > kmax ? kmax : k));
}
inline unsigned march_blocks(const Geom &g, int kseg) {
  return (unsigned)(((g.nx + TXT - 1) / TXT) * ((g.ny + TYT - 1) / TYT) * ((g.nz + kseg - 1) / kseg));
}

// the specialised kernels: every axis periodic and uniform metrics
inline bool all_periodic(const Block &B) {
  return B.S.iperiodic[0] == 1 && B.S.iperiodic[1] == 1 && B.S.iperiodic[2] == 1 && B.M.uniform;
}

template <typename RT>
hipError_t visc_flux_as(const Block &B, hipStream_t st) {
  const int kseg = march_kseg<RT>(B.g);
  const dim3 grid(march_blocks(B.g, kseg));
  const bool cd8 = B.S.scheme_vis == SCH_CD8, per = all_periodic(B);
  if (cd8 && per) hipLaunchKernelGGL((k_visc_flux<true, true, RT>), grid, dim3(NT), 0, st, B, kseg);
  else if (cd8) hipLaunchKernelGGL((k_visc_flux<true, false, RT>), grid, dim3(NT), 0, st, B, kseg);
  else if (per) hipLaunchKernelGGL((k_visc_flux<false, true, RT>), grid, dim3(NT), 0, st, B, kseg);
  else hipLaunchKernelGGL((k_visc_flux<false, false, RT>), grid, dim3(NT), 0, st, B, kseg);
  return hipGetLastError();
}

template <typename RT>
hipError_t visc_div_as(const Block &B, hipStream_t st) {
  hipLaunchKernelGGL(k_visc_div<RT>, dim3(nblocks(B.g.ni)), dim3(NT), 0, st, B);
  return hipGetLastError();
}

template <typename RT>
hipError_t visc_div_update_as(const Block &B, int krk, int refresh_planes, int *neg, hipStream_t st) {
  const int kseg = march_kseg<RT>(B.g);
  const dim3 grid(march_blocks(B.g, kseg));
  const bool cd8 = B.S.scheme_vis == SCH_CD8, per = all_periodic(B);
#define OCFD_DIVUPD(V, C, P) \
  hipLaunchKernelGGL((k_visc_div_update<V, C, P, RT>), grid, dim3(NT), 0, st, B, krk, refresh_planes, kseg, neg)
  if (B.ev == nullptr) OCFD_DIVUPD(false, false, false);
  else if (cd8 && per) OCFD_DIVUPD(true, true, true);
  else if (cd8) OCFD_DIVUPD(true, true, false);
  else if (per) OCFD_DIVUPD(true, false, true);
  else OCFD_DIVUPD(true, false, false);
#undef OCFD_DIVUPD
  return hipGetLastError();
}

template <typename RT>
hipError_t visc_flux_batch_as(const BoxSet &S, bool cd8, int ntiles, hipStream_t st) {
  if (cd8) hipLaunchKernelGGL((k_visc_flux_batch<true, RT>), dim3((unsigned)ntiles), dim3(NT), 0, st, S);
  else hipLaunchKernelGGL((k_visc_flux_batch<false, RT>), dim3((unsigned)ntiles), dim3(NT), 0, st, S);
  return hipGetLastError();
}

template <typename RT>
hipError_t visc_div_update_batch_as(const BoxSet &S, bool visc, bool cd8, int krk, int *neg, int ntiles,
                                    hipStream_t st) {
#define OCFD_DIVUPD_B(V, C) \
  hipLaunchKernelGGL((k_visc_div_update_batch<V, C, RT>), dim3((unsigned)ntiles), dim3(NT), 0, st, S, krk, neg)
  if (!visc) OCFD_DIVUPD_B(false, false);
  else if (cd8) OCFD_DIVUPD_B(true, true);
  else OCFD_DIVUPD_B(true, false);
#undef OCFD_DIVUPD_B
  return hipGetLastError();
}

}  // namespace

}  // namespace ocfd
