// ocfd_inviscid.hip -- inviscid right-hand side: one pencil sweep per axis.
//
// Replaces du_inviscous (OCFD_inviscous.f90:53-203) and du_inviscous_Character
// (OCFD_inviscous_character.f90:3-159).  One launch per axis; a workgroup owns
// a tile of TL cells along the sweep axis on TW neighbouring lines:
//   phase 1  every point of the tile plus its 4-point halo on both sides:
//            load d,u,v,w,T (+ normalised metrics), sound speed, flux-vector
//            splitting -> fp(5), fm(5) in LDS (coalesced global loads);
//   phase 2  every face of the tile: reconstruction (WENO5/WENO7/OMP6/UD7L/
//            hybrid, characteristic projection when requested, block-edge
//            closures) -> face fluxes in LDS;
//   phase 3  every cell: flux difference / h written to du (x: assign,
//            y: accumulate, z: accumulate, times -Ajac, plus mass force).
// Axis x tiles run along i (contiguous); y and z tiles put TW consecutive i on
// the lines so every global access is a contiguous run of i.
#include "ocfd_kernels.h"

namespace ocfd {

namespace {

constexpr int NTHREADS = 256;

template <int AXIS>
struct TileIndex {
  // halo'd index of point s on line w, third index c
  __device__ static inline int64_t hidx(const Geom &g, int s, int w, int c) {
    if (AXIS == 0) return g.hi(s, w, c);
    if (AXIS == 1) return g.hi(w, s, c);
    return g.hi(w, c, s);
  }
  __device__ static inline int64_t iidx(const Geom &g, int s, int w, int c) {
    if (AXIS == 0) return g.ii(s, w, c);
    if (AXIS == 1) return g.ii(w, s, c);
    return g.ii(w, c, s);
  }
  __device__ static inline int nwidth(const Geom &g) { return AXIS == 0 ? g.ny : g.nx; }
  __device__ static inline int nthird(const Geom &g) { return AXIS == 2 ? g.ny : g.nz; }
};

// LDS element (component m, point or face p, line wl) of a tile array with NPT
// entries per line: x tiles keep the points of one line contiguous, y and z
// tiles the lines of one point, so that consecutive lanes touch consecutive
// doubles in every phase.
template <int AXIS, int NPT, int TW>
__device__ constexpr inline int lidx(int m, int p, int wl) {
  return AXIS == 0 ? (m * TW + wl) * NPT + p : (m * NPT + p) * TW + wl;
}

template <int AXIS, int SCH, bool CHAR, int TL, int TW>
__global__ __launch_bounds__(NTHREADS) void k_inviscid(const Block B, const int mode) {
  typedef TileIndex<AXIS> TI;
  constexpr int NP = TL + 8;  // points s0-4 .. s0+TL+3
  constexpr int NF = TL + 1;  // faces  s0-1 .. s0+TL-1
  constexpr bool HYB = (SCH == SCH_HYBRID);
  constexpr int NH = CHAR ? 5 : 10;
  constexpr int NI1 = (NP * TW + NTHREADS - 1) / NTHREADS;      // points per thread
  constexpr int NI3 = (5 * TL * TW + NTHREADS - 1) / NTHREADS;  // du entries per thread
  // characteristic inputs per point: u, v, w, c (+ the three normalised
  // metrics when they vary in space)
  const bool unif = B.M.uniform != 0;
  const int nq = unif ? 4 : 7;
  extern __shared__ double smem[];
  // hybrid scheme: WENO7 faces listed from the front, WENO5 from the back
  constexpr int NITEM = CHAR ? NF * TW : 5 * NF * TW;  // list capacity
  __shared__ int s_cnt[2];

  const Geom &g = B.g;
  const int n = g.n(AXIS);
  const int nw = TI::nwidth(g), nc = TI::nthird(g);
  const int nts = (n + TL - 1) / TL, ntw = (nw + TW - 1) / TW;
  int b = blockIdx.x;
  const int ts = b % nts; b /= nts;
  const int tw = b % ntw;
  const int c = b / ntw + 1;
  const int s0 = 1 + ts * TL, w0 = 1 + tw * TW;
  const int tid = threadIdx.x;
  const double gamma = B.P.gamma, rma = B.P.ma;
  const int q1 = M_AKX1 + 3 * AXIS;
  const SchemeParams &S = B.S;
  const int lo = S.bidx[AXIS][0], hi = S.bidx[AXIS][1];
  const int sl = S.sb[2 * AXIS], sr = S.sb[2 * AXIS + 1];

  // point handled by phase-1 slot r of this thread
  auto p1 = [&](int r, int &p, int &wl) {
    const int it = tid + r * NTHREADS;
    if (AXIS == 0) { p = it % NP; wl = it / NP; } else { wl = it % TW; p = it / TW; }
    return it < NP * TW && s0 - 4 + p <= n + LAP && w0 + wl <= nw;
  };
  // du entry handled by phase-3 slot r
  auto p3 = [&](int r, int &sl_, int &wl, int &m) {
    const int it = tid + r * NTHREADS;
    if (AXIS == 0) { sl_ = it % TL; wl = (it / TL) % TW; m = it / (TL * TW); }
    else { wl = it % TW; sl_ = (it / TW) % TL; m = it / (TL * TW); }
    return it < 5 * TL * TW && s0 + sl_ <= n && w0 + wl <= nw;
  };

  // primitives (+ sensor) of the tile's points
  double pq[NI1][5], pr[NI1];
  {
#pragma unroll
    for (int r = 0; r < NI1; r++) {
      int p, wl;
      if (!p1(r, p, wl)) continue;
      const int64_t h = TI::hidx(g, s0 - 4 + p, w0 + wl, c);
      pq[r][0] = B.d[h];
      pq[r][1] = B.u[h];
      pq[r][2] = B.v[h];
      pq[r][3] = B.w[h];
      pq[r][4] = B.T[h];
      if (HYB) pr[r] = B.rhyb[h];
    }
  }
  // du entries of the tile (accumulating modes), fetched before the face
  // fluxes are computed
  double pdu[NI3];

  double *sF = smem;                               // fp(5), fm(5)      [10][NP][TW]
  double *sQ = sF + 10 * NP * TW;                  // char inputs       [nq][NP][TW]
  double *sR = sQ + (CHAR ? nq * NP * TW : 0);     // hybrid sensor     [NP][TW]
  double *sH = sR + (HYB ? NP * TW : 0);           // face fluxes       [NH][NF][TW]
  int *sL = reinterpret_cast<int *>(sH + NH * NF * TW);  // [NITEM] (hybrid only)
  if (HYB && tid < 2) s_cnt[tid] = 0;
  // ---- phase 1: split fluxes at the tile's points ----
#pragma unroll
  for (int r = 0; r < NI1; r++) {
    int p, wl;
    if (!p1(r, p, wl)) continue;
    const double dd = pq[r][0], uu = pq[r][1], vv = pq[r][2], ww = pq[r][3];
    const double cc = sqrt(pq[r][4]) / rma;  // OCFD_inviscous.f90:67
    double A1, A2, A3;
    if (unif) {
      A1 = B.M.u[q1]; A2 = B.M.u[q1 + 1]; A3 = B.M.u[q1 + 2];
    } else {
      const int64_t h = TI::hidx(g, s0 - 4 + p, w0 + wl, c);
      A1 = B.M.p[q1][h]; A2 = B.M.p[q1 + 1][h]; A3 = B.M.p[q1 + 2][h];
    }
    double fp[5], fm[5];
    if (S.flux_split == 1) split_sw(gamma, cc, dd, uu, vv, ww, A1, A2, A3, fp, fm);
    else split_llf(gamma, cc, dd, uu, vv, ww, A1, A2, A3, fp, fm);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      sF[lidx<AXIS, NP, TW>(m, p, wl)] = fp[m];
      sF[lidx<AXIS, NP, TW>(5 + m, p, wl)] = fm[m];
    }
    if (CHAR) {
      sQ[lidx<AXIS, NP, TW>(0, p, wl)] = uu;
      sQ[lidx<AXIS, NP, TW>(1, p, wl)] = vv;
      sQ[lidx<AXIS, NP, TW>(2, p, wl)] = ww;
      sQ[lidx<AXIS, NP, TW>(3, p, wl)] = cc;
      if (!unif) {
        sQ[lidx<AXIS, NP, TW>(4, p, wl)] = A1;
        sQ[lidx<AXIS, NP, TW>(5, p, wl)] = A2;
        sQ[lidx<AXIS, NP, TW>(6, p, wl)] = A3;
      }
    }
    if (HYB) sR[lidx<AXIS, NP, TW>(0, p, wl)] = pr[r];
  }
  __syncthreads();

  if (mode != INV_ASSIGN) {
#pragma unroll
    for (int r = 0; r < NI3; r++) {
      int sl_, wl, m;
      if (!p3(r, sl_, wl, m)) continue;
      pdu[r] = B.du[TI::iidx(g, s0 + sl_, w0 + wl, c) + (int64_t)m * g.ni];
    }
  }

  // ---- phase 2: face fluxes ----
  // flux+/- of component m at face f (non-characteristic)
  auto face_nc = [&](int m, int f, int wl, int s, int scm) {
    double xp[8], xm[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      xp[k] = sF[lidx<AXIS, NP, TW>(m, f + k, wl)];
      xm[k] = sF[lidx<AXIS, NP, TW>(5 + m, f + k, wl)];
    }
    sH[lidx<AXIS, NF, TW>(m, f, wl)] = face_p_nonchar<SCH>(S, xp, s, n, AXIS, scm);
    sH[lidx<AXIS, NF, TW>(5 + m, f, wl)] = face_m_nonchar<SCH>(S, xm, s, n, AXIS, scm);
  };
#define FPV(m, k) sF[lidx<AXIS, NP, TW>((m), f + (k), wl)]
#define FMV(m, k) sF[lidx<AXIS, NP, TW>(5 + (m), f + (k), wl)]
  // all five components of face f, projected on the characteristic basis of
  // the face unless the face is linear (flux_charteric,
  // OCFD_inviscous_character.f90:190-290)
  // characteristic basis of face f from the face-averaged state
  // (OCFD_inviscous_character.f90:200-236)
  auto basis_of = [&](int f, int wl, CharBasis &Bs) {
#define QV(q, k) sQ[lidx<AXIS, NP, TW>((q), f + (k), wl)]
    const double u1 = (QV(0, 3) + QV(0, 4)) * 0.5, v1 = (QV(1, 3) + QV(1, 4)) * 0.5;
    const double w1 = (QV(2, 3) + QV(2, 4)) * 0.5, c1 = (QV(3, 3) + QV(3, 4)) * 0.5;
    double a1, a2, a3;
    if (unif) {
      // average of two equal values, as the reference computes it
      a1 = (B.M.u[q1] + B.M.u[q1]) * 0.5;
      a2 = (B.M.u[q1 + 1] + B.M.u[q1 + 1]) * 0.5;
      a3 = (B.M.u[q1 + 2] + B.M.u[q1 + 2]) * 0.5;
    } else {
      a1 = (QV(4, 3) + QV(4, 4)) * 0.5;
      a2 = (QV(5, 3) + QV(5, 4)) * 0.5;
      a3 = (QV(6, 3) + QV(6, 4)) * 0.5;
    }
#undef QV
    char_basis(gamma, u1, v1, w1, c1, a1, a2, a3, Bs);
  };
  // characteristic flux h0(r) = h+(r) + h-(r) of face f: row r of S applied
  // to the split fluxes on the scheme's stencil, then reconstructed
  // (OCFD_inviscous_character.f90:250-282)
  auto char_row = [&](const CharBasis &Bs, int f, int wl, int s, int scm, int r) {
    int a1h = S.ka1, b1h = S.kb1, a2h = S.ka2, b2h = S.kb2;
    if (HYB) {
      if (scm == 2) { a1h = -3; b1h = 3; a2h = -2; b2h = 4; }
      else if (scm == 3) { a1h = -2; b1h = 2; a2h = -1; b2h = 3; }
    }
    double xp[8], xm[8], sr5[5];
    char_S_row(Bs, r, sr5);
    // xp(k) = sum over components c of S(r,c) * fp(c,k), summed in the
    // reference's order c = 1..5; outside the scheme's stencil the value is 0
#pragma unroll
    for (int k = 0; k < 8; k++) {
      xp[k] = sr5[0] * FPV(0, k);
      xm[k] = sr5[0] * FMV(0, k);
    }
#pragma unroll
    for (int c = 1; c < 5; c++) {
      // keeps the scheduler from fetching all 80 stencil values at once
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        xp[k] = xp[k] + sr5[c] * FPV(c, k);
        xm[k] = xm[k] + sr5[c] * FMV(c, k);
      }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (k < a1h + 3 || k > b1h + 3) xp[k] = 0.0;
      if (k < a2h + 3 || k > b2h + 3) xm[k] = 0.0;
    }
    const double h1 = face_p_char<SCH>(S, xp, s, n, AXIS, scm);
    const double h2 = face_m_char<SCH>(S, xm, s, n, AXIS, scm);
    return h1 + h2;
  };
  // component m of a linear (hybrid UD7L) face: no projection
  auto linear_comp = [&](int f, int wl, int s, int m) {
    double xp[8], xm[8];
#pragma unroll
    for (int k = 0; k < 8; k++) { xp[k] = FPV(m, k); xm[k] = FMV(m, k); }
    const double h1 = face_p_char<SCH>(S, xp, s, n, AXIS, 1);
    const double h2 = face_m_char<SCH>(S, xm, s, n, AXIS, 1);
    return h1 + h2;
  };
  // non-reflecting closure of component m (OCFD_inviscous_character.f90:295-309);
  // point t is k = t - s + 3
  auto closure_comp = [&](int f, int wl, int s, int m) {
    if (s == 0)
      return (2.0 * FPV(m, 4) + 5.0 * FPV(m, 5) - FPV(m, 6) + 2.0 * FMV(m, 6) - 7.0 * FMV(m, 5) +
              11.0 * FMV(m, 4)) / 6.0;
    return (2.0 * FPV(m, 1) - 7.0 * FPV(m, 2) + 11.0 * FPV(m, 3) + 2.0 * FMV(m, 3) + 5.0 * FMV(m, 2) -
            FMV(m, 1)) / 6.0;
  };
  // all five components of face f, projected on the characteristic basis of
  // the face (flux_charteric, OCFD_inviscous_character.f90:190-290)
  auto face_ch = [&](int f, int wl, int s, int scm) {
    CharBasis Bs;
    basis_of(f, wl, Bs);
    double h0[5], hh[5];
    for (int r = 0; r < 5; r++) h0[r] = char_row(Bs, f, wl, s, scm, r);
    char_back(Bs, h0, hh);
#pragma unroll
    for (int m = 0; m < 5; m++) sH[lidx<AXIS, NF, TW>(m, f, wl)] = hh[m];
  };
  auto push = [&](int it, int scm) {
    if (scm == 2) sL[atomicAdd(&s_cnt[0], 1)] = it;
    else sL[NITEM - 1 - atomicAdd(&s_cnt[1], 1)] = it;
  };
  auto scheme_of = [&](int f, int wl) {
    return hybrid_index(dmax2(sR[lidx<AXIS, NP, TW>(0, f + 3, wl)], sR[lidx<AXIS, NP, TW>(0, f + 4, wl)]),
                        S.hyb2, S.hyb3);
  };
  // face work item -> (component, face, line)
  auto item_nc = [&](int it, int &m, int &f, int &wl) {
    if (AXIS == 0) { f = it % NF; wl = (it / NF) % TW; } else { wl = it % TW; f = (it / TW) % NF; }
    m = it / (TW * NF);
  };
  auto item_ch = [&](int it, int &f, int &wl) {
    if (AXIS == 0) { f = it % NF; wl = it / NF; } else { wl = it % TW; f = it / TW; }
  };
  const bool closure_lo = lo == 1 && sl == 0, closure_hi = hi == 1 && sr == 0;
  if (!CHAR) {
    for (int it = tid; it < 5 * NF * TW; it += NTHREADS) {
      int m, f, wl;
      item_nc(it, m, f, wl);
      const int s = s0 - 1 + f;
      if (s > n || w0 + wl > nw) continue;
      int scm = 0;
      if (HYB) scm = scheme_of(f, wl);
      if (!HYB || scm == 1) face_nc(m, f, wl, s, scm);
      else push(it, scm);
    }
    if (HYB) {
      // second pass: the listed WENO7 faces, then the WENO5 faces, spread
      // over all threads
      __syncthreads();
      const int n7 = s_cnt[0], n5 = s_cnt[1];
      for (int q = tid; q < n7 + n5; q += NTHREADS) {
        const int it = q < n7 ? sL[q] : sL[NITEM - 1 - (q - n7)];
        int m, f, wl;
        item_nc(it, m, f, wl);
        face_nc(m, f, wl, s0 - 1 + f, q < n7 ? 2 : 3);
      }
    }
  } else if (!HYB) {
    for (int it = tid; it < NF * TW; it += NTHREADS) {
      int f, wl;
      item_ch(it, f, wl);
      const int s = s0 - 1 + f;
      if (s > n || w0 + wl > nw) continue;
      if ((closure_lo && s == 0) || (closure_hi && s == n)) {
        for (int m = 0; m < 5; m++) sH[lidx<AXIS, NF, TW>(m, f, wl)] = closure_comp(f, wl, s, m);
        continue;
      }
      face_ch(f, wl, s, 0);
    }
  } else {
    // hybrid, characteristic: linear and closure faces per component; the
    // WENO faces are listed once and then computed one characteristic row
    // per thread, followed by the back-projection one face per thread
    for (int it = tid; it < 5 * NF * TW; it += NTHREADS) {
      int m, f, wl;
      item_nc(it, m, f, wl);
      const int s = s0 - 1 + f;
      if (s > n || w0 + wl > nw) continue;
      if ((closure_lo && s == 0) || (closure_hi && s == n)) {
        sH[lidx<AXIS, NF, TW>(m, f, wl)] = closure_comp(f, wl, s, m);
        continue;
      }
      const int scm = scheme_of(f, wl);
      if (scm == 1) sH[lidx<AXIS, NF, TW>(m, f, wl)] = linear_comp(f, wl, s, m);
      else if (m == 0) push(it, scm);
    }
    __syncthreads();
    const int n7 = s_cnt[0], n5 = s_cnt[1], nl = n7 + n5;
    for (int q = tid; q < 5 * nl; q += NTHREADS) {
      const int lf = q / 5, r = q - 5 * lf;
      const int it = lf < n7 ? sL[lf] : sL[NITEM - 1 - (lf - n7)];
      int f, wl;
      item_ch(it, f, wl);
      CharBasis Bs;
      basis_of(f, wl, Bs);
      sH[lidx<AXIS, NF, TW>(r, f, wl)] = char_row(Bs, f, wl, s0 - 1 + f, lf < n7 ? 2 : 3, r);
    }
    __syncthreads();
    for (int lf = tid; lf < nl; lf += NTHREADS) {
      const int it = lf < n7 ? sL[lf] : sL[NITEM - 1 - (lf - n7)];
      int f, wl;
      item_ch(it, f, wl);
      CharBasis Bs;
      basis_of(f, wl, Bs);
      double h0[5], hh[5];
#pragma unroll
      for (int r = 0; r < 5; r++) h0[r] = sH[lidx<AXIS, NF, TW>(r, f, wl)];
      char_back(Bs, h0, hh);
#pragma unroll
      for (int m = 0; m < 5; m++) sH[lidx<AXIS, NF, TW>(m, f, wl)] = hh[m];
    }
  }
#undef FPV
#undef FMV
  __syncthreads();

  // ---- phase 3: flux difference into du ----
  const double hinv = B.P.hinv[AXIS];
#pragma unroll
  for (int r = 0; r < NI3; r++) {
    int sl_, wl, m;
    if (!p3(r, sl_, wl, m)) continue;
    const int s = s0 + sl_, w = w0 + wl;
    const int f = sl_ + 1;
    double val;
    if (!CHAR) {
      val = (sH[lidx<AXIS, NF, TW>(m, f, wl)] - sH[lidx<AXIS, NF, TW>(m, f - 1, wl)] +
             sH[lidx<AXIS, NF, TW>(5 + m, f, wl)] - sH[lidx<AXIS, NF, TW>(5 + m, f - 1, wl)]) * hinv;
    } else {
      val = (sH[lidx<AXIS, NF, TW>(m, f, wl)] - sH[lidx<AXIS, NF, TW>(m, f - 1, wl)]) * hinv;
    }
    const int64_t q = TI::iidx(g, s, w, c) + (int64_t)m * g.ni;
    if (mode == INV_ASSIGN) {
      B.du[q] = val;
    } else if (mode == INV_ACCUM) {
      B.du[q] = pdu[r] + val;
    } else {
      const int64_t h = TI::hidx(g, s, w, c);
      double rr = -(pdu[r] + val) * B.M.get(M_AJAC, h);
      if (B.P.if_mass_force == 1 && m > 0) {
        // OCFD_inviscous.f90:189-201
        const double *mf = B.P.mass_force;
        if (m < 4) rr = rr + B.d[h] * mf[m - 1];
        else rr = rr + B.d[h] * (B.u[h] * mf[0] + B.v[h] * mf[1] + B.w[h] * mf[2]);
      }
      B.du[q] = rr;
    }
  }
}

template <int TL, int TW, bool CHAR, bool HYB>
constexpr size_t sweep_lds(int nq) {
  return sizeof(double) * (10 * (TL + 8) * TW + (CHAR ? nq * (TL + 8) * TW : 0) + (HYB ? (TL + 8) * TW : 0) +
                           (CHAR ? 5 : 10) * (TL + 1) * TW) +
         (HYB ? sizeof(int) * (CHAR ? (TL + 1) * TW : 5 * (TL + 1) * TW) : 0);
}

template <int AXIS, int SCH, bool CHAR, int TL, int TW>
hipError_t launch_one(const Block &B, int mode, hipStream_t st) {
  constexpr bool HYB = (SCH == SCH_HYBRID);
  static_assert(sweep_lds<TL, TW, CHAR, HYB>(7) <= 64 * 1024, "tile too large for LDS");
  const size_t lds = sweep_lds<TL, TW, CHAR, HYB>(B.M.uniform ? 4 : 7);
  const Geom &g = B.g;
  const int n = g.n(AXIS);
  const int nw = AXIS == 0 ? g.ny : g.nx;
  const int nc = AXIS == 2 ? g.ny : g.nz;
  const int64_t blocks = (int64_t)((n + TL - 1) / TL) * ((nw + TW - 1) / TW) * nc;
  hipLaunchKernelGGL((k_inviscid<AXIS, SCH, CHAR, TL, TW>), dim3((unsigned)blocks), dim3(NTHREADS), lds, st, B,
                     mode);
  return hipGetLastError();
}

template <int AXIS, bool CHAR>
hipError_t dispatch_scheme(const Block &B, int mode, hipStream_t st) {
  constexpr int TL = AXIS == 0 ? 64 : 32;
  constexpr int TW = AXIS == 0 ? 4 : 8;
  switch (B.S.scheme_invis) {
    case SCH_WENO5: return launch_one<AXIS, SCH_WENO5, CHAR, TL, TW>(B, mode, st);
    case SCH_WENO7: return launch_one<AXIS, SCH_WENO7, CHAR, TL, TW>(B, mode, st);
    case SCH_OMP6: return launch_one<AXIS, SCH_OMP6, CHAR, TL, TW>(B, mode, st);
    case SCH_UD7L: return launch_one<AXIS, SCH_UD7L, CHAR, TL, TW>(B, mode, st);
    case SCH_HYBRID: return launch_one<AXIS, SCH_HYBRID, CHAR, TL, TW>(B, mode, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_inviscid(const Block &B, int axis, int mode, SweepTile tile, hipStream_t st) {
  const bool ch = B.S.is_char != 0;
  (void)tile;
  if (axis == 0) return ch ? dispatch_scheme<0, true>(B, mode, st) : dispatch_scheme<0, false>(B, mode, st);
  if (axis == 1) return ch ? dispatch_scheme<1, true>(B, mode, st) : dispatch_scheme<1, false>(B, mode, st);
  return ch ? dispatch_scheme<2, true>(B, mode, st) : dispatch_scheme<2, false>(B, mode, st);
}

}  // namespace ocfd
