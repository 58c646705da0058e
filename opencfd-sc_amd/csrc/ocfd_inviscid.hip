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

template <int AXIS, int SCH, bool CHAR, int TL, int TW>
__global__ __launch_bounds__(NTHREADS) void k_inviscid(const Block B, const int mode) {
  typedef TileIndex<AXIS> TI;
  constexpr int NP = TL + 8;  // points s0-4 .. s0+TL+3
  constexpr int NF = TL + 1;  // faces  s0-1 .. s0+TL-1
  constexpr bool HYB = (SCH == SCH_HYBRID);
  constexpr int NH = CHAR ? 5 : 10;
  extern __shared__ double smem[];
  double *sF = smem;                       // [10][NP][TW]
  double *sQ = sF + 10 * NP * TW;          // [7][NP][TW]  (characteristic only)
  double *sR = sQ + (CHAR ? 7 * NP * TW : 0);  // [NP][TW] (hybrid only)
  double *sH = sR + (HYB ? NP * TW : 0);   // [NH][NF][TW]
  // hybrid scheme: WENO7 faces listed from the front, WENO5 from the back
  constexpr int NITEM = CHAR ? NF * TW : 5 * NF * TW;
  int *sL = reinterpret_cast<int *>(sH + NH * NF * TW);  // [NITEM] (hybrid only)
  __shared__ int s_cnt[2];
  if (HYB && threadIdx.x < 2) s_cnt[threadIdx.x] = 0;

  const Geom &g = B.g;
  const int n = g.n(AXIS);
  const int nw = TI::nwidth(g);
  const int nts = (n + TL - 1) / TL, ntw = (nw + TW - 1) / TW;
  int b = blockIdx.x;
  const int ts = b % nts; b /= nts;
  const int tw = b % ntw;
  const int c = b / ntw + 1;
  const int s0 = 1 + ts * TL, w0 = 1 + tw * TW;
  const int tid = threadIdx.x;
  const double gamma = B.P.gamma, rma = B.P.ma;
  const int q1 = M_AKX1 + 3 * AXIS;

  // ---- phase 1: split fluxes at the tile's points ----
  for (int it = tid; it < NP * TW; it += NTHREADS) {
    int p, wl;
    if (AXIS == 0) { p = it % NP; wl = it / NP; } else { wl = it % TW; p = it / TW; }
    const int s = s0 - 4 + p, w = w0 + wl;
    if (s > n + LAP || w > nw) continue;
    const int64_t h = TI::hidx(g, s, w, c);
    const double dd = B.d[h], uu = B.u[h], vv = B.v[h], ww = B.w[h];
    const double cc = sqrt(B.T[h]) / rma;  // OCFD_inviscous.f90:67
    const double A1 = B.M.get(q1, h), A2 = B.M.get(q1 + 1, h), A3 = B.M.get(q1 + 2, h);
    double fp[5], fm[5];
    if (B.S.flux_split == 1) split_sw(gamma, cc, dd, uu, vv, ww, A1, A2, A3, fp, fm);
    else split_llf(gamma, cc, dd, uu, vv, ww, A1, A2, A3, fp, fm);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      sF[(m * NP + p) * TW + wl] = fp[m];
      sF[((5 + m) * NP + p) * TW + wl] = fm[m];
    }
    if (CHAR) {
      sQ[(0 * NP + p) * TW + wl] = uu;
      sQ[(1 * NP + p) * TW + wl] = vv;
      sQ[(2 * NP + p) * TW + wl] = ww;
      sQ[(3 * NP + p) * TW + wl] = cc;
      sQ[(4 * NP + p) * TW + wl] = A1;
      sQ[(5 * NP + p) * TW + wl] = A2;
      sQ[(6 * NP + p) * TW + wl] = A3;
    }
    if (HYB) sR[p * TW + wl] = B.rhyb[h];
  }
  __syncthreads();

  // ---- phase 2: face fluxes ----
  const SchemeParams &S = B.S;
  const int lo = S.bidx[AXIS][0], hi = S.bidx[AXIS][1];
  const int sl = S.sb[2 * AXIS], sr = S.sb[2 * AXIS + 1];
  // flux+/- of component m at face f (non-characteristic)
  auto face_nc = [&](int m, int f, int wl, int s, int scm) {
    double xp[8], xm[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      xp[k] = sF[(m * NP + f + k) * TW + wl];
      xm[k] = sF[((5 + m) * NP + f + k) * TW + wl];
    }
    sH[(m * NF + f) * TW + wl] = face_p_nonchar<SCH>(S, xp, s, n, AXIS, scm);
    sH[((5 + m) * NF + f) * TW + wl] = face_m_nonchar<SCH>(S, xm, s, n, AXIS, scm);
  };
#define FPV(m, k) sF[((m) * NP + f + (k)) * TW + wl]
#define FMV(m, k) sF[((5 + (m)) * NP + f + (k)) * TW + wl]
  // all five components of face f, projected on the characteristic basis of
  // the face unless the face is linear (flux_charteric,
  // OCFD_inviscous_character.f90:190-290)
  auto face_ch = [&](int f, int wl, int s, int scm) {
    double hh[5];
    int a1h = S.ka1, b1h = S.kb1, a2h = S.ka2, b2h = S.kb2;
    if (HYB) {
      if (scm == 2) { a1h = -3; b1h = 3; a2h = -2; b2h = 4; }
      else if (scm == 3) { a1h = -2; b1h = 2; a2h = -1; b2h = 3; }
    }
    CharBasis Bs;
    if (scm != 1) {
#define QV(q, k) sQ[((q) * NP + f + (k)) * TW + wl]
      const double u1 = (QV(0, 3) + QV(0, 4)) * 0.5, v1 = (QV(1, 3) + QV(1, 4)) * 0.5;
      const double w1 = (QV(2, 3) + QV(2, 4)) * 0.5, c1 = (QV(3, 3) + QV(3, 4)) * 0.5;
      const double a1 = (QV(4, 3) + QV(4, 4)) * 0.5, a2 = (QV(5, 3) + QV(5, 4)) * 0.5;
      const double a3 = (QV(6, 3) + QV(6, 4)) * 0.5;
#undef QV
      char_basis(gamma, u1, v1, w1, c1, a1, a2, a3, Bs);
    }
    double h0[5];
    for (int r = 0; r < 5; r++) {
      double xp[8], xm[8];
      if (scm == 1) {
#pragma unroll
        for (int k = 0; k < 8; k++) { xp[k] = FPV(r, k); xm[k] = FMV(r, k); }
      } else {
        double sr5[5];
        char_S_row(Bs, r, sr5);
#pragma unroll
        for (int k = 0; k < 8; k++) {
          xp[k] = 0.0;
          xm[k] = 0.0;
          if (k >= a1h + 3 && k <= b1h + 3)
            xp[k] = sr5[0] * FPV(0, k) + sr5[1] * FPV(1, k) + sr5[2] * FPV(2, k) + sr5[3] * FPV(3, k) +
                    sr5[4] * FPV(4, k);
          if (k >= a2h + 3 && k <= b2h + 3)
            xm[k] = sr5[0] * FMV(0, k) + sr5[1] * FMV(1, k) + sr5[2] * FMV(2, k) + sr5[3] * FMV(3, k) +
                    sr5[4] * FMV(4, k);
        }
      }
      const double h1 = face_p_char<SCH>(S, xp, s, n, AXIS, scm);
      const double h2 = face_m_char<SCH>(S, xm, s, n, AXIS, scm);
      h0[r] = h1 + h2;
    }
    if (scm == 1) {
      for (int r = 0; r < 5; r++) hh[r] = h0[r];
    } else {
      char_back(Bs, h0, hh);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) sH[(m * NF + f) * TW + wl] = hh[m];
  };
  auto push = [&](int it, int scm) {
    if (scm == 2) sL[atomicAdd(&s_cnt[0], 1)] = it;
    else sL[NITEM - 1 - atomicAdd(&s_cnt[1], 1)] = it;
  };
  if (!CHAR) {
    for (int it = tid; it < 5 * NF * TW; it += NTHREADS) {
      const int wl = it % TW;
      const int f = (it / TW) % NF;
      const int m = it / (TW * NF);
      const int s = s0 - 1 + f;
      if (s > n || w0 + wl > nw) continue;
      int scm = 0;
      if (HYB) scm = hybrid_index(dmax2(sR[(f + 3) * TW + wl], sR[(f + 4) * TW + wl]), S.hyb2, S.hyb3);
      if (!HYB || scm == 1) face_nc(m, f, wl, s, scm);
      else push(it, scm);
    }
  } else {
    for (int it = tid; it < NF * TW; it += NTHREADS) {
      const int wl = it % TW;
      const int f = it / TW;
      const int s = s0 - 1 + f;
      if (s > n || w0 + wl > nw) continue;
      if ((lo == 1 && sl == 0 && s == 0) || (hi == 1 && sr == 0 && s == n)) {
        // non-reflecting closure (OCFD_inviscous_character.f90:295-309); point t is k = t - s + 3
        for (int m = 0; m < 5; m++) {
          double hv;
          if (s == 0)
            hv = (2.0 * FPV(m, 4) + 5.0 * FPV(m, 5) - FPV(m, 6) + 2.0 * FMV(m, 6) - 7.0 * FMV(m, 5) +
                  11.0 * FMV(m, 4)) / 6.0;
          else
            hv = (2.0 * FPV(m, 1) - 7.0 * FPV(m, 2) + 11.0 * FPV(m, 3) + 2.0 * FMV(m, 3) + 5.0 * FMV(m, 2) -
                  FMV(m, 1)) / 6.0;
          sH[(m * NF + f) * TW + wl] = hv;
        }
        continue;
      }
      int scm = 0;
      if (HYB) scm = hybrid_index(dmax2(sR[(f + 3) * TW + wl], sR[(f + 4) * TW + wl]), S.hyb2, S.hyb3);
      if (!HYB || scm == 1) face_ch(f, wl, s, scm);
      else push(it, scm);
    }
  }
  if (HYB) {
    // second pass: the listed WENO7 faces, then the WENO5 faces, spread over
    // all threads
    __syncthreads();
    const int n7 = s_cnt[0], n5 = s_cnt[1];
    for (int q = tid; q < n7 + n5; q += NTHREADS) {
      const int it = q < n7 ? sL[q] : sL[NITEM - 1 - (q - n7)];
      const int scm = q < n7 ? 2 : 3;
      if (!CHAR) {
        const int wl = it % TW, f = (it / TW) % NF, m = it / (TW * NF);
        face_nc(m, f, wl, s0 - 1 + f, scm);
      } else {
        const int wl = it % TW, f = it / TW;
        face_ch(f, wl, s0 - 1 + f, scm);
      }
    }
  }
#undef FPV
#undef FMV
  __syncthreads();

  // ---- phase 3: flux difference into du ----
  const double hinv = B.P.hinv[AXIS];
  for (int it = tid; it < 5 * TL * TW; it += NTHREADS) {
    int sl_, wl, m;
    if (AXIS == 0) { sl_ = it % TL; wl = (it / TL) % TW; m = it / (TL * TW); }
    else { wl = it % TW; sl_ = (it / TW) % TL; m = it / (TL * TW); }
    const int s = s0 + sl_, w = w0 + wl;
    if (s > n || w > nw) continue;
    const int f = sl_ + 1;
    double val;
    if (!CHAR) {
      val = (sH[(m * NF + f) * TW + wl] - sH[(m * NF + f - 1) * TW + wl] + sH[((5 + m) * NF + f) * TW + wl] -
             sH[((5 + m) * NF + f - 1) * TW + wl]) * hinv;
    } else {
      val = (sH[(m * NF + f) * TW + wl] - sH[(m * NF + f - 1) * TW + wl]) * hinv;
    }
    const int64_t q = TI::iidx(g, s, w, c) + (int64_t)m * g.ni;
    if (mode == INV_ASSIGN) {
      B.du[q] = val;
    } else if (mode == INV_ACCUM) {
      B.du[q] = B.du[q] + val;
    } else {
      const int64_t h = TI::hidx(g, s, w, c);
      double r = -(B.du[q] + val) * B.M.get(M_AJAC, h);
      if (B.P.if_mass_force == 1 && m > 0) {
        // OCFD_inviscous.f90:189-201
        const double *mf = B.P.mass_force;
        if (m < 4) r = r + B.d[h] * mf[m - 1];
        else r = r + B.d[h] * (B.u[h] * mf[0] + B.v[h] * mf[1] + B.w[h] * mf[2]);
      }
      B.du[q] = r;
    }
  }
}

template <int AXIS, int SCH, bool CHAR, int TL, int TW>
hipError_t launch_one(const Block &B, int mode, hipStream_t st) {
  constexpr int NP = TL + 8, NF = TL + 1;
  constexpr bool HYB = (SCH == SCH_HYBRID);
  constexpr size_t lds = sizeof(double) * (10 * NP * TW + (CHAR ? 7 * NP * TW : 0) + (HYB ? NP * TW : 0) +
                                           (CHAR ? 5 : 10) * NF * TW) +
                         (HYB ? sizeof(int) * (CHAR ? NF * TW : 5 * NF * TW) : 0);
  static_assert(lds <= 64 * 1024, "tile too large for LDS");
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
  (void)tile;
  const bool ch = B.S.is_char != 0;
  if (axis == 0) return ch ? dispatch_scheme<0, true>(B, mode, st) : dispatch_scheme<0, false>(B, mode, st);
  if (axis == 1) return ch ? dispatch_scheme<1, true>(B, mode, st) : dispatch_scheme<1, false>(B, mode, st);
  return ch ? dispatch_scheme<2, true>(B, mode, st) : dispatch_scheme<2, false>(B, mode, st);
}

}  // namespace ocfd
