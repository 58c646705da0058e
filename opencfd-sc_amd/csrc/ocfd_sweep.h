// ocfd_sweep.h -- inviscid right-hand side: one pencil sweep per axis.
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
// R is the arithmetic of the splitting and of the faces the sweep computes
// (double; float for the mixed-precision form); primitives, du and the
// hybrid scheme's precomputed WENO faces stay double.  Each translation unit
// that includes this header instantiates one R.
#pragma once
#include "ocfd_kernels.h"
#include "ocfd_march.h"

namespace ocfd {

namespace {

constexpr int NTHREADS = 256;

// LDS element (component m, point or face p, line wl) of a tile array with NPT
// entries per line: x tiles keep the points of one line contiguous, y and z
// tiles the lines of one point, so that consecutive lanes touch consecutive
// doubles in every phase.
template <int AXIS, int NPT, int TW>
__device__ constexpr inline int lidx(int m, int p, int wl) {
  return AXIS == 0 ? (m * TW + wl) * NPT + p : (m * NPT + p) * TW + wl;
}

// One tile b of a sweep of block B.  FAST: the sweep axis is periodic (no
// block-edge closures) and the metrics are uniform, both known at compile
// time.
template <int AXIS, int SCH, bool CHAR, int TL, int TW, bool FAST, typename R>
__device__ __forceinline__ void sweep_tile(const Block &B, int b, const int mode) {
  typedef TileIndex<AXIS> TI;
  constexpr int NP = TL + 8;  // points s0-4 .. s0+TL+3
  constexpr int NF = TL + 1;  // faces  s0-1 .. s0+TL-1
  constexpr bool HYB = (SCH == SCH_HYBRID);
  constexpr int NH = CHAR ? 5 : 10;
  constexpr int NI1 = (NP * TW + NTHREADS - 1) / NTHREADS;      // points per thread
  constexpr int NI3 = (5 * TL * TW + NTHREADS - 1) / NTHREADS;  // du entries per thread
  // characteristic inputs per point: u, v, w, c (+ the three normalised
  // metrics when they vary in space)
  const bool unif = FAST || B.M.uniform != 0;
  const int nq = unif ? 4 : 7;
  extern __shared__ double smem[];

  const Geom &g = B.g;
  const int n = g.n(AXIS);
  const int nw = TI::nwidth(g), nc = TI::nthird(g);
  const int nts = (n + TL - 1) / TL, ntw = (nw + TW - 1) / TW;
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

  // primitives of the tile's points (stored in the sweep's real type)
  R pq[NI1][5];
  {
    const R *const pd = B.d.as<R>(), *const pu = B.u.as<R>(), *const pv = B.v.as<R>(), *const pw = B.w.as<R>(),
            *const pT = B.T.as<R>();
#pragma unroll
    for (int r = 0; r < NI1; r++) {
      int p, wl;
      if (!p1(r, p, wl)) continue;
      const int64_t h = TI::hidx(g, s0 - 4 + p, w0 + wl, c);
      pq[r][0] = pd[h];
      pq[r][1] = pu[h];
      pq[r][2] = pv[h];
      pq[r][3] = pw[h];
      pq[r][4] = pT[h];
    }
  }
  // du entries of the tile (accumulating modes), fetched before the face
  // fluxes are computed
  // du is stored in the sweep's real type (FP32 in the mixed-precision form)
  R *const du = reinterpret_cast<R *>(B.du);
  R pdu[NI3];

  R *sF = reinterpret_cast<R *>(smem);             // fp(5), fm(5)      [10][NP][TW]
  // the characteristic inputs are needed only where the sweep itself
  // projects, i.e. not for the hybrid scheme (its WENO faces come precomputed)
  constexpr bool PROJ = CHAR && !HYB;
  // FP64 tiles are held to 4 workgroups per CU by their LDS: there the face
  // fluxes take the place of the split fluxes once phase 2 has read them
  // (phase 2 keeps its results in registers across one more barrier).  The
  // FP32 tiles, bound by instruction issue, keep separate arrays and the
  // plain item loops (measured faster for them).
  constexpr bool ALIAS = sizeof(R) == 8;
  R *sQ = sF + 10 * NP * TW;                       // char inputs       [nq][NP][TW]
  R *const sQend = sQ + (PROJ ? nq * NP * TW : 0);
  R *sH = ALIAS ? sF : sQend;                      // face fluxes       [NH][NF][TW]
  int *sS = reinterpret_cast<int *>(ALIAS ? sQend : sH + NH * NF * TW);  // hybrid: list slot of each face [NF][TW]
  // ---- phase 1: split fluxes at the tile's points ----
  // the splitting's constants once per thread (and the metric part once per
  // sweep on a uniform grid)
  const bool sw = S.flux_split == 1;
  SplitK<R> K;
  split_k_gamma(R(gamma), K);
  if (unif) split_k_metric(R(B.M.u[q1]), R(B.M.u[q1 + 1]), R(B.M.u[q1 + 2]), sw, K);
#pragma unroll
  for (int r = 0; r < NI1; r++) {
    int p, wl;
    if (!p1(r, p, wl)) continue;
    const R dd = R(pq[r][0]), uu = R(pq[r][1]), vv = R(pq[r][2]), ww = R(pq[r][3]);
    const R cc = sqrt(R(pq[r][4])) / R(rma);  // OCFD_inviscous.f90:67
    if (!unif) {
      const int64_t h = TI::hidx(g, s0 - 4 + p, w0 + wl, c);
      split_k_metric(R(B.M.p[q1][h]), R(B.M.p[q1 + 1][h]), R(B.M.p[q1 + 2][h]), sw, K);
    }
    const R A1 = K.A1, A2 = K.A2, A3 = K.A3;
    R fp[5], fm[5];
    if (sw) split_sw<R>(K, cc, dd, uu, vv, ww, fp, fm);
    else split_llf<R>(K, cc, dd, uu, vv, ww, fp, fm);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      sF[lidx<AXIS, NP, TW>(m, p, wl)] = fp[m];
      sF[lidx<AXIS, NP, TW>(5 + m, p, wl)] = fm[m];
    }
    if (PROJ) {
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
  }
  if (HYB) {
    // WENO faces of the hybrid scheme come precomputed (k_weno_faces); with
    // ALIAS phase 2 reads them, as the face array is not free yet
    for (int it = tid; it < NF * TW; it += NTHREADS) {
      int f, wl;
      if (AXIS == 0) { f = it % NF; wl = it / NF; } else { wl = it % TW; f = it / TW; }
      const int s = s0 - 1 + f, w = w0 + wl;
      int slot = -1;
      if (s <= n && w <= nw) {
        const int64_t fid = TI::face_id(g, s, w, c);
        const unsigned long long bits = B.hbits[AXIS][fid >> 6];
        const int bit = (int)(fid & 63);
        if ((bits >> bit) & 1ull) slot = B.hgbase[AXIS][fid >> 6] + __popcll(bits & ((1ull << bit) - 1ull));
      }
      sS[lidx<AXIS, NF, TW>(0, f, wl)] = slot;
      if (!ALIAS && slot >= 0) {
        const double *wf = B.wflux[AXIS] + (int64_t)slot * NH;
#pragma unroll
        for (int m = 0; m < NH; m++) sH[lidx<AXIS, NF, TW>(m, f, wl)] = R(wf[m]);
      }
    }
  }
  __syncthreads();

  if (mode != INV_ASSIGN) {
#pragma unroll
    for (int r = 0; r < NI3; r++) {
      int sl_, wl, m;
      if (!p3(r, sl_, wl, m)) continue;
      pdu[r] = du[TI::iidx(g, s0 + sl_, w0 + wl, c) + (int64_t)m * g.ni];
    }
  }

  // ---- phase 2: face fluxes ----
  // flux+/- of component m at face f (non-characteristic)
  auto face_nc = [&](int m, int f, int wl, int s, int scm, R &hp, R &hm) {
    R xp[8], xm[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      xp[k] = sF[lidx<AXIS, NP, TW>(m, f + k, wl)];
      xm[k] = sF[lidx<AXIS, NP, TW>(5 + m, f + k, wl)];
    }
    hp = FAST ? inner_p<SCH>(S, xp, scm) : face_p_nonchar<SCH>(S, xp, s, n, AXIS, scm);
    hm = FAST ? inner_m<SCH>(S, xm, scm) : face_m_nonchar<SCH>(S, xm, s, n, AXIS, scm);
  };
#define FPV(m, k) sF[lidx<AXIS, NP, TW>((m), f + (k), wl)]
#define FMV(m, k) sF[lidx<AXIS, NP, TW>(5 + (m), f + (k), wl)]
  // all five components of face f, projected on the characteristic basis of
  // the face unless the face is linear (flux_charteric,
  // OCFD_inviscous_character.f90:190-290)
  // characteristic basis of face f from the face-averaged state
  // (OCFD_inviscous_character.f90:200-236)
  auto basis_of = [&](int f, int wl, CharBasisT<R> &Bs) {
#define QV(q, k) sQ[lidx<AXIS, NP, TW>((q), f + (k), wl)]
    const R u1 = (QV(0, 3) + QV(0, 4)) * R(0.5), v1 = (QV(1, 3) + QV(1, 4)) * R(0.5);
    const R w1 = (QV(2, 3) + QV(2, 4)) * R(0.5), c1 = (QV(3, 3) + QV(3, 4)) * R(0.5);
    R a1, a2, a3;
    if (unif) {
      // average of two equal values, as the reference computes it
      a1 = (R(B.M.u[q1]) + R(B.M.u[q1])) * R(0.5);
      a2 = (R(B.M.u[q1 + 1]) + R(B.M.u[q1 + 1])) * R(0.5);
      a3 = (R(B.M.u[q1 + 2]) + R(B.M.u[q1 + 2])) * R(0.5);
    } else {
      a1 = (QV(4, 3) + QV(4, 4)) * R(0.5);
      a2 = (QV(5, 3) + QV(5, 4)) * R(0.5);
      a3 = (QV(6, 3) + QV(6, 4)) * R(0.5);
    }
#undef QV
    char_basis<R>(R(gamma), u1, v1, w1, c1, a1, a2, a3, Bs);
  };
  // characteristic flux h0(r) = h+(r) + h-(r) of face f: row r of S applied
  // to the split fluxes on the scheme's stencil, then reconstructed
  // (OCFD_inviscous_character.f90:250-282)
  auto char_row = [&](const CharBasisT<R> &Bs, int f, int wl, int s, int scm, int r) {
    int a1h = S.ka1, b1h = S.kb1, a2h = S.ka2, b2h = S.kb2;
    if (HYB) {
      if (scm == 2) { a1h = -3; b1h = 3; a2h = -2; b2h = 4; }
      else if (scm == 3) { a1h = -2; b1h = 2; a2h = -1; b2h = 3; }
    }
    R xp[8], xm[8], sr5[5];
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
        xp[k] = madd(sr5[c], FPV(c, k), xp[k]);
        xm[k] = madd(sr5[c], FMV(c, k), xm[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (k < a1h + 3 || k > b1h + 3) xp[k] = R(0.0);
      if (k < a2h + 3 || k > b2h + 3) xm[k] = R(0.0);
    }
    const R h1 = FAST ? inner_p<SCH>(S, xp, scm) : face_p_char<SCH>(S, xp, s, n, AXIS, scm);
    const R h2 = FAST ? inner_m<SCH>(S, xm, scm) : face_m_char<SCH>(S, xm, s, n, AXIS, scm);
    return h1 + h2;
  };
  // component m of a linear (hybrid UD7L) face: no projection
  auto linear_comp = [&](int f, int wl, int s, int m) {
    R xp[8], xm[8];
#pragma unroll
    for (int k = 0; k < 8; k++) { xp[k] = FPV(m, k); xm[k] = FMV(m, k); }
    const R h1 = FAST ? inner_p<SCH>(S, xp, 1) : face_p_char<SCH>(S, xp, s, n, AXIS, 1);
    const R h2 = FAST ? inner_m<SCH>(S, xm, 1) : face_m_char<SCH>(S, xm, s, n, AXIS, 1);
    return h1 + h2;
  };
  // non-reflecting closure of component m (OCFD_inviscous_character.f90:295-309);
  // point t is k = t - s + 3
  auto closure_comp = [&](int f, int wl, int s, int m) {
    if (s == 0)
      return (R(2.0) * FPV(m, 4) + R(5.0) * FPV(m, 5) - FPV(m, 6) + R(2.0) * FMV(m, 6) - R(7.0) * FMV(m, 5) +
              R(11.0) * FMV(m, 4)) / R(6.0);
    return (R(2.0) * FPV(m, 1) - R(7.0) * FPV(m, 2) + R(11.0) * FPV(m, 3) + R(2.0) * FMV(m, 3) +
            R(5.0) * FMV(m, 2) - FMV(m, 1)) / R(6.0);
  };
  // all five components of face f, projected on the characteristic basis of
  // the face (flux_charteric, OCFD_inviscous_character.f90:190-290)
  auto face_ch = [&](int f, int wl, int s, int scm, R *hh) {
    CharBasisT<R> Bs;
    basis_of(f, wl, Bs);
    R h0[5];
    // one row at a time: unrolled, the five rows' stencils stayed live
    // together (FP32 tiles at 256 VGPRs plus AGPRs, one wave per SIMD)
#pragma unroll 1
    for (int r = 0; r < 5; r++) h0[r] = char_row(Bs, f, wl, s, scm, r);
    char_back(Bs, h0, hh);
  };
  auto listed = [&](int f, int wl) { return sS[lidx<AXIS, NF, TW>(0, f, wl)] >= 0; };
  // face work item -> (component, face, line)
  auto item_nc = [&](int it, int &m, int &f, int &wl) {
    if (AXIS == 0) { f = it % NF; wl = (it / NF) % TW; } else { wl = it % TW; f = (it / TW) % NF; }
    m = it / (TW * NF);
  };
  auto item_ch = [&](int it, int &f, int &wl) {
    if (AXIS == 0) { f = it % NF; wl = it / NF; } else { wl = it % TW; f = it / TW; }
  };
  const bool closure_lo = !FAST && lo == 1 && sl == 0, closure_hi = !FAST && hi == 1 && sr == 0;
  auto slot_of = [&](int f, int wl) { return sS[lidx<AXIS, NF, TW>(0, f, wl)]; };
  // work items of this thread: (component, face, line), or on the
  // characteristic path (face, line) with all five components, so that the
  // item's index arithmetic, slot and closure tests are done once per face
  constexpr int NIT = CHAR ? NF * TW : 5 * NF * TW;
  constexpr int NI2 = (NIT + NTHREADS - 1) / NTHREADS;
  // On the characteristic path (TL + 1) x TW faces leave TW faces past the
  // last whole round of lanes; a wavefront's round for a face with all five
  // components would cost as much as the 64 faces of a full one, so these
  // go five lanes to a face.  The hybrid scheme's faces (linear, listed or
  // closures, none projected here) take one component per lane; a
  // projected face takes one characteristic row per lane (its basis formed
  // by each), the five rows meet in LDS (sTail), and each lane projects one
  // component back.
  constexpr int FULL = (NIT / NTHREADS) * NTHREADS;
  constexpr bool SPLIT_TAIL = CHAR && NIT % NTHREADS != 0 && 5 * (NIT - FULL) <= NTHREADS;
  R *const sTail = reinterpret_cast<R *>(sS);     // projecting tiles: tail rows [5 * TW]
  // one tail item: face FULL + tid / 5, component / row tid % 5; stage 1 of
  // a projected face leaves its row in sTail, stage 2 (after a barrier)
  // returns the component
  auto tail_row = [&](CharBasisT<R> &Bs, int f, int wl, int s, int row) {
    basis_of(f, wl, Bs);
    sTail[tid] = char_row(Bs, f, wl, s, 0, row);
  };
  auto tail_back = [&](const CharBasisT<R> &Bs, int row) {
    R h0[5], hh[5];
#pragma unroll
    for (int q = 0; q < 5; q++) h0[q] = sTail[(tid / 5) * 5 + q];
    char_back(Bs, h0, hh);
    return hh[row];
  };
  constexpr int NV = CHAR ? 5 : 2;  // face-array values per item
  R hv[ALIAS ? NI2 : 1][NV];
  // item r -> its face-array entries; false for items outside the tile
  auto decode = [&](int r, int &m, int &f, int &wl) {
    if (SPLIT_TAIL && r == NI2 - 1) {
      if (tid >= 5 * (NIT - FULL)) return false;
      m = tid % 5;
      item_ch(FULL + tid / 5, f, wl);
      return s0 - 1 + f <= n && w0 + wl <= nw;
    }
    const int it = tid + r * NTHREADS;
    if (it >= NIT) return false;
    if (CHAR) { m = 0; item_ch(it, f, wl); } else item_nc(it, m, f, wl);
    return s0 - 1 + f <= n && w0 + wl <= nw;
  };
  auto put = [&](int m, int f, int wl, const R *v) {
    if constexpr (NV == 2) {
      sH[lidx<AXIS, NF, TW>(m, f, wl)] = v[0];
      sH[lidx<AXIS, NF, TW>(5 + m, f, wl)] = v[1];
    } else if constexpr (NV == 5) {
#pragma unroll
      for (int q = 0; q < 5; q++) sH[lidx<AXIS, NF, TW>(q, f, wl)] = v[q];
    } else {
      sH[lidx<AXIS, NF, TW>(m, f, wl)] = v[0];
    }
  };
  if constexpr (ALIAS) {
    // results held in registers until every stencil has been read
#pragma unroll
    for (int r = 0; r < NI2; r++) {
      int m = 0, f = 0, wl = 0;
      if constexpr (SPLIT_TAIL && !HYB) {
        if (r == NI2 - 1) {
          // projected tail faces: a row per lane, a barrier, a component back
          const bool act = decode(r, m, f, wl);
          const int s = s0 - 1 + f;
          const bool clo = (closure_lo && s == 0) || (closure_hi && s == n);
          CharBasisT<R> Bs;
          if (act && !clo) tail_row(Bs, f, wl, s, m);
          __syncthreads();
          if (act) hv[r][0] = clo ? closure_comp(f, wl, s, m) : tail_back(Bs, m);
          continue;
        }
      }
      if (!decode(r, m, f, wl)) continue;
      const int s = s0 - 1 + f;
      R v[NV];
      if constexpr (!CHAR) {
        if (HYB && slot_of(f, wl) >= 0) {
          const double *wf = B.wflux[AXIS] + (int64_t)slot_of(f, wl) * NH;
          v[0] = R(wf[m]);
          v[1] = R(wf[5 + m]);
        } else {
          face_nc(m, f, wl, s, HYB ? 1 : 0, v[0], v[1]);
        }
      } else if constexpr (!HYB) {
        if ((closure_lo && s == 0) || (closure_hi && s == n)) {
#pragma unroll
          for (int q = 0; q < NV; q++) v[q] = closure_comp(f, wl, s, q);
        } else {
          face_ch(f, wl, s, 0, v);
        }
      } else if (SPLIT_TAIL && r == NI2 - 1) {
        // hybrid, characteristic, past the whole rounds: component m
        if ((closure_lo && s == 0) || (closure_hi && s == n)) v[0] = closure_comp(f, wl, s, m);
        else if (slot_of(f, wl) >= 0) v[0] = R(B.wflux[AXIS][(int64_t)slot_of(f, wl) * NH + m]);
        else v[0] = linear_comp(f, wl, s, m);
      } else {
        // hybrid, characteristic: the closure, listed and linear faces
        if ((closure_lo && s == 0) || (closure_hi && s == n)) {
#pragma unroll
          for (int q = 0; q < 5; q++) v[q] = closure_comp(f, wl, s, q);
        } else if (slot_of(f, wl) >= 0) {
          const double *wf = B.wflux[AXIS] + (int64_t)slot_of(f, wl) * NH;
#pragma unroll
          for (int q = 0; q < 5; q++) v[q] = R(wf[q]);
        } else {
#pragma unroll
          for (int q = 0; q < 5; q++) v[q] = linear_comp(f, wl, s, q);
        }
      }
#pragma unroll
      for (int q = 0; q < NV; q++) hv[r][q] = v[q];
    }
    // every stencil has been read: the face fluxes go over the split fluxes
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NI2; r++) {
      int m, f, wl;
      if (!decode(r, m, f, wl)) continue;
      if (SPLIT_TAIL && r == NI2 - 1) sH[lidx<AXIS, NF, TW>(m, f, wl)] = hv[r][0];
      else put(m, f, wl, hv[r]);
    }
  } else {
    // separate face array: each item's fluxes are stored as computed
    if constexpr (!CHAR) {
      for (int it = tid; it < 5 * NF * TW; it += NTHREADS) {
        int m, f, wl;
        item_nc(it, m, f, wl);
        const int s = s0 - 1 + f;
        if (s > n || w0 + wl > nw) continue;
        if (HYB && slot_of(f, wl) >= 0) continue;  // written in phase 1
        R v[2];
        face_nc(m, f, wl, s, HYB ? 1 : 0, v[0], v[1]);
        put(m, f, wl, v);
      }
    } else if constexpr (!HYB) {
      for (int it = tid; it < (SPLIT_TAIL ? FULL : NF * TW); it += NTHREADS) {
        int f, wl;
        item_ch(it, f, wl);
        const int s = s0 - 1 + f;
        if (s > n || w0 + wl > nw) continue;
        R v[5];
        if ((closure_lo && s == 0) || (closure_hi && s == n)) {
          for (int m = 0; m < 5; m++) v[m] = closure_comp(f, wl, s, m);
        } else {
          face_ch(f, wl, s, 0, v);
        }
        put(0, f, wl, v);
      }
      if constexpr (SPLIT_TAIL) {
        int m = 0, f = 0, wl = 0;
        const bool act = decode(NI2 - 1, m, f, wl);
        const int s = s0 - 1 + f;
        const bool clo = (closure_lo && s == 0) || (closure_hi && s == n);
        CharBasisT<R> Bs;
        if (act && !clo) tail_row(Bs, f, wl, s, m);
        __syncthreads();
        if (act) sH[lidx<AXIS, NF, TW>(m, f, wl)] = clo ? closure_comp(f, wl, s, m) : tail_back(Bs, m);
      }
    } else {
      // hybrid, characteristic: the linear and closure faces, all five
      // components per item, those past the whole rounds one component per
      // lane
      for (int it = tid; it < (SPLIT_TAIL ? FULL : NF * TW); it += NTHREADS) {
        int f, wl;
        item_ch(it, f, wl);
        const int s = s0 - 1 + f;
        if (s > n || w0 + wl > nw) continue;
        R v[5];
        if ((closure_lo && s == 0) || (closure_hi && s == n)) {
#pragma unroll
          for (int q = 0; q < 5; q++) v[q] = closure_comp(f, wl, s, q);
        } else if (slot_of(f, wl) < 0) {
#pragma unroll
          for (int q = 0; q < 5; q++) v[q] = linear_comp(f, wl, s, q);
        } else {
          continue;  // written in phase 1
        }
        put(0, f, wl, v);
      }
      if constexpr (SPLIT_TAIL) {
        if (tid < 5 * (NIT - FULL)) {
          const int m = tid % 5;
          int f, wl;
          item_ch(FULL + tid / 5, f, wl);
          const int s = s0 - 1 + f;
          if (s <= n && w0 + wl <= nw) {
            if ((closure_lo && s == 0) || (closure_hi && s == n))
              sH[lidx<AXIS, NF, TW>(m, f, wl)] = closure_comp(f, wl, s, m);
            else if (slot_of(f, wl) < 0)
              sH[lidx<AXIS, NF, TW>(m, f, wl)] = linear_comp(f, wl, s, m);
            // listed faces: written in phase 1
          }
        }
      }
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
    R val;
    if (!CHAR) {
      val = (sH[lidx<AXIS, NF, TW>(m, f, wl)] - sH[lidx<AXIS, NF, TW>(m, f - 1, wl)] +
             sH[lidx<AXIS, NF, TW>(5 + m, f, wl)] - sH[lidx<AXIS, NF, TW>(5 + m, f - 1, wl)]) * R(hinv);
    } else {
      val = (sH[lidx<AXIS, NF, TW>(m, f, wl)] - sH[lidx<AXIS, NF, TW>(m, f - 1, wl)]) * R(hinv);
    }
    const int64_t q = TI::iidx(g, s, w, c) + (int64_t)m * g.ni;
    if (mode == INV_ASSIGN) {
      du[q] = val;
    } else if (mode == INV_ACCUM) {
      du[q] = pdu[r] + val;
    } else {
      const int64_t h = TI::hidx(g, s, w, c);
      R rr = -(pdu[r] + val) * R(FAST ? B.M.u[M_AJAC] : B.M.get(M_AJAC, h));
      if (B.P.if_mass_force == 1 && m > 0) {
        // OCFD_inviscous.f90:189-201
        const double *mf = B.P.mass_force;
        if (m < 4) rr = rr + R(B.d[h] * mf[m - 1]);
        else rr = rr + R(B.d[h] * (B.u[h] * mf[0] + B.v[h] * mf[1] + B.w[h] * mf[2]));
      }
      du[q] = rr;
    }
  }
}

// one block: the tiles in XCD order
template <int AXIS, int SCH, bool CHAR, int TL, int TW, bool FAST, typename R>
__global__ __launch_bounds__(NTHREADS) void k_inviscid(const Block B, const int mode) {
  sweep_tile<AXIS, SCH, CHAR, TL, TW, FAST, R>(B, xcd_tile(blockIdx.x, gridDim.x), mode);
}

// a batch of periodic boxes with uniform metrics: the tiles of all boxes of
// the launch, in XCD order
template <int AXIS, int SCH, bool CHAR, int TL, int TW, typename R>
__global__ __launch_bounds__(NTHREADS) void k_inviscid_batch(const BoxSet S, const int mode) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int j = S.entry(t);
  // the descriptor is read in place (a local copy would live in scratch:
  // the sweep indexes some of its arrays at run time)
  sweep_tile<AXIS, SCH, CHAR, TL, TW, true, R>(S.blk[S.ids[j]], t - S.off[j], mode);
}

// LDS bytes of a sweep tile: split fluxes, characteristic inputs, face fluxes
// (in the split fluxes' place for FP64 tiles), hybrid slots
template <int TL, int TW, bool CHAR, bool HYB, typename R>
constexpr size_t sweep_lds(int nq) {
  return sizeof(R) * (10 * (TL + 8) * TW + (CHAR && !HYB ? nq * (TL + 8) * TW : 0) +
                      (sizeof(R) == 8 ? 0 : (CHAR ? 5 : 10) * (TL + 1) * TW) + (CHAR && !HYB ? 5 * TW : 0)) +
         (HYB ? sizeof(int) * (TL + 1) * TW : 0);
}

template <int AXIS, int SCH, bool CHAR, int TL, int TW, bool FAST, typename R>
hipError_t launch_one(const Block &B, int mode, hipStream_t st) {
  constexpr bool HYB = (SCH == SCH_HYBRID);
  static_assert(sweep_lds<TL, TW, CHAR, HYB, R>(7) <= 64 * 1024, "tile too large for LDS");
  const size_t lds = sweep_lds<TL, TW, CHAR, HYB, R>(B.M.uniform ? 4 : 7);
  const Geom &g = B.g;
  const int n = g.n(AXIS);
  const int nw = AXIS == 0 ? g.ny : g.nx;
  const int nc = AXIS == 2 ? g.ny : g.nz;
  const int64_t blocks = (int64_t)((n + TL - 1) / TL) * ((nw + TW - 1) / TW) * nc;
  hipLaunchKernelGGL((k_inviscid<AXIS, SCH, CHAR, TL, TW, FAST, R>), dim3((unsigned)blocks), dim3(NTHREADS), lds, st, B,
                     mode);
  return hipGetLastError();
}

// Tile shapes of a sweep (TL points along the axis x TW lines, 256 threads).
// Shape 0 is the default; shapes 1 and 2 are the alternatives the autotuner
// times, built for the periodic uniform-metric kernels only.
//   x: 64x4, 32x8, 128x2        y, z: 32x8, 16x16, 64x4
template <int AXIS>
constexpr int shape_tl(int shape) {
  return AXIS == 0 ? (shape == 1 ? 32 : (shape == 2 ? 128 : 64)) : (shape == 1 ? 16 : (shape == 2 ? 64 : 32));
}

template <int AXIS, bool CHAR, bool FAST, typename R, int SHAPE>
hipError_t dispatch_scheme(const Block &B, int mode, hipStream_t st) {
  constexpr int TL = shape_tl<AXIS>(SHAPE);
  constexpr int TW = NTHREADS / TL;
  switch (B.S.scheme_invis) {
    case SCH_WENO5: return launch_one<AXIS, SCH_WENO5, CHAR, TL, TW, FAST, R>(B, mode, st);
    case SCH_WENO7: return launch_one<AXIS, SCH_WENO7, CHAR, TL, TW, FAST, R>(B, mode, st);
    case SCH_OMP6: return launch_one<AXIS, SCH_OMP6, CHAR, TL, TW, FAST, R>(B, mode, st);
    case SCH_UD7L: return launch_one<AXIS, SCH_UD7L, CHAR, TL, TW, FAST, R>(B, mode, st);
    case SCH_HYBRID: return launch_one<AXIS, SCH_HYBRID, CHAR, TL, TW, FAST, R>(B, mode, st);
    default: return hipErrorInvalidValue;
  }
}

// the march (shapes 3, 4) of one scheme
template <int AXIS, int SCH, bool CHAR, typename R>
hipError_t launch_march(const Block &B, int mode, int shape, hipStream_t st) {
  if constexpr (!march_supported<SCH, CHAR>()) {
    return hipErrorInvalidValue;
  } else {
    const int sl = march_segment(AXIS, shape);
    hipLaunchKernelGGL((k_inviscid_march<AXIS, SCH, CHAR, R>), dim3((unsigned)sweep_tiles(B.g, AXIS, shape)),
                       dim3(64 * MARCH_WAVES), 0, st, B, mode, sl);
    return hipGetLastError();
  }
}

template <int AXIS, bool CHAR, typename R>
hipError_t dispatch_march(const Block &B, int mode, int shape, hipStream_t st) {
  switch (B.S.scheme_invis) {
    case SCH_WENO5: return launch_march<AXIS, SCH_WENO5, CHAR, R>(B, mode, shape, st);
    case SCH_WENO7: return launch_march<AXIS, SCH_WENO7, CHAR, R>(B, mode, shape, st);
    case SCH_OMP6: return launch_march<AXIS, SCH_OMP6, CHAR, R>(B, mode, shape, st);
    case SCH_UD7L: return launch_march<AXIS, SCH_UD7L, CHAR, R>(B, mode, shape, st);
    case SCH_HYBRID: return launch_march<AXIS, SCH_HYBRID, CHAR, R>(B, mode, shape, st);
    default: return hipErrorInvalidValue;
  }
}

template <int AXIS, typename R>
hipError_t dispatch_axis(const Block &B, int mode, int shape, hipStream_t st) {
  const bool ch = B.S.is_char != 0;
  const bool fast = B.S.iperiodic[AXIS] == 1 && B.M.uniform;
  shape = sweep_shape_effective(AXIS, shape, fast, B.S.scheme_invis, ch);
  if (fast) {
    if (shape >= 3)
      return ch ? dispatch_march<AXIS, true, R>(B, mode, shape, st) : dispatch_march<AXIS, false, R>(B, mode, shape, st);
    if (shape == 1)
      return ch ? dispatch_scheme<AXIS, true, true, R, 1>(B, mode, st)
                : dispatch_scheme<AXIS, false, true, R, 1>(B, mode, st);
    if (shape == 2)
      return ch ? dispatch_scheme<AXIS, true, true, R, 2>(B, mode, st)
                : dispatch_scheme<AXIS, false, true, R, 2>(B, mode, st);
    return ch ? dispatch_scheme<AXIS, true, true, R, 0>(B, mode, st)
              : dispatch_scheme<AXIS, false, true, R, 0>(B, mode, st);
  }
  return ch ? dispatch_scheme<AXIS, true, false, R, 0>(B, mode, st)
            : dispatch_scheme<AXIS, false, false, R, 0>(B, mode, st);
}

// the three axes of one real type; shape selects the tile (periodic
// uniform-metric sweeps only; the others always use shape 0)
template <typename R>
hipError_t launch_sweep(const Block &B, int axis, int mode, int shape, hipStream_t st) {
  if (axis == 0) return dispatch_axis<0, R>(B, mode, shape, st);
  if (axis == 1) return dispatch_axis<1, R>(B, mode, shape, st);
  return dispatch_axis<2, R>(B, mode, shape, st);
}

// ---- batches (ocfd_batch): periodic boxes with uniform metrics, no hybrid --
template <int AXIS, typename R, int SHAPE>
hipError_t batch_axis(const BoxSet &S, int sch, bool ch, int mode, int ntiles, hipStream_t st) {
  constexpr int TL = shape_tl<AXIS>(SHAPE);
  constexpr int TW = NTHREADS / TL;
  const size_t lds = ch ? sweep_lds<TL, TW, true, false, R>(4) : sweep_lds<TL, TW, false, false, R>(4);
#define OCFD_BATCH_SWEEP(SCH_, CH_)                                                                         \
  hipLaunchKernelGGL((k_inviscid_batch<AXIS, SCH_, CH_, TL, TW, R>), dim3((unsigned)ntiles), dim3(NTHREADS), lds, st, \
                     S, mode)
  switch (sch) {
    case SCH_WENO5: if (ch) OCFD_BATCH_SWEEP(SCH_WENO5, true); else OCFD_BATCH_SWEEP(SCH_WENO5, false); break;
    case SCH_WENO7: if (ch) OCFD_BATCH_SWEEP(SCH_WENO7, true); else OCFD_BATCH_SWEEP(SCH_WENO7, false); break;
    case SCH_OMP6: if (ch) OCFD_BATCH_SWEEP(SCH_OMP6, true); else OCFD_BATCH_SWEEP(SCH_OMP6, false); break;
    case SCH_UD7L: if (ch) OCFD_BATCH_SWEEP(SCH_UD7L, true); else OCFD_BATCH_SWEEP(SCH_UD7L, false); break;
    default: return hipErrorInvalidValue;
  }
#undef OCFD_BATCH_SWEEP
  return hipGetLastError();
}

// the march of a batch group (no hybrid in batches)
template <int AXIS, typename R>
hipError_t batch_march(const BoxSet &S, int sch, int mode, int shape, int ntiles, hipStream_t st) {
  {
    const int sl = march_segment(AXIS, shape);
#define OCFD_BATCH_MARCH(SCH_)                                                                              \
  hipLaunchKernelGGL((k_inviscid_march_batch<AXIS, SCH_, false, R>), dim3((unsigned)ntiles), dim3(64 * MARCH_WAVES), \
                     0, st, S, mode, sl)
    switch (sch) {
      case SCH_WENO5: OCFD_BATCH_MARCH(SCH_WENO5); break;
      case SCH_WENO7: OCFD_BATCH_MARCH(SCH_WENO7); break;
      case SCH_OMP6: OCFD_BATCH_MARCH(SCH_OMP6); break;
      case SCH_UD7L: OCFD_BATCH_MARCH(SCH_UD7L); break;
      default: return hipErrorInvalidValue;
    }
#undef OCFD_BATCH_MARCH
    return hipGetLastError();
  }
}

// shape: the effective shape of the group (sweep_shape_effective)
template <int AXIS, typename R>
hipError_t batch_shape(const BoxSet &S, int sch, bool ch, int mode, int shape, int ntiles, hipStream_t st) {
  if (shape >= 3) return ch ? hipErrorInvalidValue : batch_march<AXIS, R>(S, sch, mode, shape, ntiles, st);
  if (shape == 1) return batch_axis<AXIS, R, 1>(S, sch, ch, mode, ntiles, st);
  if (shape == 2) return batch_axis<AXIS, R, 2>(S, sch, ch, mode, ntiles, st);
  return batch_axis<AXIS, R, 0>(S, sch, ch, mode, ntiles, st);
}

template <typename R>
hipError_t launch_sweep_batch(const BoxSet &S, int axis, int sch, bool ch, int mode, int shape, int ntiles,
                              hipStream_t st) {
  if (axis == 0) return batch_shape<0, R>(S, sch, ch, mode, shape, ntiles, st);
  if (axis == 1) return batch_shape<1, R>(S, sch, ch, mode, shape, ntiles, st);
  return batch_shape<2, R>(S, sch, ch, mode, shape, ntiles, st);
}

}  // namespace

}  // namespace ocfd
