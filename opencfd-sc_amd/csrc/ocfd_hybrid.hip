// ocfd_hybrid.hip -- the nonlinear faces of the hybrid scheme, gathered.
//
// The hybrid scheme (OCFD_Schemes.f90:319-358, OCFD_SetHybrid.f90:2-15)
// reconstructs most faces with the linear UD7L blend and a few with WENO7 or
// WENO5, chosen per face from the sensor Rhybrid that comput_Rhybrid forms
// once per time step.  Computed inside the sweep tiles, the few WENO faces
// leave most lanes of a tile idle; here they are
//   k_hyb_list    listed once per step, per sweep axis (a range of the
//                 list reserved per workgroup, faces in order inside it),
//                 with a face -> slot map for the sweeps;
//   k_weno_faces  computed densely before each stage's sweeps: batches of
//                 listed faces per workgroup, split fluxes of the 8 stencil
//                 points in LDS, then one (face, characteristic row) -- or
//                 (face, component) without projection -- per lane.
// The sweep kernels read the listed faces' fluxes by slot.  Every flux is the
// same expression as in the sweep (flux_charteric,
// OCFD_inviscous_character.f90:190-290; OCFD2d_flux1/2,
// OCFD_inviscous.f90:210-361), so results are unchanged bit for bit.
#include "ocfd_kernels.h"

namespace ocfd {

namespace {

constexpr int FB = 51;   // listed faces per workgroup batch: 5 * FB rows fill the 256 lanes
constexpr int SFS = 81;  // LDS stride of one face's split fluxes (8 points x 10, padded so
                         // that the faces a wavefront reads fall in different banks)

// A workgroup lists LPT * NT consecutive faces of the axis: scheme index from
// the sensor, the WENO faces appended to the axis list in face order, and the
// face -> slot map.  The list position is reserved with one atomic per
// workgroup; inside it the slots follow the face index.
constexpr int LPT = 16;  // faces per thread
template <int AXIS>
__global__ __launch_bounds__(NT) void k_hyb_list(const Block B) {
  typedef TileIndex<AXIS> TI;
  __shared__ int soff[LPT][NT / 64];
  __shared__ int sbase;
  const Geom &g = B.g;
  const SchemeParams &S = B.S;
  const int n = g.n(AXIS);
  const int64_t nfa = TI::nfaces(g);
  const int64_t f0 = (int64_t)blockIdx.x * (NT * LPT) + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned wbits = 0, hbits = 0;  // bit it: face f0 + it * NT is WENO / WENO7-class
  for (int it = 0; it < LPT; it++) {
    const int64_t fid = f0 + (int64_t)it * NT;
    bool weno = false;
    if (fid < nfa) {
      int s, w, c;
      TI::face_of(g, fid, s, w, c);
      // set_hybrid_scheme: the larger sensor value of the two cells of the face
      const double r0 = dmax2(B.rhyb[TI::hidx(g, s, w, c)], B.rhyb[TI::hidx(g, s + 1, w, c)]);
      const int scm = hybrid_index(r0, S.hyb2, S.hyb3);
      // the characteristic path closes non-periodic block edges with the
      // non-reflecting face whatever the scheme index
      const bool closure = S.is_char && ((S.bidx[AXIS][0] == 1 && S.sb[2 * AXIS] == 0 && s == 0) ||
                                         (S.bidx[AXIS][1] == 1 && S.sb[2 * AXIS + 1] == 0 && s == n));
      weno = scm >= 2 && !closure;
      if (weno) wbits |= 1u << it;
      if (scm == 3) hbits |= 1u << it;
    }
    const unsigned long long mask = __ballot(weno);
    if (lane == 0) soff[it][wave] = __popcll(mask);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int it = 0; it < LPT; it++)
      for (int q = 0; q < NT / 64; q++) {
        const int c = soff[it][q];
        soff[it][q] = t;
        t += c;
      }
    sbase = t ? atomicAdd(&B.hcount[AXIS], t) : 0;
  }
  __syncthreads();
  const int base = sbase;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int it = 0; it < LPT; it++) {
    const int64_t fid = f0 + (int64_t)it * NT;
    const bool weno = (wbits >> it) & 1u;
    const unsigned long long mask = __ballot(weno);
    const int slot = base + soff[it][wave] + __popcll(mask & below);
    // the wavefront's 64 faces are one aligned group of the face map
    if (lane == 0 && fid < nfa) {
      B.hbits[AXIS][fid >> 6] = mask;
      B.hgbase[AXIS][fid >> 6] = base + soff[it][wave];
    }
    if (weno) B.hlist[AXIS][slot] = (int)(fid * 2 + ((hbits >> it) & 1u));
  }
}

// Fluxes of the listed faces of one axis.  A fixed grid strides over batches
// of FB faces; the list length is read from the device counter.  RR is the
// real type of the splitting and the WENO reconstruction: double, or float
// in the mixed-precision form, where the characteristic basis and the
// projections onto and back from it stay FP64.
template <int AXIS, bool CHAR, bool FAST, typename RR>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(RR) == 4 ? 4 : 1)))
void k_weno_faces(const Block B) {
  typedef TileIndex<AXIS> TI;
  constexpr int NH = CHAR ? 5 : 10;
  __shared__ RR sF[FB][SFS];        // fp(1:5), fm(1:5) at stencil points s-3 .. s+4: [p * 10 + m]
  __shared__ double sQ[FB][2][7];   // u, v, w, c, A1..A3 at points s, s+1
  __shared__ double sB[FB][21];     // characteristic basis of the face
  __shared__ double sG[FB][5];      // characteristic fluxes h0(1:5)
  const Geom &g = B.g;
  const SchemeParams &S = B.S;
  const int n = g.n(AXIS);
  const int tid = threadIdx.x;
  const double gamma = B.P.gamma, rma = B.P.ma;
  const int q1 = M_AKX1 + 3 * AXIS;
  const int count = B.hcount[AXIS];
  const int nbatch = (count + FB - 1) / FB;
  const int *list = B.hlist[AXIS];
  double *out = B.wflux[AXIS];
  const bool sw = S.flux_split == 1;
  // batches in one contiguous range per XCD (workgroup b runs on XCD
  // b % NXCD), strided over that XCD's workgroups: list-neighbouring faces,
  // whose y / z stencils share rows, then meet in the same L2
  const int nw = (int)gridDim.x / NXCD, xcd = (int)blockIdx.x % NXCD;
  const bool split = (int)gridDim.x % NXCD == 0 && nw > 0;
  const int blo = split ? (int)((int64_t)nbatch * xcd / NXCD) : 0;
  const int bhi = split ? (int)((int64_t)nbatch * (xcd + 1) / NXCD) : nbatch;
  const int b0 = split ? blo + (int)blockIdx.x / NXCD : (int)blockIdx.x;
  const int bstep = split ? nw : (int)gridDim.x;
  // The FP32 form with uniform metrics loads the primitives of its next
  // batch's stencil points right after splitting the current batch's, so the
  // loads run during the current batch's projection and reconstruction (the
  // FP64 form has no registers to spare for them).
  constexpr bool PF = FAST && sizeof(RR) == 4;
  constexpr int NPI = (8 * FB + NT - 1) / NT;  // stencil points per lane and batch
  float pv[PF ? NPI : 1][5];
  auto fetch = [&](int bt) {
    const int fn0 = bt * FB;
    const int nfn = count - fn0 < FB ? count - fn0 : FB;
#pragma unroll
    for (int r = 0; r < NPI; r++) {
      const int it = tid + r * NT;
      if (bt < bhi && it < nfn * 8) {
        const int fb = it >> 3, p = it & 7;
        int s, w, c;
        TI::face_of(g, list[fn0 + fb] >> 1, s, w, c);
        const int64_t h = TI::hidx(g, s - 3 + p, w, c);
        pv[r][0] = B.d.template as<float>()[h];
        pv[r][1] = B.u.template as<float>()[h];
        pv[r][2] = B.v.template as<float>()[h];
        pv[r][3] = B.w.template as<float>()[h];
        pv[r][4] = B.T.template as<float>()[h];
      }
    }
  };
  if constexpr (PF) fetch(b0);
  for (int batch = b0; batch < bhi; batch += bstep) {
    const int f0 = batch * FB;
    const int nf = count - f0 < FB ? count - f0 : FB;
    if constexpr (PF) {
      // split fluxes at the stencil points from the prefetched primitives
      // (the same values the loop below reads: FP32 storage, widened)
#pragma unroll
      for (int r = 0; r < NPI; r++) {
        const int it = tid + r * NT;
        if (it < nf * 8) {
          const int fb = it >> 3, p = it & 7;
          const double dd = pv[r][0], uu = pv[r][1], vv = pv[r][2], ww = pv[r][3], Tt = pv[r][4];
          const double A1 = B.M.u[q1], A2 = B.M.u[q1 + 1], A3 = B.M.u[q1 + 2];
          SplitK<RR> K;
          split_k_gamma(RR(gamma), K);
          split_k_metric(RR(A1), RR(A2), RR(A3), sw, K);
          const RR ccr = sqrt(RR(Tt)) / RR(rma);
          RR fp[5], fm[5];
          if (sw) split_sw<RR>(K, ccr, RR(dd), RR(uu), RR(vv), RR(ww), fp, fm);
          else split_llf<RR>(K, ccr, RR(dd), RR(uu), RR(vv), RR(ww), fp, fm);
#pragma unroll
          for (int m = 0; m < 5; m++) {
            sF[fb][p * 10 + m] = fp[m];
            sF[fb][p * 10 + 5 + m] = fm[m];
          }
          if (CHAR && (p == 3 || p == 4)) {
            double *q = sQ[fb][p - 3];
            q[0] = uu; q[1] = vv; q[2] = ww; q[3] = sqrt(Tt) / rma;
            q[4] = A1; q[5] = A2; q[6] = A3;
          }
        }
      }
      fetch(batch + bstep);
    }
    // split fluxes at the stencil points (OCFD_inviscous.f90:64-70, split_*)
    for (int it = tid; !PF && it < nf * 8; it += NT) {
      const int fb = it >> 3, p = it & 7;
      int s, w, c;
      TI::face_of(g, list[f0 + fb] >> 1, s, w, c);
      const int64_t h = TI::hidx(g, s - 3 + p, w, c);
      const double dd = B.d[h], uu = B.u[h], vv = B.v[h], ww = B.w[h], Tt = B.T[h];
      const double cc = sqrt(Tt) / rma;
      const double A1 = (FAST || B.M.uniform) ? B.M.u[q1] : B.M.p[q1][h];
      const double A2 = (FAST || B.M.uniform) ? B.M.u[q1 + 1] : B.M.p[q1 + 1][h];
      const double A3 = (FAST || B.M.uniform) ? B.M.u[q1 + 2] : B.M.p[q1 + 2][h];
      // the splitting's constants per point (held across the batch loop they
      // would cost registers, i.e. occupancy)
      SplitK<RR> K;
      split_k_gamma(RR(gamma), K);
      split_k_metric(RR(A1), RR(A2), RR(A3), sw, K);
      const RR ccr = sqrt(RR(Tt)) / RR(rma);  // OCFD_inviscous.f90:67, in the splitting's type
      RR fp[5], fm[5];
      if (sw) split_sw<RR>(K, ccr, RR(dd), RR(uu), RR(vv), RR(ww), fp, fm);
      else split_llf<RR>(K, ccr, RR(dd), RR(uu), RR(vv), RR(ww), fp, fm);
#pragma unroll
      for (int m = 0; m < 5; m++) {
        sF[fb][p * 10 + m] = fp[m];
        sF[fb][p * 10 + 5 + m] = fm[m];
      }
      if (CHAR && (p == 3 || p == 4)) {
        double *q = sQ[fb][p - 3];
        q[0] = uu; q[1] = vv; q[2] = ww; q[3] = cc;
        q[4] = A1; q[5] = A2; q[6] = A3;
      }
    }
    __syncthreads();
    if (CHAR) {
      // basis from the face-averaged state (OCFD_inviscous_character.f90:200-236)
      for (int fb = tid; fb < nf; fb += NT) {
        const double *qa = sQ[fb][0], *qb = sQ[fb][1];
        const double u1 = (qa[0] + qb[0]) * 0.5, v1 = (qa[1] + qb[1]) * 0.5;
        const double w1 = (qa[2] + qb[2]) * 0.5, c1 = (qa[3] + qb[3]) * 0.5;
        const double a1 = (qa[4] + qb[4]) * 0.5, a2 = (qa[5] + qb[5]) * 0.5, a3 = (qa[6] + qb[6]) * 0.5;
        CharBasis Bs;
        char_basis(gamma, u1, v1, w1, c1, a1, a2, a3, Bs);
        double *o = sB[fb];
        o[0] = Bs.u1; o[1] = Bs.v1; o[2] = Bs.w1; o[3] = Bs.c1; o[4] = Bs.n1; o[5] = Bs.n2; o[6] = Bs.n3;
        o[7] = Bs.l1; o[8] = Bs.l2; o[9] = Bs.l3; o[10] = Bs.m1; o[11] = Bs.m2; o[12] = Bs.m3;
        o[13] = Bs.un; o[14] = Bs.ul; o[15] = Bs.um; o[16] = Bs.V2; o[17] = Bs.KK; o[18] = Bs.KK1;
        o[19] = Bs.X1; o[20] = Bs.H;
      }
      __syncthreads();
      auto basis = [&](int fb) {
        const double *o = sB[fb];
        CharBasis Bs;
        Bs.u1 = o[0]; Bs.v1 = o[1]; Bs.w1 = o[2]; Bs.c1 = o[3]; Bs.n1 = o[4]; Bs.n2 = o[5]; Bs.n3 = o[6];
        Bs.l1 = o[7]; Bs.l2 = o[8]; Bs.l3 = o[9]; Bs.m1 = o[10]; Bs.m2 = o[11]; Bs.m3 = o[12];
        Bs.un = o[13]; Bs.ul = o[14]; Bs.um = o[15]; Bs.V2 = o[16]; Bs.KK = o[17]; Bs.KK1 = o[18];
        Bs.X1 = o[19]; Bs.H = o[20];
        return Bs;
      };
      // one characteristic row per lane: project, reconstruct (flux_x1/x2)
      for (int it = tid; it < nf * 5; it += NT) {
        const int fb = it / 5, r = it - 5 * fb;
        const int code = list[f0 + fb];
        const int scm = (code & 1) ? 3 : 2;
        int s, w, c;
        TI::face_of(g, code >> 1, s, w, c);
        const CharBasis Bs = basis(fb);
        double sr5[5];
        char_S_row(Bs, r, sr5);
        // stencils of the hybrid's WENO7 / WENO5 faces (Ka1_H2.. / Ka1_H3..)
        const int a1h = scm == 2 ? -3 : -2, b1h = scm == 2 ? 3 : 2;
        const int a2h = scm == 2 ? -2 : -1, b2h = scm == 2 ? 4 : 3;
        // projection in FP64, reconstruction in RR
        double xp[8], xm[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          xp[k] = sr5[0] * double(sF[fb][k * 10]);
          xm[k] = sr5[0] * double(sF[fb][k * 10 + 5]);
        }
        // (FP64 path: the reference's two roundings; the mixed form's FP64
        // projection fuses them)
        auto acc = [](double a, double b, double c) { return sizeof(RR) == 8 ? madd(a, b, c) : fma(a, b, c); };
#pragma unroll
        for (int cmp = 1; cmp < 5; cmp++) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            xp[k] = acc(sr5[cmp], double(sF[fb][k * 10 + cmp]), xp[k]);
            xm[k] = acc(sr5[cmp], double(sF[fb][k * 10 + 5 + cmp]), xm[k]);
          }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
          if (k < a1h + 3 || k > b1h + 3) xp[k] = 0.0;
          if (k < a2h + 3 || k > b2h + 3) xm[k] = 0.0;
        }
        auto recon = [&](const auto *yp, const auto *ym) {
          const auto h1 = FAST ? inner_p<SCH_HYBRID>(S, yp, scm) : face_p_char<SCH_HYBRID>(S, yp, s, n, AXIS, scm);
          const auto h2 = FAST ? inner_m<SCH_HYBRID>(S, ym, scm) : face_m_char<SCH_HYBRID>(S, ym, s, n, AXIS, scm);
          return double(h1 + h2);
        };
        if constexpr (sizeof(RR) == 8) {
          sG[fb][r] = recon(xp, xm);
        } else {
          RR yp[8], ym[8];
#pragma unroll
          for (int k = 0; k < 8; k++) {
            yp[k] = RR(xp[k]);
            ym[k] = RR(xm[k]);
          }
          sG[fb][r] = recon(yp, ym);
        }
      }
      __syncthreads();
      // back to conservative components (S1 * h0): one (face, component)
      // per lane, so the stores of a wavefront are contiguous
      for (int it = tid; it < nf * 5; it += NT) {
        const int fb = it / 5, m = it - 5 * fb;
        const CharBasis Bs = basis(fb);
        double hh[5];
        char_back(Bs, sG[fb], hh);
        out[(int64_t)(f0 + fb) * NH + m] = hh[m];
      }
    } else {
      // one component per lane: flux+ and flux- of the face
      for (int it = tid; it < nf * 5; it += NT) {
        const int fb = it / 5, m = it - 5 * fb;
        const int code = list[f0 + fb];
        const int scm = (code & 1) ? 3 : 2;
        int s, w, c;
        TI::face_of(g, code >> 1, s, w, c);
        RR xp[8], xm[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          xp[k] = sF[fb][k * 10 + m];
          xm[k] = sF[fb][k * 10 + 5 + m];
        }
        double *o = out + (int64_t)(f0 + fb) * NH;
        o[m] = double(FAST ? inner_p<SCH_HYBRID>(S, xp, scm) : face_p_nonchar<SCH_HYBRID>(S, xp, s, n, AXIS, scm));
        o[5 + m] = double(FAST ? inner_m<SCH_HYBRID>(S, xm, scm)
                               : face_m_nonchar<SCH_HYBRID>(S, xm, s, n, AXIS, scm));
      }
    }
    __syncthreads();  // the next batch overwrites the LDS arrays
  }
}

// The grid of the striding face kernel: as many workgroups as the device
// holds at once (CUs x resident workgroups per CU), so that every workgroup
// runs from the start and the batches spread evenly; a larger grid leaves a
// tail of workgroups that start late and run nearly alone.  `cache` holds
// the value per device for one kernel.
template <typename K>
dim3 resident_grid(K kernel, int (&cache)[64]) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cache[dev] == 0) {
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, NT, 0) != hipSuccess || per < 1) per = 1;
    cache[dev] = cus * per;
  }
  return dim3((unsigned)cache[dev]);
}

template <int AXIS, bool CHAR, bool FAST, typename RR>
hipError_t weno_launch(const Block &B, hipStream_t st) {
  static int cache[64] = {};
  hipLaunchKernelGGL((k_weno_faces<AXIS, CHAR, FAST, RR>), resident_grid(k_weno_faces<AXIS, CHAR, FAST, RR>, cache),
                     dim3(NT), 0, st, B);
  return hipGetLastError();
}

template <int AXIS, typename RR>
hipError_t weno_axis_as(const Block &B, hipStream_t st) {
  const bool fast = B.S.iperiodic[AXIS] == 1 && B.M.uniform;
  if (B.S.is_char)
    return fast ? weno_launch<AXIS, true, true, RR>(B, st) : weno_launch<AXIS, true, false, RR>(B, st);
  return fast ? weno_launch<AXIS, false, true, RR>(B, st) : weno_launch<AXIS, false, false, RR>(B, st);
}

// FP32 splitting and reconstruction in the mixed-precision form
template <int AXIS>
hipError_t weno_axis(const Block &B, hipStream_t st) {
  if (B.prec == OCFD_PREC_MIXED_CODE) return weno_axis_as<AXIS, float>(B, st);
  return weno_axis_as<AXIS, double>(B, st);
}

}  // namespace

hipError_t launch_hybrid_lists(const Block &B, hipStream_t st) {
  const Geom &g = B.g;
  constexpr int64_t FPB = NT * LPT;
  hipLaunchKernelGGL(k_hyb_list<0>, dim3((unsigned)((TileIndex<0>::nfaces(g) + FPB - 1) / FPB)), dim3(NT), 0, st, B);
  hipLaunchKernelGGL(k_hyb_list<1>, dim3((unsigned)((TileIndex<1>::nfaces(g) + FPB - 1) / FPB)), dim3(NT), 0, st, B);
  hipLaunchKernelGGL(k_hyb_list<2>, dim3((unsigned)((TileIndex<2>::nfaces(g) + FPB - 1) / FPB)), dim3(NT), 0, st, B);
  return hipGetLastError();
}

hipError_t launch_weno_faces(const Block &B, int axis, hipStream_t st) {
  if (axis == 0) return weno_axis<0>(B, st);
  if (axis == 1) return weno_axis<1>(B, st);
  return weno_axis<2>(B, st);
}

}  // namespace ocfd
