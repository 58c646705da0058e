// ocfd_march.h -- the inviscid sweeps as per-lane marches.
//
// Same arithmetic as the tile sweep of ocfd_sweep.h (du_inviscous,
// OCFD_inviscous.f90:53-203; the linear faces of du_inviscous_Character,
// OCFD_inviscous_character.f90:190-194), laid out differently: each lane owns
// one line of a periodic axis with uniform metrics and marches along it over
// a segment of SL cells.  For y and z the 64 lanes of a wavefront hold 64
// consecutive i, so every load and store is one contiguous run; for x each
// lane marches a short run of i (march_tile_x).  The split fluxes of the last 8 points live in a register
// queue, rotated by unrolling the march 8 points at a time, so there is no
// LDS tile, no barrier, and each point's primitives are read once per
// segment (plus the 8-point overlap of neighbouring segments).  Face and
// cell results are the tile sweep's bit for bit: the same expressions in the
// same order.
//
// Schemes: every non-characteristic scheme, and the hybrid scheme with
// characteristic projection (its linear faces skip the projection; its WENO
// faces come precomputed from the face list).  The characteristic WENO/OMP6
// schemes keep the tile sweep.
#pragma once
#include "ocfd_kernels.h"

namespace ocfd {

namespace {

constexpr int MARCH_WAVES = 4;  // wavefronts per workgroup: consecutive segments of the same lines

template <int SCH, bool CHAR>
constexpr bool march_supported() {
  return !CHAR || SCH == SCH_HYBRID;
}

// One lane's line: line index i (i for y and z, j for x), third index c,
// cells s0..s1 of the axis.  `act` is false for lanes past the block (their
// loads use line 1, nothing is stored).
template <int AXIS, int SCH, bool CHAR, typename R>
__device__ __forceinline__ void march_line(const Block &B, const int i, const int c, const int s0, const int s1,
                                           const int mode, const bool act) {
  typedef TileIndex<AXIS> TI;
  constexpr bool HYB = (SCH == SCH_HYBRID);
  constexpr int NH = CHAR ? 5 : 10;  // face values: h(5), or h+(5) and h-(5)
  const Geom &g = B.g;
  const SchemeParams &S = B.S;
  const int ii = act ? i : 1;
  const R *const pd = B.d.as<R>(), *const pu = B.u.as<R>(), *const pv = B.v.as<R>(), *const pw = B.w.as<R>(),
          *const pT = B.T.as<R>();
  R *const du = reinterpret_cast<R *>(B.du);
  const int q1 = M_AKX1 + 3 * AXIS;
  const bool sw = S.flux_split == 1;
  SplitK<R> K;
  split_k_gamma(R(B.P.gamma), K);
  split_k_metric(R(B.M.u[q1]), R(B.M.u[q1 + 1]), R(B.M.u[q1 + 2]), sw, K);
  const R rma = R(B.P.ma);
  const R hinv = R(B.P.hinv[AXIS]);
  const R ajac = R(B.M.u[M_AJAC]);
  const int64_t hstep = g.hstride(AXIS), istep = g.istride(AXIS);
  const int64_t ni = g.ni;

  R qp[8][5], qm[8][5];  // split fluxes of the last 8 points, slot = point index % 8
  R hprev[NH];           // fluxes of the previous face
  // Loads run ahead of their use: the primitives PD points ahead, the du
  // values the accumulating modes add to DD cells ahead, in rings indexed
  // statically by the 8-point unrolling (PD and DD divide 8; FP64 keeps them
  // short for its registers).  The march covers whole groups of 8 points:
  // loads past the last point repeat it, faces past the last one are
  // computed from those and dropped, and nothing past the segment's last
  // cell is stored.
  constexpr int PD = sizeof(R) == 4 ? 4 : 2;
  constexpr int DD = sizeof(R) == 4 ? 4 : 1;
  const int npts = s1 - s0 + 9;  // points s0-4 .. s1+4 (halo points on the periodic axis)
  const int ncell = s1 - s0 + 1;
  const int64_t h0 = TI::hidx(g, s0 - 4, ii, c);
  const int64_t i0 = TI::iidx(g, s0, ii, c);
  R pr[PD][5];  // primitives of points pi .. pi + PD - 1
  R dr[DD][5];  // du of cells pi - 8 .. pi - 8 + DD - 1 (accumulating modes)
  auto load_point = [&](int p, R *o) {
    const int64_t h = h0 + (int64_t)(p < npts ? p : npts - 1) * hstep;
    o[0] = pd[h];
    o[1] = pu[h];
    o[2] = pv[h];
    o[3] = pw[h];
    o[4] = pT[h];
  };
  auto load_du = [&](int cell, R *o) {  // cell index from s0 (0-based)
    const int64_t q = i0 + (int64_t)(cell < ncell ? cell : ncell - 1) * istep;
#pragma unroll
    for (int m = 0; m < 5; m++) o[m] = du[q + (int64_t)m * ni];
  };
#pragma unroll
  for (int r = 0; r < PD; r++) load_point(r, pr[r]);
  if (mode != INV_ASSIGN)
#pragma unroll
    for (int r = 0; r < DD; r++) load_du(r, dr[r]);
  for (int base = 0; base < npts; base += 8) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int pi = base + j;
      const int slot = j % PD, dslot = j % DD;
      // ---- split fluxes of point s0 - 4 + pi, then the loads PD points ahead ----
      {
        const R dd = pr[slot][0], uu = pr[slot][1], vv = pr[slot][2], ww = pr[slot][3];
        const R cc = sqrt(pr[slot][4]) / rma;  // OCFD_inviscous.f90:67
        if (sw) split_sw<R>(K, cc, dd, uu, vv, ww, qp[j], qm[j]);
        else split_llf<R>(K, cc, dd, uu, vv, ww, qp[j], qm[j]);
      }
      load_point(pi + PD, pr[slot]);
      if (pi < 7) continue;
      // ---- face f = s0 - 1 + (pi - 7): stencil points f-3 .. f+4 ----
      const int f = s0 - 8 + pi;
      // the face-list bits of this face (of face s1 for the dropped faces past it)
      const int64_t fid = HYB ? TI::face_id(g, f <= s1 ? f : s1, ii, c) : 0;
      const unsigned long long bits = HYB ? B.hbits[AXIS][fid >> 6] : 0ull;
      const int bit = (int)(fid & 63);
      R hv[NH];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        R xp[8], xm[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          xp[k] = qp[(j + 1 + k) & 7][m];
          xm[k] = qm[(j + 1 + k) & 7][m];
        }
        if (CHAR) {
          // hybrid linear face: no projection (OCFD_inviscous_character.f90:192-194)
          const R h1 = inner_p<SCH>(S, xp, 1);
          const R h2 = inner_m<SCH>(S, xm, 1);
          hv[m] = h1 + h2;
        } else {
          hv[m] = inner_p<SCH>(S, xp, HYB ? 1 : 0);
          hv[5 + m] = inner_m<SCH>(S, xm, HYB ? 1 : 0);
        }
      }
      if (HYB && ((bits >> bit) & 1ull)) {
        // a listed WENO face (k_weno_faces) replaces the linear one
        const int lslot = B.hgbase[AXIS][fid >> 6] + __popcll(bits & ((1ull << bit) - 1ull));
        const double *wf = B.wflux[AXIS] + (int64_t)lslot * NH;
#pragma unroll
        for (int q = 0; q < NH; q++) hv[q] = R(wf[q]);
      }
      // ---- cell f: flux difference into du ----
      if (pi >= 8) {
        const int cell = pi - 8;
        if (act && cell < ncell) {
          const int64_t qc = i0 + (int64_t)cell * istep;
#pragma unroll
          for (int m = 0; m < 5; m++) {
            R val;
            if (!CHAR) val = (hv[m] - hprev[m] + hv[5 + m] - hprev[5 + m]) * hinv;
            else val = (hv[m] - hprev[m]) * hinv;
            const int64_t q = qc + (int64_t)m * ni;
            if (mode == INV_ASSIGN) {
              du[q] = val;
            } else if (mode == INV_ACCUM) {
              du[q] = dr[dslot][m] + val;
            } else {
              R rr = -(dr[dslot][m] + val) * ajac;
              if (B.P.if_mass_force == 1 && m > 0) {
                // OCFD_inviscous.f90:189-201
                const int64_t h = TI::hidx(g, f, ii, c);
                const double *mf = B.P.mass_force;
                if (m < 4) rr = rr + R(B.d[h] * mf[m - 1]);
                else rr = rr + R(B.d[h] * (B.u[h] * mf[0] + B.v[h] * mf[1] + B.w[h] * mf[2]));
              }
              du[q] = rr;
            }
          }
        }
        if (mode != INV_ASSIGN) load_du(cell + DD, dr[dslot]);
      }
#pragma unroll
      for (int q = 0; q < NH; q++) hprev[q] = hv[q];
    }
  }
}

// workgroup b of the march over block B: wavefront w takes segment
// sg * MARCH_WAVES + w of 64 consecutive lines
template <int AXIS, int SCH, bool CHAR, typename R>
__device__ __forceinline__ void march_tile(const Block &B, int b, const int mode, const int sl) {
  const Geom &g = B.g;
  const int n = g.n(AXIS);
  const int nseg = (n + sl - 1) / sl;
  const int nsg = (nseg + MARCH_WAVES - 1) / MARCH_WAVES;
  const int niw = (g.nx + 63) / 64;
  const int iw = b % niw;
  b /= niw;
  const int sg = b % nsg;
  const int c = b / nsg + 1;
  const int seg = sg * MARCH_WAVES + (int)(threadIdx.x >> 6);
  if (seg >= nseg) return;
  const int i = iw * 64 + (int)(threadIdx.x & 63) + 1;
  const int s0 = seg * sl + 1;
  const int s1 = s0 + sl - 1 < n ? s0 + sl - 1 : n;
  march_line<AXIS, SCH, CHAR, R>(B, i, c, s0, s1, mode, i <= g.nx);
}

// x: the lines run along i, so a lane takes one run of sl consecutive cells
// of one (j, k) line and consecutive lanes take consecutive runs (a wavefront
// covers 64 runs: whole lines when nx <= 64 sl); the lane's loads are strided
// by sl points across the wavefront and served from L1/L2 after the first.
template <int SCH, bool CHAR, typename R>
__device__ __forceinline__ void march_tile_x(const Block &B, int b, const int mode, const int sl) {
  const Geom &g = B.g;
  const int nseg = (g.nx + sl - 1) / sl;
  const int64_t lane = (int64_t)b * (64 * MARCH_WAVES) + threadIdx.x;
  const int seg = (int)(lane % nseg);
  const int64_t line = lane / nseg;
  const bool act = line < (int64_t)g.ny * g.nz;
  const int j = act ? (int)(line % g.ny) + 1 : 1, k = act ? (int)(line / g.ny) + 1 : 1;
  const int s0 = seg * sl + 1;
  const int s1 = s0 + sl - 1 < g.nx ? s0 + sl - 1 : g.nx;
  march_line<0, SCH, CHAR, R>(B, j, k, s0, s1, mode, act);
}

template <int AXIS, int SCH, bool CHAR, typename R>
__global__ __launch_bounds__(64 * MARCH_WAVES) void k_inviscid_march(const Block B, const int mode, const int sl) {
  if (AXIS == 0) march_tile_x<SCH, CHAR, R>(B, xcd_tile(blockIdx.x, gridDim.x), mode, sl);
  else march_tile<AXIS, SCH, CHAR, R>(B, xcd_tile(blockIdx.x, gridDim.x), mode, sl);
}

// a batch of periodic boxes with uniform metrics: the march workgroups of all
// boxes of the launch
template <int AXIS, int SCH, bool CHAR, typename R>
__global__ __launch_bounds__(64 * MARCH_WAVES) void k_inviscid_march_batch(const BoxSet S, const int mode,
                                                                            const int sl) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int j = S.entry(t);
  if (AXIS == 0) march_tile_x<SCH, CHAR, R>(S.blk[S.ids[j]], t - S.off[j], mode, sl);
  else march_tile<AXIS, SCH, CHAR, R>(S.blk[S.ids[j]], t - S.off[j], mode, sl);
}

}  // namespace

}  // namespace ocfd
