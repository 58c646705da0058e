// ocfd_visc_fused.h -- the viscous right-hand side fused into the stage
// update for periodic blocks with uniform metrics, CD6 (the FP32 instance
// serves the mixed-precision form).
//
// k_visc_flux writes the 12 viscous fluxes Ev of every cell and
// k_visc_div_update reads them back: 2 x 12 values per cell and stage
// through HBM.  k_visc_fused computes Ev where the divergence needs it and
// keeps it on the chip.  A workgroup owns a 32 x 8 tile of (i, j) columns
// and marches k over a segment; at plane p it forms Ev at
//   - its own 256 points (Ev1, Ev2, Ev3),
//   - the 2 x 3 x 8 points beside the tile in i (Ev1) and the 32 x 2 x 3
//     points beside it in j (Ev2),
// one or two points per lane, from the derivatives of u, v, w, T: along i
// and j from an LDS tile of plane p (44 x 20 points), along k from a
// register queue of the 7 planes p-3 .. p+3 of each of the lane's points.
// Ev1 and Ev2 go to LDS tiles; d/dxi Ev1 + d/deta Ev2 at the own points go
// to a 4-plane register queue, Ev3 to a 7-plane LDS ring, so that at plane p the
// divergence and the stage update of plane p - 3 are complete.  Every value
// is the same expression, in the same order, as in the unfused kernels
// (ev_point, dxyz0_point, the stage update of visc_div_update_tile), so the
// results are bit for bit theirs.
//
// The march reads the stage-start u, v, w, T of neighbouring columns and of
// planes beyond its segment while other workgroups update theirs, so the new
// u, v, w go to a second set of buffers (Block::u2, v2, w2; T to T2 as
// always) and the host swaps the sets after the stage; d (not read by the
// viscous terms) is updated in place.
#pragma once
#include "ocfd_visc.h"

namespace ocfd {

namespace {

constexpr int FTX = 32, FTY = 8, FR = 3;                   // tile, CD6 radius
constexpr int FPX = FTX + 4 * FR, FPY = FTY + 4 * FR;      // primitive plane tile 44 x 20
constexpr int FNP = FPX * FPY;
constexpr int FNPR = (FNP + NT - 1) / NT;                  // its points per lane (4)
constexpr int FE1X = FTX + 2 * FR, FE2Y = FTY + 2 * FR;    // Ev1 tile 38 x 8, Ev2 tile 32 x 14
constexpr int FNX1 = 2 * FR * FTY, FNX2 = FTX * 2 * FR;    // halo points: 48 beside in i, 192 beside in j
constexpr int FQ = 2 * FR + 1;
constexpr int FKSEG_MAX = 64;

// workgroups and planes per workgroup of the fused march over block g
inline int fused_kseg(const Geom &g) {
  const int64_t cols = (int64_t)(g.nx / FTX) * (g.ny / FTY);
  const int64_t k = cols * g.nz / 2048;
  return (int)(k < 16 ? (g.nz < 16 ? g.nz : 16) : (k > FKSEG_MAX ? FKSEG_MAX : k));
}
inline unsigned fused_blocks(const Geom &g, int kseg) {
  return (unsigned)((g.nx / FTX) * (g.ny / FTY) * ((g.nz + kseg - 1) / kseg));
}
// planes per workgroup of a batch's fused march: the batch's segment length
// (BATCH_KSEGS), at least 16 (each segment recomputes 2 x 3 planes of Ev3
// around it), at most the box
__host__ __device__ inline int fused_batch_kseg(int nz, int kseg) {
  const int k = kseg < 16 ? 16 : kseg;
  return k < nz ? k : nz;
}

template <typename RT>
__device__ __forceinline__ void visc_fused_tile(const Block &B, const int krk, const int refresh, const int kseg,
                                                int *neg, int tile) {
  const Geom &g = B.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int ntx = nx / FTX, nty = ny / FTY;
  const int bx = tile % ntx;
  tile /= ntx;
  const int by = tile % nty;
  const int bz = tile / nty;
  const int t = threadIdx.x, tx = t % FTX, ty = t / FTX;
  const int i0 = bx * FTX + 1, j0 = by * FTY + 1;
  const int kb = bz * kseg + 1, ke = kb + kseg - 1 < nz ? kb + kseg - 1 : nz;
  const int i = i0 + tx, j = j0 + ty;  // the lane's own column (every tile is whole)
  auto wrapx = [&](int a) { return a < 1 ? a + nx : (a > nx ? a - nx : a); };
  auto wrapy = [&](int a) { return a < 1 ? a + ny : (a > ny ? a - ny : a); };
  auto wrapz = [&](int a) { return a < 1 ? a + nz : (a > nz ? a - nz : a); };
  const int64_t ni = g.ni, skz = g.istride(2);
  const int64_t plane = (int64_t)g.sx * g.sy;
  const int64_t col = g.ii(i, j, 1);
  if (halted(B)) {
    // a whole-step run stopped at too many T <= 0 cells: the state stays and
    // is carried into the buffers the host swaps in after the stage
    for (int k = kb; k <= ke; k++) {
      const int64_t h = g.hi(i, j, k);
      B.T2.as<RT>()[h] = B.T.as<RT>()[h];
      B.u2.as<RT>()[h] = B.u.as<RT>()[h];
      B.v2.as<RT>()[h] = B.v.as<RT>()[h];
      B.w2.as<RT>()[h] = B.w.as<RT>()[h];
    }
    return;
  }

  __shared__ RT sP[2][4 * FNP];                        // u, v, w, T of one plane: [4][FPY][FPX]
  __shared__ RT sE[4 * (FE1X * FTY + FTX * FE2Y)];     // Ev1(1:4) [4][8][38], then Ev2(1:4) [4][14][32]
  // Ev3(1:4) of each lane's own point at planes p-6 .. p, slot = plane mod 7
  // (read only by the lane that writes it)
  __shared__ RT sE3[4 * FQ * NT];
  RT *const sE1 = sE, *const sE2 = sE + 4 * FE1X * FTY;
  const RT *F[4] = {B.u.as<RT>(), B.v.as<RT>(), B.w.as<RT>(), B.T.as<RT>()};

  // ---- the plane tile: one share of its points per lane ----
  int toff[FNPR];
#pragma unroll
  for (int r = 0; r < FNPR; r++) {
    const int e = t + r * NT < FNP ? t + r * NT : FNP - 1;
    const int ly = e / FPX, lx = e - ly * FPX;
    const int ii = wrapx(i0 - 2 * FR + lx), jj = wrapy(j0 - 2 * FR + ly);
    toff[r] = (ii + LAP - 1) + g.sx * (jj + LAP - 1);
  }
  RT pre[4][FNPR];
  auto fetch_tile = [&](int p) {
    const int64_t base = plane * (wrapz(p) + LAP - 1);
#pragma unroll
    for (int r = 0; r < FNPR; r++)
#pragma unroll
      for (int f = 0; f < 4; f++) pre[f][r] = F[f][base + toff[r]];
  };
  auto store_tile = [&](RT *dst) {
#pragma unroll
    for (int r = 0; r < FNPR; r++) {
      const int e = t + r * NT < FNP ? t + r * NT : FNP - 1;
#pragma unroll
      for (int f = 0; f < 4; f++) dst[f * FNP + e] = pre[f][r];
    }
  };

  // ---- the lane's points: 0 = own column; 1 = a point beside the tile in i
  // (lanes 0..47, Ev1) or in j (lanes 48..239, Ev2); lanes 240..255 repeat
  // their own point and store nothing for it ----
  int px[2], py[2];  // tile-relative position (own tile 0..31 x 0..7)
  px[0] = tx;
  py[0] = ty;
  int kind1 = 0;  // 1: Ev1 point, 2: Ev2 point, 0: none
  if (t < FNX1) {
    const int hx = t % (2 * FR), hy = t / (2 * FR);
    px[1] = hx < FR ? hx - FR : FTX + hx - FR;
    py[1] = hy;
    kind1 = 1;
  } else if (t < FNX1 + FNX2) {
    const int e = t - FNX1;
    const int hy = e / FTX;
    px[1] = e % FTX;
    py[1] = hy < FR ? hy - FR : FTY + hy - FR;
    kind1 = 2;
  } else {
    px[1] = tx;
    py[1] = ty;
  }
  int64_t pcol[2];  // halo'd column offsets of the two points (plane 0 excluded)
#pragma unroll
  for (int q = 0; q < 2; q++) pcol[q] = (wrapx(i0 + px[q]) + LAP - 1) + (int64_t)g.sx * (wrapy(j0 + py[q]) + LAP - 1);
  // u, v, w, T of the two points at planes p-3 .. p+3
  RT zq[2][4][FQ];
#pragma unroll
  for (int q = 0; q < 2; q++)
#pragma unroll
    for (int d = 0; d < FQ; d++) {
      const int64_t o = pcol[q] + plane * (wrapz(kb - FR - FR + d) + LAP - 1);
#pragma unroll
      for (int f = 0; f < 4; f++) zq[q][f][d] = F[f][o];
    }
  RT aq[4][FR + 1];  // d/dxi Ev1 + d/deta Ev2 at the own point, planes p-3 .. p (0 outside the segment)
  fetch_tile(kb - FR);
  store_tile(sP[0]);

  // ---- the stage update's cell values, loaded at the top of the iteration
  // that uses them ----
  const RT *const dus = reinterpret_cast<const RT *>(B.du);
  struct CellVals {
    RT du[5];
    double f[5], fn[5];
  };
  auto fetch_cell = [&](int k) {
    CellVals c;
    const int64_t tt = col + (int64_t)(k - 1) * skz;
#pragma unroll
    for (int m = 0; m < 5; m++) {
      c.du[m] = __builtin_nontemporal_load(&dus[tt + m * ni]);
      c.f[m] = __builtin_nontemporal_load(&B.f[tt + m * ni]);
      c.fn[m] = krk != 1 ? __builtin_nontemporal_load(&B.fn[tt + m * ni]) : 0.0;
    }
    return c;
  };
  const double dt = B.P.dt;
  const double ajac = B.M.u[M_AJAC];
  unsigned long long negmask = 0;  // bit k - kb: T <= 0 at plane k
  const int oe1 = ty * FE1X + tx + FR;    // own point in the Ev1 tile
  const int oe2 = (ty + FR) * FTX + tx;   // own point in the Ev2 tile
  __syncthreads();

#pragma unroll 1
  for (int p = kb - FR; p <= ke + FR; p++) {
    const bool inner = p >= kb && p <= ke;  // the tile's own planes: Ev1 / Ev2 needed
    // next plane's tile and queue entries, and the next update plane's cells
    fetch_tile(p + 1);
    RT zn[2][4];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int64_t o = pcol[q] + plane * (wrapz(p + 1 + FR) + LAP - 1);
#pragma unroll
      for (int f = 0; f < 4; f++) zn[q][f] = F[f][o];
    }
    const int k = p - FR;  // the plane updated in this iteration
    CellVals cv;
    if (k >= kb) cv = fetch_cell(k);

    // ---- Ev at plane p ----
    const RT *cur = sP[(p - (kb - FR)) & 1];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if (q == 1 && !inner) break;  // outside the segment only Ev3 of the own points
      const int c0 = (py[q] + 2 * FR) * FPX + px[q] + 2 * FR;
      RT dx[4], dy[4], dz[4];
#pragma unroll
      for (int f = 0; f < 4; f++) {
        const RT *cf = cur + f * FNP + c0;
        dx[f] = dxyz0_point([&](int d) { return cf[d]; }, 1, nx, B.P.dc[0], false, true);
        dy[f] = dxyz0_point([&](int d) { return cf[d * FPX]; }, 1, ny, B.P.dc[1], false, true);
        dz[f] = dxyz0_point([&](int d) { return zq[q][f][FR + d]; }, 1, nz, B.P.dc[2], false, true);
        __builtin_amdgcn_sched_barrier(0);
      }
      RT e12[12];
      ev_point<true, RT>(B, dx, dy, dz, zq[q][0][FR], zq[q][1][FR], zq[q][2][FR], zq[q][3][FR], 0, e12);
      if (q == 0) {
        const int s3 = (p - (kb - FR)) % FQ;
#pragma unroll
        for (int m = 0; m < 4; m++) sE3[(m * FQ + s3) * NT + t] = e12[8 + m];
        if (inner) {
#pragma unroll
          for (int m = 0; m < 4; m++) {
            sE1[m * FE1X * FTY + oe1] = e12[m];
            sE2[m * FTX * FE2Y + oe2] = e12[4 + m];
          }
        }
      } else if (kind1 == 1) {
        const int o = py[1] * FE1X + px[1] + FR;
#pragma unroll
        for (int m = 0; m < 4; m++) sE1[m * FE1X * FTY + o] = e12[m];
      } else if (kind1 == 2) {
        const int o = (py[1] + FR) * FTX + px[1];
#pragma unroll
        for (int m = 0; m < 4; m++) sE2[m * FTX * FE2Y + o] = e12[4 + m];
      }
    }
    __syncthreads();

    // ---- d/dxi Ev1 + d/deta Ev2 at plane p; divergence and update of plane k ----
    // (the queue moves every iteration, so that its oldest entry is plane k)
#pragma unroll
    for (int m = 0; m < 4; m++) {
      RT a = RT(0);
      if (inner) {
        const RT *c1 = sE1 + m * FE1X * FTY + oe1;
        const RT *c2 = sE2 + m * FTX * FE2Y + oe2;
        const RT d1 = dxyz0_point([&](int d) { return c1[d]; }, 1, nx, B.P.dc[0], false, true);
        const RT d2 = dxyz0_point([&](int d) { return c2[d * FTX]; }, 1, ny, B.P.dc[1], false, true);
        a = d1 + d2;
      }
#pragma unroll
      for (int d = 0; d < FR; d++) aq[m][d] = aq[m][d + 1];
      aq[m][FR] = a;
    }
    if (k >= kb) {
      const int64_t tt = col + (int64_t)(k - 1) * skz;
      const int64_t h = g.hi(i, j, k);
      RT vd[5] = {RT(0), RT(0), RT(0), RT(0), RT(0)};
      // Ev3 of planes k-3 .. k+3 = p-6 .. p in slots (p - 6 + d) mod 7
      const int s0 = (p - (kb - FR) + 1) % FQ;  // slot of plane p - 6
#pragma unroll
      for (int m = 1; m <= 4; m++) {
        const RT *e3 = sE3 + (m - 1) * FQ * NT + t;
        auto at = [&](int d) {
          const int sl = s0 + FR + d;
          return e3[(sl >= FQ ? sl - FQ : sl) * NT];
        };
        const RT d3 = dxyz0_point(at, 1, nz, B.P.dc[2], false, true);
        vd[m] = aq[m - 1][0] + d3;
      }
      double fq[5];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        double dum = double(cv.du[m]);
        if (m > 0) dum = dum + double(vd[m]) * ajac;
        const double f = cv.f[m];
        double r;
        if (krk == 1) {
          __builtin_nontemporal_store(f, &B.fn[tt + m * ni]);
          r = f + dt * dum;
        } else {
          const double ra = krk == 2 ? 3.0 / 4.0 : 1.0 / 3.0;
          const double rb = krk == 2 ? 1.0 / 4.0 : 2.0 / 3.0;
          r = ra * cv.fn[m] + rb * (f + dt * dum);
        }
        fq[m] = r;
      }
      double d, u, v, w, T;
      prims_of(B.P.cv, fq, d, u, v, w, T);
      __builtin_nontemporal_store(RT(d), B.d.as<RT>() + h);
      __builtin_nontemporal_store(RT(u), B.u2.as<RT>() + h);
      __builtin_nontemporal_store(RT(v), B.v2.as<RT>() + h);
      __builtin_nontemporal_store(RT(w), B.w2.as<RT>() + h);
      __builtin_nontemporal_store(RT(T), B.T2.as<RT>() + h);
      negmask |= (T <= 0) ? (1ull << (k - kb)) : 0ull;
      if (refresh && (i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz))
        refresh_of(B.P.cv, d, u, v, w, T, fq);
#pragma unroll
      for (int m = 0; m < 5; m++) __builtin_nontemporal_store(fq[m], &B.f[tt + m * ni]);
    }

    // ---- next plane ----
    store_tile(sP[(p + 1 - (kb - FR)) & 1]);
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int f = 0; f < 4; f++) {
#pragma unroll
        for (int d = 0; d < FQ - 1; d++) zq[q][f][d] = zq[q][f][d + 1];
        zq[q][f][FQ - 1] = zn[q][f];
      }
    __syncthreads();
  }
  // T <= 0 cells of this column, in k order
  while (negmask) {
    const int kk = __builtin_ctzll(negmask);
    negmask &= negmask - 1;
    const int slot = atomicAdd(neg, 1);
    if (slot < 64) neg[1 + slot] = (int)(col + (int64_t)(kb + kk - 1) * skz);
  }
}

template <typename RT>
__global__ __launch_bounds__(NT) void k_visc_fused(const Block B, const int krk, const int refresh, const int kseg,
                                                   int *neg) {
  visc_fused_tile<RT>(B, krk, refresh, kseg, neg, xcd_tile(blockIdx.x, gridDim.x));
}

// Batches of periodic boxes (ocfd_batch): the tiles of the launch's boxes in
// XCD order; box b marches segments of fused_batch_kseg planes and counts
// its T <= 0 cells in neg + NEG_STRIDE * b.  Every box of the launch has
// whole 32 x 8 tiles, CD6 and its second u, v, w buffers.
template <typename RT>
__global__ __launch_bounds__(NT) void k_visc_fused_batch(const BoxSet S, const int krk, int *neg) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int j = S.entry(t);
  const int box = S.ids[j];
  const Block B = S.blk[box];  // copied before any store, so its fields stay in scalar registers
  visc_fused_tile<RT>(B, krk, 1, fused_batch_kseg(B.g.nz, B.kseg), neg + NEG_STRIDE * box, t - S.off[j]);
}

template <typename RT>
hipError_t visc_fused_batch_as(const BoxSet &S, int krk, int *neg, int ntiles, hipStream_t st) {
  hipLaunchKernelGGL(k_visc_fused_batch<RT>, dim3((unsigned)ntiles), dim3(NT), 0, st, S, krk, neg);
  return hipGetLastError();
}

template <typename RT>
hipError_t visc_fused_as(const Block &B, int krk, int refresh, int *neg, hipStream_t st) {
  const Geom &g = B.g;
  if (g.nx % FTX != 0 || g.ny % FTY != 0 || !all_periodic(B) || B.S.scheme_vis == SCH_CD8 || !B.u2.p ||
      !B.v2.p || !B.w2.p)
    return hipErrorInvalidValue;
  const int kseg = fused_kseg(g);
  hipLaunchKernelGGL(k_visc_fused<RT>, dim3(fused_blocks(g, kseg)), dim3(NT), 0, st, B, krk, refresh, kseg, neg);
  return hipGetLastError();
}

}  // namespace

}  // namespace ocfd
