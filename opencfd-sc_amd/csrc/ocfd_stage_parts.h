// ocfd_stage_parts.h -- pieces of a stage shared by the per-phase kernels
// and the persistent batch kernel: the periodic halo fill of one point, the
// T <= 0 repair of one block, the copy of T's halo into T2.
#pragma once
#include "ocfd_kernels.h"

namespace ocfd {

namespace {

// ---------------------------------------------------------------------------
// periodic halo fill: for every periodic axis copy interior planes 1..L to
// n+1..n+L and n-L+1..n to 1-L..0, over the interior range of the other two
// indices (the extent of TYPE_LAP*2, OCFD_mpi.f90:114-122).
// ---------------------------------------------------------------------------
struct HaloFields {
  void *p[8];
  int f32;  // float fields (primitives of the mixed-precision form)
};
template <typename E>
__device__ inline void halo_point(const Geom &g, void *const *P, const int nf, const int px, const int py,
                                  const int pz, const int64_t t) {
  const int64_t nxh = px ? (int64_t)2 * LAP * g.ny * g.nz : 0;
  const int64_t nyh = py ? (int64_t)2 * LAP * g.nx * g.nz : 0;
  const int64_t nzh = pz ? (int64_t)2 * LAP * g.nx * g.ny : 0;
  const int64_t tot = nxh + nyh + nzh;
  if (t >= tot) return;
  int64_t dst, src;
  if (t < nxh) {
    int64_t r = t;
    const int l = (int)(r % (2 * LAP)); r /= 2 * LAP;
    const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
    if (l < LAP) { dst = g.hi(g.nx + 1 + l, j, k); src = g.hi(1 + l, j, k); }
    else { const int q = l - LAP; dst = g.hi(1 - LAP + q, j, k); src = g.hi(g.nx - LAP + 1 + q, j, k); }
  } else if (t < nxh + nyh) {
    int64_t r = t - nxh;
    const int i = (int)(r % g.nx) + 1; r /= g.nx;
    const int l = (int)(r % (2 * LAP)); r /= 2 * LAP;
    const int k = (int)r + 1;
    if (l < LAP) { dst = g.hi(i, g.ny + 1 + l, k); src = g.hi(i, 1 + l, k); }
    else { const int q = l - LAP; dst = g.hi(i, 1 - LAP + q, k); src = g.hi(i, g.ny - LAP + 1 + q, k); }
  } else {
    int64_t r = t - nxh - nyh;
    const int i = (int)(r % g.nx) + 1; r /= g.nx;
    const int j = (int)(r % g.ny) + 1; r /= g.ny;
    const int l = (int)r;
    if (l < LAP) { dst = g.hi(i, j, g.nz + 1 + l); src = g.hi(i, j, 1 + l); }
    else { const int q = l - LAP; dst = g.hi(i, j, 1 - LAP + q); src = g.hi(i, j, g.nz - LAP + 1 + q); }
  }
  for (int q = 0; q < nf; q++) static_cast<E *>(P[q])[dst] = static_cast<E *>(P[q])[src];
}

// T <= 0 repair (handle_NegativeT, OCFD_handle_NegativeT.f90:11, called from
// the i-j-k loop of comput_duvwT, OCFD_Time_Adv.f90:45-58): T = mean of the
// six neighbours.  The listed cells are processed in ascending index order,
// the reference's loop order, by one thread.  A -i/-j/-k neighbour inside the
// block already holds this stage's value (repaired, if it was listed), read
// from T2; a +i/+j/+k neighbour still holds the previous stage's, read from T;
// so does every halo point (comput_duvwT writes none), also read from T
// because the copy of T's halo into T2 runs concurrently.  With `halt` the
// first stage with more than NEG_MAX cells sets the flag and records itself
// (later stages find it set and repair nothing).
__device__ void negt_repair_body(const Block &B, int *neg, const int refresh) {
  if (halted(B)) {
    neg[0] = 0;
    return;
  }
  neg[NEG_STAGES] += 1;
  int cnt = neg[0];
  if (cnt <= 0) return;
  if (cnt > 64) cnt = 64;
  int *lst = neg + 1;
  for (int a = 1; a < cnt; a++) {  // insertion sort of <= 64 entries
    int x = lst[a], b = a - 1;
    while (b >= 0 && lst[b] > x) { lst[b + 1] = lst[b]; b--; }
    lst[b + 1] = x;
  }
  const Geom &g = B.g;
  for (int a = 0; a < cnt; a++) {
    int i, j, k;
    cell_of(g, lst[a], i, j, k);
    const int64_t h = g.hi(i, j, k);
    const double tip = B.T[g.hi(i + 1, j, k)], tjp = B.T[g.hi(i, j + 1, k)], tkp = B.T[g.hi(i, j, k + 1)];
    const double tim = i > 1 ? double(B.T2[g.hi(i - 1, j, k)]) : double(B.T[g.hi(i - 1, j, k)]);
    const double tjm = j > 1 ? double(B.T2[g.hi(i, j - 1, k)]) : double(B.T[g.hi(i, j - 1, k)]);
    const double tkm = k > 1 ? double(B.T2[g.hi(i, j, k - 1)]) : double(B.T[g.hi(i, j, k - 1)]);
    // the reference's summation order: i+1, i-1, j+1, j-1, k+1, k-1
    B.T2[h] = (tip + tim + tjp + tjm + tkp + tkm) / 6.0;
    if (refresh && (i == 1 || i == g.nx || j == 1 || j == g.ny || k == 1 || k == g.nz)) {
      const int64_t ni = g.ni, q = lst[a];
      const double d = B.d[h], u = B.u[h], v = B.v[h], w = B.w[h];
      B.f[q + 4 * ni] = d * (B.P.cv * double(B.T2[h]) + (u * u + v * v + w * w) * 0.5);
    }
  }
  const int total = neg[0];
  neg[65] += total;
  if (total > neg[66]) neg[66] = total;
  if (total > NEG_MAX && B.halt) {
    *B.halt = 1;
    neg[NEG_HALT_STAGE] = neg[NEG_STAGES];
  }
  neg[0] = 0;
}

// Copy of T's halo points into T2 (the stage update and comput_duvwT write
// the interior of T2 only).  Tile q of a box covers HCOPY_ROWS (j, k) rows
// of the halo'd array: rows outside the interior j, k range are copied whole,
// the others at their 2 * LAP i-halo points.
constexpr int HCOPY_ROWS = 16;
__device__ inline void halo_copy_tile(const Block &B, const int64_t q, const int uvw) {
  const Geom &g = B.g;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t nrows = (int64_t)g.sy * g.sz;
  const bool f32 = B.T.f32 != 0;
  for (int r = wave; r < HCOPY_ROWS; r += NT / 64) {
    const int64_t row = q * HCOPY_ROWS + r;
    if (row >= nrows) break;
    const int jh = (int)(row % g.sy), kh = (int)(row / g.sy);  // 0-based halo'd indices
    const int64_t base = row * g.sx;
    const bool whole = jh < LAP || jh >= LAP + g.ny || kh < LAP || kh >= LAP + g.nz;
    const int npts = whole ? g.sx : 2 * LAP;
    for (int p = lane; p < npts; p += 64) {
      const int ih = whole ? p : (p < LAP ? p : g.nx + p);
      const int64_t x = base + ih;
      if (f32) {
        B.T2.as<float>()[x] = B.T.as<float>()[x];
        if (uvw) {
          B.u.as<float>()[x] = B.u2.as<float>()[x];
          B.v.as<float>()[x] = B.v2.as<float>()[x];
          B.w.as<float>()[x] = B.w2.as<float>()[x];
        }
      } else {
        B.T2.as<double>()[x] = B.T.as<double>()[x];
        if (uvw) {
          B.u.as<double>()[x] = B.u2.as<double>()[x];
          B.v.as<double>()[x] = B.v2.as<double>()[x];
          B.w.as<double>()[x] = B.w2.as<double>()[x];
        }
      }
    }
  }
}
inline int64_t halo_copy_tiles_of(const Geom &g) { return ((int64_t)g.sy * g.sz + HCOPY_ROWS - 1) / HCOPY_ROWS; }

}  // namespace

}  // namespace ocfd
