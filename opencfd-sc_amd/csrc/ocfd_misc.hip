// ocfd_misc.hip -- element-wise and small kernels of the step:
//   periodic halo fill        exchange_boundary_x/y/z on one block (OCFD_mpi.f90:198-233)
//   shock sensor              comput_Rhybrid (OCFD_SetHybrid.f90:17-152)
//   RK3 / primitives          OCFD_time_adv_RK3, comput_duvwT (OCFD_Time_Adv.f90:4-63)
//   boundary conditions       OCFD_bc + Flow_Ghost_Extent_2nd (OCFD_BC.f90:2-99,
//                             OCFD_GhostCell.f90:164-286)
//   filters                   filter_[xyz]3d[_shock] (OCFD_filtering.f90:73-363)
//   diagnostics               show_flow_msg sums, ana_box sums (two-level,
//                             fixed-order reductions: run-to-run deterministic)
//   grid metrics              init_mesh (OCFD_mesh.f90:18-107, OCFD_GhostCell.f90)
#include "ocfd_kernels.h"
#include "ocfd_stage_parts.h"

namespace ocfd {

namespace {
inline unsigned nb(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

__global__ void k_halo(const Geom g, const HaloFields F, const int nf, const int px, const int py, const int pz) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (F.f32) halo_point<float>(g, F.p, nf, px, py, pz, t);
  else halo_point<double>(g, F.p, nf, px, py, pz, t);
}

// the five primitives of every box of a batch (all axes periodic)
__global__ __launch_bounds__(NT) void k_halo_batch(const BoxSet S) {
  const int j = S.entry(blockIdx.x);
  const Block &B = S.blk[S.ids[j]];
  void *const P[5] = {B.d.p, B.u.p, B.v.p, B.w.p, B.T.p};
  const int64_t t = (int64_t)(blockIdx.x - S.off[j]) * NT + threadIdx.x;
  if (B.d.f32) halo_point<float>(B.g, P, 5, 1, 1, 1, t);
  else halo_point<double>(B.g, P, 5, 1, 1, 1, t);
}

// ---------------------------------------------------------------------------
// shock sensor 1 (OCFD_SetHybrid.f90:32-52): R = sum over axes of
// |p(-1) - 2p + p(+1)| / |p(-1) + 2p + p(+1)|, p = p00 d T
// ---------------------------------------------------------------------------
// E: element type of the primitive fields
template <typename E>
__global__ __launch_bounds__(NT) void k_sensor1(const Block B, const double p00) {
  const Geom &g = B.g;
  const E *const pd = B.d.as<E>(), *const pT = B.T.as<E>();
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int i, j, k;
  cell_of(g, t, i, j, k);
  const int64_t h = g.hi(i, j, k);
  auto P = [&](int64_t q) { return p00 * double(pd[q]) * double(pT[q]); };
  const double pc = P(h);
  double r = 0.0;
  double acc[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int64_t st = g.hstride(a);
    const double pm = P(h - st), pp = P(h + st);
    acc[a] = fabs((pm - 2.0 * pc + pp) / (pm + 2.0 * pc + pp));
  }
  r = acc[0] + acc[1] + acc[2];
  B.rhyb[h] = r;
}

// sensor 2 (OCFD_SetHybrid.f90:54-85), pass 1: |grad p| and block sums
__global__ __launch_bounds__(NT) void k_sensor2a(const Block B, const double p00, double *partial) {
  const Geom &g = B.g;
  __shared__ double red[NT];
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  double r = 0.0;
  if (t < g.ni) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const int64_t h = g.hi(i, j, k);
    double dp[3];
    for (int a = 0; a < 3; a++) {
      const int64_t st = g.hstride(a);
      const int s = a == 0 ? i : (a == 1 ? j : k);
      auto at = [&](int q) { const int64_t x = h + (int64_t)q * st; return p00 * B.d[x] * B.T[x]; };
      dp[a] = dxyz0_point(at, s, g.n(a), B.P.dc[a], B.S.scheme_vis == SCH_CD8, B.S.iperiodic[a] == 1);
    }
    const Metrics &M = B.M;
    const double px = dp[0] * M.get(M_AKX, h) + dp[1] * M.get(M_AIX, h) + dp[2] * M.get(M_ASX, h);
    const double py = dp[0] * M.get(M_AKY, h) + dp[1] * M.get(M_AIY, h) + dp[2] * M.get(M_ASY, h);
    const double pz = dp[0] * M.get(M_AKZ, h) + dp[1] * M.get(M_AIZ, h) + dp[2] * M.get(M_ASZ, h);
    r = sqrt(px * px + py * py + pz * pz);
    B.rhyb[h] = r;
  }
  red[threadIdx.x] = r;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// fixed-order sum of `n` partials by one block into out[0]
__global__ __launch_bounds__(NT) void k_sum_partials(const double *partial, int64_t n, int stride, int nq,
                                                     double *out) {
  __shared__ double red[NT];
  for (int q = 0; q < nq; q++) {
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < n; b += NT) s += partial[b * stride + q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
    __syncthreads();
  }
}

// sensor 2, pass 2: R = R / (mean + 1e-30)
__global__ __launch_bounds__(NT) void k_sensor2b(const Block B, const double *sum) {
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int i, j, k;
  cell_of(g, t, i, j, k);
  const double dp_av = sum[0] / (1.0 * g.nx * g.ny * g.nz);
  const double dp_av1 = 1.0 / (dp_av + 1.0e-30);
  const int64_t h = g.hi(i, j, k);
  B.rhyb[h] = B.rhyb[h] * dp_av1;
}

// scheme shares over x-faces (OCFD_SetHybrid.f90:115-131)
__global__ __launch_bounds__(NT) void k_hyb_counts(const Block B, unsigned long long *counts) {
  const Geom &g = B.g;
  const int64_t nf = (int64_t)(g.nx + 1) * g.ny * g.nz;
  // a bounded grid strides over the faces and adds one total per scheme and
  // workgroup: one global atomic per workgroup of 256 faces serialised
  // half a million atomics on three addresses at 512^3
  unsigned c[3] = {0, 0, 0};
  for (int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x; t < nf; t += (int64_t)gridDim.x * NT) {
    const int i = (int)(t % (g.nx + 1));
    const int64_t r = t / (g.nx + 1);
    const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
    const double r0 = dmax2(B.rhyb[g.hi(i, j, k)], B.rhyb[g.hi(i + 1, j, k)]);
    const int sc = hybrid_index(r0, B.S.hyb2, B.S.hyb3);
    c[0] += sc == 1;
    c[1] += sc == 2;
    c[2] += sc == 3;
  }
  __shared__ unsigned long long cnt[3];
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 3; q++)
    if (c[q]) atomicAdd(&cnt[q], (unsigned long long)c[q]);
  __syncthreads();
  if (threadIdx.x < 3 && cnt[threadIdx.x]) atomicAdd(&counts[threadIdx.x], cnt[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// CFL time step: the largest over the cells of
//   sum over axes a of (|u . A_a| + c |A_a|) / h_a,   c = sqrt(T) / Ma,
// with A_a = (dxi_a/dx, dxi_a/dy, dxi_a/dz) the metric row of axis a.  The
// reference advances with a fixed Para%dt (OCFD_Time_Adv.f90:20, 30); this is
// the dt = CFL / rate a case generator would fix.  Positive doubles order as
// their bit patterns, so the block maxima combine with an integer atomicMax.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_cfl_rate(const Block B, unsigned long long *out) {
  const Geom &g = B.g;
  double m = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x; t < g.ni; t += (int64_t)gridDim.x * NT) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const int64_t h = g.hi(i, j, k);
    const double u = B.u[h], v = B.v[h], w = B.w[h];
    const double c = sqrt(double(B.T[h])) / B.P.ma;
    double r = 0.0;
    for (int a = 0; a < 3; a++) {
      const double ax = B.M.get(3 * a, h), ay = B.M.get(3 * a + 1, h), az = B.M.get(3 * a + 2, h);
      r += (fabs(u * ax + v * ay + w * az) + c * sqrt(ax * ax + ay * ay + az * az)) * B.P.hinv[a];
    }
    m = fmax(m, r);
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  __shared__ double red[NT / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < NT / 64; q++) m = fmax(m, red[q]);
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

// ---------------------------------------------------------------------------
// RK3 stage (OCFD_Time_Adv.f90:4-36) / fn = f / init3d f / comput_duvwT
// ---------------------------------------------------------------------------
// du is held in DU (double; float in the mixed-precision form); the stage
// combination is FP64
template <typename DU>
__global__ __launch_bounds__(NT) void k_rk3(const Block B, const int krk) {
  const int64_t n5 = 5 * B.g.ni;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= n5) return;
  const double dt = B.P.dt;
  const double du = double(reinterpret_cast<const DU *>(B.du)[t]);
  if (krk == 1) {
    B.f[t] = B.fn[t] + dt * du;
  } else {
    const double ra = krk == 2 ? 3.0 / 4.0 : 1.0 / 3.0;
    const double rb = krk == 2 ? 1.0 / 4.0 : 2.0 / 3.0;
    B.f[t] = ra * B.fn[t] + rb * (B.f[t] + dt * du);
  }
}

__global__ __launch_bounds__(NT) void k_copy(const double *src, double *dst, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

// f from primitives (init3d, OCFD_init.f90:10-20)
__global__ __launch_bounds__(NT) void k_init_f(const Block B) {
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int i, j, k;
  cell_of(g, t, i, j, k);
  const int64_t h = g.hi(i, j, k);
  const double d = B.d[h], u = B.u[h], v = B.v[h], w = B.w[h];
  const int64_t ni = g.ni;
  B.f[t] = d;
  B.f[t + ni] = d * u;
  B.f[t + 2 * ni] = d * v;
  B.f[t + 3 * ni] = d * w;
  B.f[t + 4 * ni] = d * ((u * u + v * v + w * w) * 0.5 + B.P.cv * B.T[h]);
}

__global__ __launch_bounds__(NT) void k_duvwt(const Block B, int *neg) {
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int i, j, k;
  cell_of(g, t, i, j, k);
  const int64_t h = g.hi(i, j, k), ni = g.ni;
  const double f1 = B.f[t], f2 = B.f[t + ni], f3 = B.f[t + 2 * ni], f4 = B.f[t + 3 * ni], f5 = B.f[t + 4 * ni];
  const double u = f2 / f1, v = f3 / f1, w = f4 / f1;
  const double T = (f5 - (f2 * u + f3 * v + f4 * w) * 0.5) / (f1 * B.P.cv);
  // T goes to T2: the repair reads the previous stage's T of the +i/+j/+k
  // neighbours from B.T
  B.d[h] = f1; B.u[h] = u; B.v[h] = v; B.w[h] = w; B.T2[h] = T;
  if (T <= 0) {
    const int slot = atomicAdd(neg, 1);
    if (slot < 64) neg[1 + slot] = (int)t;
  }
}


// workgroup 0: the repair; workgroups 1..: the halo copy
__global__ __launch_bounds__(NT) void k_negt_repair(const Block B, int *neg, const int refresh, const int uvw) {
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) negt_repair_body(B, neg, refresh);
    return;
  }
  halo_copy_tile(B, blockIdx.x - 1, uvw);
}

// batches: workgroup b < nbox repairs box b (its counters at neg + NEG_STRIDE
// * b); the others copy halos over the tiles of S
__global__ __launch_bounds__(NT) void k_negt_repair_batch(const BoxSet S, const int nbox, int *neg) {
  if ((int)blockIdx.x < nbox) {
    if (threadIdx.x == 0) negt_repair_body(S.blk[blockIdx.x], neg + NEG_STRIDE * blockIdx.x, 1);
    return;
  }
  const int t = blockIdx.x - nbox;
  const int j = S.entry(t);
  const Block &B = S.blk[S.ids[j]];
  halo_copy_tile(B, t - S.off[j], B.u2.p != nullptr ? 1 : 0);
}

// ---------------------------------------------------------------------------
// boundary conditions
// ---------------------------------------------------------------------------
// BC hooks on the j=1 / j=ny planes: ocfd_bc_blunt2d (OCFD_BC_Def.f90:4-33),
// or the isothermal/adiabatic wall + free stream of the BC_USER_DEF hook.
__global__ __launch_bounds__(NT) void k_bc_hook(const Block B, const int ibc, const double Tw) {
  if (halted(B)) return;  // a whole-step run stopped at too many T <= 0 cells
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)g.nx * g.nz) return;
  const int i = (int)(t % g.nx) + 1, k = (int)(t / g.nx) + 1;
  const int64_t h1 = g.hi(i, 1, k), h2 = g.hi(i, 2, k), h3 = g.hi(i, 3, k), hn = g.hi(i, g.ny, k);
  if (ibc == 1) {
    B.d[h1] = B.d[h2]; B.u[h1] = 0.0; B.v[h1] = 0.0; B.w[h1] = 0.0; B.T[h1] = B.T[h2];
  } else {
    B.u[h1] = 0.0; B.v[h1] = 0.0; B.w[h1] = 0.0;
    if (Tw > 0.0) {
      B.d[h1] = (4.0 * B.d[h2] * B.T[h2] - B.d[h3] * B.T[h3]) / (3.0 * Tw);
      B.T[h1] = Tw;
    } else {
      B.d[h1] = (4.0 * B.d[h2] - B.d[h3]) / 3.0;
      B.T[h1] = (4.0 * B.T[h2] - B.T[h3]) / 3.0;
    }
  }
  B.d[hn] = 1.0; B.u[hn] = 1.0; B.v[hn] = 0.0; B.w[hn] = 0.0; B.T[hn] = 1.0;
}

// ---------------------------------------------------------------------------
// boundary-layer / swept-corner conditions (ocfd_bc_BoundaryLayer,
// OCFD_BC_Boundarylayers.f90:17-167; ocfd_bc_swept_corner, :443-617), one
// kernel per part, launched in the reference's order: inlet, upper, wall,
// outlet, spanwise sides (a later part overwrites the edges of an earlier one)
// ---------------------------------------------------------------------------
__device__ inline void set_prims(const Block &B, int64_t h, double d, double u, double v, double w, double T) {
  B.d[h] = d; B.u[h] = u; B.v[h] = v; B.w[h] = w; B.T[h] = T;
}

// i = 1: free stream, or the inlet data
__global__ __launch_bounds__(NT) void k_bl_inlet(const Block B, const BLArgs a) {
  if (halted(B)) return;  // a whole-step run stopped at too many T <= 0 cells
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)g.ny * g.nz) return;
  const int64_t h = g.hi(1, (int)(t % g.ny) + 1, (int)(t / g.ny) + 1);
  if (!a.inlet) {
    set_prims(B, h, 1.0, a.u0, a.v0, 0.0, a.inletT);
  } else {
    const int64_t n = (int64_t)g.ny * g.nz;
    set_prims(B, h, a.inlet[t], a.inlet[t + n], a.inlet[t + 2 * n], a.inlet[t + 3 * n], a.inlet[t + 4 * n]);
  }
}

// j = ny: Dirichlet values where the point is not non-reflecting
__global__ __launch_bounds__(NT) void k_bl_upper(const Block B, const BLArgs a) {
  if (halted(B)) return;  // a whole-step run stopped at too many T <= 0 cells
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)g.nx * g.nz || a.nonref[t] != 0) return;
  const int64_t h = g.hi((int)(t % g.nx) + 1, g.ny, (int)(t / g.nx) + 1);
  if (!a.upper) {
    set_prims(B, h, 1.0, a.u0, a.v0, 0.0, 1.0);
  } else {
    const int64_t n = (int64_t)g.nx * g.nz;
    set_prims(B, h, a.upper[t], a.upper[t + n], a.upper[t + 2 * n], a.upper[t + 3 * n], a.upper[t + 4 * n]);
  }
}

// j = 1: blowing/suction along the wall normal, isothermal (Tw > 0) or
// adiabatic wall with second-order pressure; swept corner: points off the
// wall extrapolate at first order
__global__ __launch_bounds__(NT) void k_bl_wall(const Block B, const BLArgs a) {
  if (halted(B)) return;  // a whole-step run stopped at too many T <= 0 cells
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)g.nx * g.nz) return;
  const int i = (int)(t % g.nx) + 1, k = (int)(t / g.nx) + 1;
  const int64_t h1 = g.hi(i, 1, k), h2 = g.hi(i, 2, k), h3 = g.hi(i, 3, k);
  if (a.wall && a.wall[t] == 0) {
    set_prims(B, h1, B.d[h2], B.u[h2], B.v[h2], B.w[h2], B.T[h2]);
    return;
  }
  const double aix = B.M.get(M_AIX, h1), aiy = B.M.get(M_AIY, h1), aiz = B.M.get(M_AIZ, h1);
  const double tmp = 1.0 / (sqrt(aix * aix + aiy * aiy + aiz * aiz));
  const double A0 = a.A * a.ht[0] * a.pertb[t];
  B.u[h1] = A0 * aix * tmp;
  B.v[h1] = A0 * aiy * tmp;
  B.w[h1] = A0 * aiz * tmp;
  if (a.Tw > 0.0) {
    B.T[h1] = a.Tw;
    B.d[h1] = (4.0 * B.d[h2] * B.T[h2] - B.d[h3] * B.T[h3]) / (3.0 * a.Tw);
  } else {
    B.T[h1] = (4.0 * B.T[h2] - B.T[h3]) / 3.0;
    B.d[h1] = (4.0 * B.d[h2] - B.d[h3]) / 3.0;
  }
}

// i = nx: first- or second-order extrapolation
__global__ __launch_bounds__(NT) void k_bl_outlet(const Block B, const BLArgs a) {
  if (halted(B)) return;  // a whole-step run stopped at too many T <= 0 cells
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)g.ny * g.nz) return;
  const int j = (int)(t % g.ny) + 1, k = (int)(t / g.ny) + 1;
  const int64_t h = g.hi(g.nx, j, k), h1 = g.hi(g.nx - 1, j, k), h2 = g.hi(g.nx - 2, j, k);
  const PrimField F[5] = {B.d, B.u, B.v, B.w, B.T};
  for (int q = 0; q < 5; q++) F[q][h] = a.outlet == 1 ? F[q][h1] : 2.0 * F[q][h1] - F[q][h2];
}

// swept corner, k = 1 / k = nz: first-order extrapolation where requested
__global__ __launch_bounds__(NT) void k_bl_sides(const Block B, const BLArgs a) {
  if (halted(B)) return;  // a whole-step run stopped at too many T <= 0 cells
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(t % g.nx) + 1, j = (int)(t / g.nx) + 1;
  const PrimField F[5] = {B.d, B.u, B.v, B.w, B.T};
  if (a.z1 == 1)
    for (int q = 0; q < 5; q++) F[q][g.hi(i, j, 1)] = F[q][g.hi(i, j, 2)];
  if (a.z2 == 1)
    for (int q = 0; q < 5; q++) F[q][g.hi(i, j, g.nz)] = F[q][g.hi(i, j, g.nz - 1)];
}

// f recomputed from the primitives on the six block faces (OCFD_BC.f90:26-95)
__global__ __launch_bounds__(NT) void k_bc_planes(const Block B) {
  if (halted(B)) return;  // a whole-step run stopped at too many T <= 0 cells
  const Geom &g = B.g;
  const int64_t ax = (int64_t)g.ny * g.nz, ay = (int64_t)g.nx * g.nz, az = (int64_t)g.nx * g.ny;
  int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= 2 * (ax + ay + az)) return;
  int i, j, k;
  if (t < 2 * ax) {
    const int side = (int)(t / ax); const int64_t r = t % ax;
    j = (int)(r % g.ny) + 1; k = (int)(r / g.ny) + 1; i = side ? g.nx : 1;
  } else if (t < 2 * (ax + ay)) {
    t -= 2 * ax;
    const int side = (int)(t / ay); const int64_t r = t % ay;
    i = (int)(r % g.nx) + 1; k = (int)(r / g.nx) + 1; j = side ? g.ny : 1;
  } else {
    t -= 2 * (ax + ay);
    const int side = (int)(t / az); const int64_t r = t % az;
    i = (int)(r % g.nx) + 1; j = (int)(r / g.nx) + 1; k = side ? g.nz : 1;
  }
  const int64_t h = g.hi(i, j, k), q = g.ii(i, j, k), ni = g.ni;
  const double d = B.d[h], u = B.u[h], v = B.v[h], w = B.w[h];
  B.f[q] = d;
  B.f[q + ni] = d * u;
  B.f[q + 2 * ni] = d * v;
  B.f[q + 3 * ni] = d * w;
  B.f[q + 4 * ni] = d * (B.P.cv * B.T[h] + (u * u + v * v + w * w) * 0.5);
}

// Flow_Ghost_Extent_2nd (OCFD_GhostCell.f90:164-286)
__global__ __launch_bounds__(NT) void k_ghost(const Block B, const int side) {
  if (halted(B)) return;  // a whole-step run stopped at too many T <= 0 cells
  const Geom &g = B.g;
  const int axis = side / 2, hi = side % 2;
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;  // the two other axes
  const int n1 = g.n(a1), n2 = g.n(a2), n = g.n(axis);
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)LAP * n1 * n2) return;
  const int l = (int)(t % LAP);
  const int64_t r = t / LAP;
  int idx[3], wi[3], mi[3];
  idx[a1] = (int)(r % n1) + 1;
  idx[a2] = (int)(r / n1) + 1;
  const int s = hi ? n + 1 + l : 1 - LAP + l;
  idx[axis] = s;
  for (int q = 0; q < 3; q++) { wi[q] = idx[q]; mi[q] = idx[q]; }
  wi[axis] = hi ? n : 1;
  mi[axis] = hi ? 2 * n - s : 2 - s;
  const int64_t gq = g.hi(idx[0], idx[1], idx[2]), wq = g.hi(wi[0], wi[1], wi[2]), mq = g.hi(mi[0], mi[1], mi[2]);
  const double s1 = 0.8, s2 = 1.2;
  B.u[gq] = 2.0 * B.u[wq] - B.u[mq];
  B.v[gq] = 2.0 * B.v[wq] - B.v[mq];
  B.w[gq] = 2.0 * B.w[wq] - B.w[mq];
  double d = 2.0 * B.d[wq] - B.d[mq];
  double T = 2.0 * B.T[wq] - B.T[mq];
  if (d < s1 * B.d[wq]) d = s1 * B.d[wq];
  if (d > s2 * B.d[wq]) d = s2 * B.d[wq];
  if (T < s1 * B.T[wq]) T = s1 * B.T[wq];
  if (T > s2 * B.T[wq]) T = s2 * B.T[wq];
  B.d[gq] = d;
  B.T[gq] = T;
}

// Flow_Ghost_Userdef (OCFD_GhostCell_UserDef.f90:14-21): flat plate j- (the
// boundary layer; mirror about j = 1 with the 0.5 / 2.0 limiter,
// Flow_Ghost_flatplate :110-136) or swept-corner symmetry plane k- (w
// changes sign, Flow_Ghost_Sweptcorner :56-76)
__global__ __launch_bounds__(NT) void k_ghost_user(const Block B, const int ibc) {
  if (halted(B)) return;  // a whole-step run stopped at too many T <= 0 cells
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (ibc == 2) {
    if (t >= (int64_t)LAP * g.nx * g.nz) return;
    const int l = (int)(t % LAP);
    const int64_t r = t / LAP;
    const int i = (int)(r % g.nx) + 1, k = (int)(r / g.nx) + 1;
    const int j = 1 - LAP + l, j1 = 2 - j;
    const int64_t h = g.hi(i, j, k), hw = g.hi(i, 1, k), hm = g.hi(i, j1, k);
    const double s1 = 0.5, s2 = 2.0;
    B.u[h] = 2.0 * B.u[hw] - B.u[hm];
    B.v[h] = 2.0 * B.v[hw] - B.v[hm];
    B.w[h] = 2.0 * B.w[hw] - B.w[hm];
    double d = 2.0 * B.d[hw] - B.d[hm];
    double T = 2.0 * B.T[hw] - B.T[hm];
    if (d < s1 * B.d[hw]) d = s1 * B.d[hw];
    if (d > s2 * B.d[hw]) d = s2 * B.d[hw];
    if (T < s1 * B.T[hw]) T = s1 * B.T[hw];
    if (T > s2 * B.T[hw]) T = s2 * B.T[hw];
    B.d[h] = d;
    B.T[h] = T;
  } else {
    if (t >= (int64_t)LAP * g.nx * g.ny) return;
    const int l = (int)(t % LAP);
    const int64_t r = t / LAP;
    const int i = (int)(r % g.nx) + 1, j = (int)(r / g.nx) + 1;
    const int k = 1 - LAP + l, k1 = 1 - k;
    const int64_t h = g.hi(i, j, k), hm = g.hi(i, j, k1);
    B.u[h] = B.u[hm];
    B.v[h] = B.v[hm];
    B.w[h] = -B.w[hm];
    B.d[h] = B.d[hm];
    B.T[h] = B.T[hm];
  }
}

// ---------------------------------------------------------------------------
// filters: out-of-place along one axis; cells outside the filtered index box
// are copied unchanged
// ---------------------------------------------------------------------------
struct FilterArgs {
  int axis, scheme;
  double s0, rth;
  int lo[3], hi[3];
};
__global__ __launch_bounds__(NT) void k_filter(const Block B, const FilterArgs A, double *fout) {
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int idx[3];
  cell_of(g, t, idx[0], idx[1], idx[2]);
  const int axis = A.axis, n = g.n(axis);
  const int s = idx[axis];
  const int64_t st = g.istride(axis), ni = g.ni;
  // a whole-step run stopped at too many T <= 0 cells carries f over unchanged
  const bool in = !halted(B) && idx[0] >= A.lo[0] && idx[0] <= A.hi[0] && idx[1] >= A.lo[1] &&
                  idx[1] <= A.hi[1] && idx[2] >= A.lo[2] && idx[2] <= A.hi[2];
  auto wrap = [&](int q) { return q < 1 ? q + n : (q > n ? q - n : q); };
  for (int m = 0; m < 5; m++) {
    const double *f0 = B.f + (int64_t)m * ni;
    double r = f0[t];
    if (in) {
      auto G = [&](int o) { return f0[t + (int64_t)(wrap(s + o) - s) * st]; };
      if (A.scheme == 1) {
        // Bogey-Bailly Fo9p (OCFD_filtering.f90:80-81, 101-102)
        const double d0 = 0.243527493120, d1 = -0.204788880640, d2 = 0.120007591680, d3 = -0.045211119360,
                     d4 = 0.008228661760;
        r = G(0) - A.s0 * (d0 * G(0) + d1 * (G(-1) + G(1)) + d2 * (G(-2) + G(2)) + d3 * (G(-3) + G(3)) +
                           d4 * (G(-4) + G(4)));
      } else {
        // shock-capturing filter (OCFD_filtering.f90:195-239)
        const double c1 = -0.210383, c2 = 0.039617, gm1 = B.P.gamma - 1.0;
        auto P = [&](int o) {
          const int64_t q = t + (int64_t)(wrap(s + o) - s) * st;
          const double f1 = B.f[q], f2 = B.f[q + ni], f3 = B.f[q + 2 * ni], f4 = B.f[q + 3 * ni], f5 = B.f[q + 4 * ni];
          return (f5 - 0.5 * (f2 * f2 + f3 * f3 + f4 * f4) / f1) * gm1;
        };
        double pv[7];
        for (int o = -3; o <= 3; o++) pv[o + 3] = P(o);
        double sc0[3];
        for (int c = -1; c <= 1; c++) {
          const double *pp = pv + 3 + c;  // pp[0] = p(s+c)
          const double dp0 = 0.25 * (-pp[1] + 2.0 * pp[0] - pp[-1]);
          const double dp1 = 0.25 * (-pp[2] + 2.0 * pp[1] - pp[0]);
          const double dp2 = 0.25 * (-pp[0] + 2.0 * pp[-1] - pp[-2]);
          const double x = dp0 - dp1, y = dp0 - dp2;
          const double ri = 0.5 * (x * x + y * y) / (pp[0] * pp[0]) + 1.0e-16;
          sc0[c + 1] = 0.5 * (1.0 - A.rth / ri + fabs(1.0 - A.rth / ri));
        }
        const double sc1 = 0.5 * (sc0[1] + sc0[2]);
        const double sc2 = 0.5 * (sc0[1] + sc0[0]);
        r = G(0) - A.s0 * (sc1 * (c1 * (G(1) - G(0)) + c2 * (G(2) - G(-1))) -
                           sc2 * (c1 * (G(0) - G(-1)) + c2 * (G(1) - G(-2))));
      }
    }
    fout[t + (int64_t)m * ni] = r;
  }
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
// show_flow_msg sums (OCFD_Time_Adv.f90:77-86): 3 per block
__global__ __launch_bounds__(NT) void k_flow_msg(const Block B, double *partial) {
  const Geom &g = B.g;
  __shared__ double red[3][NT];
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  double e[3] = {0.0, 0.0, 0.0};
  if (t < g.ni) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const int64_t h = g.hi(i, j, k);
    const double p00 = 1.0 / (B.P.gamma * B.P.ma * B.P.ma), tmp1 = 1.0 - B.P.gamma;
    const double tmp0 = 1.0 / B.M.get(M_AJAC, h);
    const double d = B.d[h], u = B.u[h], v = B.v[h], w = B.w[h];
    e[0] = B.f[t + 4 * g.ni] * tmp0;
    e[1] = d * (u * u + v * v + w * w) * 0.5 * tmp0;
    e[2] = p00 * pow(d, tmp1) * B.T[h] * tmp0;
  }
  for (int q = 0; q < 3; q++) red[q][threadIdx.x] = e[q];
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; q++) partial[blockIdx.x * 3 + q] = red[q][0];
}

// ana_box sums (OCFD_ANA.f90:215-243): 14 per block; pass 2 (dmean != null):
// sum of (d - dmean)^2
__global__ __launch_bounds__(NT) void k_ana_box(const Block B, double *partial, const double *dsum) {
  const Geom &g = B.g;
  __shared__ double red[NT];
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  double s[14];
  for (int q = 0; q < 14; q++) s[q] = 0.0;
  const int nq = dsum ? 1 : 14;
  if (t < g.ni) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const int64_t h = g.hi(i, j, k);
    if (dsum) {
      const double d_aver = dsum[0] / (double)(g.nx * g.ny * g.nz);
      const double x = B.d[h] - d_aver;
      s[0] = x * x;
    } else {
      double gr[9];
      const PrimField fld[3] = {B.u, B.v, B.w};
      for (int c = 0; c < 3; c++)
        for (int a = 0; a < 3; a++) {
          const int64_t st = g.hstride(a);
          const int sidx = a == 0 ? i : (a == 1 ? j : k);
          const PrimField arr = fld[c];
          auto at = [&](int q) { return double(arr[h + (int64_t)q * st]); };
          gr[c * 3 + a] =
              dxyz0_point(at, sidx, g.n(a), B.P.dc[a], B.S.scheme_vis == SCH_CD8, B.S.iperiodic[a] == 1);
        }
      const Metrics &M = B.M;
      auto X = [&](int c) { return gr[c * 3] * M.get(M_AKX, h) + gr[c * 3 + 1] * M.get(M_AIX, h) + gr[c * 3 + 2] * M.get(M_ASX, h); };
      auto Y = [&](int c) { return gr[c * 3] * M.get(M_AKY, h) + gr[c * 3 + 1] * M.get(M_AIY, h) + gr[c * 3 + 2] * M.get(M_ASY, h); };
      auto Z = [&](int c) { return gr[c * 3] * M.get(M_AKZ, h) + gr[c * 3 + 1] * M.get(M_AIZ, h) + gr[c * 3 + 2] * M.get(M_ASZ, h); };
      const double ux = X(0), uy = Y(0), uz = Z(0), vx = X(1), vz = Z(1), wx = X(2), wy = Y(2);
      const double u = B.u[h], v = B.v[h], w = B.w[h], d = B.d[h], T = B.T[h];
      const double Tsb = 110.4 / B.P.ref_amu_t0, TmpR = 1.0 / B.P.re;
      const double amu = TmpR * (1.0 + Tsb) * sqrt(T * T * T) / (Tsb + T);
      s[0] = u * u; s[1] = v * v; s[2] = w * w; s[3] = ux * ux;
      s[4] = ux * ux * ux; s[5] = (ux * ux) * (ux * ux);
      s[6] = d; s[7] = sqrt(T) / B.P.ma; s[8] = amu;
      s[9] = (u * u + v * v + w * w) * d * 0.5; s[10] = T;
      s[11] = u * u * u; s[12] = (u * u) * (u * u);
      const double e1 = wy - vz, e2 = uz - wx, e3 = vx - uy;
      s[13] = e1 * e1 + e2 * e2 + e3 * e3;
    }
  }
  for (int q = 0; q < nq; q++) {
    red[threadIdx.x] = s[q];
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x * nq + q] = red[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// grid metrics
// ---------------------------------------------------------------------------
// updata_Axyz_periodic (OCFD_mesh.f90:109-185)
__global__ __launch_bounds__(NT) void k_coord_shift(const Geom g, double *X, double *Y, double *Z, const int axis,
                                                    const double s0, const double s1, const double s2) {
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const int n1 = g.n(a1), n2 = g.n(a2), n = g.n(axis);
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)2 * LAP * n1 * n2) return;
  const int l = (int)(t % (2 * LAP));
  const int64_t r = t / (2 * LAP);
  int idx[3];
  idx[a1] = (int)(r % n1) + 1;
  idx[a2] = (int)(r / n1) + 1;
  double sg;
  if (l < LAP) { idx[axis] = 1 - LAP + l; sg = -1.0; } else { idx[axis] = n + 1 + (l - LAP); sg = 1.0; }
  const int64_t h = g.hi(idx[0], idx[1], idx[2]);
  if (sg < 0) { X[h] = X[h] - s0; Y[h] = Y[h] - s1; Z[h] = Z[h] - s2; }
  else { X[h] = X[h] + s0; Y[h] = Y[h] + s1; Z[h] = Z[h] + s2; }
}

__device__ inline void jac_point(double *const *met, int64_t h, double xi1, double xj1, double xk1, double yi1,
                                 double yj1, double yk1, double zi1, double zj1, double zk1) {
  const double Jac1 = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 -
                             yi1 * xj1 * zk1 - xi1 * zj1 * yk1);
  met[M_AJAC][h] = Jac1;
  met[M_AKX][h] = Jac1 * (yj1 * zk1 - zj1 * yk1);
  met[M_AKY][h] = Jac1 * (zj1 * xk1 - xj1 * zk1);
  met[M_AKZ][h] = Jac1 * (xj1 * yk1 - yj1 * xk1);
  met[M_AIX][h] = Jac1 * (yk1 * zi1 - zk1 * yi1);
  met[M_AIY][h] = Jac1 * (zk1 * xi1 - xk1 * zi1);
  met[M_AIZ][h] = Jac1 * (xk1 * yi1 - yk1 * xi1);
  met[M_ASX][h] = Jac1 * (yi1 * zj1 - zi1 * yj1);
  met[M_ASY][h] = Jac1 * (zi1 * xj1 - xi1 * zj1);
  met[M_ASZ][h] = Jac1 * (xi1 * yj1 - yi1 * xj1);
}

struct MetPtrs {
  double *p[M_N];
};

// comput_Jacobian3d (OCFD_mesh.f90:61-107) on the interior
__global__ __launch_bounds__(NT) void k_jacobian(const Block B, const double *X, const double *Y, const double *Z,
                                                 const MetPtrs mp) {
  const Geom &g = B.g;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.ni) return;
  int i, j, k;
  cell_of(g, t, i, j, k);
  const int64_t h = g.hi(i, j, k);
  double dd[9];
  const double *C[3] = {X, Y, Z};
  for (int c = 0; c < 3; c++)
    for (int a = 0; a < 3; a++) {
      const int64_t st = g.hstride(a);
      const int s = a == 0 ? i : (a == 1 ? j : k);
      const double *arr = C[c];
      auto at = [&](int q) { return arr[h + (int64_t)q * st]; };
      dd[c * 3 + a] = dxyz0_point(at, s, g.n(a), B.P.dc[a], B.S.scheme_vis == SCH_CD8, B.S.iperiodic[a] == 1);
    }
  jac_point(mp.p, h, dd[0], dd[1], dd[2], dd[3], dd[4], dd[5], dd[6], dd[7], dd[8]);
}

// Jac_Ghost_Extent_1st (OCFD_GhostCell.f90:18-160)
__global__ __launch_bounds__(NT) void k_jac_ghost1(const Geom g, const int side, double *X, double *Y, double *Z,
                                                   const MetPtrs mp) {
  const int axis = side / 2, hi = side % 2;
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const int n1 = g.n(a1), n2 = g.n(a2), n = g.n(axis);
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)LAP * n1 * n2) return;
  const int l = (int)(t % LAP);
  const int64_t r = t / LAP;
  int idx[3], w[3], w2[3];
  idx[a1] = (int)(r % n1) + 1;
  idx[a2] = (int)(r / n1) + 1;
  const int s = hi ? n + 1 + l : 1 - LAP + l;
  idx[axis] = s;
  for (int q = 0; q < 3; q++) { w[q] = idx[q]; w2[q] = idx[q]; }
  w[axis] = hi ? n : 1;
  w2[axis] = hi ? n - 1 : 2;
  const int64_t gq = g.hi(idx[0], idx[1], idx[2]), hw = g.hi(w[0], w[1], w[2]), hw2 = g.hi(w2[0], w2[1], w2[2]);
  for (int m = 0; m <= M_AJAC; m++) mp.p[m][gq] = mp.p[m][hw];
  const double fac = hi ? (double)(s - n) : (1.0 - s);
  double *C[3] = {X, Y, Z};
  for (int c = 0; c < 3; c++) C[c][gq] = C[c][hw] + fac * (C[c][hw] - C[c][hw2]);
}

// Jac_Ghost_Extent_2nd, coordinates: odd reflection (OCFD_GhostCell.f90:294-378)
__global__ __launch_bounds__(NT) void k_jac_ghost2_coords(const Geom g, const int side, double *X, double *Y,
                                                          double *Z) {
  const int axis = side / 2, hi = side % 2;
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const int n1 = g.n(a1), n2 = g.n(a2), n = g.n(axis);
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)LAP * n1 * n2) return;
  const int l = (int)(t % LAP);
  const int64_t r = t / LAP;
  int idx[3], w[3], mr[3];
  idx[a1] = (int)(r % n1) + 1;
  idx[a2] = (int)(r / n1) + 1;
  const int s = hi ? n + 1 + l : 1 - LAP + l;
  idx[axis] = s;
  for (int q = 0; q < 3; q++) { w[q] = idx[q]; mr[q] = idx[q]; }
  w[axis] = hi ? n : 1;
  mr[axis] = hi ? 2 * n - s : 2 - s;
  const int64_t gq = g.hi(idx[0], idx[1], idx[2]), hw = g.hi(w[0], w[1], w[2]), hm = g.hi(mr[0], mr[1], mr[2]);
  double *C[3] = {X, Y, Z};
  for (int c = 0; c < 3; c++) C[c][gq] = 2.0 * C[c][hw] - C[c][hm];
}

// comput_Jacobian3d_Ghost (OCFD_GhostCell.f90:394-500)
__device__ inline double ghost_deriv(const Geom &g, const double *A, int axis, int64_t h, int s, int lo, int hi,
                                     double hh) {
  const int64_t st = g.hstride(axis);
  auto X = [&](int o) { return A[h + (int64_t)o * st]; };
  if (s == lo) return (-(3.0 * X(0)) + 4.0 * X(1) - X(2)) / (2.0 * hh);
  if (s == hi) return (X(-2) - 4.0 * X(-1) + 3.0 * X(0)) / (2.0 * hh);
  if (s == lo + 1 || s == hi - 1) return (X(1) - X(-1)) / (2.0 * hh);
  return (8.0 * (X(1) - X(-1)) - (X(2) - X(-2))) / (12.0 * hh);
}
__global__ __launch_bounds__(NT) void k_jac_ghost2(const Block B, const int side, const double *X, const double *Y,
                                                   const double *Z, const MetPtrs mp) {
  const Geom &g = B.g;
  const int axis = side / 2, hi = side % 2;
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const int n1 = g.n(a1), n2 = g.n(a2), n = g.n(axis);
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)LAP * n1 * n2) return;
  const int l = (int)(t % LAP);
  const int64_t r = t / LAP;
  int idx[3];
  idx[a1] = (int)(r % n1) + 1;
  idx[a2] = (int)(r / n1) + 1;
  idx[axis] = hi ? n + 1 + l : 1 - LAP + l;
  int b1[3] = {1, 1, 1}, e1[3] = {g.nx, g.ny, g.nz};
  if (hi) { b1[axis] = n; e1[axis] = n + LAP; } else { b1[axis] = 1 - LAP; e1[axis] = 1; }
  const int64_t h = g.hi(idx[0], idx[1], idx[2]);
  const double *C[3] = {X, Y, Z};
  double dd[3][3];
  for (int c = 0; c < 3; c++)
    for (int a = 0; a < 3; a++) dd[c][a] = ghost_deriv(g, C[c], a, h, idx[a], b1[a], e1[a], B.P.h[a]);
  jac_point(mp.p, h, dd[0][0], dd[0][1], dd[0][2], dd[1][0], dd[1][1], dd[1][2], dd[2][0], dd[2][1], dd[2][2]);
}

// Jac_Ghost_Userdef (OCFD_GhostCell_UserDef.f90:4-12): the metrics and
// coordinates of the flat-plate j- ghost layers mirrored about j = 1
// (Jac_Ghost_flatplate :79-106) or of the swept-corner k- layers about the
// plane between k = 0 and 1 (Jac_Ghost_Sweptcorner :25-52), with the
// components normal to the plane negated
__global__ __launch_bounds__(NT) void k_jac_ghost_user(const Geom g, const int ibc, double *X, double *Y, double *Z,
                                                       const MetPtrs mp) {
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  int64_t h, hm;
  // sign of Akx Aky Akz Aix Aiy Aiz Asx Asy Asz Ajac, then X Y Z
  double sg[13];
  for (int q = 0; q < 13; q++) sg[q] = 1.0;
  if (ibc == 2) {
    if (t >= (int64_t)LAP * g.nx * g.nz) return;
    const int l = (int)(t % LAP);
    const int64_t r = t / LAP;
    const int i = (int)(r % g.nx) + 1, k = (int)(r / g.nx) + 1;
    const int j = 1 - LAP + l;
    h = g.hi(i, j, k);
    hm = g.hi(i, 2 - j, k);
    sg[1] = sg[3] = sg[5] = sg[7] = sg[11] = -1.0;  // Aky Aix Aiz Asy, Y
  } else {
    if (t >= (int64_t)LAP * g.nx * g.ny) return;
    const int l = (int)(t % LAP);
    const int64_t r = t / LAP;
    const int i = (int)(r % g.nx) + 1, j = (int)(r / g.nx) + 1;
    const int k = 1 - LAP + l;
    h = g.hi(i, j, k);
    hm = g.hi(i, j, 1 - k);
    sg[2] = sg[5] = sg[6] = sg[7] = sg[12] = -1.0;  // Akz Aiz Asx Asy, Z
  }
  for (int q = 0; q < 10; q++) mp.p[q][h] = sg[q] < 0.0 ? -mp.p[q][hm] : mp.p[q][hm];
  double *C[3] = {X, Y, Z};
  for (int c = 0; c < 3; c++) C[c][h] = sg[10 + c] < 0.0 ? -C[c][hm] : C[c][hm];
}

// Akx1 = Akx/Ajac ... over the whole halo'd block (OCFD_mesh.f90:41-55)
__global__ __launch_bounds__(NT) void k_normalise(const Geom g, const MetPtrs mp) {
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= g.nh) return;
  const double J = mp.p[M_AJAC][t];
  for (int m = 0; m < 9; m++) mp.p[M_AKX1 + m][t] = mp.p[M_AKX + m][t] / J;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
hipError_t halo_fill(const Block &B, void *const *fields, int nfields, int f32, hipStream_t st) {
  const Geom &g = B.g;
  const int px = B.S.iperiodic[0] == 1, py = B.S.iperiodic[1] == 1, pz = B.S.iperiodic[2] == 1;
  const int64_t tot = (px ? (int64_t)2 * LAP * g.ny * g.nz : 0) + (py ? (int64_t)2 * LAP * g.nx * g.nz : 0) +
                      (pz ? (int64_t)2 * LAP * g.nx * g.ny : 0);
  if (tot == 0) return hipSuccess;
  for (int f0 = 0; f0 < nfields; f0 += 8) {
    HaloFields F;
    const int nf = nfields - f0 < 8 ? nfields - f0 : 8;
    for (int q = 0; q < nf; q++) F.p[q] = fields[f0 + q];
    F.f32 = f32;
    hipLaunchKernelGGL(k_halo, dim3(nb(tot)), dim3(NT), 0, st, g, F, nf, px, py, pz);
  }
  return hipGetLastError();
}
}  // namespace

hipError_t launch_halo_fill(const Block &B, double *const *fields, int nfields, hipStream_t st) {
  void *p[M_N + 8];
  if (nfields > M_N + 8) return hipErrorInvalidValue;
  for (int q = 0; q < nfields; q++) p[q] = fields[q];
  return halo_fill(B, p, nfields, 0, st);
}

hipError_t launch_halo_prims(const Block &B, const PrimField *fields, int nfields, hipStream_t st) {
  void *p[5];
  if (nfields > 5) return hipErrorInvalidValue;
  for (int q = 0; q < nfields; q++) {
    if (fields[q].f32 != fields[0].f32) return hipErrorInvalidValue;
    p[q] = fields[q].p;
  }
  return halo_fill(B, p, nfields, fields[0].f32, st);
}

hipError_t launch_sensor(const Block &B, hipStream_t st) {
  const double p00 = 1.0 / (B.P.gamma * B.P.ma * B.P.ma);
  if (B.d.f32) hipLaunchKernelGGL(k_sensor1<float>, dim3(nb(B.g.ni)), dim3(NT), 0, st, B, p00);
  else hipLaunchKernelGGL(k_sensor1<double>, dim3(nb(B.g.ni)), dim3(NT), 0, st, B, p00);
  return hipGetLastError();
}

hipError_t launch_sensor2(const Block &B, double *work, hipStream_t st) {
  const double p00 = 1.0 / (B.P.gamma * B.P.ma * B.P.ma);
  const unsigned nblk = nb(B.g.ni);
  double *partial = work, *sum = work + nblk;
  hipLaunchKernelGGL(k_sensor2a, dim3(nblk), dim3(NT), 0, st, B, p00, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(NT), 0, st, partial, (int64_t)nblk, 1, 1, sum);
  hipLaunchKernelGGL(k_sensor2b, dim3(nblk), dim3(NT), 0, st, B, sum);
  return hipGetLastError();
}

hipError_t launch_hybrid_counts(const Block &B, unsigned long long *counts, hipStream_t st) {
  const int64_t nf = (int64_t)(B.g.nx + 1) * B.g.ny * B.g.nz;
  const unsigned nblk = nb(nf);
  hipLaunchKernelGGL(k_hyb_counts, dim3(nblk < 2048 ? nblk : 2048), dim3(NT), 0, st, B, counts);
  return hipGetLastError();
}

hipError_t launch_rk3(const Block &B, int krk, hipStream_t st) {
  if (B.prec == OCFD_PREC_MIXED_CODE) hipLaunchKernelGGL(k_rk3<float>, dim3(nb(5 * B.g.ni)), dim3(NT), 0, st, B, krk);
  else hipLaunchKernelGGL(k_rk3<double>, dim3(nb(5 * B.g.ni)), dim3(NT), 0, st, B, krk);
  return hipGetLastError();
}

hipError_t launch_copy_f(const Block &B, hipStream_t st) {
  hipLaunchKernelGGL(k_copy, dim3(nb(5 * B.g.ni)), dim3(NT), 0, st, (const double *)B.f, B.fn, 5 * B.g.ni);
  return hipGetLastError();
}

hipError_t launch_init_f(const Block &B, hipStream_t st) {
  hipLaunchKernelGGL(k_init_f, dim3(nb(B.g.ni)), dim3(NT), 0, st, B);
  return hipGetLastError();
}

hipError_t launch_duvwt(const Block &B, int *neg, hipStream_t st) {
  hipLaunchKernelGGL(k_duvwt, dim3(nb(B.g.ni)), dim3(NT), 0, st, B, neg);
  return hipGetLastError();
}

int64_t halo_tiles(const Geom &g) {
  return nb(2 * LAP * ((int64_t)g.ny * g.nz + (int64_t)g.nx * g.nz + (int64_t)g.nx * g.ny));
}

hipError_t launch_halo_batch(const BoxSet &S, int ntiles, hipStream_t st) {
  hipLaunchKernelGGL(k_halo_batch, dim3((unsigned)ntiles), dim3(NT), 0, st, S);
  return hipGetLastError();
}

int64_t halo_copy_tiles(const Geom &g) { return halo_copy_tiles_of(g); }

hipError_t launch_negt_repair_batch(const BoxSet &S, int nbox, int *neg, int ntiles, hipStream_t st) {
  hipLaunchKernelGGL(k_negt_repair_batch, dim3((unsigned)(nbox + ntiles)), dim3(NT), 0, st, S, nbox, neg);
  return hipGetLastError();
}

hipError_t launch_cfl_rate(const Block &B, unsigned long long *out, hipStream_t st) {
  const int64_t nblk = (B.g.ni + NT - 1) / NT;
  hipLaunchKernelGGL(k_cfl_rate, dim3((unsigned)(nblk < 2048 ? nblk : 2048)), dim3(NT), 0, st, B, out);
  return hipGetLastError();
}

hipError_t launch_negt_repair(const Block &B, int *neg, hipStream_t st, int uvw) {
  if (uvw && (!B.u2.p || !B.v2.p || !B.w2.p)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_negt_repair, dim3((unsigned)(1 + halo_copy_tiles_of(B.g))), dim3(NT), 0, st, B, neg, 1, uvw);
  return hipGetLastError();
}

hipError_t launch_bc(const Block &B, int ibc, double tw, const BLArgs *bl, const int *ghost, int planes,
                     hipStream_t st) {
  const Geom &g = B.g;
  if (ibc == 1 || ibc == 99) hipLaunchKernelGGL(k_bc_hook, dim3(nb((int64_t)g.nx * g.nz)), dim3(NT), 0, st, B, ibc, tw);
  if ((ibc == 2 || ibc == 3) && bl) {
    const int64_t nyz = (int64_t)g.ny * g.nz, nxz = (int64_t)g.nx * g.nz;
    hipLaunchKernelGGL(k_bl_inlet, dim3(nb(nyz)), dim3(NT), 0, st, B, *bl);
    hipLaunchKernelGGL(k_bl_upper, dim3(nb(nxz)), dim3(NT), 0, st, B, *bl);
    hipLaunchKernelGGL(k_bl_wall, dim3(nb(nxz)), dim3(NT), 0, st, B, *bl);
    if (bl->outlet == 1 || bl->outlet == 2) hipLaunchKernelGGL(k_bl_outlet, dim3(nb(nyz)), dim3(NT), 0, st, B, *bl);
    if (ibc == 3 && (bl->z1 == 1 || bl->z2 == 1))
      hipLaunchKernelGGL(k_bl_sides, dim3(nb((int64_t)g.nx * g.ny)), dim3(NT), 0, st, B, *bl);
  }
  const int64_t np = 2 * ((int64_t)g.ny * g.nz + (int64_t)g.nx * g.nz + (int64_t)g.nx * g.ny);
  if (planes) hipLaunchKernelGGL(k_bc_planes, dim3(nb(np)), dim3(NT), 0, st, B);
  for (int side = 0; side < 6; side++) {
    if (ghost[side] != 1 && ghost[side] != 2) continue;
    const int axis = side / 2;
    const int64_t cnt = (int64_t)LAP * (axis == 0 ? (int64_t)g.ny * g.nz : (axis == 1 ? (int64_t)g.nx * g.nz : (int64_t)g.nx * g.ny));
    hipLaunchKernelGGL(k_ghost, dim3(nb(cnt)), dim3(NT), 0, st, B, side);
  }
  // user ghost layers (Ghost_Cell = -1): flat plate j- / swept corner k-
  const int ibc_any = ibc < 0 ? -ibc : ibc;
  if (ibc_any == 2 && ghost[2] == -1)
    hipLaunchKernelGGL(k_ghost_user, dim3(nb((int64_t)LAP * g.nx * g.nz)), dim3(NT), 0, st, B, 2);
  if (ibc_any == 3 && ghost[4] == -1)
    hipLaunchKernelGGL(k_ghost_user, dim3(nb((int64_t)LAP * g.nx * g.ny)), dim3(NT), 0, st, B, 3);
  return hipGetLastError();
}

hipError_t launch_filter(const Block &B, int axis, int scheme, double s0, double rth, const int rng[6],
                         double *fout, double *p, hipStream_t st) {
  (void)p;
  FilterArgs A;
  A.axis = axis;
  A.scheme = scheme;
  A.s0 = s0;
  A.rth = rth;
  for (int q = 0; q < 3; q++) { A.lo[q] = rng[2 * q]; A.hi[q] = rng[2 * q + 1]; }
  const int n = B.g.n(axis);
  if (B.S.iperiodic[axis] != 1) {
    const int lo_min = scheme == 1 ? 6 : 4, hi_off = scheme == 1 ? 5 : 3;
    if (A.lo[axis] < lo_min) A.lo[axis] = lo_min;
    if (A.hi[axis] > n - hi_off) A.hi[axis] = n - hi_off;
  }
  hipLaunchKernelGGL(k_filter, dim3(nb(B.g.ni)), dim3(NT), 0, st, B, A, fout);
  return hipGetLastError();
}

int64_t flow_msg_partials(const Geom &g) { return nb(g.ni); }

hipError_t launch_reduce_flow_msg(const Block &B, double *partial, double *out, hipStream_t st) {
  const unsigned nblk = nb(B.g.ni);
  hipLaunchKernelGGL(k_flow_msg, dim3(nblk), dim3(NT), 0, st, B, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(NT), 0, st, (const double *)partial, (int64_t)nblk, 3, 3, out);
  return hipGetLastError();
}

hipError_t launch_ana_box(const Block &B, double *grad, double *amu, double *partial, double *out, hipStream_t st) {
  (void)grad;
  (void)amu;
  const unsigned nblk = nb(B.g.ni);
  hipLaunchKernelGGL(k_ana_box, dim3(nblk), dim3(NT), 0, st, B, partial, (const double *)nullptr);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(NT), 0, st, (const double *)partial, (int64_t)nblk, 14, 14, out);
  // pass 2: sum (d - <d>)^2 with <d> = out[6] / N
  hipLaunchKernelGGL(k_ana_box, dim3(nblk), dim3(NT), 0, st, B, partial, (const double *)(out + 6));
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(NT), 0, st, (const double *)partial, (int64_t)nblk, 1, 1, out + 14);
  return hipGetLastError();
}

hipError_t launch_metrics(const Block &B, int ibc, double *coords[3], double *met[M_N], const double spans[9],
                          const int ghost[6], double *grad, hipStream_t st) {
  (void)grad;
  const Geom &g = B.g;
  // halos of the coordinates, then the periodic shift
  (void)launch_halo_fill(B, coords, 3, st);
  for (int a = 0; a < 3; a++) {
    if (B.S.iperiodic[a] != 1) continue;
    const int64_t cnt = (int64_t)2 * LAP * (a == 0 ? (int64_t)g.ny * g.nz : (a == 1 ? (int64_t)g.nx * g.nz : (int64_t)g.nx * g.ny));
    hipLaunchKernelGGL(k_coord_shift, dim3(nb(cnt)), dim3(NT), 0, st, g, coords[0], coords[1], coords[2], a,
                       spans[3 * a], spans[3 * a + 1], spans[3 * a + 2]);
  }
  MetPtrs mp;
  for (int m = 0; m < M_N; m++) mp.p[m] = met[m];
  hipLaunchKernelGGL(k_jacobian, dim3(nb(g.ni)), dim3(NT), 0, st, B, (const double *)coords[0],
                     (const double *)coords[1], (const double *)coords[2], mp);
  (void)launch_halo_fill(B, met, M_AJAC + 1, st);
  // Jac_Ghost_Extent_1st for every side first, then _2nd (OCFD_GhostCell.f90:1-7)
  for (int side = 0; side < 6; side++) {
    if (ghost[side] != 1) continue;
    const int axis = side / 2;
    const int64_t cnt = (int64_t)LAP * (axis == 0 ? (int64_t)g.ny * g.nz : (axis == 1 ? (int64_t)g.nx * g.nz : (int64_t)g.nx * g.ny));
    hipLaunchKernelGGL(k_jac_ghost1, dim3(nb(cnt)), dim3(NT), 0, st, g, side, coords[0], coords[1], coords[2], mp);
  }
  for (int side = 0; side < 6; side++) {
    if (ghost[side] != 2) continue;
    const int axis = side / 2;
    const int64_t cnt = (int64_t)LAP * (axis == 0 ? (int64_t)g.ny * g.nz : (axis == 1 ? (int64_t)g.nx * g.nz : (int64_t)g.nx * g.ny));
    hipLaunchKernelGGL(k_jac_ghost2_coords, dim3(nb(cnt)), dim3(NT), 0, st, g, side, coords[0], coords[1], coords[2]);
    hipLaunchKernelGGL(k_jac_ghost2, dim3(nb(cnt)), dim3(NT), 0, st, B, side, (const double *)coords[0],
                       (const double *)coords[1], (const double *)coords[2], mp);
  }
  // Jac_Ghost_Userdef after the extrapolated layers
  if (ibc == 2 && ghost[2] == -1)
    hipLaunchKernelGGL(k_jac_ghost_user, dim3(nb((int64_t)LAP * g.nx * g.nz)), dim3(NT), 0, st, g, 2, coords[0],
                       coords[1], coords[2], mp);
  if (ibc == 3 && ghost[4] == -1)
    hipLaunchKernelGGL(k_jac_ghost_user, dim3(nb((int64_t)LAP * g.nx * g.ny)), dim3(NT), 0, st, g, 3, coords[0],
                       coords[1], coords[2], mp);
  return launch_normalise_metrics(B, met, st);
}

hipError_t launch_normalise_metrics(const Block &B, double *met[M_N], hipStream_t st) {
  MetPtrs mp;
  for (int m = 0; m < M_N; m++) mp.p[m] = met[m];
  hipLaunchKernelGGL(k_normalise, dim3(nb(B.g.nh)), dim3(NT), 0, st, B.g, mp);
  return hipGetLastError();
}

}  // namespace ocfd
