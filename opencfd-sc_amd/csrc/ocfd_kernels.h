// ocfd_kernels.h -- launch-parameter blocks and host-side launchers of the
// gfx950 kernels (definitions in the .hip files of this directory).
#pragma once
#include "ocfd_common.h"

namespace ocfd {

// A halo'd primitive field (d, u, v, w or T): FP64, or FP32 in the
// mixed-precision form, where the primitives are derived values read only by
// the FP32 kernels (the conserved variables f stay FP64).  The sweeps and
// the viscous kernels take the typed pointer, as<R>() with R their real
// type; other code indexes the field and reads and writes doubles.
struct PrimField {
  void *p;
  int f32;
  template <typename R>
  __host__ __device__ R *as() const { return static_cast<R *>(p); }
  struct Ref {
    void *p;
    int f32;
    int64_t i;
    __device__ operator double() const {
      return f32 ? double(static_cast<const float *>(p)[i]) : static_cast<const double *>(p)[i];
    }
    __device__ Ref &operator=(double x) {
      if (f32) static_cast<float *>(p)[i] = float(x);
      else static_cast<double *>(p)[i] = x;
      return *this;
    }
    __device__ Ref &operator=(const Ref &r) { return *this = double(r); }
  };
  __device__ Ref operator[](int64_t i) const { return Ref{p, f32, i}; }
};

// Everything a kernel may need about the block; passed by value.
struct Block {
  Geom g;
  SchemeParams S;
  Phys P;
  Metrics M;
  PrimField d, u, v, w, T;    // halo'd primitives
  // The temperature the stage being computed writes.  comput_duvwT repairs a
  // cell with T <= 0 from its six neighbours in the reference's i-j-k loop,
  // so the -i/-j/-k neighbours hold this stage's T and the +i/+j/+k ones the
  // previous stage's (OCFD_Time_Adv.f90:45-58, OCFD_handle_NegativeT.f90:11).
  // The stage update therefore writes T2 and leaves T untouched; the repair
  // reads both, copies T's halo into T2, and the host swaps the two after
  // every stage (no extra HBM traffic per cell).
  PrimField T2;
  // The second set of u, v, w of the fused viscous stage update
  // (k_visc_fused, mixed-precision periodic blocks; p == nullptr elsewhere):
  // it writes the new values there while other workgroups still read the
  // stage-start ones, and the host swaps the sets after every fused stage.
  PrimField u2, v2, w2;
  double *f, *fn, *du;        // 5 planes each
  double *rhyb;               // halo'd sensor (hybrid scheme only)
  double *ev;                 // 12 interior planes: Ev1(1:4), Ev2(1:4), Ev3(1:4)
  double *trash;              // 16 x 256 doubles: stores of lanes outside the block
  // hybrid scheme, per sweep axis: the faces that use WENO7/WENO5 (face id
  // * 2 + (WENO5)), their count, the face -> list slot map and the fluxes of
  // the listed faces (5 per face characteristic, 10 otherwise).  The map is
  // kept per group of 64 consecutive face ids: a bit mask of the listed
  // faces and the slot of the group's first listed face (slots follow the
  // face order), i.e. 12 bytes per 64 faces.
  int *hlist[3], *hcount;
  unsigned long long *hbits[3];
  int *hgbase[3];
  double *wflux[3];
  int prec;  // OCFD_PREC_FP64_CODE / OCFD_PREC_MIXED_CODE (host-side dispatch)
  int kseg;  // boxes of a batch: planes per workgroup of the k-marching kernels
  // Whole-step launches only (NULL otherwise): set by the repair of the first
  // stage with more than 10 cells T <= 0 (handle_NegativeT stops the run
  // there, OCFD_handle_NegativeT.f90:13-16).  From then on the stage update
  // only carries T over, and the repair, boundary and filter kernels do
  // nothing, so the state stays that of the failing stage.
  int *halt;
};
// the kernels of a precision read the primitives in their real type: FP32
// fields belong to the mixed-precision form, FP64 fields to the FP64 path
inline bool prims_match(const Block &B) {
  const int f32 = B.prec == OCFD_PREC_MIXED_CODE;
  return B.d.f32 == f32 && B.u.f32 == f32 && B.v.f32 == f32 && B.w.f32 == f32 && B.T.f32 == f32 &&
         B.T2.f32 == f32 && (!B.u2.p || B.u2.f32 == f32) && (!B.v2.p || B.v2.f32 == f32) &&
         (!B.w2.p || B.w2.f32 == f32);
}
__device__ inline bool halted(const Block &B) { return B.halt != nullptr && *B.halt != 0; }

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  xcd_tile
// turns the dispatch index into a tile index such that every XCD works
// through one contiguous range of tiles, so neighbouring tiles, which read
// each other's halo points, share an L2.  A bijection on [0, nb).
constexpr int NXCD = 8;
__device__ inline int xcd_tile(int b, int nb) {
  const int q = nb / NXCD, r = nb % NXCD;
  const int x = b % NXCD, i = b / NXCD;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// ints per block of the T <= 0 bookkeeping: [0] cells found in the current
// stage, [1..64] their interior indices, [65] total, [66] largest per stage,
// [NEG_HALT] halt flag (Block::halt), [NEG_HALT_STAGE] the stage that set it,
// [NEG_STAGES] stages completed since the counters were cleared
constexpr int NEG_STRIDE = 72;
constexpr int NEG_HALT = 67, NEG_HALT_STAGE = 68, NEG_STAGES = 69;
constexpr int NEG_MAX = 10;  // Max_NegT (OCFD_handle_NegativeT.f90:6)

// The boxes of one batched launch (ocfd_batch): workgroup tiles of box
// ids[j] are off[j] .. off[j+1]-1 of the launch.  box() finds the box of a
// tile by binary search over the m entries (the tile index is uniform over
// the workgroup, so the search runs on scalar registers).
struct BoxSet {
  const Block *blk;  // the batch's blocks (device table)
  const int *ids;    // boxes of this launch
  const int *off;    // m + 1 tile offsets
  int m;
  __device__ inline int entry(int t) const {
    int lo = 0, hi = m - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }
};

// Index arithmetic of a sweep along AXIS: point s of line w, third index c
// (x: lines j, third k; y: lines i, third k; z: lines i, third j).
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
  __host__ __device__ static inline int nwidth(const Geom &g) { return AXIS == 0 ? g.ny : g.nx; }
  __host__ __device__ static inline int nthird(const Geom &g) { return AXIS == 2 ? g.ny : g.nz; }
  // faces s = 0..n of every line, in the order the sweep tiles read them
  // (x: s fastest; y, z: the line index fastest)
  __host__ __device__ static inline int64_t nfaces(const Geom &g) {
    return (int64_t)(g.n(AXIS) + 1) * nwidth(g) * nthird(g);
  }
  __host__ __device__ static inline int64_t ngroups(const Geom &g) { return (nfaces(g) + 63) / 64; }
  __device__ static inline int64_t face_id(const Geom &g, int s, int w, int c) {
    if (AXIS == 0) return ((int64_t)(c - 1) * g.ny + (w - 1)) * (g.nx + 1) + s;
    return ((int64_t)(c - 1) * (g.n(AXIS) + 1) + s) * g.nx + (w - 1);
  }
  // in 32-bit unsigned arithmetic: only the hybrid scheme's face lists use
  // it, and ocfd_create refuses hybrid blocks with 2^30 or more faces per
  // axis (the 64-bit divisions took a tenth of the list kernel's time)
  __device__ static inline void face_of(const Geom &g, int64_t fid, int &s, int &w, int &c) {
    const unsigned f = (unsigned)fid;
    if (AXIS == 0) {
      const unsigned nl = (unsigned)g.nx + 1u, r = f / nl;
      s = (int)(f - r * nl);
      const unsigned r2 = r / (unsigned)g.ny;
      w = (int)(r - r2 * (unsigned)g.ny) + 1;
      c = (int)r2 + 1;
    } else {
      const unsigned r = f / (unsigned)g.nx;
      w = (int)(f - r * (unsigned)g.nx) + 1;
      const unsigned nl = (unsigned)g.n(AXIS) + 1u, r2 = r / nl;
      s = (int)(r - r2 * nl);
      c = (int)r2 + 1;
    }
  }
};

enum InvMode : int { INV_ASSIGN = 0, INV_ACCUM = 1, INV_FINAL = 2 };

// Launch shape of one sweep: 0 is the built-in tile of each axis (x 64x4,
// y and z 32x8 points x lines); 1 (x 32x8, y/z 16x16) and 2 (x 128x2, y/z
// 64x4) exist for periodic axes with uniform metrics; 3 and 4 are the
// marches (one lane per line segment, ocfd_march.h: segments of 32 / 64
// cells for y and z, runs of 8 / 16 cells for x) of periodic axes with
// uniform metrics, for the non-characteristic schemes and the characteristic
// hybrid scheme.  The autotuner (ocfd_autotune) chooses
// among the shapes a sweep supports; a shape a sweep does not support runs
// as shape 0.
constexpr int NSWEEP_SHAPES = 5;
struct SweepTile {
  int shape;
};
// cells per segment of march shape `shape` (3, 4) along `axis`
inline int march_segment(int axis, int shape) {
  return axis == 0 ? (shape == 4 ? 16 : 8) : (shape == 4 ? 64 : 32);
}
// whether `shape` runs on a sweep of `axis` (fast: periodic axis, uniform
// metrics; sch, ch: scheme and characteristic flag)
inline bool sweep_shape_supported(int axis, int shape, bool fast, int sch, bool ch) {
  if (shape == 0) return true;
  if (!fast) return false;
  if (shape <= 2) return true;
  return !ch || sch == SCH_HYBRID;
}
// the shape that actually runs
inline int sweep_shape_effective(int axis, int shape, bool fast, int sch, bool ch) {
  return sweep_shape_supported(axis, shape, fast, sch, ch) ? shape : 0;
}
// points along the axis of a tile of `shape` 0-2 (256 threads: lines = 256 / TL)
inline int sweep_shape_tl(int axis, int shape) {
  return axis == 0 ? (shape == 1 ? 32 : (shape == 2 ? 128 : 64)) : (shape == 1 ? 16 : (shape == 2 ? 64 : 32));
}
// workgroups of one sweep of a block in an effective shape
inline int64_t sweep_tiles(const Geom &g, int axis, int shape) {
  const int n = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
  if (shape >= 3) {
    const int sl = march_segment(axis, shape);
    const int nseg = (n + sl - 1) / sl;
    // x: one lane per run, 256 lanes per workgroup
    if (axis == 0) return ((int64_t)nseg * g.ny * g.nz + 255) / 256;
    // y, z: 4 segments of 64 lines per workgroup
    const int nsg = (nseg + 3) / 4;
    return (int64_t)((g.nx + 63) / 64) * nsg * (axis == 2 ? g.ny : g.nz);
  }
  const int tl = sweep_shape_tl(axis, shape), tw = 256 / tl;
  const int nw = axis == 0 ? g.ny : g.nx;
  const int nc = axis == 2 ? g.ny : g.nz;
  return (int64_t)((n + tl - 1) / tl) * ((nw + tw - 1) / tw) * nc;
}

// Data of the boundary-layer / swept-corner conditions (ocfd_bc_BoundaryLayer,
// ocfd_bc_swept_corner, OCFD_BC_Boundarylayers.f90): device arrays the
// caller provides (ocfd_set_bc_boundarylayer) and the parameters taken from
// BC_para.  inlet [5][nz][ny] and upper [5][nz][nx] (NULL: free stream),
// nonref [nz][nx] (1: non-reflecting upper point, no Dirichlet value),
// pertb [nz][nx] wall blow-and-suction shape, wall [nz][nx] (swept corner:
// 1 = no-slip wall point, 0 = extrapolated; NULL: all wall), ht the current
// time factor of the perturbation (one value, read by the kernels so that
// captured steps see updates).
struct BLArgs {
  const double *inlet, *upper, *pertb, *ht;
  const int *nonref, *wall;
  double u0, v0, Tw, A;  // free-stream direction, wall temperature, perturbation amplitude
  double inletT;         // T of the free-stream inlet (the reference sets 0)
  int outlet, z1, z2;    // outlet extrapolation order (0 none), side extrapolation flags
};

// ---- launchers (return hipError_t of the launch) ----
hipError_t launch_inviscid(const Block &B, int axis, int mode, SweepTile tile, hipStream_t st);
hipError_t launch_hybrid_lists(const Block &B, hipStream_t st);
// batched launches over the boxes of S (ntiles workgroups in all)
hipError_t launch_inviscid_batch(const BoxSet &S, int prec, int axis, int sch, bool ch, int mode, int shape,
                                 int ntiles, hipStream_t st);
hipError_t launch_visc_flux_batch(const BoxSet &S, int prec, bool cd8, int ntiles, hipStream_t st);
hipError_t launch_visc_div_update_batch(const BoxSet &S, int prec, bool visc, bool cd8, int krk, int *neg,
                                        int ntiles, hipStream_t st);
hipError_t launch_halo_batch(const BoxSet &S, int ntiles, hipStream_t st);
// the T <= 0 repair of every box (workgroup b repairs box b of the table) and
// the copy of T's halo into T2, over the halo-copy tiles of S (for boxes
// with second u, v, w buffers, i.e. a fused stage, also the halos of u2, v2,
// w2 into u, v, w)
hipError_t launch_negt_repair_batch(const BoxSet &S, int nbox, int *neg, int ntiles, hipStream_t st);
// workgroups of the T -> T2 halo copy of one box
int64_t halo_copy_tiles(const Geom &g);
// workgroups of the k-marching kernels of a block marched in segments of
// kseg planes (a batch fills the chip with many boxes; its autotuner picks
// the segment length from BATCH_KSEGS)
int64_t march_tiles(const Geom &g, int kseg);
constexpr int NBATCH_KSEG = 4;
constexpr int BATCH_KSEGS[NBATCH_KSEG] = {4, 8, 16, 32};
// workgroups of the periodic halo fill of one box
int64_t halo_tiles(const Geom &g);

// One persistent launch per batched time step (ocfd_batch, queue mode): the
// per-phase launches become segments of a work queue of (segment, tile)
// items; Q_NKIND phase kinds per stage.
enum QueueKind : int { Q_HALO = 0, Q_X, Q_Y, Q_Z, Q_VFLUX, Q_UPDATE, Q_REPAIR, Q_NKIND };
struct QSeg {
  int kind, stage;     // phase kind, RK stage 1..3
  int sch, visc, cd8;  // sweeps: scheme; flux / update: viscous terms, CD8
  int ids, off, m;     // box list and tile offsets in the batch's table (BoxSet)
  int first, ntiles;   // first ticket, items
};
struct QPlan {
  const Block *blk0, *blk1;  // descriptor tables; stage s reads table q0 ^ ((s - 1) & 1)
  int q0;
  const int *tab;
  const QSeg *seg;
  int nseg, nitems;
  int *head;        // next ticket
  int *err;         // set when a wait ran out (not cleared between steps)
  int *done;        // [box][stage][Q_NKIND] finished items
  const int *need;  // [box][Q_NKIND] items of each kind in one stage
  int *neg;         // T <= 0 counters, NEG_STRIDE per box
};
// the sweep shape (0-4) the queue kernel runs along `axis`
int queue_shape(int axis);
hipError_t launch_batch_queue(const QPlan &P, int prec, int nwg, hipStream_t st);
// workgroups of the queue kernel resident at once (< 0: query failed)
int batch_queue_grid(int prec, int device);
hipError_t launch_weno_faces(const Block &B, int axis, hipStream_t st);
hipError_t launch_halo_fill(const Block &B, double *const *fields, int nfields, hipStream_t st);
// the same for primitive fields (at most 5, all of one precision)
hipError_t launch_halo_prims(const Block &B, const PrimField *fields, int nfields, hipStream_t st);
hipError_t launch_sensor(const Block &B, hipStream_t st);
hipError_t launch_sensor2(const Block &B, double *work, hipStream_t st);
hipError_t launch_hybrid_counts(const Block &B, unsigned long long *counts, hipStream_t st);
hipError_t launch_visc_flux(const Block &B, hipStream_t st);
hipError_t launch_visc_div(const Block &B, hipStream_t st);
hipError_t launch_visc_div_update(const Block &B, int krk, int refresh_planes, int *neg, hipStream_t st);
hipError_t launch_rk3(const Block &B, int krk, hipStream_t st);
hipError_t launch_copy_f(const Block &B, hipStream_t st);
hipError_t launch_duvwt(const Block &B, int *neg, hipStream_t st);
// T <= 0 repair of the cells listed in neg (T2 from T2 / T neighbours) and
// the copy of T's halo into T2, in one launch
// (uvw: also copy the halos of u2, v2, w2 into u, v, w -- after a fused
// stage, whose new u, v, w the host has swapped in)
hipError_t launch_negt_repair(const Block &B, int *neg, hipStream_t st, int uvw = 0);
// the viscous terms fused into the stage update (k_visc_fused): whether this
// block can take it, and the launch (writes u2, v2, w2 and T2)
bool visc_fused_supported(const Block &B);
hipError_t launch_visc_fused(const Block &B, int krk, int refresh, int *neg, hipStream_t st);
// the same over the boxes of a batch (mixed precision; boxes whose handles
// run it alone), and its workgroups for one box at segment length kseg
hipError_t launch_visc_fused_batch(const BoxSet &S, int krk, int *neg, int ntiles, hipStream_t st);
int64_t fused_batch_tiles(const Geom &g, int kseg);
// workgroups of the show_flow_msg reduction (3 partial sums each)
int64_t flow_msg_partials(const Geom &g);
// largest CFL rate over the cells, as the bit pattern of a positive double,
// combined into *out with atomicMax (*out must hold 0 before)
hipError_t launch_cfl_rate(const Block &B, unsigned long long *out, hipStream_t st);
hipError_t launch_bc(const Block &B, int ibc, double tw, const BLArgs *bl, const int *ghost, int planes,
                     hipStream_t st);
hipError_t launch_init_f(const Block &B, hipStream_t st);
hipError_t launch_filter(const Block &B, int axis, int scheme, double s0, double rth, const int rng[6],
                         double *fout, double *p, hipStream_t st);
hipError_t launch_reduce_flow_msg(const Block &B, double *partial, double *out, hipStream_t st);
hipError_t launch_ana_box(const Block &B, double *grad, double *amu, double *partial, double *out,
                          hipStream_t st);
hipError_t launch_metrics(const Block &B, int ibc, double *coords[3], double *met[M_N], const double spans[9],
                          const int ghost[6], double *grad, hipStream_t st);
hipError_t launch_normalise_metrics(const Block &B, double *met[M_N], hipStream_t st);

// threads per workgroup of the element-wise, viscous and stage kernels
constexpr int NT = 256;

// interior (i, j, k), 1-based, of cell t of the nx * ny * nz block (i fastest)
__device__ inline void cell_of(const Geom &g, int64_t t, int &i, int &j, int &k) {
  i = (int)(t % g.nx) + 1;
  int64_t r = t / g.nx;
  j = (int)(r % g.ny) + 1;
  k = (int)(r / g.ny) + 1;
}

}  // namespace ocfd
