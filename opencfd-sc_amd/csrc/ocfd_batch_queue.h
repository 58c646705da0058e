// ocfd_batch_queue.h -- one persistent launch per time step of a batch of
// periodic boxes (ocfd_batch with the queue launch mode).
//
// The per-phase launches of a batched step (halo fill, x/y/z sweeps per
// scheme group, viscous flux, stage update, T <= 0 repair; three stages)
// become segments of one work queue.  An item is (segment, tile): the tile
// of one box in one phase of one stage, the unit the per-phase kernel gives
// a workgroup.  A fixed grid of workgroups takes tickets from a counter in
// queue order; before running an item, a workgroup waits until every item
// of the same box in the phase before it has finished (per-box, per-phase
// completion counters).  Items of other boxes never wait for each other, so
// a box whose phase has finished moves on while the last tiles of a larger
// box still run: no grid-wide barrier remains between phases or stages.
//
// Progress: a wait is only ever for items with smaller tickets, which were
// taken by running workgroups or have finished, so the item with the
// smallest unfinished ticket always has its inputs and completes.  Every
// wait is bounded (QUEUE_SPIN_MAX polls, about a second); one that runs out
// sets the error flag, which the host reports, after which no item waits,
// so the kernel always ends soon.
//
// Each item calls the same tile function, with the same arguments, as its
// per-phase launch, so results are bit-identical to those launches.
#pragma once
#include "ocfd_stage_parts.h"
#include "ocfd_sweep.h"
#include "ocfd_visc.h"

namespace ocfd {

namespace {

constexpr int QUEUE_SPIN_MAX = 1 << 19;

// waves per SIMD the queue kernel is compiled for: those of its heaviest
// phases (the stage update and the marches) in their own launches
template <typename R>
constexpr int QUEUE_WAVES = sizeof(R) == 8 ? 1 : 2;

// the sweep shapes the queue kernel runs (x: 64 x 4 tiles; y, z: marches of
// 32-point segments); the per-phase launches take the autotuned ones
constexpr int QSHAPE_X = 0, QSHAPE_Y = 3, QSHAPE_Z = 3;

template <int AXIS, int SCH, int SHAPE, typename R>
__device__ __forceinline__ void q_sweep_scheme(const Block &B, int lt, int mode) {
  if constexpr (SHAPE >= 3) {
    constexpr int SL = AXIS == 0 ? (SHAPE == 4 ? 16 : 8) : (SHAPE == 4 ? 64 : 32);
    if constexpr (AXIS == 0) march_tile_x<SCH, false, R>(B, lt, mode, SL);
    else march_tile<AXIS, SCH, false, R>(B, lt, mode, SL);
  } else {
    constexpr int TL = shape_tl<AXIS>(SHAPE), TW = NTHREADS / TL;
    sweep_tile<AXIS, SCH, false, TL, TW, true, R>(B, lt, mode);
  }
}

template <int AXIS, int SHAPE, typename R>
__device__ __forceinline__ void q_sweep(int sch, const Block &B, int lt, int mode) {
  switch (sch) {
    case SCH_WENO5: q_sweep_scheme<AXIS, SCH_WENO5, SHAPE, R>(B, lt, mode); break;
    case SCH_WENO7: q_sweep_scheme<AXIS, SCH_WENO7, SHAPE, R>(B, lt, mode); break;
    case SCH_OMP6: q_sweep_scheme<AXIS, SCH_OMP6, SHAPE, R>(B, lt, mode); break;
    default: q_sweep_scheme<AXIS, SCH_UD7L, SHAPE, R>(B, lt, mode); break;
  }
}

// counter of (box, stage, kind)
__device__ inline int *q_counter(const QPlan &P, int box, int stage, int kind) {
  return P.done + ((int64_t)box * 3 + (stage - 1)) * Q_NKIND + kind;
}

// the kind an item of `kind` waits for in its box (Q_NKIND: the previous
// stage's repair; -1: nothing)
__device__ inline int q_dependency(const QPlan &P, int box, int kind) {
  if (kind == Q_HALO) return Q_NKIND;
  if (kind == Q_UPDATE && P.need[box * Q_NKIND + Q_VFLUX] == 0) return Q_Z;
  return kind - 1;
}

template <typename R>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(QUEUE_WAVES<R>, QUEUE_WAVES<R>)))
void k_batch_queue(const QPlan P) {
  __shared__ int s_ticket;
  // Thread 0 takes the tickets, the workgroup reads them as a scalar: the
  // loop test and every branch on the item are uniform, and the only
  // divergent regions (thread 0's ticket, wait and counter) close inside the
  // loop body.  With the fetch as a divergent region at the top of the loop
  // and the ticket in a vector register, the compiler splits the loop so
  // that all lanes but thread 0 return to the ticket barrier without thread
  // 0's next fetch, and the workgroup repeats its first item for ever.
  if (threadIdx.x == 0) s_ticket = __hip_atomic_fetch_add(P.head, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  int tk = __builtin_amdgcn_readfirstlane(s_ticket);
  while (tk < P.nitems) {
    int si = 0;
    while (si + 1 < P.nseg && P.seg[si + 1].first <= tk) si++;
    const QSeg &sg = P.seg[si];
    const int kind = sg.kind, stage = sg.stage;
    const BoxSet S{(P.q0 ^ ((stage - 1) & 1)) ? P.blk1 : P.blk0, P.tab + sg.ids, P.tab + sg.off, sg.m};
    const int t = tk - sg.first;
    const int j = S.entry(t);
    const int box = S.ids[j];
    const int lt = t - S.off[j];

    // wait for the box's previous phase
    if (threadIdx.x == 0) {
      const int dk = q_dependency(P, box, kind);
      if (!(dk == Q_NKIND && stage == 1)) {
        const int ds = dk == Q_NKIND ? stage - 1 : stage, dkind = dk == Q_NKIND ? Q_REPAIR : dk;
        const int need = P.need[box * Q_NKIND + dkind];
        const int *c = q_counter(P, box, ds, dkind);
        // relaxed polls (an acquire on every poll would invalidate the L2
        // each time, for every waiting workgroup), one acquire fence after
        int spins = 0;
        while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          __builtin_amdgcn_s_sleep(8);
          // after one wait has run out, none waits any more: the kernel
          // then drains at once
          if ((++spins & 63) == 0 && __hip_atomic_load(P.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
          if (spins > QUEUE_SPIN_MAX) {
            __hip_atomic_store(P.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();

    switch (kind) {
      case Q_HALO: {
        const Block &B = S.blk[box];
        void *const F[5] = {B.d.p, B.u.p, B.v.p, B.w.p, B.T.p};
        const int64_t pt = (int64_t)lt * NT + threadIdx.x;
        if (B.d.f32) halo_point<float>(B.g, F, 5, 1, 1, 1, pt);
        else halo_point<double>(B.g, F, 5, 1, 1, 1, pt);
        break;
      }
      case Q_X: q_sweep<0, QSHAPE_X, R>(sg.sch, S.blk[box], lt, INV_ASSIGN); break;
      case Q_Y: q_sweep<1, QSHAPE_Y, R>(sg.sch, S.blk[box], lt, INV_ACCUM); break;
      case Q_Z: q_sweep<2, QSHAPE_Z, R>(sg.sch, S.blk[box], lt, INV_FINAL); break;
      case Q_VFLUX: {  // CD6 (the host keeps CD8 batches on per-phase launches)
        const Block B = S.blk[box];
        visc_flux_tile<false, true, R>(B, batch_kseg_dev(B), lt);
        break;
      }
      case Q_UPDATE: {
        const Block B = S.blk[box];
        int *neg = P.neg + NEG_STRIDE * box;
        const int ks = batch_kseg_dev(B);
        if (sg.visc) visc_div_update_tile<true, false, true, R>(B, stage, 1, ks, neg, lt);
        else visc_div_update_tile<false, false, true, R>(B, stage, 1, ks, neg, lt);
        break;
      }
      default: {  // Q_REPAIR: item 0 of a box repairs, the others copy T's halo into T2
        const Block &B = S.blk[box];
        if (lt == 0) {
          if (threadIdx.x == 0) negt_repair_body(B, P.neg + NEG_STRIDE * box, 1);
        } else {
          halo_copy_tile(B, lt - 1, 0);
        }
        break;
      }
    }

    // publish the item: every wave's stores (a release per wave, so that
    // the stores of all four are complete and written back), then the
    // counter; then the next ticket
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
      __hip_atomic_fetch_add(q_counter(P, box, stage, kind), 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      s_ticket = __hip_atomic_fetch_add(P.head, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    tk = __builtin_amdgcn_readfirstlane(s_ticket);
  }
}

// dynamic LDS of the queue kernel: the largest sweep tile it runs
template <typename R>
constexpr size_t queue_lds() {
  size_t m = 0;
  if (QSHAPE_X < 3) {
    constexpr int TL = shape_tl<0>(QSHAPE_X);
    m = sweep_lds<TL, NTHREADS / TL, false, false, R>(4);
  }
  if (QSHAPE_Y < 3) {
    constexpr int TL = shape_tl<1>(QSHAPE_Y < 3 ? QSHAPE_Y : 0);
    const size_t b = sweep_lds<TL, NTHREADS / TL, false, false, R>(4);
    m = b > m ? b : m;
  }
  if (QSHAPE_Z < 3) {
    constexpr int TL = shape_tl<2>(QSHAPE_Z < 3 ? QSHAPE_Z : 0);
    const size_t b = sweep_lds<TL, NTHREADS / TL, false, false, R>(4);
    m = b > m ? b : m;
  }
  return m;
}

template <typename R>
hipError_t batch_queue_as(const QPlan &P, int nwg, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_queue<R>, dim3((unsigned)nwg), dim3(NT), queue_lds<R>(), st, P);
  return hipGetLastError();
}

// workgroups of the queue kernel that fit on the device at once
template <typename R>
int batch_queue_grid_as(int device) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_queue<R>, NT, queue_lds<R>()) != hipSuccess)
    return -1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return -1;
  return per_cu * cus;
}

}  // namespace

}  // namespace ocfd
