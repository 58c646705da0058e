// ocfd_kernels.h -- launch-parameter blocks and host-side launchers of the
// gfx950 kernels (definitions in the .hip files of this directory).
#pragma once
#include "ocfd_common.h"

namespace ocfd {

// Everything a kernel may need about the block; passed by value.
struct Block {
  Geom g;
  SchemeParams S;
  Phys P;
  Metrics M;
  double *d, *u, *v, *w, *T;  // halo'd primitives
  double *f, *fn, *du;        // 5 planes each
  double *rhyb;               // halo'd sensor (hybrid scheme only)
  double *ev;                 // 12 interior planes: Ev1(1:4), Ev2(1:4), Ev3(1:4)
};

enum InvMode : int { INV_ASSIGN = 0, INV_ACCUM = 1, INV_FINAL = 2 };

// Launch shape of one sweep (0: the built-in tile of each axis).
struct SweepTile {
  int shape;
};

// ---- launchers (return hipError_t of the launch) ----
hipError_t launch_inviscid(const Block &B, int axis, int mode, SweepTile tile, hipStream_t st);
hipError_t launch_halo_fill(const Block &B, double *const *fields, int nfields, hipStream_t st);
hipError_t launch_sensor(const Block &B, hipStream_t st);
hipError_t launch_sensor2(const Block &B, double *work, hipStream_t st);
hipError_t launch_hybrid_counts(const Block &B, unsigned long long *counts, hipStream_t st);
hipError_t launch_visc_flux(const Block &B, hipStream_t st);
hipError_t launch_visc_div(const Block &B, hipStream_t st);
hipError_t launch_visc_div_update(const Block &B, int krk, int refresh_planes, int *neg, hipStream_t st);
hipError_t launch_rk3(const Block &B, int krk, hipStream_t st);
hipError_t launch_copy_f(const Block &B, hipStream_t st);
hipError_t launch_duvwt(const Block &B, int *neg, hipStream_t st);
hipError_t launch_negt_repair(const Block &B, int *neg, hipStream_t st);
hipError_t launch_bc(const Block &B, int ibc, double tw, const int *ghost, int planes, hipStream_t st);
hipError_t launch_init_f(const Block &B, hipStream_t st);
hipError_t launch_filter(const Block &B, int axis, int scheme, double s0, double rth, const int rng[6],
                         double *fout, double *p, hipStream_t st);
hipError_t launch_reduce_flow_msg(const Block &B, double *partial, double *out, hipStream_t st);
hipError_t launch_ana_box(const Block &B, double *grad, double *amu, double *partial, double *out,
                          hipStream_t st);
hipError_t launch_metrics(const Block &B, double *coords[3], double *met[M_N], const double spans[9],
                          const int ghost[6], double *grad, hipStream_t st);
hipError_t launch_normalise_metrics(const Block &B, double *met[M_N], hipStream_t st);

}  // namespace ocfd
