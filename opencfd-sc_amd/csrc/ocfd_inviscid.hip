// ocfd_inviscid.hip -- the FP64 sweep kernels (ocfd_sweep.h) and the launcher
// of du_inviscous / du_inviscous_Character (OCFD_inviscous.f90:53-203,
// OCFD_inviscous_character.f90:3-159), which picks the FP64 or the
// mixed-precision instantiation.
#include "ocfd_sweep.h"

namespace ocfd {

hipError_t launch_inviscid_mixed(const Block &B, int axis, int mode, int shape, hipStream_t st);

hipError_t launch_inviscid(const Block &B, int axis, int mode, SweepTile tile, hipStream_t st) {
  if (tile.shape < 0 || tile.shape >= NSWEEP_SHAPES || !prims_match(B)) return hipErrorInvalidValue;
  if (B.prec == OCFD_PREC_MIXED_CODE) return launch_inviscid_mixed(B, axis, mode, tile.shape, st);
  return launch_sweep<double>(B, axis, mode, tile.shape, st);
}

hipError_t launch_inviscid_batch_mixed(const BoxSet &S, int axis, int sch, bool ch, int mode, int shape, int ntiles,
                                       hipStream_t st);

hipError_t launch_inviscid_batch(const BoxSet &S, int prec, int axis, int sch, bool ch, int mode, int shape,
                                 int ntiles, hipStream_t st) {
  if (shape < 0 || shape >= NSWEEP_SHAPES) return hipErrorInvalidValue;
  if (prec == OCFD_PREC_MIXED_CODE) return launch_inviscid_batch_mixed(S, axis, sch, ch, mode, shape, ntiles, st);
  return launch_sweep_batch<double>(S, axis, sch, ch, mode, shape, ntiles, st);
}

}  // namespace ocfd
