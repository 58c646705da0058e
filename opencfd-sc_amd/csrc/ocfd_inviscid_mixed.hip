// ocfd_inviscid_mixed.hip -- the sweep kernels (ocfd_sweep.h) with FP32
// splitting and reconstruction, for the mixed-precision right-hand side.
#include "ocfd_sweep.h"

namespace ocfd {

hipError_t launch_inviscid_mixed(const Block &B, int axis, int mode, int shape, hipStream_t st) {
  return launch_sweep<float>(B, axis, mode, shape, st);
}

hipError_t launch_inviscid_batch_mixed(const BoxSet &S, int axis, int sch, bool ch, int mode, int shape, int ntiles,
                                       hipStream_t st) {
  return launch_sweep_batch<float>(S, axis, sch, ch, mode, shape, ntiles, st);
}

}  // namespace ocfd
