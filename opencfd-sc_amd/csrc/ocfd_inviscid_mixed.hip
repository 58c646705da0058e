// ocfd_inviscid_mixed.hip -- the sweep kernels (ocfd_sweep.h) with FP32
// splitting and reconstruction, for the mixed-precision right-hand side.
// Built with FMA contraction allowed: this form is not bit-matched to the
// reference.
#include "ocfd_sweep.h"

namespace ocfd {

hipError_t launch_inviscid_mixed(const Block &B, int axis, int mode, int shape, hipStream_t st) {
  return launch_sweep<float>(B, axis, mode, shape, st);
}

}  // namespace ocfd
