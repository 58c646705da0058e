// ocfd_inviscid_mixed.hip -- the sweep kernels (ocfd_sweep.h) with FP32
// splitting and reconstruction, for the mixed-precision right-hand side.
#include "ocfd_sweep.h"

namespace ocfd {

hipError_t launch_inviscid_mixed(const Block &B, int axis, int mode, hipStream_t st) {
  return launch_sweep<float>(B, axis, mode, st);
}

}  // namespace ocfd
