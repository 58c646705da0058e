// ocfd_batchq.hip -- the persistent queue kernel of batched steps, FP64
// instantiation, and its dispatchers (the FP32 one is in ocfd_batchq_mixed.hip).
#include "ocfd_batch_queue.h"

namespace ocfd {

hipError_t launch_batch_queue_mixed(const QPlan &P, int nwg, hipStream_t st);
int batch_queue_grid_mixed(int device);

int queue_shape(int axis) { return axis == 0 ? QSHAPE_X : (axis == 1 ? QSHAPE_Y : QSHAPE_Z); }

hipError_t launch_batch_queue(const QPlan &P, int prec, int nwg, hipStream_t st) {
  if (prec == OCFD_PREC_MIXED_CODE) return launch_batch_queue_mixed(P, nwg, st);
  return batch_queue_as<double>(P, nwg, st);
}

int batch_queue_grid(int prec, int device) {
  if (prec == OCFD_PREC_MIXED_CODE) return batch_queue_grid_mixed(device);
  return batch_queue_grid_as<double>(device);
}

}  // namespace ocfd
