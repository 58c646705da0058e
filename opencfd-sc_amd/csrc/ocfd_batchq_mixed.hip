// ocfd_batchq_mixed.hip -- the persistent queue kernel of batched steps,
// FP32 instantiation (the mixed-precision form).
#include "ocfd_batch_queue.h"

namespace ocfd {

hipError_t launch_batch_queue_mixed(const QPlan &P, int nwg, hipStream_t st) { return batch_queue_as<float>(P, nwg, st); }

int batch_queue_grid_mixed(int device) { return batch_queue_grid_as<float>(device); }

}  // namespace ocfd
