// Batched solves (placeholder until the persistent batched sweep lands).
#include <hip/hip_runtime.h>

#include "fdtd_mi355x.h"

extern "C" {

int fdtd_batch_create(int, fdtd_grid* const*, fdtd_batch** out) {
    if (out) *out = nullptr;
    return FDTD_ERR_UNSUPPORTED;
}
int fdtd_batch_destroy(fdtd_batch*) { return FDTD_OK; }
int fdtd_batch_update_fields(fdtd_batch*, int) { return FDTD_ERR_UNSUPPORTED; }
int fdtd_batch_update_fields_graph(fdtd_batch*, int, int) { return FDTD_ERR_UNSUPPORTED; }
int fdtd_batch_synchronize(fdtd_batch*) { return FDTD_ERR_UNSUPPORTED; }
int fdtd_batch_time(fdtd_batch*, int, float*, float*) { return FDTD_ERR_UNSUPPORTED; }

}  // extern "C"
