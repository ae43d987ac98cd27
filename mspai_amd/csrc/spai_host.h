/*
 * Host-side launchers of the column kernels (one per kernel family).
 */
#ifndef MSPAI_SPAI_HOST_H
#define MSPAI_SPAI_HOST_H

#include <hip/hip_runtime.h>

#include "device_common.h"

namespace mspai {

/* Generic FP64 kernel: one wave per column, per-wave scratch in HBM. */
void launch_generic_fp64(const BuildArgs& a, int nwaves, hipStream_t stream);

}  // namespace mspai

#endif
