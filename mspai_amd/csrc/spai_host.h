/*
 * Host-side launchers of the column kernels (one per kernel family).
 */
#ifndef MSPAI_SPAI_HOST_H
#define MSPAI_SPAI_HOST_H

#include <hip/hip_runtime.h>

#include "device_common.h"

namespace mspai {

/* Column kernel with the working set in LDS: one column per group of
 * `group` lanes (16, 32 or 64), L.per_wave bytes of LDS per group; columns
 * that do not fit go to a.overflow.  mixed: FP32 QR + FP64 refinement, and
 * columns whose selection margins are too small go to a.fallback. */
void launch_binned(const BuildArgs& a, const ScratchLayout& L, int group, bool mixed, const int* queue,
                   const int* nqueue, int* head, int nwaves, hipStream_t stream);

/* Route columns to bin queues: queue b < nbins at queues + b*n with its
 * length in counts[b]; the rest to gqueue (length gcount). */
void launch_classify(const BuildArgs& a, int nbins, const BinEdges& edges, int* queues, int* counts,
                     int* gqueue, int* gcount, int n, hipStream_t stream);

/* dst[0..count) = src[0..count) on the stream (count <= 64). */
void launch_copy_ints(const int* src, int* dst, int count, hipStream_t stream);

/* Same kernel with the working set in HBM (a.scratch, L.per_wave per wave). */
void launch_fp64_generic(const BuildArgs& a, const ScratchLayout& L, const int* queue, const int* nqueue,
                         int* head, int nwaves, hipStream_t stream);

}  // namespace mspai

#endif
