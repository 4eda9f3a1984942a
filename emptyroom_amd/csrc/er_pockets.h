// er_pockets.h -- launchers of the pocket-blocking and grid-Widom kernels
// (er_pockets.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "er_common.h"

namespace er {

hipError_t launch_cc_init(const float *e, long long nv, double thr, int *parent, hipStream_t s);
hipError_t launch_cc_hook(const int n[3], int *parent, hipStream_t s);
hipError_t launch_cc_flatten(long long nv, int *parent, uint32_t *is_root, hipStream_t s);
hipError_t launch_cc_segments(long long nv, const int *parent, const uint32_t *root_rank,
                              int *seg, hipStream_t s);
hipError_t launch_cc_faces(const int n[3], const int *seg, int *links, hipStream_t s);
hipError_t launch_cc_block(long long nv, const int *seg, const uint8_t *pocket, float *e,
                           hipStream_t s);
hipError_t launch_widom_grid(const float *g, const int n[3], uint64_t seed, uint32_t stream,
                             uint64_t nins, double *e_out, hipStream_t s);
// Exclusive scan of n uint32 values into out[0..n] (out[n] = total) with
// kScanChunk-element blocks; block_sums needs n/kScanChunk + 2 entries.
constexpr int kScanChunk = 2048;
hipError_t launch_scan_large(const uint32_t *in, long long n, uint32_t *out,
                             uint32_t *block_sums, uint32_t *block_off, hipStream_t s);

}  // namespace er
