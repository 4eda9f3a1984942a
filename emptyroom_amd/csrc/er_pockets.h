// er_pockets.h -- launchers of the pocket-blocking and grid-Widom kernels
// (er_pockets.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "er_common.h"

namespace er {

constexpr int kCCVox = 2048;  // voxels per block of the per-voxel labelling passes
constexpr int kBrick = 512;   // voxels of an 8x8x8 brick (local union-find in LDS)
constexpr int kFaceChunk = 4096;  // face pairs per block of the link pass

// One FP32 grid of a labelling launch.  Plain-old-data, copied to the device.
struct CCGrid {
    float *e;                  // energies, index k + j*nz + i*ny*nz (blocked in place)
    int n[3];
    int pad;
    long long nv;              // voxels
    long long nw;              // root-mask words, ceil(nv / 64)
    int *parent;               // nv: union-find parent, then root (-1 inaccessible)
    unsigned long long *mask;  // nw: bit v set iff voxel v is a root
    uint32_t *wpre;            // nw + 1: roots before each word; wpre[nw] = segments
    int *links;                // 3 * faces: compacted (from, to, dir) face links
    const uint8_t *pocket;     // segments: 1 if pocket (blocking pass), or null
    long long blk0;            // first block of this grid in per-voxel launches
    long long bblk0;           // first block of this grid in per-brick launches
    long long fblk0;           // first block of this grid in the face-link launch
    long long seg0;            // offset of this grid in a segment-id output
};

long long cc_blocks(long long nv);
long long cc_bricks(const int n[3]);
long long cc_face_blocks(const int n[3]);
// Brick-local labelling, face merges, flatten, word scan and face links of
// every grid.  packed receives grid g's links (3 ints each) at offset
// meta[4g+2]; meta[4g] = links, meta[4g+1] = segments.
hipError_t launch_cc_label(const CCGrid *grids, int ng, long long total_blocks,
                           long long total_bricks, long long total_face_blocks, double thr,
                           int *packed, unsigned int *meta, hipStream_t s);
// seg_out[seg0 + v] = segment id of voxel v (-1 inaccessible), after labelling.
hipError_t launch_cc_segments(const CCGrid *grids, int ng, long long total_blocks, int *seg_out,
                              hipStream_t s);
// Overwrite the voxels of pocket segments with 1e13 (grids with pocket != null).
hipError_t launch_cc_block(const CCGrid *grids, int ng, long long total_blocks, hipStream_t s);
// Grid-interpolated Widom with fused Boltzmann sums: partials (kTriBlocks
// entries) receive fixed-order per-block (sum b, sum E b); e_out nullable.
constexpr int kTriBlocks = 4096;
// Energies (K) at npos fractional positions f[3 npos] (wrapped into [0, 1)),
// interpolated on an FP32 (precision 0) or FP64 (precision 1) grid.
hipError_t launch_grid_energies(const void *g, int precision, const int n[3], const double *f,
                                long long npos, double *out, hipStream_t s);
hipError_t launch_widom_grid(const float *g, const int n[3], uint64_t seed, uint32_t stream,
                             uint64_t nins, double temperature, double *e_out, double2 *partials,
                             hipStream_t s);

}  // namespace er
