// er_pockets.hip -- pocket blocking (src/pocketblocking.cc:126-391) and
// grid-interpolated Widom insertion (src/gcmc.cc:145-177) on gfx950.
//
// Segments: voxels with E < 15 T, 6-connected, no periodic images.  A
// lock-free union-find links every accessible voxel to its accessible -x, -y
// and -z neighbours, always hooking the larger root under the smaller, so the
// final root of a segment is its smallest linear index k + j*nz + i*ny*nz --
// the voxel where the reference's raster-order flood fill starts it.  The
// segment id is then the rank of that root among all roots (an exclusive scan
// of root flags), which is exactly the reference's numbering.  The face links
// (x = N-1 -> x = 0 etc.) go to the host, which classifies segments; a last
// kernel overwrites pocket voxels with 1e13.
#include "er_pockets.h"

namespace er {

namespace {

__device__ __forceinline__ int ld(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int find_root(int *parent, int v) {
    // parent[x] <= x for every accessible x and parents only ever decrease,
    // so the walk strictly descends and ends at a root.
    int p = ld(&parent[v]);
    while (p != v) {
        const int gp = ld(&parent[p]);
        if (gp != p) st(&parent[v], gp);  // path halving: gp is an ancestor of v
        v = p;
        p = gp;
    }
    return v;
}

// Root of v without path compression.  The flatten pass must not compress:
// a halving store of another thread could land after v's own thread has
// stored v's root and replace it with a stale (non-root) ancestor.
__device__ __forceinline__ int find_root_ro(const int *parent, int v) {
    int p = ld(&parent[v]);
    while (p != v) {
        v = p;
        p = ld(&parent[v]);
    }
    return v;
}

__device__ __forceinline__ void unite(int *parent, int a, int b) {
    int ra = find_root(parent, a), rb = find_root(parent, b);
    while (ra != rb) {
        if (ra < rb) {
            const int t = ra;
            ra = rb;
            rb = t;
        }
        // hook the larger root ra under rb if ra is still a root
        const int old = atomicCAS(&parent[ra], ra, rb);
        if (old == ra) return;
        ra = find_root(parent, old);
        rb = find_root(parent, rb);
    }
}

__global__ void k_cc_init(const float *__restrict__ e, long long nv, double thr,
                          int *__restrict__ parent) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (long long)gridDim.x * blockDim.x)
        parent[v] = ((double)e[v] >= thr) ? -1 : (int)v;  // src/pocketblocking.cc:138
}

__global__ void k_cc_hook(int nx, int ny, int nz, int *__restrict__ parent) {
    const long long nv = (long long)nx * ny * nz;
    const long long syz = (long long)ny * nz;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (long long)gridDim.x * blockDim.x) {
        if (parent[v] < 0) continue;
        const int k = (int)(v % nz), j = (int)((v / nz) % ny), i = (int)(v / syz);
        if (k > 0 && parent[v - 1] >= 0) unite(parent, (int)v, (int)(v - 1));
        if (j > 0 && parent[v - nz] >= 0) unite(parent, (int)v, (int)(v - nz));
        if (i > 0 && parent[v - syz] >= 0) unite(parent, (int)v, (int)(v - syz));
    }
}

__global__ void k_cc_flatten(long long nv, int *__restrict__ parent,
                             uint32_t *__restrict__ is_root) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (long long)gridDim.x * blockDim.x) {
        const int p = parent[v];
        if (p < 0) {
            is_root[v] = 0;
            continue;
        }
        // Only v's own thread writes parent[v] here, so it ends as the root.
        const int r = find_root_ro(parent, (int)v);
        st(&parent[v], r);
        is_root[v] = (r == (int)v) ? 1u : 0u;
    }
}

// seg[v] = rank of the root of v among all roots, -1 if inaccessible.
__global__ void k_cc_segments(long long nv, const int *__restrict__ parent,
                              const uint32_t *__restrict__ root_rank, int *__restrict__ seg) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (long long)gridDim.x * blockDim.x) {
        const int p = parent[v];
        seg[v] = p < 0 ? -1 : (int)root_rank[p];
    }
}

// Face links of src/pocketblocking.cc:273-321: for each face pair the
// segment at index N-1 links to the segment at index 0 in direction +1/+2/+3.
// links[3*q] = from, [3*q+1] = to, [3*q+2] = direction, or from = -1.
__global__ void k_cc_faces(int nx, int ny, int nz, const int *__restrict__ seg,
                           int *__restrict__ links) {
    const long long fx = (long long)ny * nz, fy = (long long)nx * nz, fz = (long long)nx * ny;
    const long long total = fx + fy + fz;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        long long here, there;
        int dir;
        if (q < fx) {
            const int j = (int)(q / nz), k = (int)(q % nz);
            here = (long long)k + (long long)j * nz;
            there = here + (long long)(nx - 1) * ny * nz;
            dir = 1;
        } else if (q < fx + fy) {
            const long long r = q - fx;
            const int i = (int)(r / nz), k = (int)(r % nz);
            here = (long long)k + (long long)i * ny * nz;
            there = here + (long long)(ny - 1) * nz;
            dir = 2;
        } else {
            const long long r = q - fx - fy;
            const int i = (int)(r / ny), j = (int)(r % ny);
            here = (long long)j * nz + (long long)i * ny * nz;
            there = here + (nz - 1);
            dir = 3;
        }
        const int a = seg[there], b = seg[here];
        links[3 * q] = (a >= 0 && b >= 0) ? a : -1;
        links[3 * q + 1] = b;
        links[3 * q + 2] = dir;
    }
}

// src/pocketblocking.cc:373-382 on an FP32 grid.
__global__ void k_cc_block(long long nv, const int *__restrict__ seg,
                           const uint8_t *__restrict__ pocket, float *__restrict__ e) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (long long)gridDim.x * blockDim.x) {
        const int s = seg[v];
        if (s >= 0 && pocket[s]) e[v] = 1e13f;
    }
}

// Trilinear energy of src/gcmc.cc:145-177 after the wrap of src/gcmc.cc:74-77,
// with the lower index clamped to N-2 (f = 1 reads node N-1).
__device__ __forceinline__ double trilinear(const float *__restrict__ g, int nx, int ny, int nz,
                                            double xf, double yf, double zf) {
    // Separate multiplies and adds, as the reference's host code computes
    // them: next to blocked (1e13 K) nodes a fused multiply-add changes the
    // interpolated energy far beyond rounding of the result.
#pragma clang fp contract(off)
    const double dx = 1.0 / (nx - 1), dy = 1.0 / (ny - 1), dz = 1.0 / (nz - 1);
    int ix = (int)floor(xf / dx), iy = (int)floor(yf / dy), iz = (int)floor(zf / dz);
    ix = ix > nx - 2 ? nx - 2 : (ix < 0 ? 0 : ix);
    iy = iy > ny - 2 ? ny - 2 : (iy < 0 ? 0 : iy);
    iz = iz > nz - 2 ? nz - 2 : (iz < 0 ? 0 : iz);
    const double xd = (xf - ix * dx) / dx;
    const double yd = (yf - iy * dy) / dy;
    const double zd = (zf - iz * dz) / dz;
    const long long syz = (long long)ny * nz;
    const long long b = (long long)iz + (long long)iy * nz + (long long)ix * syz;
    const double g000 = g[b], g100 = g[b + syz], g010 = g[b + nz], g110 = g[b + syz + nz];
    const double g001 = g[b + 1], g101 = g[b + syz + 1], g011 = g[b + nz + 1],
                 g111 = g[b + syz + nz + 1];
    const double c00 = g000 * (1.0 - xd) + xd * g100;
    const double c10 = g010 * (1.0 - xd) + xd * g110;
    const double c01 = g001 * (1.0 - xd) + xd * g101;
    const double c11 = g011 * (1.0 - xd) + xd * g111;
    const double c0 = c00 * (1.0 - yd) + c10 * yd;
    const double c1 = c01 * (1.0 - yd) + c11 * yd;
    return c0 * (1 - zd) + c1 * zd;
}

__global__ void k_widom_grid(const float *__restrict__ g, int nx, int ny, int nz, uint64_t seed,
                             uint32_t stream, uint64_t n, double *__restrict__ e_out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        double f[3];
        widom_point(seed, stream, i, f);
        e_out[i] = trilinear(g, nx, ny, nz, f[0] - floor(f[0]), f[1] - floor(f[1]),
                             f[2] - floor(f[2]));
    }
}

__global__ __launch_bounds__(256) void k_scan_block_sums(const uint32_t *in, long long n,
                                                       uint32_t *sums) {
    __shared__ uint32_t s_w[4];
    const long long base = (long long)blockIdx.x * kScanChunk;
    uint32_t v = 0;
    for (int k = threadIdx.x; k < kScanChunk; k += 256) {
        const long long i = base + k;
        if (i < n) v += in[i];
    }
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(256) void k_scan_sums(const uint32_t *sums, int nb, uint32_t *off) {
    if (threadIdx.x != 0) return;
    uint32_t run = 0;
    for (int b = 0; b < nb; ++b) {
        off[b] = run;
        run += sums[b];
    }
    off[nb] = run;
}

__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t *in, long long n,
                                                  const uint32_t *off, uint32_t *out, int nb) {
    __shared__ uint32_t s_part[256];
    const long long base = (long long)blockIdx.x * kScanChunk;
    const int per = kScanChunk / 256;
    uint32_t v[kScanChunk / 256];
    uint32_t sum = 0;
    for (int k = 0; k < per; ++k) {
        const long long i = base + (long long)threadIdx.x * per + k;
        v[k] = i < n ? in[i] : 0u;
        sum += v[k];
    }
    s_part[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const uint32_t add = threadIdx.x >= d ? s_part[threadIdx.x - d] : 0u;
        __syncthreads();
        s_part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = off[blockIdx.x] + s_part[threadIdx.x] - sum;
    for (int k = 0; k < per; ++k) {
        const long long i = base + (long long)threadIdx.x * per + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = off[nb];
}

int blocks_for(long long n) {
    long long b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace

hipError_t launch_cc_init(const float *e, long long nv, double thr, int *parent, hipStream_t s) {
    hipLaunchKernelGGL(k_cc_init, dim3(blocks_for(nv)), dim3(256), 0, s, e, nv, thr, parent);
    return hipGetLastError();
}

hipError_t launch_cc_hook(const int n[3], int *parent, hipStream_t s) {
    const long long nv = (long long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_cc_hook, dim3(blocks_for(nv)), dim3(256), 0, s, n[0], n[1], n[2],
                       parent);
    return hipGetLastError();
}

hipError_t launch_cc_flatten(long long nv, int *parent, uint32_t *is_root, hipStream_t s) {
    hipLaunchKernelGGL(k_cc_flatten, dim3(blocks_for(nv)), dim3(256), 0, s, nv, parent, is_root);
    return hipGetLastError();
}

hipError_t launch_cc_segments(long long nv, const int *parent, const uint32_t *root_rank,
                              int *seg, hipStream_t s) {
    hipLaunchKernelGGL(k_cc_segments, dim3(blocks_for(nv)), dim3(256), 0, s, nv, parent,
                       root_rank, seg);
    return hipGetLastError();
}

hipError_t launch_cc_faces(const int n[3], const int *seg, int *links, hipStream_t s) {
    const long long total = (long long)n[1] * n[2] + (long long)n[0] * n[2] + (long long)n[0] * n[1];
    hipLaunchKernelGGL(k_cc_faces, dim3(blocks_for(total)), dim3(256), 0, s, n[0], n[1], n[2],
                       seg, links);
    return hipGetLastError();
}

hipError_t launch_cc_block(long long nv, const int *seg, const uint8_t *pocket, float *e,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_cc_block, dim3(blocks_for(nv)), dim3(256), 0, s, nv, seg, pocket, e);
    return hipGetLastError();
}

hipError_t launch_scan_large(const uint32_t *in, long long n, uint32_t *out,
                             uint32_t *block_sums, uint32_t *block_off, hipStream_t s) {
    const long long nb = (n + kScanChunk - 1) / kScanChunk;
    if (nb <= 0 || nb > (1LL << 30)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_block_sums, dim3((unsigned)nb), dim3(256), 0, s, in, n, block_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(64), 0, s, block_sums, (int)nb, block_off);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(256), 0, s, in, n, block_off, out,
                       (int)nb);
    return hipGetLastError();
}

hipError_t launch_widom_grid(const float *g, const int n[3], uint64_t seed, uint32_t stream,
                             uint64_t nins, double *e_out, hipStream_t s) {
    if (nins == 0) return hipSuccess;
    hipLaunchKernelGGL(k_widom_grid, dim3(blocks_for((long long)nins)), dim3(256), 0, s, g, n[0],
                       n[1], n[2], seed, stream, nins, e_out);
    return hipGetLastError();
}

}  // namespace er
