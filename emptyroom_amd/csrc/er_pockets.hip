// er_pockets.hip -- pocket blocking (src/pocketblocking.cc:126-391) and
// grid-interpolated Widom insertion (src/gcmc.cc:145-177) on gfx950.
//
// Segments: voxels with E < 15 T, 6-connected, no periodic images.  Any
// number of grids run together, one launch per phase, blocks mapped to grids
// through a table.  Each 8x8x8 brick is first labelled by a union-find in
// LDS; a lock-free union-find in global memory then joins accessible voxels
// across brick faces.  Both always hook the larger root under the smaller,
// so the final root of a segment is its smallest linear index
// k + j*nz + i*ny*nz -- the voxel where the reference's raster-order flood
// fill starts it.  The segment id is the rank of that root among all
// roots; roots are kept as a bitmask (one ballot per 64 voxels) with a
// prefix count per 64-bit word, so rank(r) = prefix[r/64] +
// popcount(mask[r/64] below r) and no per-voxel rank array exists.  Face
// links (x = N-1 -> x = 0 etc.) are run-length compacted on the device and
// packed into one buffer for the host, which classifies segments; a last
// pass overwrites pocket voxels with 1e13.
#include "er_pockets.h"

#include <algorithm>

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

// Block -> grid of a per-voxel launch (blocks of kCCVox voxels per grid).
__device__ __forceinline__ int cc_grid_of(const CCGrid *__restrict__ g, int ng, long long blk) {
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g[mid].blk0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Segment id of voxel v after the flatten pass: rank of its root, -1 if
// inaccessible.
__device__ __forceinline__ int seg_of(const CCGrid &G, long long v) {
    const int r = G.parent[v];
    if (r < 0) return -1;
    const long long w = r >> 6;
    const unsigned long long below = G.mask[w] & ((1ull << (r & 63)) - 1ull);
    return (int)(G.wpre[w] + (uint32_t)__popcll(below));
}

// Brick of 8x8x8 voxels (grid indices i, j, k) labelled by one block.
__device__ __forceinline__ int cc_brick_of(const CCGrid *__restrict__ g, int ng, long long blk) {
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g[mid].bblk0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int lds_ld(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int lds_find(int *p, int v) {
    int q = lds_ld(&p[v]);
    while (q != v) {
        const int g = lds_ld(&p[q]);
        if (g != q)  // halving; g is an ancestor of v
            __hip_atomic_store(&p[v], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        v = q;
        q = g;
    }
    return v;
}

__device__ __forceinline__ void lds_unite(int *p, int a, int b) {
    int ra = lds_find(p, a), rb = lds_find(p, b);
    while (ra != rb) {
        if (ra < rb) {
            const int t = ra;
            ra = rb;
            rb = t;
        }
        const int old = atomicCAS(&p[ra], ra, rb);
        if (old == ra) return;
        ra = lds_find(p, old);
        rb = lds_find(p, rb);
    }
}

// Accessibility (src/pocketblocking.cc:138) and union-find inside one brick
// in LDS.  Local index (ii*8 + jj)*8 + kk orders a brick's voxels as their
// grid indices do, so a local root is the smallest grid index of its local
// component; parent[v] = that grid index (-1 if inaccessible).
__global__ __launch_bounds__(kWG) void k_cc_local(const CCGrid *__restrict__ grids, int ng,
                                                  double thr) {
    __shared__ int s_p[kBrick];
    const CCGrid &G = grids[cc_brick_of(grids, ng, blockIdx.x)];
    const int nx = G.n[0], ny = G.n[1], nz = G.n[2];
    const int nbz = (nz + 7) >> 3, nby = (ny + 7) >> 3;
    const long long b = blockIdx.x - G.bblk0;
    const int bk = (int)(b % nbz), bj = (int)((b / nbz) % nby), bi = (int)(b / ((long long)nbz * nby));
    const long long syz = (long long)ny * nz;
    long long vg[kBrick / kWG];
    bool acc[kBrick / kWG];
#pragma unroll
    for (int h = 0; h < kBrick / kWG; ++h) {
        const int l = threadIdx.x + h * kWG;
        const int i = bi * 8 + (l >> 6), j = bj * 8 + ((l >> 3) & 7), k = bk * 8 + (l & 7);
        const bool in = i < nx && j < ny && k < nz;
        vg[h] = in ? (long long)k + (long long)j * nz + (long long)i * syz : -1;
        acc[h] = in && (double)G.e[vg[h]] < thr;
        s_p[l] = acc[h] ? l : -1;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < kBrick / kWG; ++h) {
        const int l = threadIdx.x + h * kWG;
        if (!acc[h]) continue;
        if ((l & 7) > 0 && lds_ld(&s_p[l - 1]) >= 0) lds_unite(s_p, l, l - 1);
        if (((l >> 3) & 7) > 0 && lds_ld(&s_p[l - 8]) >= 0) lds_unite(s_p, l, l - 8);
        if ((l >> 6) > 0 && lds_ld(&s_p[l - 64]) >= 0) lds_unite(s_p, l, l - 64);
    }
    __syncthreads();
    int r[kBrick / kWG];
#pragma unroll
    for (int h = 0; h < kBrick / kWG; ++h) {
        const int l = threadIdx.x + h * kWG;
        int q = acc[h] ? l : -1;
        if (q >= 0)
            while (s_p[q] != q) q = s_p[q];
        r[h] = q;
    }
#pragma unroll
    for (int h = 0; h < kBrick / kWG; ++h) {
        if (vg[h] < 0) continue;
        int out = -1;
        if (r[h] >= 0) {
            const int rl = r[h];
            const int i = bi * 8 + (rl >> 6), j = bj * 8 + ((rl >> 3) & 7), k = bk * 8 + (rl & 7);
            out = (int)((long long)k + (long long)j * nz + (long long)i * syz);
        }
        G.parent[vg[h]] = out;
    }
}

// Unions across brick faces: every accessible voxel on a brick's -x, -y or
// -z face with its accessible neighbour in the adjacent brick.  One wave per
// brick walks the 169 face voxels: the -z face (64), then the rest of the -y
// face (8 x 7), then the rest of the -x face (7 x 7).
__global__ __launch_bounds__(64) void k_cc_merge(const CCGrid *__restrict__ grids, int ng) {
    const CCGrid &G = grids[cc_brick_of(grids, ng, blockIdx.x)];
    const int nx = G.n[0], ny = G.n[1], nz = G.n[2];
    const int nbz = (nz + 7) >> 3, nby = (ny + 7) >> 3;
    const long long b = blockIdx.x - G.bblk0;
    const int bk = (int)(b % nbz), bj = (int)((b / nbz) % nby), bi = (int)(b / ((long long)nbz * nby));
    const long long syz = (long long)ny * nz;
    int *parent = G.parent;
    const int t = threadIdx.x;
#pragma unroll
    for (int face = 0; face < 3; ++face) {
        int ii, jj, kk;
        if (face == 0) {
            ii = t >> 3;
            jj = t & 7;
            kk = 0;
        } else if (face == 1) {
            if (t >= 56) continue;
            ii = t / 7;
            jj = 0;
            kk = 1 + t % 7;
        } else {
            if (t >= 49) continue;
            ii = 0;
            jj = 1 + t / 7;
            kk = 1 + t % 7;
        }
        const int i = bi * 8 + ii, j = bj * 8 + jj, k = bk * 8 + kk;
        if (i >= nx || j >= ny || k >= nz) continue;
        const long long v = (long long)k + (long long)j * nz + (long long)i * syz;
        const int pv = ld(&parent[v]);
        if (pv < 0) continue;
        // A union is skipped when the face voxel just before this one along
        // one of the face's two axes has the same (parent, neighbour's parent)
        // pair: both are ancestors of both pairs, so that union (or the one
        // it in turn relies on, always at a smaller face position) joins
        // these two sets as well.
        if (kk == 0 && k > 0) {
            const int pn = ld(&parent[v - 1]);
            bool dup = false;
            if (jj > 0) dup |= ld(&parent[v - nz]) == pv && ld(&parent[v - nz - 1]) == pn;
            if (ii > 0) dup |= ld(&parent[v - syz]) == pv && ld(&parent[v - syz - 1]) == pn;
            if (pn >= 0 && !dup) unite(parent, (int)v, (int)(v - 1));
        }
        if (jj == 0 && j > 0) {
            const int pn = ld(&parent[v - nz]);
            bool dup = false;
            if (kk > 0) dup |= ld(&parent[v - 1]) == pv && ld(&parent[v - nz - 1]) == pn;
            if (ii > 0) dup |= ld(&parent[v - syz]) == pv && ld(&parent[v - syz - nz]) == pn;
            if (pn >= 0 && !dup) unite(parent, (int)v, (int)(v - nz));
        }
        if (ii == 0 && i > 0) {
            const int pn = ld(&parent[v - syz]);
            bool dup = false;
            if (kk > 0) dup |= ld(&parent[v - 1]) == pv && ld(&parent[v - syz - 1]) == pn;
            if (jj > 0) dup |= ld(&parent[v - nz]) == pv && ld(&parent[v - syz - nz]) == pn;
            if (pn >= 0 && !dup) unite(parent, (int)v, (int)(v - syz));
        }
    }
}

// parent[v] = root of v; bit v of mask set iff v is a root.  Each wave covers
// 64 consecutive voxels starting at a multiple of 64, so one ballot is one
// mask word.
__global__ __launch_bounds__(kWG) void k_cc_flatten(const CCGrid *__restrict__ grids, int ng) {
    const CCGrid &G = grids[cc_grid_of(grids, ng, blockIdx.x)];
    const long long base = (long long)(blockIdx.x - G.blk0) * kCCVox;
    for (int k = threadIdx.x; k < kCCVox; k += kWG) {
        const long long v = base + k;
        bool root = false;
        if (v < G.nv) {
            const int p = G.parent[v];
            if (p >= 0) {
                // only v's own thread writes parent[v] here, so it ends as the root
                const int r = find_root_ro(G.parent, (int)v);
                st(&G.parent[v], r);
                root = r == (int)v;
            }
        }
        const unsigned long long m = __ballot(root);
        const long long w = v >> 6;
        if ((threadIdx.x & 63) == 0 && w < G.nw) G.mask[w] = m;
    }
}

// wpre[w] = roots in words before w; wpre[nw] = number of segments.  One
// block per grid.
__global__ __launch_bounds__(kWG) void k_cc_word_scan(const CCGrid *__restrict__ grids) {
    const CCGrid &G = grids[blockIdx.x];
    __shared__ uint32_t s_part[kWG];
    __shared__ uint32_t s_carry;
    const int tid = threadIdx.x;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    constexpr int per = 8;
    for (long long b = 0; b < G.nw; b += (long long)kWG * per) {
        uint32_t v[per], sum = 0;
#pragma unroll
        for (int k = 0; k < per; ++k) {
            const long long w = b + (long long)tid * per + k;
            v[k] = w < G.nw ? (uint32_t)__popcll(G.mask[w]) : 0u;
            sum += v[k];
        }
        s_part[tid] = sum;
        __syncthreads();
        for (int d = 1; d < kWG; d <<= 1) {
            const uint32_t add = tid >= d ? s_part[tid - d] : 0u;
            __syncthreads();
            s_part[tid] += add;
            __syncthreads();
        }
        uint32_t run = s_carry + s_part[tid] - sum;
#pragma unroll
        for (int k = 0; k < per; ++k) {
            const long long w = b + (long long)tid * per + k;
            if (w < G.nw) G.wpre[w] = run;
            run += v[k];
        }
        __syncthreads();
        if (tid == kWG - 1) s_carry += s_part[kWG - 1];
        __syncthreads();
    }
    if (tid == 0) G.wpre[G.nw] = s_carry;
}

__device__ __forceinline__ int cc_face_grid_of(const CCGrid *__restrict__ g, int ng,
                                                long long blk) {
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g[mid].fblk0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_cc_zero(unsigned int *__restrict__ meta, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) meta[i] = 0u;
}

// Face links of src/pocketblocking.cc:273-321: for each face pair the segment
// at index N-1 links to the segment at index 0 in direction +1/+2/+3.  Pairs
// with an inaccessible side are dropped, and so is every pair equal to the
// one before it in face order within a block's chunk (repeats change no
// classification, nor does the order of the links).  A block takes
// kFaceChunk face pairs of one grid, compacts them in LDS, then reserves its
// slots in the grid's link region with one atomic on meta[4g].
__global__ __launch_bounds__(kWG) void k_cc_faces(const CCGrid *__restrict__ grids, int ng,
                                                  unsigned int *__restrict__ meta) {
    __shared__ int s_l[3 * kFaceChunk];
    __shared__ int s_a[kWG], s_b[kWG], s_d[kWG];
    __shared__ int s_cnt[kWG / 64];
    __shared__ unsigned int s_base;
    const int g = cc_face_grid_of(grids, ng, blockIdx.x);
    const CCGrid &G = grids[g];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nx = G.n[0], ny = G.n[1], nz = G.n[2];
    const long long fx = (long long)ny * nz, fy = (long long)nx * nz, fz = (long long)nx * ny;
    const long long total = fx + fy + fz;
    const long long c0 = (long long)(blockIdx.x - G.fblk0) * kFaceChunk;
    const long long c1 = c0 + kFaceChunk < total ? c0 + kFaceChunk : total;
    int pa = -2, pb = -2, pd = -2;  // previous pair (block-uniform)
    int cnt = 0;
    for (long long q0 = c0; q0 < c1; q0 += kWG) {
        const long long q = q0 + tid;
        int a = -1, b = -1, dir = 0;
        if (q < c1) {
            long long here, there;
            if (q < fx) {
                here = q;
                there = here + (long long)(nx - 1) * ny * nz;
                dir = 1;
            } else if (q < fx + fy) {
                const long long r = q - fx;
                here = (r % nz) + (r / nz) * (long long)ny * nz;
                there = here + (long long)(ny - 1) * nz;
                dir = 2;
            } else {
                const long long r = q - fx - fy;
                here = (r % ny) * nz + (r / ny) * (long long)ny * nz;
                there = here + (nz - 1);
                dir = 3;
            }
            a = seg_of(G, there);
            b = seg_of(G, here);
        }
        s_a[tid] = a;
        s_b[tid] = b;
        s_d[tid] = dir;
        __syncthreads();
        const int qa = tid > 0 ? s_a[tid - 1] : pa;
        const int qb = tid > 0 ? s_b[tid - 1] : pb;
        const int qd = tid > 0 ? s_d[tid - 1] : pd;
        const bool keep = q < c1 && a >= 0 && b >= 0 && !(a == qa && b == qb && dir == qd);
        const unsigned long long bk = __ballot(keep);
        if (lane == 0) s_cnt[wave] = __popcll(bk);
        pa = s_a[kWG - 1];
        pb = s_b[kWG - 1];
        pd = s_d[kWG - 1];
        __syncthreads();
        int off = cnt, tot = 0;
#pragma unroll
        for (int w = 0; w < kWG / 64; ++w) {
            if (w < wave) off += s_cnt[w];
            tot += s_cnt[w];
        }
        if (keep) {
            const int o = off + __builtin_amdgcn_mbcnt_hi(
                                    (uint32_t)(bk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bk, 0u));
            s_l[3 * o] = a;
            s_l[3 * o + 1] = b;
            s_l[3 * o + 2] = dir;
        }
        cnt += tot;
        __syncthreads();
    }
    if (tid == 0) s_base = cnt > 0 ? atomicAdd(&meta[4 * g], (unsigned int)cnt) : 0u;
    __syncthreads();
    const unsigned int base = s_base;
    for (int k = tid; k < 3 * cnt; k += kWG) G.links[3 * (long long)base + k] = s_l[k];
}

// One block per grid: its links go to packed[3 * off_g ...] with off_g the
// links of all earlier grids; meta[4g+1] = segments, meta[4g+2] = off_g.
__global__ __launch_bounds__(kWG) void k_cc_pack(const CCGrid *__restrict__ grids,
                                                 int *__restrict__ packed,
                                                 unsigned int *__restrict__ meta) {
    const int g = blockIdx.x;
    const CCGrid &G = grids[g];
    unsigned int off = 0;
    for (int h = 0; h < g; ++h) off += meta[4 * h];
    const unsigned int cnt = meta[4 * g];
    for (long long k = threadIdx.x; k < 3LL * cnt; k += kWG)
        packed[3 * (long long)off + k] = G.links[k];
    if (threadIdx.x == 0) {
        meta[4 * g + 1] = G.wpre[G.nw];
        meta[4 * g + 2] = off;
    }
}

__global__ __launch_bounds__(kWG) void k_cc_segments(const CCGrid *__restrict__ grids, int ng,
                                                     int *__restrict__ seg_out) {
    const CCGrid &G = grids[cc_grid_of(grids, ng, blockIdx.x)];
    const long long base = (long long)(blockIdx.x - G.blk0) * kCCVox;
    for (int k = threadIdx.x; k < kCCVox; k += kWG) {
        const long long v = base + k;
        if (v < G.nv) seg_out[G.seg0 + v] = seg_of(G, v);
    }
}

// src/pocketblocking.cc:373-382 on FP32 grids.
__global__ __launch_bounds__(kWG) void k_cc_block(const CCGrid *__restrict__ grids, int ng) {
    const CCGrid &G = grids[cc_grid_of(grids, ng, blockIdx.x)];
    if (!G.pocket) return;
    const long long base = (long long)(blockIdx.x - G.blk0) * kCCVox;
    for (int k = threadIdx.x; k < kCCVox; k += kWG) {
        const long long v = base + k;
        if (v >= G.nv) continue;
        const int s = seg_of(G, v);
        if (s >= 0 && G.pocket[s]) G.e[v] = 1e13f;
    }
}

// Trilinear energy of src/gcmc.cc:145-177 after the wrap of src/gcmc.cc:74-77,
// with the lower index clamped to N-2 (f = 1 reads node N-1).
template <class T>
__device__ __forceinline__ double trilinear(const T *__restrict__ g, int nx, int ny, int nz,
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

// Guest-framework energy at caller-given fractional positions: the wrap of
// src/gcmc.cc:74-77 (x - floor(x)) and the interpolation of
// src/gcmc.cc:145-177 on an FP32 or FP64 grid, one position per thread.
template <class T>
__global__ __launch_bounds__(kWG) void k_grid_energies(const T *__restrict__ g, int nx, int ny,
                                                       int nz, const double *__restrict__ f,
                                                       long long n, double *__restrict__ out) {
    for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < n;
         i += (long long)gridDim.x * kWG) {
        const double x = f[3 * i], y = f[3 * i + 1], z = f[3 * i + 2];
        out[i] = trilinear(g, nx, ny, nz, x - floor(x), y - floor(y), z - floor(z));
    }
}

// Grid-interpolated insertions with the Boltzmann sums fused in: block b of
// kTriBlocks takes the contiguous insertion range [b n / B, (b+1) n / B),
// its threads stride through it, and the block writes one (sum b, sum E b)
// partial in a fixed order, so the result does not depend on scheduling.
// e_out (nullable) receives the energy of every insertion.
__global__ __launch_bounds__(kWG) void k_widom_grid(const float *__restrict__ g, int nx, int ny,
                                                    int nz, uint64_t seed, uint32_t stream,
                                                    uint64_t n, double temperature,
                                                    double *__restrict__ e_out,
                                                    double2 *__restrict__ partials) {
    __shared__ double s_b[kWG / 64], s_eb[kWG / 64];
    const uint64_t B = gridDim.x, b = blockIdx.x;
    const uint64_t lo = b * n / B, hi = (b + 1) * n / B;
    double sb = 0.0, seb = 0.0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kWG) {
        double f[3];
        widom_point(seed, stream, i, f);
        const double E = trilinear(g, nx, ny, nz, f[0] - floor(f[0]), f[1] - floor(f[1]),
                                   f[2] - floor(f[2]));
        if (e_out) e_out[i] = E;
        const double bz = exp(-E / temperature);
        sb += bz;
        seb += (bz == 0.0) ? 0.0 : E * bz;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        sb += __shfl_xor(sb, o, 64);
        seb += __shfl_xor(seb, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_b[threadIdx.x >> 6] = sb;
        s_eb[threadIdx.x >> 6] = seb;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        partials[b] = make_double2((s_b[0] + s_b[1]) + (s_b[2] + s_b[3]),
                                   (s_eb[0] + s_eb[1]) + (s_eb[2] + s_eb[3]));
}

}  // namespace

hipError_t launch_grid_energies(const void *g, int precision, const int n[3], const double *f,
                                long long npos, double *out, hipStream_t s) {
    if (npos <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<long long>((npos + kWG - 1) / kWG, 65536);
    if (precision == 1)
        hipLaunchKernelGGL(k_grid_energies<double>, dim3(blocks), dim3(kWG), 0, s,
                           (const double *)g, n[0], n[1], n[2], f, npos, out);
    else
        hipLaunchKernelGGL(k_grid_energies<float>, dim3(blocks), dim3(kWG), 0, s,
                           (const float *)g, n[0], n[1], n[2], f, npos, out);
    return hipGetLastError();
}

long long cc_blocks(long long nv) { return (nv + kCCVox - 1) / kCCVox; }

long long cc_bricks(const int n[3]) {
    return (long long)((n[0] + 7) / 8) * ((n[1] + 7) / 8) * ((n[2] + 7) / 8);
}

long long cc_face_blocks(const int n[3]) {
    const long long f = (long long)n[1] * n[2] + (long long)n[0] * n[2] + (long long)n[0] * n[1];
    return (f + kFaceChunk - 1) / kFaceChunk;
}

hipError_t launch_cc_label(const CCGrid *grids, int ng, long long total_blocks,
                           long long total_bricks, long long total_face_blocks, double thr,
                           int *packed, unsigned int *meta, hipStream_t s) {
    if (ng <= 0) return hipSuccess;
    if (total_blocks <= 0 || total_blocks > (1LL << 31) - 1 || total_bricks <= 0 ||
        total_bricks > (1LL << 31) - 1 || total_face_blocks <= 0 ||
        total_face_blocks > (1LL << 31) - 1 || 4LL * ng > (1LL << 30))
        return hipErrorInvalidValue;
    const dim3 nb((unsigned)total_blocks);
    hipLaunchKernelGGL(k_cc_local, dim3((unsigned)total_bricks), dim3(kWG), 0, s, grids, ng, thr);
    hipLaunchKernelGGL(k_cc_merge, dim3((unsigned)total_bricks), dim3(64), 0, s, grids, ng);
    hipLaunchKernelGGL(k_cc_flatten, nb, dim3(kWG), 0, s, grids, ng);
    hipLaunchKernelGGL(k_cc_word_scan, dim3(ng), dim3(kWG), 0, s, grids);
    hipLaunchKernelGGL(k_cc_zero, dim3(1), dim3(kWG), 0, s, meta, 4 * ng);
    hipLaunchKernelGGL(k_cc_faces, dim3((unsigned)total_face_blocks), dim3(kWG), 0, s, grids, ng,
                       meta);
    hipLaunchKernelGGL(k_cc_pack, dim3(ng), dim3(kWG), 0, s, grids, packed, meta);
    return hipGetLastError();
}

hipError_t launch_cc_segments(const CCGrid *grids, int ng, long long total_blocks, int *seg_out,
                              hipStream_t s) {
    if (ng <= 0 || total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cc_segments, dim3((unsigned)total_blocks), dim3(kWG), 0, s, grids, ng,
                       seg_out);
    return hipGetLastError();
}

hipError_t launch_cc_block(const CCGrid *grids, int ng, long long total_blocks, hipStream_t s) {
    if (ng <= 0 || total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cc_block, dim3((unsigned)total_blocks), dim3(kWG), 0, s, grids, ng);
    return hipGetLastError();
}

hipError_t launch_widom_grid(const float *g, const int n[3], uint64_t seed, uint32_t stream,
                             uint64_t nins, double temperature, double *e_out, double2 *partials,
                             hipStream_t s) {
    if (nins == 0) return hipSuccess;
    hipLaunchKernelGGL(k_widom_grid, dim3(kTriBlocks), dim3(kWG), 0, s, g, n[0], n[1], n[2], seed,
                       stream, nins, temperature, e_out, partials);
    return hipGetLastError();
}

}  // namespace er
