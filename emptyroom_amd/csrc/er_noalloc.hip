// er_noalloc.hip -- the plain grid computation without any device allocation,
// for callers that cannot call hipMalloc (code under stream capture, tight
// serving loops, harnesses that forbid allocation).
//
// The device memory the tiled kernel needs beyond its input (the resident
// framework) and its output (the caller's grid) is a fixed array of this
// library, `scratch`, one copy per device:
//   [0, 256)          work counter of the persistent launch (cleared by a
//                     fill kernel before every launch)
//   [256, 4096)       the job descriptor
//   [4096, ...)       region sub-lists: kNoAllocSlices slices of
//                     kNoAllocRecords near-list indices (4 B), one per
//                     workgroup
// 32 MiB + 4 KiB in all.  No input size changes what a call needs: regions
// are sized to fit a slice (a region whose sub-list still overflows streams
// the framework's whole near list instead), and a launch uses at most
// kNoAllocSlices workgroups.  The device address of the array is looked up
// once per device (hipGetSymbolAddress, that device current) and used for
// every kernel argument and copy; the array's name is never passed to the
// device.  Every call on a device shares its array, so calls that overlap on
// one device (other threads, other contexts) race; calls on one context are
// ordered by its stream and each call waits for its work before returning.
#include <mutex>

#include "er_runtime.h"

namespace er {

namespace {

constexpr int kNoAllocSlices = 1024;
constexpr int kNoAllocRecords = 8192;
constexpr size_t kNoAllocJobOffset = 256;
constexpr size_t kNoAllocListOffset = 4096;
constexpr size_t kNoAllocBytes =
    kNoAllocListOffset + (size_t)kNoAllocSlices * kNoAllocRecords * sizeof(uint32_t);
static_assert(kNoAllocJobOffset + sizeof(JobDev) <= kNoAllocListOffset,
              "job descriptor does not fit its scratch slot");

constexpr int kMaxDevices = 64;

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// Device address of this device's scratch array (looked up once, with the
// device current).
int scratch_address(int device, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("grid_compute_noalloc: device index beyond the supported range");
        return ST_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            set_error(std::string("grid_compute_noalloc: hipGetSymbolAddress failed: ") +
                      hipGetErrorString(e));
            return ST_HIP;
        }
        g_scratch[device] = p;
    }
    *out = static_cast<unsigned char *>(g_scratch[device]);
    return ST_OK;
}

int hip_fail(const char *what, hipError_t e) {
    (void)hipGetLastError();
    set_error(std::string("grid_compute_noalloc: ") + what + ": " + hipGetErrorString(e));
    return ST_HIP;
}

}  // namespace

int grid_compute_noalloc(Context *ctx, FrameworkDev *fw, const int n[3], int precision,
                         void *d_out) {
    if (!ctx || !fw || !n || !d_out) {
        set_error("grid_compute_noalloc: null argument");
        return ST_ARG;
    }
    int st = check_grid_dims(n);
    if (st != ST_OK) return st;
    if ((st = check_tuning(ctx->tuning)) != ST_OK) return st;
    if (precision != 0 && precision != 1) {
        set_error("grid_compute_noalloc: precision must be 0 (FP32) or 1 (FP64)");
        return ST_ARG;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail("hipSetDevice", e);
    unsigned char *base = nullptr;
    if ((st = scratch_address(ctx->device, &base)) != ST_OK) return st;
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(base);
    JobDev *d_job = reinterpret_cast<JobDev *>(base + kNoAllocJobOffset);
    uint32_t *slices = reinterpret_cast<uint32_t *>(base + kNoAllocListOffset);

    const int math = precision == 1 ? MATH_FP64 : MATH_MIXED;
    // regions sized for this array's slices; a tile's list, and so every
    // value, does not depend on the region it is in
    const JobDev job = make_grid_job(fw, n, precision, d_out, ctx->tuning, math, kNoAllocRecords);
    hipStream_t s = ctx->stream;
    if ((e = hipMemcpyAsync(d_job, &job, sizeof(job), hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail("job copy", e);
    if ((e = launch_fill_u64(counter, 0ull, 1, s)) != hipSuccess)
        return hip_fail("counter clear", e);
    const int wgs = tiles_workgroups(ctx, math, job.n_items, kNoAllocSlices);
    if ((e = launch_tiles(math, ctx->tuning.ppl, d_job, 1, job.n_items, counter, wgs,
                          ctx->tuning.cap, slices, kNoAllocRecords, s)) != hipSuccess)
        return hip_fail("tile kernel launch", e);
    // the job descriptor was copied from this frame's stack: wait for the
    // whole call before returning
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail("synchronize", e);
    return ST_OK;
}

}  // namespace er
