// mcb_noalloc.hip - mcb_points_dev_noalloc: the device-buffer points path with no allocation, for callers
// that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// mcb_points_dev keeps its per-seed noise state and slot map in engine buffers it grows with hipMalloc.
// This variant keeps them in one file-scope device array, `scratch`: up to MCB_NOALLOC_MAX_SEEDS seed
// slots (15,472 bytes each) followed by the identity slot map, 15.8 MB in all, held by every device the
// library runs on until the process ends.  In host code the array's name is only a stand-in for its
// device address: the address is looked up once per device with hipGetSymbolAddress (the device
// current) and only that pointer is ever handed to kernels.  The slot map is rewritten by a fill kernel
// on every call (no memset node, so a captured call replays correctly).  Every call on a device shares
// that device's array: calls that overlap on one device (different streams, or different threads
// without ordering) race on it.
#include <mutex>

#include "mcb_internal.h"

using namespace mcb;

namespace {

constexpr size_t kSlotBytes = MCB_NOALLOC_MAX_SEEDS * sizeof(McbSlot);
constexpr size_t kMapBytes = MCB_NOALLOC_MAX_SEEDS * sizeof(int32_t);
static_assert(kSlotBytes % 256 == 0, "slot map must start 256-byte aligned");

__device__ __align__(256) unsigned char scratch[kSlotBytes + kMapBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};           // device address of `scratch` per device (null = not looked up)

// address of this device's scratch; looked up on the first call on the device, with it current
int scratch_for(int dev, unsigned char **out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(MCB_E_DEVICE, "device %d beyond the %d this entry point tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(MCB_E_DEVICE, "scratch address on device %d: %s", dev, hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    *out = (unsigned char *)g_ptr[dev];
    return MCB_OK;
}

__global__ void k_iota(int32_t *__restrict__ map, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) map[i] = i;
}

}  // namespace

extern "C" int mcb_points_dev_noalloc(McbEngine *e, const int64_t *d_seeds, int ns, const int32_t *d_pts, int np,
                                      uint8_t *d_out, int32_t *d_tg, void *stream) {
    if (!e || ns < 0 || np < 0 || (ns && np && (!d_seeds || !d_pts || !d_out))) return fail(MCB_E_ARG, "bad arguments");
    if (ns > MCB_NOALLOC_MAX_SEEDS)
        return fail(MCB_E_LIMIT, "%d seeds: mcb_points_dev_noalloc takes at most %d per call (its static scratch)", ns,
                    MCB_NOALLOC_MAX_SEEDS);
    if (ns == 0 || np == 0) return MCB_OK;
    int dev = -1;
    MCB_HIPCHK(hipGetDevice(&dev));
    if (dev != e->device) return fail(MCB_E_DEVICE, "engine lives on device %d, device %d is current", e->device, dev);
    unsigned char *base = nullptr;
    if (int rc = scratch_for(dev, &base)) return rc;
    McbSlot *slots = (McbSlot *)base;
    int32_t *map = (int32_t *)(base + kSlotBytes);
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_iota, dim3((ns + 255) / 256), dim3(256), 0, st, map, ns);
    MCB_HIPCHK(hipGetLastError());
    MCB_HIPCHK(launch_seed_init(e->w, d_seeds, map, ns, slots, st));
    MCB_HIPCHK(launch_points(e->w, slots, map, ns, d_pts, np, d_out, d_tg, st));
    return MCB_OK;
}
