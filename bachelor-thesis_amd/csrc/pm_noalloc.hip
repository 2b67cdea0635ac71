// pm_noalloc.hip -- the scratch memory of pm_march_device_noalloc.
//
// pm_march_device_noalloc runs the configured isotropic march into the
// caller's device memory without allocating anything: the buffers the march
// needs beyond its inputs and outputs -- the per-pixel hit records, the tile
// pre-pass's flags, scan and list, the counters and the tile-band counters --
// live in one file-scope device array, `scratch`, of kNoAllocScratchBytes.
// Each GPU holds its own copy of the array for the life of the process.  Its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) when the first context opens the device; host code then
// uses only that address, never the array's name.
//
// Layout of the array (offsets in bytes, all 256-aligned):
//   hit records   PM_NOALLOC_MAX_PIXELS uint4
//   tile buffer   3 (T + 1) + scan scratch u32, T <= kNoAllocMaxTiles
//   counters      kCtrWords u64
//   band counters 64 u32
// Every field is rewritten by the launch before it is read (records and tile
// lists by the kernels that produce them, the counters and band counters by
// fill kernels at the start of the launch), so contents left by an earlier
// call do not matter.  Calls on one GPU share the array: two marches that
// overlap on one GPU (two contexts or two threads) race.
#include <mutex>

#include "pm_internal.h"

namespace pm {

// Tiles of the largest image the limits allow (W, H <= 16384, W H <= 2^21):
// ceil(W/8) ceil(H/8) <= W H / 64 + (W + H) / 8 + 1 < 36864.
constexpr uint32_t kNoAllocMaxTiles = 36864;
constexpr size_t kNoAllocHitBytes = sizeof(uint4) * (size_t)PM_NOALLOC_MAX_PIXELS;
constexpr size_t kNoAllocTileWords = 3 * (kNoAllocMaxTiles + 1) + 2048;   // scan scratch <= (T + 1) / 256 + 2
constexpr size_t kNoAllocTileOff = kNoAllocHitBytes;
constexpr size_t kNoAllocCtrOff = kNoAllocTileOff + ((sizeof(uint32_t) * kNoAllocTileWords + 255) & ~(size_t)255);
constexpr size_t kNoAllocBandOff = kNoAllocCtrOff + sizeof(unsigned long long) * kCtrWords;
constexpr size_t kNoAllocScratchBytes = kNoAllocBandOff + 256;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
void *g_ptr[kMaxDevices] = {};
hipError_t g_rc[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::call_once(g_once[device], [device]() {
        void *p = nullptr;
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_rc[device] = e;
        g_ptr[device] = e == hipSuccess ? p : nullptr;
    });
    *out = g_ptr[device];
    return g_rc[device];
}

bool noalloc_layout(void *base, uint32_t W, uint32_t H, NoAllocScratch &s)
{
    const uint64_t tiles = (uint64_t)((W + 7) / 8) * ((H + 7) / 8);
    if (!base || (uint64_t)W * H > PM_NOALLOC_MAX_PIXELS || tiles > kNoAllocMaxTiles) return false;
    if (3 * (tiles + 1) + scan_tmp_size((uint32_t)(tiles + 1)) > kNoAllocTileWords) return false;
    uint8_t *b = (uint8_t *)base;
    s.hit_rec = (uint4 *)b;
    s.tiles = (uint32_t *)(b + kNoAllocTileOff);
    s.counters = (unsigned long long *)(b + kNoAllocCtrOff);
    s.band = (uint32_t *)(b + kNoAllocBandOff);
    return true;
}

size_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

}  // namespace pm
