// Device-resident scratch for the allocation-free ragged-product entry point
// (dmx_ring_mul_batch_device_noalloc): the queue of a launch (the degrees, offsets, order, items
// and header of the LDS-resident product kernel) lives in one file-scope device array instead of
// memory the library allocates. Each device holds its own copy of the array for the life of the
// process; its device address is looked up once per device (noalloc_scratch_init, called when a
// context is created with that device current) and only that address is used afterwards.
#include <mutex>

#include "dmx_kernels.hpp"

namespace dmx {

// degrees (4 bytes) + ring_scratch_bytes (20 bytes per product + a 16-byte header)
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)NOALLOC_MAX_PRODUCTS * 4 + 256 +
                                         ((size_t)NOALLOC_MAX_PRODUCTS * 20 + 16 + 255) / 256 * 256;
static_assert(NOALLOC_SCRATCH_BYTES <= (size_t)2 << 20, "the allocation-free scratch stays modest");

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

// The device address of `scratch` on the current device (which must be `device`), looked up once
// per device. hipSuccess and *out set, or the lookup's error.
hipError_t noalloc_scratch_init(int device, void** out) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!ptr) return hipErrorInvalidSymbol;
        g_noalloc_ptr[device] = ptr;
    }
    *out = g_noalloc_ptr[device];
    return hipSuccess;
}

}  // namespace dmx
