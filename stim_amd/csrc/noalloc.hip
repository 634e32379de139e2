// The allocation-free sampling entry point (mfs_sample_noalloc; no counterpart in the reference).
//
// It runs exactly what mfs_sample runs for device-resident outputs, but every piece of device staging
// (packed rows before they are expanded to one byte per entry, or unpacked rows before they are packed
// across a non-byte-aligned observable boundary) comes from `scratch`, a fixed region of device memory
// that each GPU holds for the life of the process, instead of buffers grown with hipMalloc. Calls are
// split into batches of whole RNG tiles that fit the region, so the rows equal mfs_sample's bit for
// bit. Calls that overlap on one GPU share the region and race.
#include <mutex>
#include <vector>

#include "sampler_impl.h"

namespace {

constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;

// Device staging shared by every call on a GPU (each GPU has its own copy at its own address). Host
// code never passes this name to a kernel or a copy: it uses the address hipGetSymbolAddress reports
// for the current device.
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::vector<uint8_t *> g_scratch_ptr;   // per device ordinal

// The current device's address of `scratch`, looked up once per device.
uint8_t *scratch_region() {
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr.size() <= (size_t)dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = static_cast<uint8_t *>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int mfs_sample_noalloc(mfs_sampler *s, uint64_t shots, uint32_t flags, uint8_t *dets, size_t dets_stride,
                       uint8_t *obs, size_t obs_stride) {
    return guarded([&] {
        if (!(flags & MFS_DEVICE_OUTPUT))
            throw std::invalid_argument("mfs_sample_noalloc writes device-resident rows only (MFS_DEVICE_OUTPUT)");
        mfs_sampler::Arena region{scratch_region(), SCRATCH_BYTES, 0};
        struct Detach {
            mfs_sampler *s;
            ~Detach() { s->arena = nullptr; }
        } detach{s};
        s->arena = &region;
        s->sample(shots, flags, dets, dets_stride, obs, obs_stride);
    });
}

}  // extern "C"
