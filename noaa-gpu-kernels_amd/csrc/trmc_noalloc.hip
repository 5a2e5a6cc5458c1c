// Allocation-free diffusive-wave step: trmc_diffusive_tridiag_f32_noalloc.
//
// The diffusive step needs scratch beyond its inputs and outputs: the forward
// elimination factors of every node (500 floats per reach, node-major) and a
// count of invalid reach records.  trmc_diffusive_tridiag_f32 allocates them
// per call; this entry point takes device-resident inputs and outputs from the
// caller and keeps its scratch in the file-scope array below, so it calls no
// allocator (for callers that cannot allocate: serving loops, harnesses that
// forbid hipMalloc).
//
// The array exists once per GPU for the life of the process.  Host code never
// passes the array's name to a kernel or a copy: the name is a host-side
// stand-in, not a device address.  Its address on each device is taken once,
// with that device current, by hipGetSymbolAddress, and only that address is
// used.  Every call on one GPU shares that GPU's array and the array keeps the
// last call's contents: the counter is cleared at the start of each call, and
// calls that overlap on one GPU (different host threads or streams) race.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "trmc.h"
#include "trmc_device.h"
#include "trmc_util.h"

namespace {

constexpr int kNoAllocMaxDevices = 64;
// words [0, 64): control (word 0: invalid reach records); then the factors
constexpr size_t kScratchHeadWords = 64;
constexpr size_t kScratchWords = kScratchHeadWords + (size_t)TRMC_DIFF_MAX_NODES * TRMC_DIFF_NOALLOC_MAX_REACHES;

static __device__ __align__(256) float scratch[kScratchWords];

std::mutex g_mu;
void *g_ptr[kNoAllocMaxDevices] = {};

// the address of this device's copy of scratch (taken once per device)
int scratch_on(int dev, float **out)
{
    if (dev < 0 || dev >= kNoAllocMaxDevices) return trmc::fail(TRMC_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return trmc::fail(TRMC_ERR_HIP, "scratch address on device %d: %s", dev, hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    *out = static_cast<float *>(g_ptr[dev]);
    return TRMC_OK;
}

}  // namespace

extern "C" int trmc_diffusive_tridiag_f32_noalloc(const trmc_diff_reach *d_rp, const float *d_dx,
                                                  const float *d_diffusivity, const float *d_celerity,
                                                  const float *d_qp_prev, float *d_qp_out, int nreach, void *stream)
{
    (void)d_celerity;   // unused by the step, as in the reference
    if (nreach < 0) return trmc::fail(TRMC_ERR_ARG, "negative nreach");
    if (nreach > TRMC_DIFF_NOALLOC_MAX_REACHES)
        return trmc::fail(TRMC_ERR_ARG, "nreach %d exceeds the allocation-free limit of %d reaches", nreach,
                          TRMC_DIFF_NOALLOC_MAX_REACHES);
    if (nreach == 0) return TRMC_OK;
    if (!d_rp || !d_dx || !d_diffusivity || !d_qp_prev || !d_qp_out) return trmc::fail(TRMC_ERR_ARG, "null array");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return trmc::fail(TRMC_ERR_NO_DEVICE, "no gfx950 device");
    float *ptr = nullptr;
    int rc = scratch_on(dev, &ptr);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned *bad = reinterpret_cast<unsigned *>(ptr);
    hipError_t e = trmc::launch_reset(bad, 1, s);   // a fill kernel: the counter starts at 0
    if (e == hipSuccess)
        e = trmc::launch_diffusive_tridiag(reinterpret_cast<const trmc::DiffReach *>(d_rp), d_dx, d_diffusivity,
                                           d_qp_prev, d_qp_out, ptr + kScratchHeadWords, nreach, s, bad);
    unsigned n_bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&n_bad, bad, sizeof(n_bad), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return trmc::fail(TRMC_ERR_HIP, "diffusive step: %s", hipGetErrorString(e));
    if (n_bad)
        return trmc::fail(TRMC_ERR_ARG, "%u reach records have ncomp outside 1..%d (their nodes are untouched)",
                          n_bad, TRMC_DIFF_MAX_NODES);
    return TRMC_OK;
}
