// Plan-free transform that allocates nothing: fft3d_mx_execute_noalloc (fft3d_mi355x.h).
//
// The plan path (plan.cpp) allocates, at create, the device constant bank and the claim
// counters its persistent kernels draw blocks from.  Here both live in file-scope device
// arrays instead: the kernels run unchanged, taking the arrays' device addresses as they
// take the plan's buffers.  Set-up happens once per device, on the first call made with
// that device current: it looks up the arrays' device addresses (a host-side use of an
// array's name is not a device address), uploads the bank images computed on the host in
// long double, and clears the counters.  That first call is synchronous and must not be
// made while `stream` is being captured; a call with batch 0 performs only the set-up.
//
// Every call on a device shares that device's counters: kNoallocSlots counters taken
// round-robin, each handed back at zero by the launch that used it, so up to kNoallocSlots
// calls may be in flight on one device at once; calls beyond that overlap on a counter
// and race (blocks skipped or repeated).  The kernel for a geometry is the first one
// registered for it, the pick of fft3d_mx_create with FFT3D_MX_NO_TUNE, so results are
// bit-identical to such a plan's.
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>
#include <cstring>
#include <mutex>
#include <vector>

#include "cx.hpp"
#include "fft3d_mi355x.h"
#include "plan_internal.hpp"
#include "registry.hpp"

using namespace fx;

constexpr int kNoallocSlots = 64;
constexpr int kLineU = kClaimWords;  // claim counters of one call (claim_layout.hpp)
constexpr int kMaxDevices = 64;

// Claim counters of every call on this device (kNoallocSlots sets of 2 KiB, 128 KiB).
static __device__ __align__(256) unsigned scratch[kNoallocSlots * kLineU];
// Constant banks, one per precision, uploaded at set-up.
static __device__ __align__(256) Bank<double> bank_c128;
static __device__ __align__(256) Bank<float> bank_c64;

namespace {

struct DeviceState {
    std::once_flag once;
    std::atomic<bool> ready{false};
    int status = FFT3D_MX_EHIP;    // result of the set-up
    unsigned* counters = nullptr;  // device address of scratch on this device
    void* bank[2] = {nullptr, nullptr};
    std::atomic<unsigned> next{0};
};
DeviceState g_dev[kMaxDevices];

int setup(DeviceState& d)
{
    void* sp = nullptr;
    void* b128 = nullptr;
    void* b64 = nullptr;
    hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e == hipSuccess) e = hipGetSymbolAddress(&b128, HIP_SYMBOL(bank_c128));
    if (e == hipSuccess) e = hipGetSymbolAddress(&b64, HIP_SYMBOL(bank_c64));
    if (e != hipSuccess)
        return fail(FFT3D_MX_EHIP, "fft3d_mx_execute_noalloc: hipGetSymbolAddress: %s",
                    hipGetErrorString(e));
    std::vector<unsigned char> img;
    bank_host_image(kC128, img);
    if (img.size() != sizeof(Bank<double>))
        return fail(FFT3D_MX_EHIP, "fft3d_mx_execute_noalloc: bank image size mismatch");
    e = hipMemcpy(b128, img.data(), img.size(), hipMemcpyHostToDevice);
    bank_host_image(kC64, img);
    if (e == hipSuccess && img.size() != sizeof(Bank<float>))
        return fail(FFT3D_MX_EHIP, "fft3d_mx_execute_noalloc: bank image size mismatch");
    if (e == hipSuccess) e = hipMemcpy(b64, img.data(), img.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(sp, 0, sizeof(scratch));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess)
        return fail(FFT3D_MX_EHIP, "fft3d_mx_execute_noalloc: set-up: %s", hipGetErrorString(e));
    d.counters = (unsigned*)sp;
    d.bank[kC128] = b128;
    d.bank[kC64] = b64;
    d.ready.store(true, std::memory_order_release);
    return FFT3D_MX_OK;
}

}  // namespace

extern "C" int fft3d_mx_execute_noalloc(int n0, int n1, int n2, long batch, int direction,
                                        int precision, unsigned flags, const void* in,
                                        void* out, void* stream)
{
    if (n0 <= 0 || n1 <= 0 || n2 <= 0)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_noalloc: bad extents %dx%dx%d", n0, n1,
                    n2);
    if (batch < 0 || batch > (long)INT_MAX)
        return fail(FFT3D_MX_EINVAL,
                    "fft3d_mx_execute_noalloc: batch %ld outside [0, %d] (claim tickets are "
                    "32-bit)", batch, INT_MAX);
    if (direction != -1 && direction != 1)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_noalloc: direction must be -1 or +1");
    if (precision != kC128 && precision != kC64)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_noalloc: unknown precision %d",
                    precision);
    if (flags & ~(unsigned)FFT3D_MX_NORMALIZE)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_noalloc: only FFT3D_MX_NORMALIZE is "
                                     "accepted in flags");
    const KernelDesc* k = nullptr;
    for (const KernelDesc& c : registry())
        if (c.n0 == n0 && c.n1 == n1 && c.n2 == n2 && c.prec == precision &&
            c.dir == direction && c.algo == 0) {
            k = &c;
            break;
        }
    if (!k)
        return fail(FFT3D_MX_EUNSUPPORTED, "fft3d_mx_execute_noalloc: no kernel for %dx%dx%d %s",
                    n0, n1, n2, precision == kC128 ? "complex128" : "complex64");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return fail(FFT3D_MX_EHIP, "fft3d_mx_execute_noalloc: hipGetDevice: %s",
                    hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(FFT3D_MX_EUNSUPPORTED, "fft3d_mx_execute_noalloc: device %d beyond %d", dev,
                    kMaxDevices);
    DeviceState& d = g_dev[dev];
    if (!d.ready.load(std::memory_order_acquire)) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing((hipStream_t)stream, &cs) == hipSuccess &&
            cs != hipStreamCaptureStatusNone)
            return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_noalloc: the first call on a device "
                                         "sets up and must not be captured (call it once with "
                                         "batch 0 first)");
        std::call_once(d.once, [&d]() { d.status = setup(d); });
        if (!d.ready.load(std::memory_order_acquire))
            return fail(d.status, "fft3d_mx_execute_noalloc: set-up failed on device %d", dev);
    }
    if (batch == 0) return FFT3D_MX_OK;
    if (!in || !out || in == (const void*)out)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_noalloc: in and out must be distinct "
                                     "device buffers");
    if (!elem_aligned(in, precision) || !elem_aligned(out, precision))
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_noalloc: buffers must be aligned to "
                                     "%zu bytes", elem_bytes(precision));
    const double scale =
        (flags & FFT3D_MX_NORMALIZE) ? 1.0 / ((double)n0 * n1 * n2) : 1.0;
    const unsigned slot = d.next.fetch_add(1u, std::memory_order_relaxed) % kNoallocSlots;
    e = k->launch(in, out, batch, d.bank[precision], scale, d.counters + (size_t)slot * kLineU,
                  (hipStream_t)stream);
    if (e != hipSuccess)
        return fail(FFT3D_MX_EHIP, "fft3d_mx_execute_noalloc: %s: %s", k->name,
                    hipGetErrorString(e));
    return FFT3D_MX_OK;
}
