// The reference's two plugin ABIs, served by the extended plan (plan.cpp).
//
//   fft3d_*      host pointers (bench/geom/fft3d_api.h:49-66): create() owns device
//                buffers for the batch, pinned staging and two streams; execute() runs the
//                batch in chunks through a pipeline (host copies into pinned staging by a
//                few threads, H2D, transform, D2H, copies out), two chunks in flight, and
//                returns with `out` complete, as bench/geom/driver.c:207-218 requires.
//   fft3d_gpu_*  device pointers (bench/gpu/fft3d_gpu_api.h:61-79): execute() enqueues on
//                the NULL stream and returns; the caller synchronizes.
// Both are forward, unnormalised, complex128 cubes.  A HIP failure inside execute() is
// recorded for fft3d_mx_last_error() and printed once to stderr -- never swallowed.
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "fft3d_api.h"
#include "fft3d_gpu_api.h"
#include "fft3d_mi355x.h"
#include "plan_internal.hpp"

using namespace fx;

static const int kCubes[] = {6, 8, 12, 17, 24, 36};

static int cube_supported(int L)
{
    for (int c : kCubes)
        if (c == L) return fft3d_mx_supports(L, L, L, FFT3D_MX_C128);
    return 0;
}

static void report(const char* where)
{
    static bool printed = false;
    if (!printed) {
        fprintf(stderr, "%s: %s\n", where, fft3d_mx_last_error());
        printed = true;
    }
}

// ------------------------------------------------------------------- host-pointer ABI
// One large memcpy split between `workers` threads and the caller (pageable caller buffers
// to and from pinned staging: a single thread's memcpy, or the runtime's own staging of a
// pageable hipMemcpy, runs at a fraction of the host link).
class CopyPool {
public:
    explicit CopyPool(int workers)
    {
        for (int k = 0; k < workers; ++k) threads_.emplace_back([this, k] { run(k); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        wake_.notify_all();
        for (auto& t : threads_) t.join();
    }
    void copy(void* dst, const void* src, size_t bytes)
    {
        const int parts = (int)threads_.size() + 1;
        if (parts == 1 || bytes < (1u << 20)) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            dst_ = (char*)dst;
            src_ = (const char*)src;
            bytes_ = bytes;
            parts_ = parts;
            pending_ = parts - 1;
            ++gen_;
        }
        wake_.notify_all();
        part(parts - 1);
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [&] { return pending_ == 0; });
    }

private:
    void part(int k) const
    {
        const size_t per = ((bytes_ + parts_ - 1) / parts_ + 63) & ~(size_t)63;
        const size_t lo = (size_t)k * per, hi = std::min(bytes_, lo + per);
        if (lo < hi) std::memcpy(dst_ + lo, src_ + lo, hi - lo);
    }
    void run(int k)
    {
        unsigned long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu_);
                wake_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            part(k);
            std::lock_guard<std::mutex> g(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable wake_, done_;
    bool stop_ = false;
    unsigned long gen_ = 0;
    int pending_ = 0, parts_ = 1;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
};

struct fft3d_plan {
    fft3d_mx_plan* p;
    void* din;
    void* dout;
    size_t bytes;
    long batch;
    size_t vol_bytes;
    long chunk;               // volumes per pipeline chunk
    void* hin[2];             // pinned staging, one chunk each
    void* hout[2];
    hipStream_t st[2];
    hipEvent_t done[2];       // a slot's D2H has landed in hout
    CopyPool* pool;
};

// About 8 MB per chunk: small enough for a pipeline of several chunks at the configurations'
// sizes, large enough that the per-chunk launches and waits stay a few percent.
static const size_t kChunkBytes = (size_t)8 << 20;

static int copy_workers()
{
    cpu_set_t set;
    int n = 1;
    if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
    return std::max(0, std::min(7, n / 2 - 1));
}

static void host_plan_free(fft3d_plan* h)
{
    for (int s = 0; s < 2; ++s) {
        if (h->st[s]) (void)hipStreamSynchronize(h->st[s]);
        if (h->done[s]) (void)hipEventDestroy(h->done[s]);
        if (h->st[s]) (void)hipStreamDestroy(h->st[s]);
        if (h->hin[s]) (void)hipHostFree(h->hin[s]);
        if (h->hout[s]) (void)hipHostFree(h->hout[s]);
    }
    (void)hipFree(h->din);
    (void)hipFree(h->dout);
    delete h->pool;
    if (h->p) fft3d_mx_destroy(h->p);
    delete h;
}

extern "C" const char* fft3d_name(void) { return "mi355x_hip"; }

extern "C" const char* fft3d_description(void)
{
    return "MI355X gfx950 HIP: LDS-resident fused 3-axis kernels (L<=17, Rader-17), "
           "residue-split single pass or two-pass pencil (L=24, 36); "
           "pipelined H2D+kernel+D2H per execute through pinned staging";
}

extern "C" int fft3d_supports(int L) { return cube_supported(L); }

extern "C" fft3d_plan* fft3d_create(int L, int batch)
{
    if (!cube_supported(L) || batch <= 0) {
        fail(FFT3D_MX_EUNSUPPORTED, "fft3d_create: unsupported L=%d batch=%d", L, batch);
        return nullptr;
    }
    fft3d_plan* h = new (std::nothrow) fft3d_plan();
    if (!h) return nullptr;
    h->batch = batch;
    h->vol_bytes = (size_t)L * L * L * 16;
    h->bytes = h->vol_bytes * (size_t)batch;
    h->chunk = std::max<long>(1, std::min<long>(batch, (long)(kChunkBytes / h->vol_bytes)));
    h->p = fft3d_mx_create(L, L, L, batch, FFT3D_MX_FORWARD, FFT3D_MX_C128, 0);
    if (!h->p) {
        delete h;
        return nullptr;
    }
    const size_t cb = (size_t)h->chunk * h->vol_bytes;
    hipError_t e = hipMalloc(&h->din, h->bytes);
    if (e == hipSuccess) e = hipMalloc(&h->dout, h->bytes);
    for (int s = 0; s < 2 && e == hipSuccess; ++s) {
        e = hipHostMalloc(&h->hin[s], cb, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc(&h->hout[s], cb, hipHostMallocDefault);
        if (e == hipSuccess) e = hipStreamCreate(&h->st[s]);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h->done[s], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        fail(FFT3D_MX_ENOMEM, "fft3d_create: %s", hipGetErrorString(e));
        host_plan_free(h);
        return nullptr;
    }
    h->pool = new (std::nothrow) CopyPool(copy_workers());
    if (!h->pool) {
        fail(FFT3D_MX_ENOMEM, "fft3d_create: copy threads");
        host_plan_free(h);
        return nullptr;
    }
    return h;
}

// Chunk c (two slots, c % 2): its caller-side copy into pinned staging, then H2D, transform
// and D2H on the slot's stream; before a slot is reused, its previous chunk's output is
// copied out (that chunk's D2H, and so its H2D, are complete).  While the host copies chunk
// c in and chunk c - 2 out, the device runs chunk c - 1, whose D2H overlaps chunk c's H2D.
extern "C" void fft3d_execute(fft3d_plan* h, const void* in, void* out)
{
    if (!h || !in || !out) {
        fail(FFT3D_MX_EINVAL, "fft3d_execute: NULL argument");
        report("fft3d_execute");
        return;
    }
    const long nch = (h->batch + h->chunk - 1) / h->chunk;
    auto span = [&](long c, size_t& off, size_t& len) {
        const long n = std::min(h->chunk, h->batch - c * h->chunk);
        off = (size_t)(c * h->chunk) * h->vol_bytes;
        len = (size_t)n * h->vol_bytes;
        return n;
    };
    auto drain = [&](long c) {
        const int s = (int)(c & 1);
        size_t off, len;
        span(c, off, len);
        const hipError_t e = hipEventSynchronize(h->done[s]);
        if (e == hipSuccess) h->pool->copy((char*)out + off, h->hout[s], len);
        return e;
    };
    hipError_t e = hipSuccess;
    long issued = 0;
    for (long c = 0; c < nch && e == hipSuccess; ++c) {
        const int s = (int)(c & 1);
        if (c >= 2 && (e = drain(c - 2)) != hipSuccess) break;
        size_t off, len;
        const long n = span(c, off, len);
        h->pool->copy(h->hin[s], (const char*)in + off, len);
        e = hipMemcpyAsync((char*)h->din + off, h->hin[s], len, hipMemcpyHostToDevice, h->st[s]);
        if (e != hipSuccess) break;
        if (fft3d_mx_execute_n(h->p, (char*)h->din + off, (char*)h->dout + off, n, h->st[s]) !=
            FFT3D_MX_OK) {
            report("fft3d_execute");
            e = hipErrorUnknown;
            (void)hipStreamSynchronize(h->st[0]);
            (void)hipStreamSynchronize(h->st[1]);
            return;
        }
        e = hipMemcpyAsync(h->hout[s], (char*)h->dout + off, len, hipMemcpyDeviceToHost, h->st[s]);
        if (e == hipSuccess) e = hipEventRecord(h->done[s], h->st[s]);
        if (e == hipSuccess) issued = c + 1;
    }
    for (long c = std::max(0L, issued - 2); c < issued && e == hipSuccess; ++c) e = drain(c);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(h->st[0]);
        (void)hipStreamSynchronize(h->st[1]);
        fail(FFT3D_MX_EHIP, "fft3d_execute: %s", hipGetErrorString(e));
        report("fft3d_execute");
    }
}

extern "C" void fft3d_destroy(fft3d_plan* h)
{
    if (h) host_plan_free(h);
}

// ----------------------------------------------------------------- device-pointer ABI
struct fft3d_gpu_plan {
    fft3d_mx_plan* p;
};

extern "C" const char* fft3d_gpu_name(void) { return "mi355x_hip_gpu"; }

extern "C" const char* fft3d_gpu_description(void)
{
    return "MI355X gfx950 HIP, device-resident: LDS-resident fused 3-axis kernels (L<=17, "
           "Rader-17), two-pass LDS/pencil (L=24,36)";
}

extern "C" int fft3d_gpu_supports(int L) { return cube_supported(L); }

extern "C" fft3d_gpu_plan* fft3d_gpu_create(int L, int batch)
{
    if (!cube_supported(L) || batch <= 0) {
        fail(FFT3D_MX_EUNSUPPORTED, "fft3d_gpu_create: unsupported L=%d batch=%d", L, batch);
        return nullptr;
    }
    fft3d_gpu_plan* g = new (std::nothrow) fft3d_gpu_plan();
    if (!g) return nullptr;
    g->p = fft3d_mx_create(L, L, L, batch, FFT3D_MX_FORWARD, FFT3D_MX_C128, 0);
    if (!g->p) {
        delete g;
        return nullptr;
    }
    return g;
}

extern "C" void fft3d_gpu_execute(fft3d_gpu_plan* g, const fft3d_double2* in,
                                  fft3d_double2* out)
{
    if (!g) {
        fail(FFT3D_MX_EINVAL, "fft3d_gpu_execute: NULL plan");
        report("fft3d_gpu_execute");
        return;
    }
    if (fft3d_mx_execute(g->p, in, out, nullptr) != FFT3D_MX_OK) report("fft3d_gpu_execute");
}

extern "C" void fft3d_gpu_destroy(fft3d_gpu_plan* g)
{
    if (!g) return;
    fft3d_mx_destroy(g->p);
    delete g;
}
