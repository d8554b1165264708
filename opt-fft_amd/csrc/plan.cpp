// Plan step and extended C ABI (include/fft3d_mi355x.h).
//
// create():
//   1. filters the kernel registry by geometry, precision and direction;
//   2. computes the constant bank in long double -- twiddle tables W_N^m for every line
//      length, the Rader-17 spectra Bhat = DFT16(b')/16 with b'[q] = W17^(3^-q)
//      (bench/geom/strategies/L17_rader.md:21-55; Bhat[0] = -1/16 and Bhat[8] = sqrt(17)/16
//      written in closed form) and the folded-17 cos/sin rows -- rounds it to the plan's
//      precision and uploads it;
//   3. when several kernels qualify, races them on the device and keeps the fastest: the
//      race-at-create of bench/geom/best/fft3d_best.c:131-193 (trial: warm executes, then
//      best of several timed groups), with the verdict cached per process and optionally
//      in a wisdom file (the per-host cache of bench/gen/impl_14/gen_race.c:45-49).
// execute() is a single asynchronous launch sequence on the caller's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "claims_hip.hpp"
#include "cx.hpp"
#include "fft3d_mi355x.h"
#include "plan_internal.hpp"
#include "registry.hpp"

namespace fx {

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char* last_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

// ------------------------------------------------------------------------ race + wisdom
static std::mutex g_wisdom_mu;
static std::map<std::string, std::string> g_wisdom;
static bool g_wisdom_loaded = false;

static void wisdom_load_locked()
{
    if (g_wisdom_loaded) return;
    g_wisdom_loaded = true;
    const char* path = getenv("FFT3D_MX_WISDOM");
    if (!path || !*path) return;
    FILE* f = fopen(path, "r");
    if (!f) return;
    char key[256], name[128];
    while (fscanf(f, "%255s %127s", key, name) == 2) g_wisdom[key] = name;
    fclose(f);
}

static void wisdom_store_locked(const std::string& key, const std::string& name)
{
    g_wisdom[key] = name;
    const char* path = getenv("FFT3D_MX_WISDOM");
    if (!path || !*path) return;
    FILE* f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s %s\n", key.c_str(), name.c_str());
    fclose(f);
}

static std::string wisdom_key(int n0, int n1, int n2, int prec, int dir, long batch, int algo)
{
    int dev = 0;
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        arch = prop.gcnArchName;
        cus = prop.multiProcessorCount;
        for (auto& ch : arch)
            if (ch == ':' || ch == ' ') ch = '_';
    }
    int bclass = 0;
    while ((1L << (bclass + 1)) <= batch) ++bclass;
    char buf[256];
    snprintf(buf, sizeof buf, "%s/%dcu/%dx%dx%d/%s/%s/b2^%d%s", arch.c_str(), cus, n0, n1, n2,
             prec == kC128 ? "c128" : "c64", dir < 0 ? "fwd" : "bwd", bclass,
             algo == 1 ? "/mfma-bf16" : algo == 2 ? "/mfma-fp16" : "");
    return buf;
}

// Time one candidate: 2 warm launches, one group of 5 to size the groups, then the best of
// 3 groups of n launches, n chosen so a group lasts about 300 us (5..64): event timing
// jitter of a few microseconds would otherwise decide between kernels of ~10 us.
// Successive launches use successive windows of the race buffers, so with enough windows
// no launch finds its data in the 256 MB Infinity Cache and the race ranks kernels as they
// run from HBM.  Seconds per launch, or -1 if it failed to launch.
static double trial(const KernelDesc& k, const char* din, char* dout, size_t wbytes, int nwin,
                    long nvol, const void* bank, unsigned* work, hipStream_t s, hipEvent_t e0,
                    hipEvent_t e1)
{
    int w = 0;
    auto launch = [&]() {
        const size_t off = (size_t)(w++ % nwin) * wbytes;
        return k.launch(din + off, dout + off, nvol, bank, 1.0, work, s);
    };
    for (int i = 0; i < 2; ++i)
        if (launch() != hipSuccess) return -1.0;
    auto group = [&](int n) -> double {
        if (hipEventRecord(e0, s) != hipSuccess) return -1.0;
        for (int i = 0; i < n; ++i)
            if (launch() != hipSuccess) return -1.0;
        if (hipEventRecord(e1, s) != hipSuccess) return -1.0;
        if (hipEventSynchronize(e1) != hipSuccess) return -1.0;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
        return (double)ms * 1e-3 / n;
    };
    const double first = group(5);
    if (first <= 0) return -1.0;
    int n = (int)(300e-6 / first);
    n = n < 5 ? 5 : (n > 64 ? 64 : n);
    double best = 1e30;
    for (int g = 0; g < 3; ++g) {
        const double per = group(n);
        if (per <= 0) return -1.0;
        if (per < best) best = per;
    }
    return best;
}

// *best_s = 0 when no trial completed (no race buffers, or every launch failed): the
// first candidate is returned unmeasured and the caller must not record it as wisdom.
static const KernelDesc* race(const std::vector<const KernelDesc*>& cands, int n0, int n1,
                              int n2, int prec, long batch, const void* bank, unsigned* work,
                              double* best_s)
{
    *best_s = 0.0;
    const size_t vol_bytes = (size_t)n0 * n1 * n2 * elem_bytes(prec);
    // Each trial launch transforms rb volumes: ~256 MiB of input so planning stays bounded,
    // at least one volume, at most the planned batch.
    long rb = (long)((256ull << 20) / vol_bytes);
    if (rb < 1) rb = 1;
    if (rb > batch) rb = batch;
    const size_t wbytes = (size_t)rb * vol_bytes;
    // Windows per buffer: enough for ~768 MiB, at most 8; fewer if memory is short.
    int nwin = (int)(((768ull << 20) + wbytes - 1) / wbytes);
    if (nwin > 8) nwin = 8;
    void *din = nullptr, *dout = nullptr;
    for (; nwin >= 1; nwin = nwin > 1 ? nwin / 2 : 0) {
        if (hipMalloc(&din, nwin * wbytes) != hipSuccess) {
            din = nullptr;
            continue;
        }
        if (hipMalloc(&dout, nwin * wbytes) == hipSuccess) break;
        (void)hipFree(din);
        din = dout = nullptr;
    }
    (void)hipGetLastError();
    if (!din || !dout) return cands[0];
    // 0x3F80_0000 per 32-bit word: 1.0f for complex64, ~0.0078 per double for complex128.
    (void)hipMemsetD32((hipDeviceptr_t)din, 0x3F800000u, nwin * wbytes / 4);
    hipStream_t s;
    hipEvent_t e0, e1;
    const KernelDesc* win = cands[0];
    double wt = 1e30;
    if (hipStreamCreate(&s) == hipSuccess) {
        if (hipEventCreate(&e0) == hipSuccess) {
            if (hipEventCreate(&e1) == hipSuccess) {
                // A later candidate must beat the incumbent by more than 2% (about the
                // run-to-run spread of one trial), so near-ties resolve the same way each
                // time, to the earlier-registered variant.
                for (const KernelDesc* k : cands) {
                    const double t = trial(*k, (const char*)din, (char*)dout, wbytes, nwin,
                                           rb, bank, work, s, e0, e1);
                    if (t > 0 && (wt > 1e29 || t < 0.98 * wt)) {
                        wt = t;
                        win = k;
                    }
                }
                (void)hipEventDestroy(e1);
            }
            (void)hipEventDestroy(e0);
        }
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    (void)hipFree(din);
    (void)hipFree(dout);
    (void)hipGetLastError();
    *best_s = wt < 1e29 ? wt * (double)batch / (double)rb : 0.0;
    return win;
}

}  // namespace fx

using namespace fx;

struct fft3d_mx_plan {
    int n0, n1, n2, batch, dir, prec;
    unsigned flags;
    const KernelDesc* k;
    // Small-batch class: executes of at most small_max volumes launch k_small (raced at an
    // eighth of the planned batch) when it differs from k; NULL otherwise.
    const KernelDesc* k_small;
    long small_max;
    void* bank;
    unsigned* work;     // device: the claim counters `claims` hands out (claims.hpp)
    ClaimSlots claims;
    double scale;
    double tune_us;
    int device;
    int cus;            // compute units of the device
};

// True when a launch of kernel k over `batch` volumes may draw blocks from claim counters.
// A persistent launch claims only when its blocks number at least twice its grid
// (claim_counter), and its grid is at least min(CUs, blocks), so fewer than 2 x CUs blocks
// never claim: such launches skip the counter bookkeeping (and, on several streams, its
// event).
static bool may_claim(const fft3d_mx_plan* p, const KernelDesc* k, long batch)
{
    if (k->persist <= 0) return false;
    const long units = k->kind == kPencil ? batch * p->n0 : batch;
    const long nblocks = (units + k->vpw - 1) / k->vpw;
    return nblocks >= 2L * p->cus;
}

// The plan's kernel for a launch of `batch` volumes.
static const KernelDesc* kernel_for(const fft3d_mx_plan* p, long batch)
{
    return p->k_small && batch <= p->small_max ? p->k_small : p->k;
}

// One launch of the plan's kernel with a claim counter no running launch shares.
static int plan_launch(fft3d_mx_plan* p, const char* who, const void* in, void* out, long batch,
                       double scale, hipStream_t s)
{
    const KernelDesc* k = kernel_for(p, batch);
    const hipError_t e =
        may_claim(p, k, batch)
            ? p->claims.run(s, [&](unsigned* w) {
                  return k->launch(in, out, batch, p->bank, scale, w, s);
              })
            : k->launch(in, out, batch, p->bank, scale, nullptr, s);
    if (e == HipClaimRT::busy)
        return fail(FFT3D_MX_EINVAL, "%s: %d launches of this plan were already captured into "
                    "graphs; each keeps its own claim counter and no counter is left", who,
                    ClaimSlots::kSlots);
    if (e != hipSuccess)
        return fail(FFT3D_MX_EHIP, "%s: %s: %s", who, k->name, hipGetErrorString(e));
    return FFT3D_MX_OK;
}

extern "C" const char* fft3d_mx_last_error(void) { return last_error(); }
extern "C" void fft3d_mx_clear_error(void) { clear_error(); }

// algo < 0 accepts every algorithm (lookups by name).
static void candidates(int n0, int n1, int n2, int prec, int dir,
                       std::vector<const KernelDesc*>& out, int algo = 0)
{
    for (const KernelDesc& k : registry())
        if (k.n0 == n0 && k.n1 == n1 && k.n2 == n2 && k.prec == prec && k.dir == dir &&
            (algo < 0 || k.algo == algo))
            out.push_back(&k);
}

extern "C" int fft3d_mx_supports(int n0, int n1, int n2, int precision)
{
    std::vector<const KernelDesc*> c;
    candidates(n0, n1, n2, precision, -1, c);
    return !c.empty();
}

extern "C" size_t fft3d_mx_volume_bytes(int n0, int n1, int n2, int precision)
{
    if (n0 <= 0 || n1 <= 0 || n2 <= 0 || (precision != kC128 && precision != kC64)) return 0;
    return (size_t)n0 * n1 * n2 * elem_bytes(precision);
}

static fft3d_mx_plan* create_impl(int n0, int n1, int n2, int batch, int direction,
                                  int precision, unsigned flags, const char* forced)
{
    if (n0 <= 0 || n1 <= 0 || n2 <= 0 || batch < 0) {
        fail(FFT3D_MX_EINVAL, "fft3d_mx_create: bad extents %dx%dx%d or batch %d", n0, n1, n2,
             batch);
        return nullptr;
    }
    if (direction != -1 && direction != 1) {
        fail(FFT3D_MX_EINVAL, "fft3d_mx_create: direction must be -1 or +1, got %d", direction);
        return nullptr;
    }
    if (precision != kC128 && precision != kC64) {
        fail(FFT3D_MX_EINVAL, "fft3d_mx_create: unknown precision %d", precision);
        return nullptr;
    }
    if ((flags & FFT3D_MX_MFMA_BF16) && (flags & FFT3D_MX_MFMA_FP16)) {
        fail(FFT3D_MX_EINVAL, "fft3d_mx_create: MFMA_BF16 and MFMA_FP16 are exclusive");
        return nullptr;
    }
    const int algo = (flags & FFT3D_MX_MFMA_BF16) ? 1 : (flags & FFT3D_MX_MFMA_FP16) ? 2 : 0;
    std::vector<const KernelDesc*> cands, any;
    candidates(n0, n1, n2, precision, direction, cands, algo);
    candidates(n0, n1, n2, precision, direction, any, -1);
    if (cands.empty()) {
        fail(FFT3D_MX_EUNSUPPORTED, "fft3d_mx_create: no %skernel for %dx%dx%d %s",
             algo ? "matrix-core " : "", n0, n1, n2,
             precision == kC128 ? "complex128" : "complex64");
        return nullptr;
    }
    if (!forced || !*forced) forced = getenv("FFT3D_MX_KERNEL");
    const KernelDesc* pick = nullptr;
    if (forced && *forced) {
        for (const KernelDesc* k : any)
            if (!strcmp(k->name, forced)) pick = k;
        if (!pick && getenv("FFT3D_MX_KERNEL") != forced) {
            fail(FFT3D_MX_EUNSUPPORTED, "fft3d_mx_create: kernel '%s' does not serve %dx%dx%d",
                 forced, n0, n1, n2);
            return nullptr;
        }
    }

    fft3d_mx_plan* p = new (std::nothrow) fft3d_mx_plan();
    if (!p) {
        fail(FFT3D_MX_ENOMEM, "fft3d_mx_create: out of host memory");
        return nullptr;
    }
    p->n0 = n0;
    p->n1 = n1;
    p->n2 = n2;
    p->batch = batch;
    p->dir = direction;
    p->prec = precision;
    p->flags = flags;
    p->scale = (flags & FFT3D_MX_NORMALIZE) ? 1.0 / ((double)n0 * n1 * n2) : 1.0;
    p->tune_us = 0.0;
    if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
    if (hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device) !=
            hipSuccess || p->cus < 1)
        p->cus = 1;

    std::vector<unsigned char> img;
    bank_host_image(precision, img);
    hipError_t e = hipMalloc(&p->bank, img.size());
    if (e != hipSuccess) {
        fail(FFT3D_MX_ENOMEM, "fft3d_mx_create: hipMalloc(bank): %s", hipGetErrorString(e));
        delete p;
        return nullptr;
    }
    e = hipMemcpy(p->bank, img.data(), img.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        fail(FFT3D_MX_EHIP, "fft3d_mx_create: upload bank: %s", hipGetErrorString(e));
        (void)hipFree(p->bank);
        delete p;
        return nullptr;
    }
    const size_t work_bytes = ClaimSlots::bytes();
    e = hipMalloc((void**)&p->work, work_bytes);
    if (e == hipSuccess) e = hipMemset(p->work, 0, work_bytes);
    if (e != hipSuccess) {
        fail(FFT3D_MX_ENOMEM, "fft3d_mx_create: claim counters: %s", hipGetErrorString(e));
        (void)hipFree(p->bank);
        if (p->work) (void)hipFree(p->work);
        delete p;
        return nullptr;
    }
    p->claims.attach(p->work);

    p->k_small = nullptr;
    p->small_max = 0;
    if (!pick) {
        if (cands.size() == 1 || (flags & FFT3D_MX_NO_TUNE) || batch == 0) {
            pick = cands[0];
        } else {
            // The race (or the wisdom it left) at the planned batch, and at an eighth of it:
            // a kernel that fills the device at 512 volumes (one workgroup per volume) can
            // lose to one that spreads a volume over several workgroups at 64.
            std::lock_guard<std::mutex> g(g_wisdom_mu);
            wisdom_load_locked();
            auto pick_at = [&](long b, double* tune_us) -> const KernelDesc* {
                const std::string key = wisdom_key(n0, n1, n2, precision, direction, b, algo);
                auto it = g_wisdom.find(key);
                if (it != g_wisdom.end())
                    for (const KernelDesc* k : cands)
                        if (it->second == k->name) return k;
                double best_s = 0.0;
                const KernelDesc* k = race(cands, n0, n1, n2, precision, b, p->bank, p->work,
                                           &best_s);
                if (tune_us) *tune_us = best_s * 1e6;
                if (best_s > 0) wisdom_store_locked(key, k->name);
                return k;
            };
            pick = pick_at(batch, &p->tune_us);
            if (batch >= 16) {
                const long bs = batch / 8;
                const KernelDesc* ks = pick_at(bs, nullptr);
                if (ks != pick) {
                    p->k_small = ks;
                    p->small_max = (long)std::sqrt((double)batch * (double)bs);
                }
            }
        }
    }
    p->k = pick;
    return p;
}

extern "C" fft3d_mx_plan* fft3d_mx_create(int n0, int n1, int n2, int batch, int direction,
                                          int precision, unsigned flags)
{
    return create_impl(n0, n1, n2, batch, direction, precision, flags, nullptr);
}

extern "C" fft3d_mx_plan* fft3d_mx_create_named(int n0, int n1, int n2, int batch,
                                                int direction, int precision, unsigned flags,
                                                const char* kernel)
{
    if (!kernel || !*kernel) {
        fail(FFT3D_MX_EINVAL, "fft3d_mx_create_named: empty kernel name");
        return nullptr;
    }
    return create_impl(n0, n1, n2, batch, direction, precision, flags, kernel);
}

extern "C" int fft3d_mx_execute_n(fft3d_mx_plan* p, const void* in, void* out, long batch,
                                  void* stream)
{
    if (!p) return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute: NULL plan");
    if (batch < 0 || batch > p->batch)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute: batch %ld outside [0, %d]", batch,
                    p->batch);
    if (batch == 0) return FFT3D_MX_OK;
    if (!in || !out) return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute: NULL buffer");
    if (in == (const void*)out)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute: in and out must be distinct");
    if (!elem_aligned(in, p->prec) || !elem_aligned(out, p->prec))
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute: buffers must be aligned to %zu bytes",
                    elem_bytes(p->prec));
    int cur = p->device;
    (void)hipGetDevice(&cur);
    if (cur != p->device) (void)hipSetDevice(p->device);
    const int rc = plan_launch(p, "fft3d_mx_execute", in, out, batch, p->scale,
                               (hipStream_t)stream);
    if (cur != p->device) (void)hipSetDevice(cur);
    return rc;
}

extern "C" int fft3d_mx_execute(fft3d_mx_plan* p, const void* in, void* out, void* stream)
{
    if (!p) return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute: NULL plan");
    return fft3d_mx_execute_n(p, in, out, p->batch, stream);
}

extern "C" int fft3d_mx_execute_chain(fft3d_mx_plan* p, const void* in, void* out, void* pong,
                                      int m, double step_scale, void* stream)
{
    if (!p) return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_chain: NULL plan");
    if (m < 1) return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_chain: m = %d < 1", m);
    if (!in || !out || !pong || in == (const void*)out || in == (const void*)pong || out == pong)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_chain: in, out, pong must be distinct");
    if (!elem_aligned(in, p->prec) || !elem_aligned(out, p->prec) || !elem_aligned(pong, p->prec))
        return fail(FFT3D_MX_EINVAL,
                    "fft3d_mx_execute_chain: buffers must be aligned to %zu bytes",
                    elem_bytes(p->prec));
    if (!(step_scale == step_scale))
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_execute_chain: step_scale is NaN");
    int cur = p->device;
    (void)hipGetDevice(&cur);
    if (cur != p->device) (void)hipSetDevice(p->device);
    // With an even m the first step writes pong, so the m-th lands in out.
    const void* src = in;
    void* dst = (m % 2) ? out : pong;
    int rc = FFT3D_MX_OK;
    for (int s = 0; s < m && rc == FFT3D_MX_OK; ++s) {
        rc = plan_launch(p, "fft3d_mx_execute_chain", src, dst, p->batch, step_scale,
                         (hipStream_t)stream);
        src = dst;
        dst = dst == out ? pong : out;
    }
    if (cur != p->device) (void)hipSetDevice(cur);
    return rc;
}

extern "C" void fft3d_mx_destroy(fft3d_mx_plan* p)
{
    if (!p) return;
    p->claims.release_events();
    (void)hipFree(p->bank);
    (void)hipFree(p->work);
    delete p;
}

extern "C" int fft3d_mx_plan_get_info(const fft3d_mx_plan* p, fft3d_mx_plan_info* info)
{
    if (!p || !info) return fail(FFT3D_MX_EINVAL, "fft3d_mx_plan_get_info: NULL argument");
    std::memset(info, 0, sizeof *info);
    info->n0 = p->n0;
    info->n1 = p->n1;
    info->n2 = p->n2;
    info->batch = p->batch;
    info->direction = p->dir;
    info->precision = p->prec;
    info->kind = p->k->kind;
    info->volumes_per_wg = p->k->vpw;
    info->threads = p->k->nthreads;
    info->variant = p->k->variant;
    info->lds_bytes = p->k->lds_bytes;
    info->tune_us = p->tune_us;
    snprintf(info->kernel, sizeof info->kernel, "%s", p->k->name);
    info->small_max = p->k_small ? p->small_max : 0;
    snprintf(info->kernel_small, sizeof info->kernel_small, "%s",
             p->k_small ? p->k_small->name : "");
    return FFT3D_MX_OK;
}

extern "C" int fft3d_mx_kernel_count(void) { return (int)registry().size(); }

extern "C" const char* fft3d_mx_kernel_name(int i, int* n0, int* n1, int* n2, int* precision,
                                            int* direction)
{
    if (i < 0 || i >= (int)registry().size()) return nullptr;
    const KernelDesc& k = registry()[i];
    if (n0) *n0 = k.n0;
    if (n1) *n1 = k.n1;
    if (n2) *n2 = k.n2;
    if (precision) *precision = k.prec;
    if (direction) *direction = k.dir;
    return k.name;
}

// Bank image for tests: lets the CPU suite check the plan-time tables without a GPU.
extern "C" int fft3d_mx_bank_image(int precision, void* dst, size_t cap, size_t* size)
{
    if (precision != kC128 && precision != kC64)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_bank_image: unknown precision %d", precision);
    std::vector<unsigned char> img;
    bank_host_image(precision, img);
    if (size) *size = img.size();
    if (dst) {
        if (cap < img.size()) return fail(FFT3D_MX_EINVAL, "fft3d_mx_bank_image: buffer too small");
        std::memcpy(dst, img.data(), img.size());
    }
    return FFT3D_MX_OK;
}
