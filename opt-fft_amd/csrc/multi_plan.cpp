// Host side of the mixed-batch persistent launch (include/fft3d_mi355x.h, multi.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "claims_hip.hpp"
#include "fft3d_mi355x.h"
#include "multi.hpp"
#include "multi_tasks.hpp"
#include "plan_internal.hpp"


using namespace fx;

struct fft3d_mx_multi {
    int nitems, prec;
    int code[kMultiMaxItems], batch[kMultiMaxItems], conj[kMultiMaxItems];
    double scale[kMultiMaxItems];
    const MultiDesc* k;
    int2* tasks;   // device
    int ntasks;
    void* bank;    // device
    unsigned* work;  // device: the claim counters `claims` hands out (claims.hpp)
    ClaimSlots claims;
    int device;
};

template <typename T>
static void fill_args(const fft3d_mx_multi* m, const void* const* in, void* const* out,
                      MultiArgs<T>& a)
{
    std::memset(&a, 0, sizeof a);
    for (int i = 0; i < m->nitems; ++i) {
        a.item[i].in = (const cx<T>*)in[i];
        a.item[i].out = (cx<T>*)out[i];
        a.item[i].nvol = m->batch[i];
        a.item[i].code = m->code[i];
        a.item[i].conj = m->conj[i];
        a.item[i].scale = (T)m->scale[i];
    }
}

// One launch with claim counter `work` (the race passes the first counter directly; user
// launches take theirs from m->claims).
static hipError_t multi_launch(fft3d_mx_multi* m, const MultiDesc* k, const int2* tasks,
                               int ntasks, const void* const* in, void* const* out,
                               unsigned* work, hipStream_t s)
{
    if (m->prec == kC128) {
        MultiArgs<double> a;
        fill_args(m, in, out, a);
        return k->launch(&a, tasks, ntasks, m->bank, work, s);
    }
    MultiArgs<float> a;
    fill_args(m, in, out, a);
    return k->launch(&a, tasks, ntasks, m->bank, work, s);
}

static int upload_tasks(const std::vector<int2>& t, int2** dev)
{
    *dev = nullptr;
    if (t.empty()) return FFT3D_MX_OK;
    hipError_t e = hipMalloc((void**)dev, t.size() * sizeof(int2));
    if (e == hipSuccess) e = hipMemcpy(*dev, t.data(), t.size() * sizeof(int2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(*dev);
        *dev = nullptr;
        return fail(FFT3D_MX_ENOMEM, "fft3d_mx_multi_create: task table: %s", hipGetErrorString(e));
    }
    return FFT3D_MX_OK;
}

// Time every preset on the items (batches capped so the race buffers stay under ~512 MiB).
static const MultiDesc* multi_race(fft3d_mx_multi* m, const std::vector<const MultiDesc*>& cands)
{
    const size_t elem = elem_bytes(m->prec);
    size_t total = 0;
    int rb[kMultiMaxItems];
    for (int i = 0; i < m->nitems; ++i) {
        const int g = m->code[i];
        total += (size_t)kMultiShape[g][0] * kMultiShape[g][1] * kMultiShape[g][2] * elem * m->batch[i];
    }
    const double shrink = total > (256ull << 20) ? (double)(256ull << 20) / (double)total : 1.0;
    std::vector<void*> bufs;
    std::vector<const void*> ins(m->nitems);
    std::vector<void*> outs(m->nitems);
    bool ok = true;
    for (int i = 0; i < m->nitems && ok; ++i) {
        const int g = m->code[i];
        rb[i] = std::max(m->batch[i] > 0 ? 1 : 0, (int)(m->batch[i] * shrink));
        const size_t bytes = std::max<size_t>(1, (size_t)kMultiShape[g][0] * kMultiShape[g][1] *
                                                     kMultiShape[g][2] * elem * rb[i]);
        void *a = nullptr, *b = nullptr;
        ok = hipMalloc(&a, bytes) == hipSuccess && hipMalloc(&b, bytes) == hipSuccess;
        if (a) bufs.push_back(a);
        if (b) bufs.push_back(b);
        if (ok) (void)hipMemsetD32((hipDeviceptr_t)a, 0x3F800000u, bytes / 4);
        ins[i] = a;
        outs[i] = b;
    }
    const MultiDesc* win = cands[0];
    if (ok) {
        int saved[kMultiMaxItems];
        std::memcpy(saved, m->batch, sizeof saved);
        std::memcpy(m->batch, rb, sizeof(int) * m->nitems);
        hipEvent_t e0, e1;
        hipStream_t s;
        double best = 1e30;
        const bool have_s = hipStreamCreate(&s) == hipSuccess;
        const bool have_e0 = have_s && hipEventCreate(&e0) == hipSuccess;
        const bool have_e1 = have_e0 && hipEventCreate(&e1) == hipSuccess;
        if (have_e1) {
            for (const MultiDesc* k : cands) {
                std::vector<int2> t;
                multi_tasks(*k, m->code, m->batch, m->nitems, elem, t);
                int2* dt = nullptr;
                if (upload_tasks(t, &dt) != FFT3D_MX_OK) continue;
                bool good = true;
                for (int w = 0; w < 2 && good; ++w)
                    good = multi_launch(m, k, dt, (int)t.size(), ins.data(), outs.data(), m->work,
                                        s) == hipSuccess;
                double kt = 1e30;
                for (int g = 0; g < 3 && good; ++g) {
                    good = hipEventRecord(e0, s) == hipSuccess;
                    for (int r = 0; r < 3 && good; ++r)
                        good = multi_launch(m, k, dt, (int)t.size(), ins.data(), outs.data(),
                                            m->work, s) == hipSuccess;
                    good = good && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
                    float ms = 0.f;
                    if (good && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) kt = std::min(kt, (double)ms);
                }
                (void)hipStreamSynchronize(s);
                (void)hipFree(dt);
                if (good && kt < best) {
                    best = kt;
                    win = k;
                }
            }
        }
        if (have_e1) (void)hipEventDestroy(e1);
        if (have_e0) (void)hipEventDestroy(e0);
        if (have_s) (void)hipStreamDestroy(s);
        std::memcpy(m->batch, saved, sizeof saved);
    }
    for (void* b : bufs) (void)hipFree(b);
    (void)hipGetLastError();
    return win;
}

extern "C" fft3d_mx_multi* fft3d_mx_multi_create(const fft3d_mx_item* items, int nitems,
                                                 int precision, unsigned flags)
{
    if (!items || nitems < 1 || nitems > kMultiMaxItems) {
        fail(FFT3D_MX_EINVAL, "fft3d_mx_multi_create: need 1..%d items, got %d", kMultiMaxItems,
             nitems);
        return nullptr;
    }
    if (precision != kC128 && precision != kC64) {
        fail(FFT3D_MX_EINVAL, "fft3d_mx_multi_create: unknown precision %d", precision);
        return nullptr;
    }
    if ((flags & FFT3D_MX_MFMA_BF16) && (flags & FFT3D_MX_MFMA_FP16)) {
        fail(FFT3D_MX_EINVAL, "fft3d_mx_multi_create: MFMA_BF16 and MFMA_FP16 are exclusive");
        return nullptr;
    }
    fft3d_mx_multi* m = new (std::nothrow) fft3d_mx_multi();
    if (!m) {
        fail(FFT3D_MX_ENOMEM, "fft3d_mx_multi_create: out of host memory");
        return nullptr;
    }
    m->nitems = nitems;
    m->prec = precision;
    for (int i = 0; i < nitems; ++i) {
        const fft3d_mx_item& it = items[i];
        const int g = multi_code(it.n0, it.n1, it.n2);
        if (g < 0 || it.batch < 0 || (it.direction != -1 && it.direction != 1)) {
            fail(g < 0 ? FFT3D_MX_EUNSUPPORTED : FFT3D_MX_EINVAL,
                 "fft3d_mx_multi_create: item %d (%dx%dx%d, batch %d, direction %d) is not "
                 "supported in a mixed launch", i, it.n0, it.n1, it.n2, it.batch, it.direction);
            delete m;
            return nullptr;
        }
        m->code[i] = g;
        m->batch[i] = it.batch;
        m->conj[i] = it.direction > 0;
        m->scale[i] = (flags & FFT3D_MX_NORMALIZE) ? 1.0 / ((double)it.n0 * it.n1 * it.n2) : 1.0;
    }
    if (hipGetDevice(&m->device) != hipSuccess) m->device = 0;
    std::vector<unsigned char> img;
    bank_host_image(precision, img);
    hipError_t e = hipMalloc(&m->bank, img.size());
    if (e == hipSuccess) e = hipMemcpy(m->bank, img.data(), img.size(), hipMemcpyHostToDevice);
    const size_t work_bytes = ClaimSlots::bytes();
    if (e == hipSuccess) e = hipMalloc((void**)&m->work, work_bytes);
    if (e == hipSuccess) e = hipMemset(m->work, 0, work_bytes);
    if (e != hipSuccess) {
        fail(FFT3D_MX_ENOMEM, "fft3d_mx_multi_create: bank: %s", hipGetErrorString(e));
        (void)hipFree(m->bank);
        (void)hipFree(m->work);
        delete m;
        return nullptr;
    }
    m->claims.attach(m->work);
    const int algo = (flags & FFT3D_MX_MFMA_BF16) ? 1 : (flags & FFT3D_MX_MFMA_FP16) ? 2 : 0;
    std::vector<const MultiDesc*> cands;
    const char* forced = getenv("FFT3D_MX_MULTI_KERNEL");
    for (const MultiDesc& d : multi_registry())
        if (d.prec == precision && d.algo == algo &&
            (!forced || !*forced || !strcmp(forced, d.name)))
            cands.push_back(&d);
    if (cands.empty()) {
        fail(FFT3D_MX_EUNSUPPORTED, "fft3d_mx_multi_create: no %smixed kernel%s%s for %s",
             algo ? "matrix-core " : "", forced && *forced ? " named " : "",
             forced && *forced ? forced : "", precision == kC128 ? "complex128" : "complex64");
        (void)hipFree(m->bank);
        (void)hipFree(m->work);
        delete m;
        return nullptr;
    }
    m->k = (cands.size() > 1 && !(flags & FFT3D_MX_NO_TUNE)) ? multi_race(m, cands) : cands[0];
    std::vector<int2> t;
    multi_tasks(*m->k, m->code, m->batch, nitems, elem_bytes(precision), t);
    m->ntasks = (int)t.size();
    if (upload_tasks(t, &m->tasks) != FFT3D_MX_OK) {
        (void)hipFree(m->bank);
        (void)hipFree(m->work);
        delete m;
        return nullptr;
    }
    return m;
}

extern "C" int fft3d_mx_multi_execute(fft3d_mx_multi* m, const void* const* in, void* const* out,
                                      void* stream)
{
    if (!m || !in || !out) return fail(FFT3D_MX_EINVAL, "fft3d_mx_multi_execute: NULL argument");
    for (int i = 0; i < m->nitems; ++i) {
        if (m->batch[i] == 0) continue;
        if (!in[i] || !out[i] || in[i] == (const void*)out[i])
            return fail(FFT3D_MX_EINVAL, "fft3d_mx_multi_execute: item %d needs distinct buffers", i);
        if (!elem_aligned(in[i], m->prec) || !elem_aligned(out[i], m->prec))
            return fail(FFT3D_MX_EINVAL,
                        "fft3d_mx_multi_execute: item %d buffers must be aligned to %zu bytes", i,
                        elem_bytes(m->prec));
    }
    int cur = m->device;
    (void)hipGetDevice(&cur);
    if (cur != m->device) (void)hipSetDevice(m->device);
    const hipStream_t s = (hipStream_t)stream;
    const hipError_t e = m->claims.run(s, [&](unsigned* w) {
        return multi_launch(m, m->k, m->tasks, m->ntasks, in, out, w, s);
    });
    if (cur != m->device) (void)hipSetDevice(cur);
    if (e == HipClaimRT::busy)
        return fail(FFT3D_MX_EINVAL, "fft3d_mx_multi_execute: %d launches of this plan were "
                    "already captured into graphs; no claim counter is left",
                    ClaimSlots::kSlots);
    if (e != hipSuccess)
        return fail(FFT3D_MX_EHIP, "fft3d_mx_multi_execute: %s: %s", m->k->name, hipGetErrorString(e));
    return FFT3D_MX_OK;
}

extern "C" int fft3d_mx_multi_get_info(const fft3d_mx_multi* m, char* kernel, size_t cap,
                                       int* ntasks, int* grid)
{
    if (!m) return fail(FFT3D_MX_EINVAL, "fft3d_mx_multi_get_info: NULL plan");
    if (kernel && cap) snprintf(kernel, cap, "%s", m->k->name);
    if (ntasks) *ntasks = m->ntasks;
    if (grid) {
        int cur = m->device;
        (void)hipGetDevice(&cur);
        if (cur != m->device) (void)hipSetDevice(m->device);
        *grid = m->k->launch_grid && m->ntasks > 0 ? (int)m->k->launch_grid(m->ntasks) : 0;
        if (cur != m->device) (void)hipSetDevice(cur);
    }
    return FFT3D_MX_OK;
}

extern "C" void fft3d_mx_multi_destroy(fft3d_mx_multi* m)
{
    if (!m) return;
    m->claims.release_events();
    (void)hipFree(m->tasks);
    (void)hipFree(m->bank);
    (void)hipFree(m->work);
    delete m;
}
