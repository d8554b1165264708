// The reference's two plugin ABIs, served by the extended plan (plan.cpp).
//
//   fft3d_*      host pointers (bench/geom/fft3d_api.h:49-66): create() owns device
//                buffers for the batch and a stream; execute() is H2D + transform + D2H
//                and returns with `out` complete, as bench/geom/driver.c:207-218 requires.
//   fft3d_gpu_*  device pointers (bench/gpu/fft3d_gpu_api.h:61-79): execute() enqueues on
//                the NULL stream and returns; the caller synchronizes.
// Both are forward, unnormalised, complex128 cubes.  A HIP failure inside execute() is
// recorded for fft3d_mx_last_error() and printed once to stderr -- never swallowed.
#include <hip/hip_runtime.h>

#include <cstdio>

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
struct fft3d_plan {
    fft3d_mx_plan* p;
    void* din;
    void* dout;
    size_t bytes;
    hipStream_t stream;
};

extern "C" const char* fft3d_name(void) { return "mi355x_hip"; }

extern "C" const char* fft3d_description(void)
{
    return "MI355X gfx950 HIP: LDS-resident fused 3-axis kernels (L<=17, Rader-17), "
           "LDS+register single pass or two-pass pencil (L=24), pencil (L=36); "
           "H2D+kernel+D2H per execute";
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
    h->bytes = (size_t)L * L * L * 16 * (size_t)batch;
    h->p = fft3d_mx_create(L, L, L, batch, FFT3D_MX_FORWARD, FFT3D_MX_C128, 0);
    if (!h->p) {
        delete h;
        return nullptr;
    }
    hipError_t e = hipMalloc(&h->din, h->bytes);
    if (e == hipSuccess) e = hipMalloc(&h->dout, h->bytes);
    if (e == hipSuccess) e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) {
        fail(FFT3D_MX_ENOMEM, "fft3d_create: %s", hipGetErrorString(e));
        (void)hipFree(h->din);
        (void)hipFree(h->dout);
        fft3d_mx_destroy(h->p);
        delete h;
        return nullptr;
    }
    return h;
}

extern "C" void fft3d_execute(fft3d_plan* h, const void* in, void* out)
{
    if (!h || !in || !out) {
        fail(FFT3D_MX_EINVAL, "fft3d_execute: NULL argument");
        report("fft3d_execute");
        return;
    }
    hipError_t e = hipMemcpyAsync(h->din, in, h->bytes, hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) {
        fail(FFT3D_MX_EHIP, "fft3d_execute: H2D: %s", hipGetErrorString(e));
        report("fft3d_execute");
        return;
    }
    if (fft3d_mx_execute(h->p, h->din, h->dout, h->stream) != FFT3D_MX_OK) {
        report("fft3d_execute");
        return;
    }
    e = hipMemcpyAsync(out, h->dout, h->bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        fail(FFT3D_MX_EHIP, "fft3d_execute: D2H/sync: %s", hipGetErrorString(e));
        report("fft3d_execute");
    }
}

extern "C" void fft3d_destroy(fft3d_plan* h)
{
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamDestroy(h->stream);
    (void)hipFree(h->din);
    (void)hipFree(h->dout);
    fft3d_mx_destroy(h->p);
    delete h;
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
