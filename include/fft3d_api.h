/* Host-pointer C ABI of the batched small-volume 3D FFT -- the six symbols of the
 * reference's plugin contract (bench/geom/fft3d_api.h:49-66), with the same semantics:
 *
 *   out[b][k0][k1][k2] = sum_j in[b][j0][j1][j2] * exp(-2 pi i (k0 j0 + k1 j1 + k2 j2) / L)
 *
 * forward, unnormalised, out-of-place; element (b,x,y,z) at ((b*L + x)*L + y)*L + z,
 * interleaved complex double; `in` is not modified.  Buffers are host memory owned by the
 * caller.  This library runs the transform on the GPU: fft3d_create() allocates device
 * buffers for `batch` volumes and picks the kernel; fft3d_execute() copies in, transforms,
 * copies out, and returns with `out` complete (the driver reads it immediately,
 * bench/geom/driver.c:207-218).
 *
 * Errors: fft3d_supports() is 0 and fft3d_create() returns NULL for an unsupported L.
 * execute() has no return channel (bench/geom/fft3d_api.h:64); a HIP failure inside it is
 * reported by fft3d_mx_last_error() (fft3d_mi355x.h), which the bundled driver checks
 * and turns into exit status 4, the reference's kernel-fault code. */
#ifndef FFT3D_API_H
#define FFT3D_API_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fft3d_plan fft3d_plan;

/* replaces bench/geom/fft3d_api.h:52 */
const char *fft3d_name(void);
/* replaces bench/geom/fft3d_api.h:55 */
const char *fft3d_description(void);
/* replaces bench/geom/fft3d_api.h:58 -- nonzero for L in {6, 8, 12, 17, 24, 36} */
int fft3d_supports(int L);
/* replaces bench/geom/fft3d_api.h:61 */
fft3d_plan *fft3d_create(int L, int batch);
/* replaces bench/geom/fft3d_api.h:64 (C++ callers pass std::complex<double> or double[2]) */
#ifdef __cplusplus
void fft3d_execute(fft3d_plan *plan, const void *in, void *out);
#else
void fft3d_execute(fft3d_plan *plan, const double _Complex *in, double _Complex *out);
#endif
/* replaces bench/geom/fft3d_api.h:66 */
void fft3d_destroy(fft3d_plan *plan);

#ifdef __cplusplus
}
#endif

#endif /* FFT3D_API_H */
