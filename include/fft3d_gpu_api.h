/* Device-pointer C ABI -- the six symbols of the reference's GPU plugin contract
 * (bench/gpu/fft3d_gpu_api.h:61-79), same transform and layout as fft3d_api.h, with `in`
 * and `out` DEVICE allocations (hipMalloc), distinct, out-of-place.  execute() enqueues on
 * the NULL stream and may return before the transform completes; the caller synchronizes
 * (bench/gpu/fft3d_gpu_api.h:75-77).  A launch failure is left in the HIP runtime's error
 * state for hipGetLastError() (bench/gpu/fft3d_gpu_api.h:46-50) and is also reported by
 * fft3d_mx_last_error(). */
#ifndef FFT3D_GPU_API_H
#define FFT3D_GPU_API_H

#ifdef __cplusplus
extern "C" {
#endif

/* interleaved complex double {re, im}, layout-identical to HIP's double2 */
typedef struct fft3d_double2 { double x, y; } fft3d_double2;

typedef struct fft3d_gpu_plan fft3d_gpu_plan;

/* replaces bench/gpu/fft3d_gpu_api.h:64 */
const char *fft3d_gpu_name(void);
/* replaces bench/gpu/fft3d_gpu_api.h:67 */
const char *fft3d_gpu_description(void);
/* replaces bench/gpu/fft3d_gpu_api.h:70 -- nonzero for L in {6, 8, 12, 17, 24, 36} */
int fft3d_gpu_supports(int L);
/* replaces bench/gpu/fft3d_gpu_api.h:73 */
fft3d_gpu_plan *fft3d_gpu_create(int L, int batch);
/* replaces bench/gpu/fft3d_gpu_api.h:77 */
void fft3d_gpu_execute(fft3d_gpu_plan *plan, const fft3d_double2 *in, fft3d_double2 *out);
/* replaces bench/gpu/fft3d_gpu_api.h:79 */
void fft3d_gpu_destroy(fft3d_gpu_plan *plan);

#ifdef __cplusplus
}
#endif

#endif /* FFT3D_GPU_API_H */
