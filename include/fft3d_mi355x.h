/* Extended C ABI of the MI355X batched small-volume 3D FFT.
 *
 * The reference's ABIs (fft3d_api.h, fft3d_gpu_api.h) are forward-only, cubes-only and
 * complex128-only (bench/geom/fft3d_api.h:5-11,57-61).  This header adds, beside them:
 * anisotropic extents, the backward direction, complex64, explicit streams, a per-call
 * batch, kernel enumeration and forcing, and an error channel.  Conventions follow the
 * reference's definition module: forward exponent -1, backward +1, unnormalised unless
 * FFT3D_MX_NORMALIZE asks for 1/V (python/slow_dft.py:11-18,27-28,109-113).
 *
 * Layout for every entry point: B contiguous volumes, element (b,x,y,z) at
 * ((b*n0 + x)*n1 + y)*n2 + z, interleaved complex (double {re,im} or float {re,im}).
 * Pointers passed to execute are DEVICE pointers; `in` is never modified.
 *
 * Error convention: functions returning int give 0 on success and a negative
 * FFT3D_MX_E* code on failure; functions returning a pointer give NULL.  The message of
 * the most recent failure on the calling thread is fft3d_mx_last_error().
 */
#ifndef FFT3D_MI355X_H
#define FFT3D_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FFT3D_MX_FORWARD = -1,
    FFT3D_MX_BACKWARD = +1
};
enum {
    FFT3D_MX_C128 = 0, /* complex double, 16 bytes per element */
    FFT3D_MX_C64 = 1   /* complex float, 8 bytes per element */
};
/* create() flags */
enum {
    FFT3D_MX_NORMALIZE = 1u, /* multiply the result by 1/(n0*n1*n2) */
    FFT3D_MX_NO_TUNE = 2u,   /* skip the create-time race; take the first registered kernel */
    /* complex64 only: each line DFT as a matrix product on the matrix cores (17-point lines
       through the conjugate-pair fold), operands split into two bf16 (or fp16) parts, three
       products accumulated in FP32.  Relative L2 error about 5e-6 (bf16) or 2e-7 (fp16),
       against about 1e-7 for the default FFT. */
    FFT3D_MX_MFMA_BF16 = 4u,
    FFT3D_MX_MFMA_FP16 = 8u
};
/* error codes */
enum {
    FFT3D_MX_OK = 0,
    FFT3D_MX_EINVAL = -1,      /* bad argument */
    FFT3D_MX_EUNSUPPORTED = -2,/* no kernel for this geometry / precision / direction */
    FFT3D_MX_EHIP = -3,        /* a HIP runtime call failed */
    FFT3D_MX_ENOMEM = -4       /* device or host allocation failed */
};

typedef struct fft3d_mx_plan fft3d_mx_plan;

/* Message of the last failure on this thread ("" if none).  Never NULL. */
const char *fft3d_mx_last_error(void);
/* Clear the last-failure message. */
void fft3d_mx_clear_error(void);

/* Nonzero if a kernel exists for (n0, n1, n2) at this precision (both directions exist). */
int fft3d_mx_supports(int n0, int n1, int n2, int precision);

/* Plan a transform of up to `batch` volumes on the current HIP device.  Computes the
 * twiddle / Rader tables in long double, uploads them, and -- when several kernels can
 * serve the geometry -- times each on the device and keeps the fastest (the choice is
 * cached per process, and in the file named by $FFT3D_MX_WISDOM when that is set).  For
 * batch >= 16 the race also runs at batch / 8, and executes of at most
 * sqrt(batch * (batch / 8)) volumes use that winner if it differs (plan_info's
 * kernel_small, small_max).  $FFT3D_MX_KERNEL=<name> forces a kernel by name.  NULL on
 * failure. */
fft3d_mx_plan *fft3d_mx_create(int n0, int n1, int n2, int batch, int direction,
                               int precision, unsigned flags);

/* Same, but with the kernel chosen by name (see fft3d_mx_kernel_name); no race. */
fft3d_mx_plan *fft3d_mx_create_named(int n0, int n1, int n2, int batch, int direction,
                                     int precision, unsigned flags, const char *kernel);

/* Enqueue the transform of the plan's batch on `stream` (a hipStream_t; NULL = the NULL
 * stream).  Asynchronous.  0 or a negative error code.  `in` and `out` are distinct device
 * buffers aligned to the element size (16 B complex128, 8 B complex64; hipMalloc gives
 * 256 B), as are the buffers of every other entry point below. */
int fft3d_mx_execute(fft3d_mx_plan *plan, const void *in, void *out, void *stream);

/* Same for the first `batch` volumes only (0 <= batch <= the planned batch). */
int fft3d_mx_execute_n(fft3d_mx_plan *plan, const void *in, void *out, long batch,
                       void *stream);

/* Chain of m transforms, each output multiplied by step_scale and fed to the next step
 * (the reference driver's --chain M mode, bench/geom/driver.c:147-164; step_scale =
 * 1/sqrt(V) is its --unitary mode, whose end state has the closed form FFT^2 = V * index
 * reversal, bench/geom/check.py:43-61).  The scaling is fused into each step's kernel.
 * Steps ping-pong between `out` and `pong` (both device buffers of the plan's batch size,
 * distinct from each other and from `in`) so that the end state always lands in `out`.
 * Enqueued on `stream`; m >= 1.  0 or a negative error code. */
int fft3d_mx_execute_chain(fft3d_mx_plan *plan, const void *in, void *out, void *pong, int m,
                           double step_scale, void *stream);

void fft3d_mx_destroy(fft3d_mx_plan *plan);

/* Introspection of a live plan. */
typedef struct fft3d_mx_plan_info {
    int n0, n1, n2, batch, direction, precision;
    int kind;            /* 0 fused (volume resident in LDS), 1 pencil (two passes) */
    int volumes_per_wg;  /* fused: volumes per workgroup; pencil: slabs per workgroup */
    int threads;         /* workgroup size of the main kernel */
    int variant;         /* line-algorithm variant (L=17: 0 Rader, 1 folded dense) */
    size_t lds_bytes;    /* LDS per workgroup of the main kernel */
    double tune_us;      /* create-time race result of the winner (0 if no race) */
    char kernel[96];     /* kernel name */
    long small_max;      /* executes of at most this many volumes use kernel_small (0: none) */
    char kernel_small[96];  /* small-batch kernel name ("" if none) */
} fft3d_mx_plan_info;
int fft3d_mx_plan_get_info(const fft3d_mx_plan *plan, fft3d_mx_plan_info *info);

/* Kernel registry enumeration: every compiled kernel variant. */
int fft3d_mx_kernel_count(void);
/* Name of kernel i and its geometry/precision/direction; NULL if i is out of range. */
const char *fft3d_mx_kernel_name(int i, int *n0, int *n1, int *n2, int *precision,
                                 int *direction);

/* The plan-time constant bank, exactly as uploaded to the device, for inspection.  Layout
 * (element type double for FFT3D_MX_C128, float for FFT3D_MX_C64, complex = {re, im}):
 *   W tables, complex, concatenated for N = 2,3,4,6,8,9,12,16,17,24,36,48 (185 entries),
 *     entry m of table N = exp(-2 pi i m / N);
 *   Rader-17 spectra, complex [2][16]: [0] forward, [1] backward, Bhat = DFT16(b')/16,
 *     b'[q] = W17^(3^-q mod 17) with W17 = exp(direction * 2 pi i / 17);
 *   cos(2 pi m/17), m = 0..16, real [17];  sin(2 pi m/17), m = 0..16, real [17].
 * Writes at most `cap` bytes to dst (dst may be NULL to query *size).  0 or error. */
int fft3d_mx_bank_image(int precision, void *dst, size_t cap, size_t *size);

/* ---------------------------------------------------------------------------------------
 * One persistent launch over a mixed batch.  Each item is a batch of volumes of one of the
 * fused geometries (6^3, 8^3, 12^3, 17^3, 8x8x16), in either direction; all items share
 * one precision.  create() cuts the items into blocks, orders them heaviest first, races the
 * compiled tile presets on the device and keeps the fastest; execute() is ONE kernel launch
 * that transforms every item (in[i] -> out[i], device buffers of items[i].batch volumes). */
typedef struct fft3d_mx_item {
    int n0, n1, n2;
    int batch;
    int direction; /* FFT3D_MX_FORWARD or FFT3D_MX_BACKWARD */
} fft3d_mx_item;
typedef struct fft3d_mx_multi fft3d_mx_multi;

/* At most 16 items.  flags: FFT3D_MX_NORMALIZE (1/V on every item), FFT3D_MX_NO_TUNE, and
 * for complex64 FFT3D_MX_MFMA_BF16 or FFT3D_MX_MFMA_FP16 (the matrix-core form in the same
 * single launch).  $FFT3D_MX_MULTI_KERNEL=<name> forces a preset.  NULL on failure. */
fft3d_mx_multi *fft3d_mx_multi_create(const fft3d_mx_item *items, int nitems, int precision,
                                      unsigned flags);
int fft3d_mx_multi_execute(fft3d_mx_multi *m, const void *const *in, void *const *out,
                           void *stream);
/* Kernel name of the plan (at most cap-1 chars written), task count and grid size. */
int fft3d_mx_multi_get_info(const fft3d_mx_multi *m, char *kernel, size_t cap, int *ntasks,
                            int *grid);
void fft3d_mx_multi_destroy(fft3d_mx_multi *m);

/* ---------------------------------------------------------------------------------------
 * Transform that allocates nothing.  No counterpart in the reference: for callers that
 * cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Same transform as a plan of (n0, n1, n2, batch, direction,
 * precision) created with FFT3D_MX_NO_TUNE, enqueued on `stream`, with no plan: the
 * constant tables and the claim counters the kernels use are static device arrays of the
 * library (about 134 KiB per device, held until the process ends).  flags: 0 or
 * FFT3D_MX_NORMALIZE.  0 <= batch <= 2^31 - 1 (claim tickets are 32-bit); other sizes are
 * refused with FFT3D_MX_EINVAL.  The first call with a device current sets that device up
 * synchronously and must not be captured; a call with batch 0 does only that.  All calls
 * on one device share its 64 sets of claim counters: up to 64 calls may be in flight on a device
 * at once, and calls beyond that overlap on a counter and race.  0 or an error code. */
int fft3d_mx_execute_noalloc(int n0, int n1, int n2, long batch, int direction, int precision,
                             unsigned flags, const void *in, void *out, void *stream);

/* Bytes of device memory one volume of this geometry/precision occupies. */
size_t fft3d_mx_volume_bytes(int n0, int n1, int n2, int precision);

#ifdef __cplusplus
}
#endif

#endif /* FFT3D_MI355X_H */
