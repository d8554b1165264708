/* Device-resident timing and correctness driver for the fft3d_gpu_api.h plugin ABI.
 *
 * Mirrors the reference's GPU driver (bench/gpu/driver.cu:74-312): input read from the
 * shared file, H2D and D2H measured separately and outside the timed region, create timed
 * with events, warmup then an error check (exit 4), inner count calibrated to >= 20 ms per
 * sample (the clock-ramp lesson of bench/gpu/driver.cu:17-32), every sample bracketed by
 * events plus a device synchronize, checked output from the same code path, and the
 * anti-memoisation poke (exit 5).  Same JSON keys, plus:
 *   --rotate K   cycle the timed executes over K (in, out) buffer pairs.  The MI355X's
 *                256 MB Infinity Cache holds every single pair of the target workloads, so
 *                a roofline figure needs K pairs totalling well over 256 MB; K = 1 is the
 *                reference's same-buffer method.
 * Written in C against the HIP runtime API directly. */
#define _POSIX_C_SOURCE 200809L
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fft3d_gpu_api.h"
#include "fft3d_mi355x.h"

#define HIP_OK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            fprintf(stderr, "fft3d_gpu_driver: %s failed at line %d: %s\n", #call, __LINE__, \
                    hipGetErrorString(e_));                                                 \
            exit(2);                                                                        \
        }                                                                                   \
    } while (0)

static int by_value(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

static void usage(const char *prog)
{
    fprintf(stderr,
            "usage: %s --L N --batch B --in FILE [--out FILE] [--json FILE] [--samples N]\n"
            "          [--warmup N] [--min-sample-ms X] [--run-index N] [--rotate K]\n",
            prog);
    exit(2);
}

static int kernel_error(const char *phase)
{
    hipError_t e = hipGetLastError();
    const char *m = fft3d_mx_last_error();
    if (e != hipSuccess || (m && *m)) {
        fprintf(stderr, "%s: kernel error during %s: %s %s\n", fft3d_gpu_name(), phase,
                hipGetErrorString(e), m ? m : "");
        return 1;
    }
    return 0;
}

int main(int argc, char **argv)
{
    int L = 0, batch = 0, samples = 30, warmup = 5, run_index = 0, rotate = 1;
    double min_ms = 20.0;
    const char *in_path = NULL, *out_path = NULL, *json_path = NULL;
    for (int i = 1; i + 1 < argc; i += 2) {
        const char *a = argv[i], *v = argv[i + 1];
        if (!strcmp(a, "--L")) L = atoi(v);
        else if (!strcmp(a, "--batch")) batch = atoi(v);
        else if (!strcmp(a, "--samples")) samples = atoi(v);
        else if (!strcmp(a, "--warmup")) warmup = atoi(v);
        else if (!strcmp(a, "--min-sample-ms")) min_ms = atof(v);
        else if (!strcmp(a, "--run-index")) run_index = atoi(v);
        else if (!strcmp(a, "--rotate")) rotate = atoi(v);
        else if (!strcmp(a, "--in")) in_path = v;
        else if (!strcmp(a, "--out")) out_path = v;
        else if (!strcmp(a, "--json")) json_path = v;
        else usage(argv[0]);
    }
    if (argc % 2 == 0 || L <= 0 || batch <= 0 || !in_path || samples <= 0 || rotate <= 0)
        usage(argv[0]);

    if (!fft3d_gpu_supports(L)) {
        fprintf(stderr, "%s: L=%d is not supported\n", fft3d_gpu_name(), L);
        if (json_path) {
            FILE *f = fopen(json_path, "w");
            if (f) {
                fprintf(f, "{\"name\":\"%s\",\"L\":%d,\"batch\":%d,\"supported\":false}\n",
                        fft3d_gpu_name(), L, batch);
                fclose(f);
            }
        }
        return 3;
    }
    const size_t V = (size_t)L * L * L, count = V * (size_t)batch;
    const size_t bytes = count * sizeof(fft3d_double2);
    fft3d_double2 *h_in = malloc(bytes), *h_out = malloc(bytes);
    if (!h_in || !h_out) { fprintf(stderr, "fft3d_gpu_driver: out of host memory\n"); return 2; }
    FILE *f = fopen(in_path, "rb");
    if (!f) { perror(in_path); return 2; }
    const size_t got = fread(h_in, sizeof(fft3d_double2), count, f);
    fclose(f);
    if (got != count) {
        fprintf(stderr, "fft3d_gpu_driver: %s holds %zu of %zu elements\n", in_path, got, count);
        return 2;
    }

    HIP_OK(hipSetDevice(0));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, 0));
    fft3d_double2 **d_in = calloc((size_t)rotate, sizeof *d_in);
    fft3d_double2 **d_out = calloc((size_t)rotate, sizeof *d_out);
    if (!d_in || !d_out) return 2;
    for (int r = 0; r < rotate; ++r) {
        HIP_OK(hipMalloc((void **)&d_in[r], bytes));
        HIP_OK(hipMalloc((void **)&d_out[r], bytes));
        HIP_OK(hipMemset(d_out[r], 0, bytes));
    }
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));

    float h2d_ms = 0.f, d2h_ms = 0.f, setup_ms = 0.f;
    HIP_OK(hipEventRecord(e0, NULL));
    HIP_OK(hipMemcpy(d_in[0], h_in, bytes, hipMemcpyHostToDevice));
    HIP_OK(hipEventRecord(e1, NULL));
    HIP_OK(hipEventSynchronize(e1));
    HIP_OK(hipEventElapsedTime(&h2d_ms, e0, e1));
    for (int r = 1; r < rotate; ++r)
        HIP_OK(hipMemcpy(d_in[r], d_in[0], bytes, hipMemcpyDeviceToDevice));

    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipEventRecord(e0, NULL));
    fft3d_gpu_plan *plan = fft3d_gpu_create(L, batch);
    HIP_OK(hipEventRecord(e1, NULL));
    HIP_OK(hipEventSynchronize(e1));
    HIP_OK(hipEventElapsedTime(&setup_ms, e0, e1));
    if (!plan) {
        fprintf(stderr, "%s: create failed for L=%d batch=%d: %s\n", fft3d_gpu_name(), L, batch,
                fft3d_mx_last_error());
        return 3;
    }

    for (int i = 0; i < warmup; ++i) fft3d_gpu_execute(plan, d_in[i % rotate], d_out[i % rotate]);
    HIP_OK(hipDeviceSynchronize());
    if (kernel_error("warmup")) return 4;

    long inner = 1, next_buf = 0;
    for (;;) {
        HIP_OK(hipEventRecord(e0, NULL));
        for (long i = 0; i < inner; ++i, ++next_buf)
            fft3d_gpu_execute(plan, d_in[next_buf % rotate], d_out[next_buf % rotate]);
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipEventRecord(e1, NULL));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        if (ms >= min_ms || inner >= (1L << 24)) break;
        const double grow = ms > 0.f ? min_ms / (double)ms : 8.0;
        const long nx = (long)((double)inner * (grow > 2.0 ? grow : 2.0));
        inner = nx > inner ? nx : 2 * inner;
    }
    double *per = malloc(sizeof(double) * (size_t)samples);
    if (!per) return 2;
    for (int k = 0; k < samples; ++k) {
        HIP_OK(hipEventRecord(e0, NULL));
        for (long i = 0; i < inner; ++i, ++next_buf)
            fft3d_gpu_execute(plan, d_in[next_buf % rotate], d_out[next_buf % rotate]);
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipEventRecord(e1, NULL));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        per[k] = (double)ms * 1e-3 / (double)inner;
    }

    HIP_OK(hipMemset(d_out[0], 0, bytes));
    fft3d_gpu_execute(plan, d_in[0], d_out[0]);
    HIP_OK(hipDeviceSynchronize());
    if (kernel_error("checked execute")) return 4;
    HIP_OK(hipEventRecord(e0, NULL));
    HIP_OK(hipMemcpy(h_out, d_out[0], bytes, hipMemcpyDeviceToHost));
    HIP_OK(hipEventRecord(e1, NULL));
    HIP_OK(hipEventSynchronize(e1));
    HIP_OK(hipEventElapsedTime(&d2h_ms, e0, e1));
    if (out_path) {
        FILE *g = fopen(out_path, "wb");
        if (!g) { perror(out_path); return 2; }
        if (fwrite(h_out, sizeof(fft3d_double2), count, g) != count) {
            fprintf(stderr, "fft3d_gpu_driver: short write to %s\n", out_path);
            fclose(g);
            return 2;
        }
        fclose(g);
    }

    {
        const fft3d_double2 poke = {1.0, -2.0};
        HIP_OK(hipMemcpy(d_in[0], &poke, sizeof poke, hipMemcpyHostToDevice));
        HIP_OK(hipMemset(d_out[0], 0, bytes));
        fft3d_gpu_execute(plan, d_in[0], d_out[0]);
        HIP_OK(hipDeviceSynchronize());
        fft3d_double2 *h_poked = malloc(bytes);
        if (!h_poked) return 2;
        HIP_OK(hipMemcpy(h_poked, d_out[0], bytes, hipMemcpyDeviceToHost));
        const int changed = memcmp(h_poked, h_out, bytes) != 0;
        free(h_poked);
        if (!changed) {
            fprintf(stderr, "%s: output unchanged after the input changed\n", fft3d_gpu_name());
            return 5;
        }
    }

    double mean = 0, var = 0;
    for (int k = 0; k < samples; ++k) mean += per[k];
    mean /= samples;
    for (int k = 0; k < samples; ++k) var += (per[k] - mean) * (per[k] - mean);
    const double sd = samples > 1 ? sqrt(var / (samples - 1)) : 0.0;
    qsort(per, (size_t)samples, sizeof(double), by_value);
    const double best = per[0], median = per[samples / 2], worst = per[samples - 1];
    const double flops = 5.0 * (double)V * log2((double)V) * (double)batch;
    const double min_bytes = 2.0 * (double)bytes;
    if (json_path) {
        FILE *g = fopen(json_path, "w");
        if (!g) { perror(json_path); return 2; }
        fprintf(g,
                "{\"name\":\"%s\",\"description\":\"%s\",\"supported\":true,"
                "\"device\":\"%s\",\"L\":%d,\"batch\":%d,\"volume\":%zu,\"run_index\":%d,"
                "\"samples\":%d,\"inner\":%ld,\"warmup\":%d,\"rotate\":%d,"
                "\"setup_seconds\":%.9g,\"h2d_seconds\":%.9g,\"d2h_seconds\":%.9g,"
                "\"per_execute_seconds\":{\"min\":%.9g,\"median\":%.9g,\"mean\":%.9g,"
                "\"sd\":%.9g,\"max\":%.9g},\"per_transform_seconds_min\":%.9g,"
                "\"gflops_from_min\":%.6f,\"gflops_from_median\":%.6f,"
                "\"effective_gbytes_per_s\":%.6f}\n",
                fft3d_gpu_name(), fft3d_gpu_description(), prop.name, L, batch, V, run_index,
                samples, inner, warmup, rotate, setup_ms * 1e-3, h2d_ms * 1e-3, d2h_ms * 1e-3,
                best, median, mean, sd, worst, best / (double)batch, flops / best / 1e9,
                flops / median / 1e9, min_bytes / best / 1e9);
        fclose(g);
    }
    printf("%-24s L=%-3d B=%-6d rot=%-3d inner=%-7ld min=%10.3f us  median=%10.3f us  "
           "sd=%5.2f%%  %8.2f GF/s  %8.1f GB/s  setup=%.3f s  h2d=%.3f ms d2h=%.3f ms\n",
           fft3d_gpu_name(), L, batch, rotate, inner, best * 1e6, median * 1e6,
           mean > 0 ? 100.0 * sd / mean : 0.0, flops / best / 1e9, min_bytes / best / 1e9,
           setup_ms * 1e-3, h2d_ms, d2h_ms);
    free(per);
    fft3d_gpu_destroy(plan);
    for (int r = 0; r < rotate; ++r) {
        HIP_OK(hipFree(d_in[r]));
        HIP_OK(hipFree(d_out[r]));
    }
    free(d_in);
    free(d_out);
    free(h_in);
    free(h_out);
    return 0;
}
