/* Host-pointer timing and correctness driver for the fft3d_api.h plugin ABI.
 *
 * Same command line, exit codes and JSON record as the reference's single-process driver
 * (bench/geom/driver.c:55-75 CLI, :99-111 unsupported -> 3, :138-142 create failure -> 3,
 * :147-164 chain unit, :166-189 warmup/calibrate/samples, :192-218 chain and checked
 * output, :225-239 anti-memoisation -> 5, :266-279 JSON), so its output files feed the
 * same numpy checker.  One addition: after each phase it asks fft3d_mx_last_error()
 * whether the GPU path failed, and exits 4 (the reference GPU driver's kernel-fault code,
 * bench/gpu/driver.cu:161-168) if it did.
 *
 * What is timed here includes the host<->device copies inside fft3d_execute(); the
 * device-resident figure comes from fft3d_gpu_driver. */
#define _POSIX_C_SOURCE 200809L
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "fft3d_api.h"
#include "fft3d_mi355x.h"

typedef struct {
    int L, batch, samples, warmup, run_index, chain, unitary;
    double min_sample_ms;
    const char *in, *out, *json;
} options;

static double wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int by_value(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

static void *alloc64(size_t bytes)
{
    void *p = NULL;
    if (posix_memalign(&p, 64, bytes ? bytes : 64) != 0) {
        fprintf(stderr, "fft3d_driver: cannot allocate %zu bytes\n", bytes);
        exit(2);
    }
    return p;
}

static void usage(const char *prog)
{
    fprintf(stderr,
            "usage: %s --L N --batch B --in FILE [--out FILE] [--json FILE] [--samples N]\n"
            "          [--warmup N] [--min-sample-ms X] [--run-index N] [--chain M] [--unitary]\n",
            prog);
    exit(2);
}

static options parse(int argc, char **argv)
{
    options o = {0, 0, 30, 5, 0, 1, 0, 20.0, NULL, NULL, NULL};
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (!strcmp(a, "--unitary")) { o.unitary = 1; continue; }
        if (!v) usage(argv[0]);
        if (!strcmp(a, "--L")) o.L = atoi(v);
        else if (!strcmp(a, "--batch")) o.batch = atoi(v);
        else if (!strcmp(a, "--samples")) o.samples = atoi(v);
        else if (!strcmp(a, "--warmup")) o.warmup = atoi(v);
        else if (!strcmp(a, "--min-sample-ms")) o.min_sample_ms = atof(v);
        else if (!strcmp(a, "--run-index")) o.run_index = atoi(v);
        else if (!strcmp(a, "--chain")) o.chain = atoi(v);
        else if (!strcmp(a, "--in")) o.in = v;
        else if (!strcmp(a, "--out")) o.out = v;
        else if (!strcmp(a, "--json")) o.json = v;
        else usage(argv[0]);
        ++i;
    }
    if (o.L <= 0 || o.batch <= 0 || !o.in || o.samples <= 0) usage(argv[0]);
    return o;
}

static void gpu_check(const char *phase)
{
    const char *e = fft3d_mx_last_error();
    if (e && *e) {
        fprintf(stderr, "%s: GPU failure during %s: %s\n", fft3d_name(), phase, e);
        exit(4);
    }
}

/* One timed unit: a single transform, or a chain of `chain` transforms ping-ponging
 * between out and pong, optionally scaled by 1/sqrt(V) after each step. */
static void run_unit(fft3d_plan *plan, const options *o, const double _Complex *in,
                     double _Complex *out, double _Complex *pong, size_t count, double s)
{
    if (o->chain <= 1) {
        fft3d_execute(plan, in, out);
        return;
    }
    const double _Complex *src = in;
    double _Complex *dst = out;
    for (int k = 0; k < o->chain; ++k) {
        fft3d_execute(plan, src, dst);
        if (o->unitary) {
            double *d = (double *)dst;
            for (size_t j = 0; j < 2 * count; ++j) d[j] *= s;
        }
        src = dst;
        dst = dst == out ? pong : out;
    }
}

int main(int argc, char **argv)
{
    const options o = parse(argc, argv);
    if (!fft3d_supports(o.L)) {
        fprintf(stderr, "%s: L=%d is not supported\n", fft3d_name(), o.L);
        if (o.json) {
            FILE *f = fopen(o.json, "w");
            if (f) {
                fprintf(f, "{\"name\":\"%s\",\"L\":%d,\"batch\":%d,\"supported\":false}\n",
                        fft3d_name(), o.L, o.batch);
                fclose(f);
            }
        }
        return 3;
    }
    const size_t V = (size_t)o.L * o.L * o.L, count = V * (size_t)o.batch;
    const size_t bytes = count * sizeof(double _Complex);
    double _Complex *in = alloc64(bytes), *out = alloc64(bytes);
    double _Complex *pong = o.chain > 1 ? alloc64(bytes) : NULL;
    memset(out, 0, bytes);
    if (pong) memset(pong, 0, bytes);
    FILE *f = fopen(o.in, "rb");
    if (!f) { perror(o.in); return 2; }
    const size_t got = fread(in, sizeof(double _Complex), count, f);
    fclose(f);
    if (got != count) {
        fprintf(stderr, "fft3d_driver: %s holds %zu of %zu elements\n", o.in, got, count);
        return 2;
    }

    double t0 = wall();
    fft3d_plan *plan = fft3d_create(o.L, o.batch);
    const double setup = wall() - t0;
    if (!plan) {
        fprintf(stderr, "%s: fft3d_create(%d, %d) failed: %s\n", fft3d_name(), o.L, o.batch,
                fft3d_mx_last_error());
        return 3;
    }
    const double s = 1.0 / sqrt((double)V);
    fft3d_mx_clear_error();

    for (int i = 0; i < o.warmup; ++i) run_unit(plan, &o, in, out, pong, count, s);
    gpu_check("warmup");

    long inner = 1;
    for (;;) {
        t0 = wall();
        for (long i = 0; i < inner; ++i) run_unit(plan, &o, in, out, pong, count, s);
        const double dt = wall() - t0;
        if (dt * 1e3 >= o.min_sample_ms || inner >= (1L << 30)) break;
        const double grow = dt > 0 ? o.min_sample_ms * 1e-3 / dt : 8.0;
        const long next = (long)((double)inner * (grow > 2.0 ? grow : 2.0));
        inner = next > inner ? next : 2 * inner;
    }
    double *per = malloc(sizeof(double) * (size_t)o.samples);
    if (!per) return 2;
    for (int k = 0; k < o.samples; ++k) {
        t0 = wall();
        for (long i = 0; i < inner; ++i) run_unit(plan, &o, in, out, pong, count, s);
        per[k] = (wall() - t0) / (double)inner;
    }
    gpu_check("timed samples");

    if (o.chain > 1 && o.out) {
        run_unit(plan, &o, in, out, pong, count, s);
        char path[4096];
        snprintf(path, sizeof path, "%s.chain", o.out);
        FILE *g = fopen(path, "wb");
        if (g) {
            fwrite(o.chain % 2 ? out : pong, sizeof(double _Complex), count, g);
            fclose(g);
        }
    }

    memset(out, 0, bytes);
    fft3d_execute(plan, in, out);
    gpu_check("checked execute");
    if (o.out) {
        FILE *g = fopen(o.out, "wb");
        if (!g) { perror(o.out); return 2; }
        if (fwrite(out, sizeof(double _Complex), count, g) != count) {
            fprintf(stderr, "fft3d_driver: short write to %s\n", o.out);
            fclose(g);
            return 2;
        }
        fclose(g);
    }

    /* The output must follow the input: perturb one element and require a change. */
    double _Complex *keep = alloc64(bytes);
    memcpy(keep, out, bytes);
    in[0] += 1.0 - 2.0 * I;
    memset(out, 0, bytes);
    fft3d_execute(plan, in, out);
    gpu_check("anti-memoisation execute");
    const int changed = memcmp(keep, out, bytes) != 0;
    free(keep);
    if (!changed) {
        fprintf(stderr, "%s: output unchanged after the input changed\n", fft3d_name());
        return 5;
    }

    double mean = 0, var = 0;
    for (int k = 0; k < o.samples; ++k) mean += per[k];
    mean /= o.samples;
    for (int k = 0; k < o.samples; ++k) var += (per[k] - mean) * (per[k] - mean);
    const double sd = o.samples > 1 ? sqrt(var / (o.samples - 1)) : 0.0;
    qsort(per, (size_t)o.samples, sizeof(double), by_value);
    const double best = per[0], median = per[o.samples / 2], worst = per[o.samples - 1];
    const double xforms = (double)o.batch * (o.chain > 1 ? o.chain : 1);
    const double flops = 5.0 * (double)V * log2((double)V) * xforms;

    if (o.json) {
        FILE *g = fopen(o.json, "w");
        if (!g) { perror(o.json); return 2; }
        fprintf(g,
                "{\"name\":\"%s\",\"description\":\"%s\",\"supported\":true,"
                "\"L\":%d,\"batch\":%d,\"volume\":%zu,\"run_index\":%d,"
                "\"samples\":%d,\"inner\":%ld,\"warmup\":%d,\"setup_seconds\":%.9g,"
                "\"per_execute_seconds\":{\"min\":%.9g,\"median\":%.9g,\"mean\":%.9g,"
                "\"sd\":%.9g,\"max\":%.9g},\"chain\":%d,\"per_transform_seconds_min\":%.9g,"
                "\"gflops_from_min\":%.6f,\"gflops_from_median\":%.6f}\n",
                fft3d_name(), fft3d_description(), o.L, o.batch, V, o.run_index, o.samples,
                inner, o.warmup, setup, best, median, mean, sd, worst, o.chain,
                best / xforms, flops / best / 1e9, flops / median / 1e9);
        fclose(g);
    }
    printf("%-24s L=%-3d B=%-5d m=%-5d inner=%-6ld min=%10.3f us/xform  median=%10.3f  "
           "sd=%6.2f%%  %8.2f GF/s  setup=%.3f s\n",
           fft3d_name(), o.L, o.batch, o.chain, inner, best / xforms * 1e6,
           median / xforms * 1e6, mean > 0 ? 100.0 * sd / mean : 0.0, flops / best / 1e9,
           setup);
    free(per);
    fft3d_destroy(plan);
    free(in);
    free(out);
    free(pong);
    return 0;
}
