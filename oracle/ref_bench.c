/* ORACLE / CPU BASELINE HARNESS -- test and measurement infrastructure only.
 *
 * Times the reference's own CPU kernels (linked in by build_ref.sh) through their fft3d_api
 * semantics (bench/geom/fft3d_api.h:49-66): create once, then repeated execute on the same
 * buffers for a bounded wall time, like the timed loop of bench/geom/driver.c:182-189.
 *
 *   ref_bench_{st,mt} --L n --batch B [--seconds S] [--in FILE] [--out FILE]
 *
 * Without --in the input is a seeded hash fill.  --out writes one execute's output (raw
 * complex128) so tests can check the reference kernel against the oracle.  Prints one JSON
 * line. */
#define _POSIX_C_SOURCE 200809L
#include <complex.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef void *(*create_fn)(int, int);
typedef void (*exec_fn)(void *, const double _Complex *, double _Complex *);
typedef void (*destroy_fn)(void *);

#ifdef REF_BEST
void *fft3d_best_create(int L, int batch);
void fft3d_best_execute(void *plan, const double _Complex *in, double _Complex *out);
void fft3d_best_destroy(void *plan);
const char *fft3d_best_selected_name(const void *plan);
#endif
#ifdef REF_MT
#include <omp.h>
#define DECL(k)                                                                   \
    void *k##_create(int, int);                                                   \
    void k##_execute(void *, const double _Complex *, double _Complex *);         \
    void k##_destroy(void *);                                                     \
    const char *k##_name(void);
DECL(m6) DECL(m8) DECL(m17) DECL(m36)
#endif

static double wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char **argv)
{
    int L = 0, B = 0;
    double seconds = 5.0;
    const char *in_path = NULL, *out_path = NULL;
    for (int i = 1; i + 1 < argc; i += 2) {
        if (!strcmp(argv[i], "--L")) L = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--batch")) B = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--seconds")) seconds = atof(argv[i + 1]);
        else if (!strcmp(argv[i], "--in")) in_path = argv[i + 1];
        else if (!strcmp(argv[i], "--out")) out_path = argv[i + 1];
    }
    if (L <= 0 || B <= 0) {
        fprintf(stderr, "usage: %s --L n --batch B [--seconds S] [--in F] [--out F]\n", argv[0]);
        return 2;
    }
    create_fn create = NULL;
    exec_fn execute = NULL;
    destroy_fn destroy = NULL;
    const char *kernel = "?";
    int threads = 1;
#ifdef REF_BEST
    create = (create_fn)fft3d_best_create;
    execute = (exec_fn)fft3d_best_execute;
    destroy = (destroy_fn)fft3d_best_destroy;
#endif
#ifdef REF_MT
    threads = omp_get_max_threads();
    switch (L) {
    case 6: create = m6_create; execute = m6_execute; destroy = m6_destroy; kernel = m6_name(); break;
    case 8: create = m8_create; execute = m8_execute; destroy = m8_destroy; kernel = m8_name(); break;
    case 17: create = m17_create; execute = m17_execute; destroy = m17_destroy; kernel = m17_name(); break;
    case 36: create = m36_create; execute = m36_execute; destroy = m36_destroy; kernel = m36_name(); break;
    default: fprintf(stderr, "ref_bench_mt: no kernel for L=%d\n", L); return 3;
    }
#endif
    const size_t n = (size_t)L * L * L * (size_t)B;
    double _Complex *in = NULL, *out = NULL;
    if (posix_memalign((void **)&in, 64, n * sizeof *in) || posix_memalign((void **)&out, 64, n * sizeof *out))
        return 2;
    if (in_path) {
        FILE *f = fopen(in_path, "rb");
        if (!f || fread(in, sizeof *in, n, f) != n) { fprintf(stderr, "ref_bench: bad --in\n"); return 2; }
        fclose(f);
    } else {
        unsigned long long s = 0x9E3779B97F4A7C15ULL;
        for (size_t i = 0; i < n; ++i) {
            s = s * 6364136223846793005ULL + 1442695040888963407ULL;
            double re = (double)(s >> 11) / 9007199254740992.0 * 2.0 - 1.0;
            s = s * 6364136223846793005ULL + 1442695040888963407ULL;
            double im = (double)(s >> 11) / 9007199254740992.0 * 2.0 - 1.0;
            in[i] = re + im * I;
        }
    }
    memset(out, 0, n * sizeof *out);
    double t0 = wall();
    void *plan = create(L, B);
    const double setup = wall() - t0;
    if (!plan) { fprintf(stderr, "ref_bench: create(%d,%d) failed\n", L, B); return 3; }
#ifdef REF_BEST
    kernel = fft3d_best_selected_name(plan);
#endif
    execute(plan, in, out);
    if (out_path) {
        FILE *g = fopen(out_path, "wb");
        if (!g || fwrite(out, sizeof *out, n, g) != n) return 2;
        fclose(g);
    }
    execute(plan, in, out);
    long reps = 0;
    t0 = wall();
    double dt = 0.0;
    do {
        execute(plan, in, out);
        ++reps;
        dt = wall() - t0;
    } while (dt < seconds);
    printf("{\"L\":%d,\"batch\":%d,\"kernel\":\"%s\",\"threads\":%d,\"reps\":%ld,"
           "\"seconds\":%.6f,\"setup_seconds\":%.6f,\"us_per_transform\":%.6f,"
           "\"transforms_per_s\":%.3f}\n",
           L, B, kernel, threads, reps, dt, setup, dt / ((double)reps * B) * 1e6,
           (double)reps * B / dt);
    destroy(plan);
    free(in);
    free(out);
    return 0;
}
