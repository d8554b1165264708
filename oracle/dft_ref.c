/* ORACLE -- test infrastructure only.
 *
 * Nothing in the product links, loads or calls this file.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may use it, and only as the
 * checker or as the timed CPU port, never as the thing that produces a product result.
 *
 * What it restates
 * ----------------
 * The transform definition of bench/geom/fft3d_api.h:5-9 (forward, unnormalised,
 * out-of-place, element (b,x,y,z) at ((b*n0+x)*n1+y)*n2+z, interleaved complex), computed
 * the slow way: a dense DFT matrix applied along each axis in turn, exactly the row-column
 * algorithm of python/slow_dft.py:93-106 (dft3d_separable) and of the C self-check
 * bench/geom/best/verify_best.c:62-90 (axis_reference).  Two choices follow the reference's
 * extended-precision check python/test_fft3d.py:539-569:
 *   * accumulation and the twiddle table are long double (x87 80-bit, 64-bit mantissa);
 *   * the exponent j*k is reduced mod n BEFORE the angle is formed, so the table holds
 *     exactly n distinct angles 2*pi*m/n, m in [0,n).
 * The backward sign and the optional 1/V follow python/slow_dft.py:27-28,109-113.
 *
 * Pins (tests/test_oracle.py): numpy.fft at every shape, the six-fold definitional sum on
 * tiny shapes, delta/ones/single-mode closed forms, DFT^4 = V^2 I, Parseval, the shift
 * theorem and Hermitian symmetry of real input.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef long double ld;

/* exp(sign * 2 pi i m / n) for m in [0, n), long double. */
static void twiddles(int n, int sign, ld *wr, ld *wi)
{
    const ld two_pi = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < n; ++m) {
        ld a = two_pi * (ld)m / (ld)n;
        wr[m] = cosl(a);
        wi[m] = (ld)sign * sinl(a);
    }
}

/* Contract one axis of a block shaped (outer, n, inner) with the dense n x n DFT matrix.
 * Work arrays are long double split re/im so intermediate passes keep the extra precision. */
static void axis_pass(const ld *xr, const ld *xi, ld *yr, ld *yi,
                      long outer, int n, long inner, const ld *wr, const ld *wi)
{
    for (long o = 0; o < outer; ++o) {
        for (int k = 0; k < n; ++k) {
            for (long i = 0; i < inner; ++i) {
                ld ar = 0.0L, ai = 0.0L;
                for (int j = 0; j < n; ++j) {
                    int m = (int)(((long)j * k) % n);
                    long s = (o * n + j) * inner + i;
                    ar += xr[s] * wr[m] - xi[s] * wi[m];
                    ai += xr[s] * wi[m] + xi[s] * wr[m];
                }
                long d = (o * n + k) * inner + i;
                yr[d] = ar;
                yi[d] = ai;
            }
        }
    }
}

/* One volume: in/out are long double split arrays of length V. */
static int volume_dft(ld *ar, ld *ai, ld *br, ld *bi, int n0, int n1, int n2, int sign)
{
    int nmax = n0 > n1 ? n0 : n1;
    if (n2 > nmax) nmax = n2;
    ld *wr = malloc(sizeof(ld) * (size_t)nmax);
    ld *wi = malloc(sizeof(ld) * (size_t)nmax);
    if (!wr || !wi) { free(wr); free(wi); return -1; }
    /* axis 0 (x, slowest): outer 1, inner n1*n2 */
    twiddles(n0, sign, wr, wi);
    axis_pass(ar, ai, br, bi, 1, n0, (long)n1 * n2, wr, wi);
    /* axis 1 (y): outer n0, inner n2 */
    twiddles(n1, sign, wr, wi);
    axis_pass(br, bi, ar, ai, n0, n1, n2, wr, wi);
    /* axis 2 (z, fastest): outer n0*n1, inner 1 */
    twiddles(n2, sign, wr, wi);
    axis_pass(ar, ai, br, bi, (long)n0 * n1, n2, 1, wr, wi);
    free(wr);
    free(wi);
    return 0;
}

/* Batched 3D DFT of complex128 input, complex128 output (interleaved re,im).
 * sign = -1 forward, +1 backward.  normalize != 0 multiplies by 1/(n0*n1*n2).
 * Returns 0 on success, -1 on bad arguments or allocation failure. */
int oracle_dft3d_c128(const double *in, double *out, long batch,
                      int n0, int n1, int n2, int sign, int normalize)
{
    if (batch < 0 || n0 <= 0 || n1 <= 0 || n2 <= 0 || (sign != -1 && sign != 1)) return -1;
    long V = (long)n0 * n1 * n2;
    ld *buf = malloc(sizeof(ld) * 4 * (size_t)V);
    if (!buf) return -1;
    ld *ar = buf, *ai = buf + V, *br = buf + 2 * V, *bi = buf + 3 * V;
    ld scale = normalize ? 1.0L / (ld)V : 1.0L;
    for (long b = 0; b < batch; ++b) {
        const double *src = in + 2 * b * V;
        double *dst = out + 2 * b * V;
        for (long e = 0; e < V; ++e) { ar[e] = src[2 * e]; ai[e] = src[2 * e + 1]; }
        if (volume_dft(ar, ai, br, bi, n0, n1, n2, sign)) { free(buf); return -1; }
        for (long e = 0; e < V; ++e) {
            dst[2 * e] = (double)(br[e] * scale);
            dst[2 * e + 1] = (double)(bi[e] * scale);
        }
    }
    free(buf);
    return 0;
}

/* Same transform for complex64 input; the result is returned in complex128 so that a
 * single-precision kernel can be checked against the exact answer, not a rounded one. */
int oracle_dft3d_c64(const float *in, double *out, long batch,
                     int n0, int n1, int n2, int sign, int normalize)
{
    if (batch < 0 || n0 <= 0 || n1 <= 0 || n2 <= 0 || (sign != -1 && sign != 1)) return -1;
    long V = (long)n0 * n1 * n2;
    ld *buf = malloc(sizeof(ld) * 4 * (size_t)V);
    if (!buf) return -1;
    ld *ar = buf, *ai = buf + V, *br = buf + 2 * V, *bi = buf + 3 * V;
    ld scale = normalize ? 1.0L / (ld)V : 1.0L;
    for (long b = 0; b < batch; ++b) {
        const float *src = in + 2 * b * V;
        double *dst = out + 2 * b * V;
        for (long e = 0; e < V; ++e) { ar[e] = src[2 * e]; ai[e] = src[2 * e + 1]; }
        if (volume_dft(ar, ai, br, bi, n0, n1, n2, sign)) { free(buf); return -1; }
        for (long e = 0; e < V; ++e) {
            dst[2 * e] = (double)(br[e] * scale);
            dst[2 * e + 1] = (double)(bi[e] * scale);
        }
    }
    free(buf);
    return 0;
}

/* Rader-17 plan constants, restated from bench/geom/strategies/L17_rader.md:21-55 so the
 * product's plan step can be pinned against an independent computation:
 *   perm[u] = 3^u mod 17, u in [0,16)
 *   b'[q]   = W17^(3^-q mod 17),  W17 = exp(sign 2 pi i / 17)
 *   Bhat    = DFT16(b') / 16      (DFT16 always with exp(-2 pi i / 16))
 * Output: perm[16] ints, bhat[32] doubles (interleaved re,im). */
void oracle_rader17(int sign, int *perm, double *bhat)
{
    const ld two_pi = 6.283185307179586476925286766559005768L;
    int p = 1;
    for (int u = 0; u < 16; ++u) { perm[u] = p; p = (p * 3) % 17; }
    /* 3^-1 mod 17 = 6 */
    ld br[16], bi[16];
    int q_pow = 1;
    for (int q = 0; q < 16; ++q) {
        ld a = two_pi * (ld)q_pow / 17.0L;
        br[q] = cosl(a);
        bi[q] = (ld)sign * sinl(a);
        q_pow = (q_pow * 6) % 17;
    }
    for (int s = 0; s < 16; ++s) {
        ld ar = 0.0L, ai = 0.0L;
        for (int q = 0; q < 16; ++q) {
            int m = (s * q) % 16;
            ld c = cosl(two_pi * (ld)m / 16.0L), sn = -sinl(two_pi * (ld)m / 16.0L);
            ar += br[q] * c - bi[q] * sn;
            ai += br[q] * sn + bi[q] * c;
        }
        bhat[2 * s] = (double)(ar / 16.0L);
        bhat[2 * s + 1] = (double)(ai / 16.0L);
    }
}
