// Plan-time constant bank, computed on the host in long double (no device calls here).
//   * W_N^m = exp(-2 pi i m/N) for every line length in kLens, with the quarter-turn
//     entries written exactly;
//   * the Rader-17 spectra Bhat = DFT16(b')/16, b'[q] = W17^(3^-q mod 17), for both
//     directions (bench/geom/strategies/L17_rader.md:21-55), with Bhat[0] = -1/16 and
//     Bhat[8] = sqrt(17)/16 written in closed form;
//   * cos and sin of 2 pi m/17 for the folded-dense 17-point line.
#include <cmath>
#include <cstring>
#include <vector>

#include "cx.hpp"
#include "plan_internal.hpp"
#include "registry.hpp"

namespace fx {

typedef long double ld;
static const ld kTwoPi = 6.283185307179586476925286766559005768L;

template <typename T>
static void build_bank(Bank<T>& b)
{
    std::memset(&b, 0, sizeof b);
    for (int k = 0; k < kNumLens; ++k) {
        const int n = kLens[k], off = woff(n);
        for (int m = 0; m < n; ++m) {
            ld c, s;
            if (m == 0) { c = 1; s = 0; }
            else if (4 * m == n) { c = 0; s = -1; }
            else if (2 * m == n) { c = -1; s = 0; }
            else if (4 * m == 3 * n) { c = 0; s = 1; }
            else { const ld a = kTwoPi * (ld)m / (ld)n; c = cosl(a); s = -sinl(a); }
            b.w[off + m].r = (T)c;
            b.w[off + m].i = (T)s;
        }
    }
    // Rader-17 spectra.  perm[u] = 3^u mod 17; b'[q] = W17^(3^-q mod 17), 3^-1 = 6 mod 17.
    for (int d = 0; d < 2; ++d) {
        const ld sign = d == 0 ? -1.0L : 1.0L;
        ld br[16], bi[16];
        int p = 1;
        for (int q = 0; q < 16; ++q) {
            const ld a = kTwoPi * (ld)p / 17.0L;
            br[q] = cosl(a);
            bi[q] = sign * sinl(a);
            p = (p * 6) % 17;
        }
        for (int s = 0; s < 16; ++s) {
            ld ar = 0, ai = 0;
            for (int q = 0; q < 16; ++q) {
                const int m = (s * q) % 16;
                const ld c = cosl(kTwoPi * (ld)m / 16.0L), sn = -sinl(kTwoPi * (ld)m / 16.0L);
                ar += br[q] * c - bi[q] * sn;
                ai += br[q] * sn + bi[q] * c;
            }
            b.rader17[d][s].r = (T)(ar / 16.0L);
            b.rader17[d][s].i = (T)(ai / 16.0L);
        }
        // Closed forms: the 16 nontrivial 17th roots sum to -1; index 8 is the quadratic
        // Gauss sum, real and equal to sqrt(17) for p = 1 mod 4, in either direction.
        b.rader17[d][0].r = (T)(-1.0L / 16.0L);
        b.rader17[d][0].i = (T)0;
        b.rader17[d][8].r = (T)(sqrtl(17.0L) / 16.0L);
        b.rader17[d][8].i = (T)0;
    }
    for (int m = 0; m < 17; ++m) {
        const ld a = kTwoPi * (ld)m / 17.0L;
        b.cos17[m] = (T)cosl(a);
        b.sin17[m] = (T)sinl(a);
    }
}

void bank_host_image(int prec, std::vector<unsigned char>& img)
{
    if (prec == kC128) {
        Bank<double> b;
        build_bank(b);
        img.assign((unsigned char*)&b, (unsigned char*)&b + sizeof b);
    } else {
        Bank<float> b;
        build_bank(b);
        img.assign((unsigned char*)&b, (unsigned char*)&b + sizeof b);
    }
}

size_t elem_bytes(int prec) { return prec == kC128 ? 16 : 8; }

}  // namespace fx
