/* Checks the division sequences the integer-input bandpass kernel uses,
 *   one correction:    q = a*rb; r = fma(-q, b, a); q = fma(r, rb, q)
 *   split reciprocal:  q = fma(a, rh, RN(a*rl)), rh = RN(1/b), rl = RN(1/b - rh)
 * with rb = RN(1/b), against IEEE a / b.  Usage:
 *   div_check <divisor> <n_random> <int_limit> <binade_lo> <binade_hi> [split]
 * tests every integer a in [0, int_limit], every float in the binades
 * [2^binade_lo, 2^binade_hi), and n_random random floats in [2^-90, 2^25).
 * Prints the number of mismatches (0 expected) and exits non-zero on any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float div_const(float a, float b, float rb) {
    const float q = a * rb;
    const float r = fmaf(-q, b, a);
    return fmaf(r, rb, q);
}

static int use_split;
static float rh, rl;

static float quot(float a, float b, float rb) {
    return use_split ? fmaf(a, rh, a * rl) : div_const(a, b, rb);
}

static uint32_t bits(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    return u;
}

int main(int argc, char **argv) {
    if (argc < 6) return 2;
    const float b = strtof(argv[1], NULL);
    const long nrand = atol(argv[2]);
    const long ilim = atol(argv[3]);
    const int blo = atoi(argv[4]), bhi = atoi(argv[5]);
    const float rb = 1.0f / b;
    use_split = argc > 6 && strcmp(argv[6], "split") == 0;
    rh = (float)(1.0 / (double)b);
    rl = (float)(1.0 / (double)b - (double)rh);
    long bad = 0, checked = 0;
    for (long i = 0; i <= ilim; i++) {
        const float a = (float)i;
        bad += bits(quot(a, b, rb)) != bits(a / b);
        checked++;
    }
    for (int e = blo; e < bhi; e++) {
        const float base = ldexpf(1.0f, e);
        for (uint32_t m = 0; m < (1u << 23); m++) {
            float a;
            uint32_t u = bits(base) + m;
            memcpy(&a, &u, 4);
            bad += bits(quot(a, b, rb)) != bits(a / b);
            checked++;
        }
    }
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (long i = 0; i < nrand; i++) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        const int e = -90 + (int)((s >> 33) % 115);
        const uint32_t m = (uint32_t)(s >> 8) & 0x7FFFFFu;
        const float base = ldexpf(1.0f, e);
        float a;
        uint32_t u = bits(base) + m;
        memcpy(&a, &u, 4);
        bad += bits(quot(a, b, rb)) != bits(a / b);
        checked++;
    }
    printf("%ld %ld\n", bad, checked);
    return bad != 0;
}
