/* The substitution's division a / b from y = RN(1/b) by Markstein's
 * correction (q = RN(a y); r = a - b q by one FMA; RN(q + r y)), restated
 * with the host's IEEE operations, against a / b itself, bit for bit, over
 * random and hand-picked operands inside the range the device code takes the
 * method for ([2^-400, 2^400] double, [2^-50, 2^50] float). Prints the
 * number of pairs tried and of mismatches. Build: cc -O2 -ffp-contract=off. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint64_t s = 88172645463325252ull;
static uint64_t next(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }

static double dk(double a, double b, double y) { double q = a * y; return fma(fma(-q, b, a), y, q); }
static float fk(float a, float b, float y) { float q = a * y; return fmaf(fmaf(-q, b, a), y, q); }

static double rd(int emax) {
    uint64_t m = next() & 0xFFFFFFFFFFFFFull, e = (uint64_t)(1023 + (int)(next() % (2 * emax + 1)) - emax);
    uint64_t sg = next() & 1, bits = (sg << 63) | (e << 52) | m;
    switch (next() % 8) {  /* mantissas at the edges: 1, all ones, one bit */
    case 0: bits &= ~0xFFFFFFFFFFFFFull; break;
    case 1: bits |= 0xFFFFFFFFFFFFFull; break;
    case 2: bits = (bits & ~0xFFFFFFFFFFFFFull) | 1; break;
    default: break;
    }
    double v; memcpy(&v, &bits, 8); return v;
}
static float rf(int emax) {
    uint32_t m = (uint32_t)next() & 0x7FFFFF, e = (uint32_t)(127 + (int)(next() % (2 * emax + 1)) - emax);
    uint32_t sg = (uint32_t)next() & 1, bits = (sg << 31) | (e << 23) | m;
    switch (next() % 8) {
    case 0: bits &= ~0x7FFFFFu; break;
    case 1: bits |= 0x7FFFFFu; break;
    case 2: bits = (bits & ~0x7FFFFFu) | 1; break;
    default: break;
    }
    float v; memcpy(&v, &bits, 4); return v;
}

int main(void) {
    long tried = 0, bad = 0;
    for (long i = 0; i < 4000000; i++) {
        double a = rd(399), b = rd(399), y = 1.0 / b, q = a / b, m = dk(a, b, y);
        if (memcmp(&q, &m, 8) != 0) bad++;
        tried++;
        /* a close to a multiple of b: quotients near representable values */
        double c = b * (double)(next() % 4096 + 1), qc = c / b, mc = dk(c, b, y);
        if (fabs(c) >= 0x1p-400 && fabs(c) <= 0x1p400) { if (memcmp(&qc, &mc, 8) != 0) bad++; tried++; }
    }
    for (long i = 0; i < 4000000; i++) {
        float a = rf(49), b = rf(49), y = 1.0f / b, q = a / b, m = fk(a, b, y);
        if (memcmp(&q, &m, 4) != 0) bad++;
        tried++;
    }
    printf("%ld %ld\n", tried, bad);
    return 0;
}
