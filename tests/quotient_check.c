/*
 * CPU check of the quotient identity the ingest normalisation kernel uses
 * in place of an IEEE division (FP32, round to nearest):
 *
 *   y = RN(1/b), q = RN(a*y), q' = RN(q + RN(a - b*q) * y)  ==  RN(a/b)
 *
 * for b >= 1 when |a| >= 2^-100, |q| >= 2^-100, a finite and b finite
 * (otherwise the kernel divides).  Counts mismatches against a / b over
 * random and structured operands; fmaf is the C library's fused multiply-add.
 * Built and run by tests/test_quotient.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint32_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st >> 11); }
static float bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static float quotient(float a, float b, float y) {
    float q = a * y;
    if (fabsf(a) >= 0x1p-100f && fabsf(a) <= 3.40282347e38f && fabsf(q) >= 0x1p-100f && b <= 3.40282347e38f)
        return fmaf(fmaf(-b, q, a), y, q);
    return a / b;
}

static long long bad, total;
static void one(float a, float b) {
    volatile float y = 1.0f / b;  /* RN(1/b) */
    float q1 = quotient(a, b, y), q = a / b;
    ++total;
    if (memcmp(&q1, &q, 4) != 0 && !(isnan(q) && isnan(q1))) {
        if (bad < 5) printf("mismatch a=%a b=%a: %a vs %a\n", a, b, q1, q);
        ++bad;
    }
}

int main(int argc, char **argv) {
    const long long n = argc > 1 ? atoll(argv[1]) : 100000000LL;
    for (long long t = 0; t < n; ++t) {  /* any finite a, b in [1, 2^40) or [1, inf] */
        float a = bits((rnd() & 0x80000000u) | (rnd() % 0x7F800000u));
        float b = (t & 1) ? bits(0x3F800000u + rnd() % (40u << 23)) : bits(0x3F800000u + rnd() % 0x40000001u);
        one(a, b);
    }
    for (uint32_t e = 0; e < 24; ++e)  /* significands near all ones, integer counts */
        for (uint32_t lo = 0; lo < 512; ++lo)
            for (uint32_t k = 0; k < 65536; k += 3) one((float)k, bits(((127u + e) << 23) | (0x7FFFFFu - lo)));
    for (uint32_t bi = 1; bi < 65536; bi += 5)  /* counts over counts */
        for (uint32_t k = 0; k < 65536; k += 11) one((float)k, (float)bi);
    one(INFINITY, 3.0f); one(NAN, 3.0f); one(-0.0f, 2.0f); one(1e-45f, 1.0f); one(5.0f, INFINITY);
    printf("checked %lld quotients, %lld mismatches\n", total, bad);
    return bad != 0;
}
