"""The shared-reciprocal division of the D3Q19 interface form (div_shared in march_core.hpp):
q = RN(a y), r = fma(-q, b, a), q' = fma(r, y, q) with y = RN(1/b) must equal the IEEE quotient
a / b bit for bit wherever the kernel uses it (numerators at or above 2^-100, zero numerators
passed through). Checked here on the host in binary32 (gcc, hardware fma, no contraction) against
IEEE division: every significand of b with random numerators over 2^-27..2^29, plus random pairs
in the ranges of the colour-gradient terms (b = |F| or |F|^2, |a| <= 3 b)."""
import os
import subprocess
import tempfile

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint64_t s = 88172645463325252ull;
static uint64_t nxt(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static float shared_div(float a, float b, float y) {
    float q = a * y;
    float r = fmaf(-q, b, a);
    q = fmaf(r, y, q);
    if (a == 0.0f && y == y) q = a;
    if (fabsf(a) < 0x1p-100f && a != 0.0f) q = a / b;
    return q;
}
int main(void) {
    long bad = 0, n = 0;
    volatile float one = 1.0f;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        float b = fb(0x3F800000u | m), y = one / b;
        for (int k = 0; k < 3; ++k) {
            uint32_t ua = (uint32_t)(nxt() & 0x007FFFFFu) | ((uint32_t)(100 + (nxt() % 56)) << 23);
            if (nxt() & 1) ua |= 0x80000000u;
            float a = fb(ua);
            ++n; bad += bf(a / b) != bf(shared_div(a, b, y));
        }
    }
    for (long i = 0; i < 20000000L; ++i) {
        float b = fb((uint32_t)(nxt() & 0x007FFFFFu) | ((uint32_t)(97 + (nxt() % 34)) << 23)), y = one / b;
        float a = b * (float)((double)(nxt() >> 11) / 9007199254740992.0 * 3.0);
        if (nxt() & 1) a = -a;
        ++n; bad += bf(a / b) != bf(shared_div(a, b, y));
    }
    /* zero and tiny numerators, including -0 */
    float zs[4] = {0.0f, -0.0f, 0x1p-120f, -0x1p-140f};
    for (int k = 0; k < 4; ++k) { float b = 0.37f, y = one / b; ++n; bad += bf(zs[k] / b) != bf(shared_div(zs[k], b, y)); }
    printf("%ld %ld\n", n, bad);
    return 0;
}
"""


def test_shared_reciprocal_division_is_correctly_rounded():
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, "div.c")
        with open(c, "w") as f:
            f.write(SRC)
        exe = os.path.join(tmp, "div")
        subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", c, "-o", exe, "-lm"], check=True)
        n, bad = (int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split())
    assert n > 45_000_000
    assert bad == 0
