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
    q = copysignf(fmaf(r, y, q), a);
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


SRC_OMEGA = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t nxt(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static float markstein(float a, float b, float y) { float q = a * y; float r = fmaf(-q, b, a); return fmaf(r, y, q); }
/* weights of the moving directions (B_i = w_i for i >= 1), D3Q19 and D3Q27, as binary32 */
static const float W[5] = {(float)(1.0 / 18.0), (float)(1.0 / 36.0), (float)(2.0 / 27.0), (float)(1.0 / 54.0),
                           (float)(1.0 / 216.0)};
int main(void) {
    long n = 0, bad = 0;
    volatile float one = 1.0f;
    for (long i = 0; i < 4000000L; ++i) {
        /* |F| > 1e-4: af in [1e-4, 8], a numerator (F.c)^2 anywhere from 0 and the smallest
           subnormal up to 3 af^2 (F.c can cancel to anything) */
        float af = 1e-4f * (1.0f + (float)((double)(nxt() >> 11) / 9007199254740992.0) * 79999.0f);
        if (!(af > 1e-4f)) continue;
        float af2 = af * af, y = one / af2;
        float a;
        int kind = (int)(nxt() % 4);
        if (kind == 0) a = fb((uint32_t)(nxt() % 0x0D000000u));              /* 0 .. ~2^-101, subnormals included */
        else if (kind == 1) a = fb((uint32_t)(nxt() & 0x007FFFFFu) | ((uint32_t)(nxt() % 40) << 23));
        else { float fc = af * (float)((double)(nxt() >> 11) / 9007199254740992.0 * 1.7320508); a = fc * fc; }
        float ha = 0.5f * 0.0212f * af;
        for (int k = 0; k < 5; ++k) {
            float c2i = a / af2, c2m = markstein(a, af2, y);
            float omi = ha * (W[k] * c2i - W[k]), omm = ha * (W[k] * c2m - W[k]);
            ++n; bad += bf(omi) != bf(omm);
        }
    }
    printf("%ld %ld\n", n, bad);
    return 0;
}
"""


def test_markstein_quotient_gives_the_same_omega2():
    """div_markstein (march_core.hpp) skips div_shared's tiny-numerator and zero corrections for the
    Omega2 term c2 = (F.c)^2 / |F|^2 of sources with |F| > 1e-4: the resulting
    om = (A|F|/2)(w c2 - B) must be the IEEE-division result bit for bit, including numerators
    that are zero, subnormal or far below 2^-100 (F.c cancelling)."""
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, "om.c")
        with open(c, "w") as f:
            f.write(SRC_OMEGA)
        exe = os.path.join(tmp, "om")
        subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", c, "-o", exe, "-lm"], check=True)
        n, bad = (int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split())
    assert n > 19_000_000
    assert bad == 0
