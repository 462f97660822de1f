"""guo_y (step_common.hpp), the body-force term the marching kernels evaluate for gravity along y,
against guo_full(0, Fy, 0, u) (the general Guo term of the gather kernel and of the oracle): bitwise
equal for every direction of both stencils on random inputs (zeros of both signs, magnitudes
2^-27..2^3 of either sign), apart from the sign of an all-zero result, which the step adds to a
positive equilibrium. guo_y_pair (the pair form of the own sources) must give guo_y of direction I
and of its opposite I + 1 with the same tolerance for zero signs. Compiled for the host with hipcc
(the functions are host and device code), no contraction."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lightweight-mb-color-gradient-lbm_amd", "csrc")

SRC = r"""
#include "step_common.hpp"
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <utility>
using namespace cglbm;
static uint64_t s = 0x2545F4914F6CDD1Dull;
static uint64_t nxt() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static long total = 0;
template <int Q, int I> long check(long n) {
    long bad = 0;
    for (long k = 0; k < n; ++k) {
        float v[4];
        for (int j = 0; j < 4; ++j) {
            const int kind = (int)(nxt() % 8);
            if (kind == 0) v[j] = 0.0f;
            else if (kind == 1) v[j] = -0.0f;
            else v[j] = fb((uint32_t)(nxt() & 0x807FFFFFu) | ((uint32_t)(100 + nxt() % 31) << 23));
        }
        const float a = guo_full<Q, I>(0.0f, v[0], 0.0f, v[1], v[2], v[3]);
        const float b = guo_y<Q, I>(v[0], v[1], v[2], v[3]);
        ++total;
        if (bf(a) != bf(b) && !(a == 0.0f && b == 0.0f)) ++bad;
        if constexpr (I % 2 == 1) {
            float fi, fo;
            guo_y_pair<Q, I>(v[0], v[1], v[2], v[3], fi, fo);
            const float o = guo_y<Q, I + 1>(v[0], v[1], v[2], v[3]);
            total += 2;
            if (bf(fi) != bf(b) && !(fi == 0.0f && b == 0.0f)) ++bad;
            if (bf(fo) != bf(o) && !(fo == 0.0f && o == 0.0f)) ++bad;
        }
    }
    return bad;
}
template <int Q, int... I> long all(std::integer_sequence<int, I...>, long n) { return (check<Q, I>(n) + ...); }
int main() {
    const long bad = all<19>(std::make_integer_sequence<int, 19>{}, 100000) +
                     all<27>(std::make_integer_sequence<int, 27>{}, 100000);
    printf("%ld %ld\n", total, bad);
    return 0;
}
"""


def test_guo_y_equals_guo_full_with_zero_x_and_z_force():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "guo.hip")
        with open(src, "w") as f:
            f.write(SRC)
        exe = os.path.join(tmp, "guo")
        subprocess.run([hipcc, "-O1", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", f"-I{CSRC}",
                        f"-I{os.path.join(ROOT, 'include')}", "-x", "hip", "--cuda-host-only", src, "-o", exe],
                       check=True, capture_output=True)
        total, bad = (int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                                     check=True).stdout.split())
    assert total == (46 + 2 * 22) * 100000
    assert bad == 0
