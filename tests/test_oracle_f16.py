"""The oracle's model of FP16-deviation storage (cgref.quantize_f16 / run_f16), pinned on the CPU.

Storage keeps rhor/rhob in binary32 and the nine other moments as binary16 of value * 2^10 (u) or
value * 2^12 (Pi), rounded to nearest even. Pinned here by hand-picked binary16 values (ties to
even, the subnormal range and overflow), by the bounds that rounding to 11 significant bits
implies, and by the physics the storage must not disturb (per-phase mass, the FP64 distance).
"""
import numpy as np
import pytest

pytestmark = []


def _state(vals, f):
    A = np.zeros((11, 1, 1, len(vals)), dtype=np.float32)
    A[f, 0, 0, :] = vals
    return A


def test_known_values_per_field_scale(oracle):
    q = oracle.quantize_f16
    # u fields: scale 2^10. 1/1024 * (1 + 2^-11) is a tie between 1 and 1 + 2^-10 (units of 2^-10):
    # ties go to the even significand (1); 1 + 3 * 2^-11 rounds up to 1 + 2^-9.
    u = np.array([1 / 1024, (1 + 2.0 ** -11) / 1024, (1 + 3 * 2.0 ** -11) / 1024, 0.05, -0.05, 0.0], dtype=np.float32)
    got = q(_state(u, 3))[3, 0, 0]
    want = np.array([1 / 1024, 1 / 1024, (1 + 2.0 ** -9) / 1024, None, None, 0.0], dtype=object)
    assert got[0] == want[0] and got[1] == want[1] and got[2] == want[2] and got[5] == 0.0
    # 0.05 * 1024 = 51.2 -> binary16 spacing at 32..64 is 2^-5: 51.1875 or 51.21875; 51.2 is nearer 51.1875
    assert got[3] == np.float32(51.1875 / 1024) and got[4] == -got[3]
    # Pi fields: scale 2^12; the smallest binary16 subnormal is 2^-24, i.e. 2^-36 in Pi units
    p = np.array([2.0 ** -36, 2.0 ** -38, 3 * 2.0 ** -37, 1e-3, 17.0], dtype=np.float32)
    got = q(_state(p, 7))[7, 0, 0]
    assert got[0] == np.float32(2.0 ** -36)
    assert got[1] == 0.0  # a quarter of the smallest subnormal rounds to zero
    assert got[2] == np.float32(2 * 2.0 ** -36)  # 1.5 subnormal units: tie to even (2)
    assert abs(got[3] - 1e-3) <= 1e-3 * 2.0 ** -11
    assert np.isinf(got[4])  # 17 * 4096 > 65504: beyond the stored range
    # rhor / rhob untouched
    r = np.array([0.123456789, 1.0 - 2.0 ** -24], dtype=np.float32)
    assert np.array_equal(q(_state(r, 0))[0], _state(r, 0)[0])


def test_rounding_bounds_and_idempotence(oracle):
    rng = np.random.default_rng(0)
    A = np.empty((11, 4, 5, 6), dtype=np.float32)
    A[:2] = rng.uniform(0, 1, A[:2].shape)
    A[2:5] = rng.uniform(-0.1, 0.1, A[2:5].shape)
    A[5:] = rng.uniform(-1e-2, 1e-2, A[5:].shape)
    Q = oracle.quantize_f16(A)
    assert np.array_equal(oracle.quantize_f16(Q), Q)
    assert np.array_equal(Q[:2], A[:2])
    for f in range(2, 11):
        s = oracle.F16_SCALE[f]
        normal = np.abs(A[f] * s) >= 2.0 ** -14
        rel = np.abs(Q[f] - A[f])[normal] / np.abs(A[f])[normal]
        assert rel.max() <= 2.0 ** -11
        assert np.abs(Q[f] - A[f])[~normal].max(initial=0.0) <= 2.0 ** -25 / s
    # binary16 has 11 significant bits: every stored value times scale is an integer times 2^(e-10)
    v = (Q[2:5] * 1024).astype(np.float64).ravel()
    v = v[v != 0]
    e = np.floor(np.log2(np.abs(v)))
    assert np.all(np.mod(v / 2.0 ** (np.maximum(e, -14) - 10), 1.0) == 0.0)


def test_f16_run_conserves_phase_mass_and_stays_near_fp64(oracle):
    """FP16 storage leaves rhor/rhob in binary32, so it adds no per-phase mass drift of its own: the
    FP16 run's masses stay within 1e-7 (relative) of the FP32 run's over 100 steps. The FP32 drift
    itself (about 1e-6 per 100 steps here) is rounding in the binary32 update plus, without the
    conservative variant, the reference's |F| <= 1e-4 leak; the FP64 conservative run conserves
    both masses to 1e-12. The state stays within a few binary16 units of the FP64 run."""
    for conservative in (0, 1):
        p = oracle.Params(nx=20, ny=20, nz=20, q=19, tau=0.8, sigma=0.01, beta=0.7,
                          conservative_omega2=conservative)
        A = oracle.init_bubble(p, radius=5.0, width=4.0)
        m0 = oracle.reduce(p, A)
        S16 = oracle.run_f16(p, A, 100)
        S32 = oracle.run(p, A, 100)
        S64 = oracle.run(p, A.astype(np.float64), 100)
        m16, m32 = oracle.reduce(p, S16), oracle.reduce(p, S32)
        assert np.all(np.abs(m16[:2] - m32[:2]) / m0[:2] < 1e-7)
        m64 = np.array([S64[0].sum(), S64[1].sum()])
        if conservative:
            assert np.all(np.abs(m64 / m0[:2] - 1) < 1e-12)
        # phase densities: FP16 storage only enters through u and Pi
        assert np.abs(S16[:2] - S64[:2]).max() < 1e-5
        for f in range(2, 11):
            scale = np.sqrt(np.mean(S64[f] ** 2))
            err = np.sqrt(np.mean((S16[f] - S64[f]) ** 2)) / scale
            assert err < 2e-2, (f, err)
