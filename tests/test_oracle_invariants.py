"""Closed-form and invariant pins for the oracle paths the reference's golden vectors do not cover:
the FP64 instantiation, the build-defined body force, the conservative Omega2 option, the wall
planes, the zero-gradient outlet and the inlet perturbation. Each check is chosen so a plausible
mistake (dropped term, wrong sign or index, transposed axis) fails it."""
import numpy as np
import pytest


def _periodic(oracle, n=12, q=19, **kw):
    return oracle.Params(nx=n, ny=n, nz=n, q=q, ybc=oracle.Y_PERIODIC, **kw)


def _random_two_phase(p, rng, dtype=np.float64, u=0.04, pi=1e-3):
    A = np.zeros(p.shape, dtype=dtype)
    z, y, x = np.meshgrid(np.arange(p.nz), np.arange(p.nyt), np.arange(p.nx), indexing="ij")
    phi = np.tanh(1.2 * (np.sin(2 * np.pi * x / p.nx) + np.cos(2 * np.pi * (y + 2 * z) / p.ny)))
    rho = 1.0 + 0.05 * rng.standard_normal(phi.shape)
    A[0] = 0.5 * (1 + phi) * rho
    A[1] = 0.5 * (1 - phi) * rho
    A[2:5] = u * rng.uniform(-1, 1, (3,) + phi.shape)
    A[5:] = pi * rng.uniform(-1, 1, (6,) + phi.shape)
    return A


@pytest.mark.parametrize("q", [19, 27])
def test_uniform_flow_is_fixed_point(oracle, q):
    """feq reproduces rho and rho*u, fneqr has no mass/momentum and reproduces Pi: a uniform
    single-phase flow with Pi = 0 is unchanged by a step, for any tau."""
    p = _periodic(oracle, 6, q, tau=0.63)
    A = np.zeros(p.shape)
    A[0] = 1.07
    A[2], A[3], A[4] = 0.031, -0.017, 0.008
    B = oracle.step(p, A)
    np.testing.assert_allclose(B[:5], A[:5], rtol=0, atol=1e-14)
    np.testing.assert_allclose(B[5:], 0.0, atol=1e-14)


@pytest.mark.parametrize("q", [19, 27])
def test_stress_relaxes_by_one_minus_omega(oracle, q):
    """Uniform state with u = 0 and a stress Pi: the regularized collision relaxes Pi by
    (1 - 1/tau) per step (Sum fneqr H2 = Pi, third-order part vanishes at u = 0)."""
    tau = 0.7
    p = _periodic(oracle, 5, q, tau=tau)
    A = np.zeros(p.shape)
    A[0] = 1.0
    A[5:] = np.array([3e-3, -1e-3, 2e-3, 5e-4, -2.5e-3, 1.5e-3])[:, None, None, None]
    B = oracle.step(p, A)
    tau_f = float(np.float32(tau))  # parameters cross the C ABI as binary32
    np.testing.assert_allclose(B[5:], (1 - 1 / tau_f) * A[5:], rtol=1e-12, atol=1e-17)
    np.testing.assert_allclose(B[2:5], 0.0, atol=1e-16)


@pytest.mark.parametrize("q", [19, 27])
def test_single_phase_conserves_mass_and_momentum(oracle, q):
    rng = np.random.default_rng(3)
    p = _periodic(oracle, 10, q, tau=0.9)
    A = _random_two_phase(p, rng)
    A[0] += A[1]
    A[1] = 0.0
    B = oracle.step(p, A)
    m0, m1 = A[0].sum(), B[0].sum()
    P0 = (A[0] * A[2:5]).sum(axis=(1, 2, 3))
    P1 = (B[0] * B[2:5]).sum(axis=(1, 2, 3))
    assert abs(m1 - m0) < 1e-12 * m0
    np.testing.assert_allclose(P1, P0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("q", [19, 27])
def test_conservative_variant_conserves_each_phase(oracle, q):
    """With Omega2 skipped where |F| <= 1e-4 every remaining operator is mass-neutral per
    phase (Sum Omega2 = 0 for |F| > 1e-4, Sum Delta = 0), so both phase masses are conserved."""
    rng = np.random.default_rng(5)
    p = _periodic(oracle, 10, q, conservative_omega2=1)
    A = _random_two_phase(p, rng)
    A[:2, :, :, :3] = 0.5  # psi = 0 block: colour gradient below threshold
    A[:2, :, :, :3] *= 1.0
    B = oracle.step(p, A)
    for f in (0, 1):
        assert abs(B[f].sum() - A[f].sum()) < 1e-12 * A[f].sum()
    P0 = ((A[0] + A[1]) * A[2:5]).sum(axis=(1, 2, 3))
    P1 = ((B[0] + B[1]) * B[2:5]).sum(axis=(1, 2, 3))
    np.testing.assert_allclose(P1, P0, rtol=0, atol=1e-12)


def _colour_gradient(p, A):
    """Independent numpy evaluation of F = (1/cs2) sum_i w_i psi(x + c_i) c_i on a periodic box."""
    c = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1],
                  [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1],
                  [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1]])
    c = c[:, : p.q]
    s = (c ** 2).sum(0)
    w = {19: {0: 1 / 3, 1: 1 / 18, 2: 1 / 36}, 27: {0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}}[p.q]
    psi = (A[0] - A[1]) / (A[0] + A[1])
    F = np.zeros((3,) + psi.shape)
    for i in range(1, p.q):
        shifted = np.roll(psi, shift=(-c[2, i], -c[1, i], -c[0, i]), axis=(0, 1, 2))
        for a in range(3):
            F[a] += w[s[i]] * shifted * c[a, i]
    return 3.0 * F


@pytest.mark.parametrize("q", [19, 27])
def test_omega2_leak_closed_form(oracle, q):
    """Reference semantics: where the interface indicator exceeds 1e-4 but |F| <= 1e-4 the
    cos^2 term is zeroed while -B_i is still applied, so each such source loses
    (A/2)|F| * Sum_i B_i = A|F|/6 of mass (Sum B_i = 1/3 on both lattices). The total mass
    difference between the reference and conservative variants must equal that sum."""
    rng = np.random.default_rng(7)
    p = _periodic(oracle, 10, q, tau=0.8, sigma=0.02)
    A = _random_two_phase(p, rng)
    # Near-uniform psi block so |F| falls below 1e-4 while the indicator stays ~1.
    A[0, :, :, :4] = 0.5 + 1e-6 * rng.standard_normal(A[0, :, :, :4].shape)
    A[1, :, :, :4] = 0.5
    B_ref = oracle.step(p, A)
    B_con = oracle.step(oracle.Params(**{**p.__dict__, "conservative_omega2": 1}), A)
    F = _colour_gradient(p, A)
    absF = np.sqrt((F ** 2).sum(0))
    ind = 4 * A[0] * A[1] / (A[0] + A[1]) ** 2
    sel = (ind > 1e-4) & (absF <= 1e-4)
    assert sel.sum() > 20
    Acoef = float(np.float32(p.sigma)) / (4 / 9) / float(np.float32(p.tau))
    leak = (Acoef * absF[sel] / 6).sum()
    dm = (B_con[0] + B_con[1]).sum() - (B_ref[0] + B_ref[1]).sum()
    assert leak > 0
    np.testing.assert_allclose(dm, leak, rtol=1e-6)


@pytest.mark.parametrize("q", [19, 27])
def test_axis_permutation_symmetry(oracle, q):
    """The lattices are invariant under the cubic group, so swapping x and y in the input swaps
    them (and the matching velocity/stress components) in the output, to FP64 roundoff."""
    rng = np.random.default_rng(11)
    p = _periodic(oracle, 10, q)
    A = _random_two_phase(p, rng)
    B = oracle.step(p, A)
    perm = [0, 1, 3, 2, 4, 7, 6, 5, 10, 9, 8]  # rr rb uy ux uz Pyy Pxy Pxx Pxz Pzz Pyz
    As = np.ascontiguousarray(A[perm].transpose(0, 1, 3, 2))
    Bs = oracle.step(p, As)
    np.testing.assert_allclose(Bs, B[perm].transpose(0, 1, 3, 2), rtol=0, atol=2e-13)


@pytest.mark.parametrize("q", [19, 27])
def test_reflection_symmetry(oracle, q):
    """x -> -x flips ux, Pixy, Pixz and mirrors the field."""
    rng = np.random.default_rng(13)
    p = _periodic(oracle, 10, q)
    A = _random_two_phase(p, rng)
    B = oracle.step(p, A)
    sign = np.array([1, 1, -1, 1, 1, 1, -1, 1, 1, 1, -1], dtype=float)[:, None, None, None]

    def mirror(S):
        return np.ascontiguousarray(sign * np.roll(S[..., ::-1], 1, axis=3))

    np.testing.assert_allclose(oracle.step(p, mirror(A)), mirror(B), rtol=0, atol=2e-13)


@pytest.mark.parametrize("q", [19, 27])
def test_fp64_and_fp32_paths_agree(oracle, q):
    rng = np.random.default_rng(17)
    p = _periodic(oracle, 10, q)
    A = _random_two_phase(p, rng)
    B64 = oracle.step(p, A)
    B32 = oracle.step(p, A.astype(np.float32))
    for f in range(11):
        err = np.linalg.norm(B32[f] - B64[f]) / max(np.linalg.norm(B64[f]), 1e-30)
        assert err < (2e-6 if f < 5 else 2e-3), (f, err)


@pytest.mark.parametrize("q", [19, 27])
def test_body_force_uniform_acceleration(oracle, q):
    """Single phase at rest, force F_y = -g*At*(rb - rr): velocity grows by F/rho per step and
    the stress stays at its O(u^2) value (Guo scheme with the half-force velocity shift)."""
    g, At = float(np.float32(1e-4)), 0.5
    p = _periodic(oracle, 6, q, tau=0.7, gravity=g, atwood=At)
    A = np.zeros(p.shape)
    A[1] = 1.0  # blue everywhere
    Fy = -g * At * 1.0
    a = A.copy()
    for k in range(5):
        a = oracle.step(p, a, step_index=k + 1)
        np.testing.assert_allclose(a[3], (k + 1) * Fy, rtol=1e-12)
        np.testing.assert_allclose(a[1], 1.0, rtol=0, atol=1e-14)
        np.testing.assert_allclose(a[2], 0.0, atol=1e-16)
        np.testing.assert_allclose(a[4], 0.0, atol=1e-16)


@pytest.mark.parametrize("q", [19, 27])
def test_body_force_momentum_balance(oracle, q):
    """Periodic two-phase box: total momentum changes by (Sum F_old + Sum F_new)/2 per step."""
    rng = np.random.default_rng(19)
    g, At = float(np.float32(2e-4)), 0.5
    p = _periodic(oracle, 10, q, gravity=g, atwood=At, conservative_omega2=1)
    A = _random_two_phase(p, rng)
    B = oracle.step(p, A)
    F0 = (-g * At * (A[1] - A[0])).sum()
    F1 = (-g * At * (B[1] - B[0])).sum()
    P0 = ((A[0] + A[1]) * A[3]).sum()
    P1 = ((B[0] + B[1]) * B[3]).sum()
    np.testing.assert_allclose(P1 - P0, 0.5 * (F0 + F1), rtol=1e-9)
    for f in (2, 4):
        np.testing.assert_allclose(((B[0] + B[1]) * B[f]).sum(), ((A[0] + A[1]) * A[f]).sum(), atol=1e-12)


def test_zero_force_path_is_unforced_path(oracle):
    rng = np.random.default_rng(23)
    p = _periodic(oracle, 8, 19)
    A = _random_two_phase(p, rng).astype(np.float32)
    pf = oracle.Params(**{**p.__dict__, "gravity": 0.0, "atwood": 0.5})
    np.testing.assert_array_equal(oracle.step(pf, A), oracle.step(p, A))


def test_wall_planes_copy_and_stop(oracle):
    p = oracle.Params(nx=8, ny=10, nz=6, q=19, ybc=oracle.Y_WALLS)
    rng = np.random.default_rng(29)
    A = _random_two_phase(p, rng).astype(np.float32)
    oracle.apply_ybc(p, A, A)
    B = oracle.step(p, A)
    for yg, yf in ((0, 1), (p.ny + 1, p.ny)):
        for f in (0, 1, 5, 6, 7, 8, 9, 10):
            np.testing.assert_array_equal(B[f, :, yg], B[f, :, yf])
        np.testing.assert_array_equal(B[2:5, :, yg], 0.0)


def test_zero_gradient_outlet(oracle):
    p = oracle.Params(nx=8, ny=10, nz=8, q=19, ybc=oracle.Y_JET, outlet=oracle.OUTLET_ZERO_GRADIENT,
                      jet_radius=2.0)
    rng = np.random.default_rng(31)
    A = _random_two_phase(p, rng).astype(np.float32)
    B = oracle.step(p, A)
    top, F = p.ny + 1, p.ny
    for f in range(5):
        old = A[f, :, top].astype(np.float32)
        want = old - np.float32(1.0) * (old - B[f, :, F])
        np.testing.assert_array_equal(B[f, :, top], want)
    np.testing.assert_array_equal(B[5:, :, top], B[5:, :, F])


def test_inlet_perturbation_bounds(oracle):
    U, eps = 0.05, 0.01
    p = oracle.Params(nx=16, ny=8, nz=16, q=27, ybc=oracle.Y_JET, jet_radius=5.0, jet_velocity=U,
                      jet_noise=eps, seed=3)
    A = oracle.init_jet(p)
    B = oracle.step(p, A, step_index=7)
    z, x = np.meshgrid(np.arange(p.nz), np.arange(p.nx), indexing="ij")
    m = ((x - p.jet_x0) ** 2 + (z - p.jet_z0) ** 2) <= p.jet_radius ** 2
    uy = B[3, :, 0]
    assert np.all(np.abs(uy[m] - U) <= U * eps * 1.0001) and np.all(uy[~m] == 0)
    assert np.std(uy[m]) > 0.1 * U * eps
    assert np.all(np.abs(B[2, :, 0][m]) <= U * eps * 1.0001)
    C = oracle.step(p, A, step_index=8)
    assert not np.array_equal(C[3, :, 0], uy)  # keyed by step


def _interface_height(A, width):
    """Interface height y_i(x, z) of an RTI state from its phase field: in each column, at the
    stored row where |rb - rr| is smallest, phi = tanh((y - y_i)/w) with phi = rb - rr gives
    y_i = y - w atanh(phi)."""
    phi = (A[1].astype(np.float64) - A[0].astype(np.float64))  # (nz, nyt, nx)
    row = np.argmin(np.abs(phi), axis=1)  # (nz, nx)
    ph = np.take_along_axis(phi, row[:, None, :], axis=1)[:, 0, :]
    return row - width * np.arctanh(ph)


def test_rti_random_modes_spectrum(oracle):
    """The RTI initialiser's seeded random modes (an addition of this build; the reference's
    initialiser has the single cosine mode only): the interface height they add is a sum of
    n_random_modes cosines of the given amplitude, with wave numbers 1..8 along x and along z and
    uniform phases. Read back from the phase field, independently of how the initialiser builds it:
    the added height has zero mean, its Fourier spectrum lives on the (kx, kz) in 1..8 (and their
    conjugates) only, each isolated mode has the stated amplitude, the phases spread over the
    circle and the wave numbers over 1..8; one seed reproduces its field, another changes it."""
    n, width, ra = 32, 4.0, 0.2
    p = oracle.Params(nx=n, ny=48, nz=n, q=19, ybc=oracle.Y_WALLS, tau=0.6, sigma=0.002, beta=0.7)
    base = _interface_height(oracle.init_rti(p, 2.0, width, 0, 0.0, 0), width)
    amps, phases, waves = [], [], []
    for seed in range(40):
        A = oracle.init_rti(p, 2.0, width, 8, ra, seed)
        d = _interface_height(A, width) - base  # (nz, nx)
        assert abs(d.mean()) < 1e-4
        F = np.fft.fft2(d) / d.size  # a cos(2 pi (m x + k z)/n + ph) -> a/2 at (k, m) and (-k, -m)
        amp = 2.0 * np.abs(F)
        big = np.argwhere(amp > 0.02)
        for kz, kx in big:
            kz_s = kz if kz <= n // 2 else kz - n
            kx_s = kx if kx <= n // 2 else kx - n
            # one quadrant pair only: both wave numbers positive in 1..8, or both negative
            assert (1 <= kx_s <= 8 and 1 <= kz_s <= 8) or (-8 <= kx_s <= -1 and -8 <= kz_s <= -1), (seed, kx_s, kz_s)
        pos = [(kz, kx) for kz, kx in big if 1 <= kx <= 8 and 1 <= kz <= 8]
        assert 1 <= len(pos) <= 8
        assert sum(amp[kz, kx] for kz, kx in pos) <= 8 * ra + 1e-3
        if len(pos) == 8:  # no two draws share a wave vector: every mode stands alone
            for kz, kx in pos:
                assert abs(amp[kz, kx] - ra) < 2e-3, (seed, kx, kz, amp[kz, kx])
                amps.append(amp[kz, kx])
                phases.append(np.angle(F[kz, kx]))
                waves += [kx, kz]
        assert np.array_equal(A, oracle.init_rti(p, 2.0, width, 8, ra, seed))
    assert len(amps) >= 80  # enough isolated modes among the 40 seeds to judge their spread
    # phases uniform on the circle: every quarter holds a fair share
    q = np.histogram(np.mod(phases, 2 * np.pi), bins=4, range=(0, 2 * np.pi))[0]
    assert q.min() > 0.12 * len(phases), q
    # wave numbers cover 1..8
    w = np.bincount(np.asarray(waves), minlength=9)[1:]
    assert w.min() > 0.04 * len(waves), w
    assert not np.array_equal(oracle.init_rti(p, 2.0, width, 8, ra, 1), oracle.init_rti(p, 2.0, width, 8, ra, 2))
