"""GPU parity: the HIP update against the reference's golden vectors and the oracle.

The bar is bit identity: every field of every snapshot must carry the sha256 the reference's own
update produced (tests/golden), and at sizes without fixtures the GPU state must equal the
oracle's element for element (np.array_equal). Every path goes through the C ABI."""
import numpy as np
import pytest

import _golden
from conftest import load_product

pytestmark = pytest.mark.gpu

CASES = [n for n in _golden.names() if n != "rti_init"]
KERNELS = ("gather", "auto")


def _cfg(cg, p):
    return cg.Config(**{k: getattr(p, k) for k in p.__dataclass_fields__})


def _kernel(cg, name):
    return {"gather": cg.KERNEL_GATHER, "auto": cg.KERNEL_AUTO, "march": cg.KERNEL_MARCH}[name]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name", CASES)
def test_matches_reference_golden(oracle, name, kernel):
    cg = load_product()
    z, meta = _golden.load(name)
    p, A = _golden.initial_state(oracle, z, meta)
    print(f"case {name} kernel {kernel} shape {p.shape}", flush=True)
    with cg.Domain(_cfg(cg, p), kernel=_kernel(cg, kernel)) as dom:
        dom.set_state(A)
        k = 0
        for s in meta["steps"]:
            dom.step(s - k)
            k = s
            got = dom.get_state()
            shas = _golden.field_shas(got)
            bad = [f for f in range(11) if shas[f] != z[f"sha_{s}"][f]]
            if bad:
                vals = got.reshape(11, -1)[:, z["idx"]]
                diff = np.abs(vals.astype(np.float64) - z[f"val_{s}"]).max(axis=1)
                pytest.fail(f"step {s}: fields {bad} differ from the reference; sampled max |diff| {diff}")


@pytest.mark.parametrize("q", [19, 27])
def test_initialisers_match_oracle(oracle, q):
    cg = load_product()
    p = oracle.Params(nx=20, ny=18, nz=16, q=q, ybc=oracle.Y_JET, jet_radius=4.0)
    with cg.Domain(_cfg(cg, p)) as dom:
        dom.init_jet()
        np.testing.assert_array_equal(dom.get_state(), oracle.init_jet(p))
    p = oracle.Params(nx=20, ny=18, nz=16, q=q)
    with cg.Domain(_cfg(cg, p)) as dom:
        for sharp in (False, True):
            dom.init_bubble(radius=5.0, width=4.0, sharp=sharp)
            np.testing.assert_array_equal(dom.get_state(), oracle.init_bubble(p, 5.0, 4.0, sharp))
    p = oracle.Params(nx=16, ny=24, nz=12, q=q, ybc=oracle.Y_WALLS)
    with cg.Domain(_cfg(cg, p)) as dom:
        dom.init_rti(2.0, 4.0, 8, 0.2, 5)
        np.testing.assert_array_equal(dom.get_state(), oracle.init_rti(p, 2.0, 4.0, 8, 0.2, 5))


CONFIGS = [
    # (label, params, steps)
    ("bubble64_q19", dict(nx=64, ny=64, nz=64, q=19, tau=0.8, sigma=0.01, beta=0.7), 100),
    ("bubble48_q27", dict(nx=48, ny=40, nz=36, q=27, tau=0.8, sigma=0.01, beta=0.7), 40),
    ("jet_noise_q27", dict(nx=24, ny=40, nz=20, q=27, ybc=1, tau=0.55, sigma=0.005, beta=0.7, jet_radius=5.0,
                           jet_noise=0.01, seed=9, outlet=1), 30),
    ("rti_q19", dict(nx=24, ny=48, nz=20, q=19, ybc=2, tau=0.6, sigma=0.002, beta=0.7, gravity=1e-4, atwood=0.5),
     30),
    ("conservative_q19", dict(nx=20, ny=20, nz=20, q=19, conservative_omega2=1, sigma=0.02), 20),
    # a body force on a periodic y axis: the automatic choice takes the gather kernel with forcing
    ("forced_periodic_q19", dict(nx=16, ny=16, nz=16, q=19, tau=0.7, sigma=0.01, beta=0.7, gravity=1e-4,
                                 atwood=0.5), 20),
    ("forced_periodic_q27", dict(nx=16, ny=14, nz=12, q=27, tau=0.7, sigma=0.01, beta=0.7, gravity=1e-4,
                                 atwood=0.5), 20),
]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("label,kw,steps", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_matches_oracle(oracle, label, kw, steps, kernel):
    cg = load_product()
    p = oracle.Params(**kw)
    if p.ybc == oracle.Y_JET:
        A = oracle.init_jet(p)
    elif p.ybc == oracle.Y_WALLS:
        A = oracle.init_rti(p, 2.0, 4.0, 8, 0.2, 1)
    else:
        A = oracle.init_bubble(p, radius=min(p.nx, p.ny, p.nz) / 5.0, width=4.0)
    print(f"config {label} kernel {kernel} steps {steps}", flush=True)
    want = oracle.run(p, A, steps)
    with cg.Domain(_cfg(cg, p), kernel=_kernel(cg, kernel)) as dom:
        dom.set_state(A)
        dom.step(steps)
        got = dom.get_state()
    if not np.array_equal(got, want):
        diff = np.abs(got.astype(np.float64) - want).reshape(11, -1)
        pytest.fail(f"{label}: cells differing per field {(diff > 0).sum(axis=1)}, max {diff.max(axis=1)}")


def test_reduce_matches_oracle(oracle):
    cg = load_product()
    p = oracle.Params(nx=40, ny=36, nz=32, q=19)
    A = oracle.init_bubble(p, radius=8.0)
    A[2:5] = 0.01 * np.random.default_rng(0).standard_normal(A[2:5].shape).astype(np.float32)
    with cg.Domain(_cfg(cg, p)) as dom:
        dom.set_state(A)
        got = dom.reduce()
    want = oracle.reduce(p, A)
    np.testing.assert_allclose(got[:5], want, rtol=1e-12, atol=1e-12)
    assert got[5] == 0


def test_reduce_is_deterministic_and_flags_bad_cells(oracle):
    cg = load_product()
    p = oracle.Params(nx=32, ny=32, nz=32, q=19)
    A = oracle.init_bubble(p, radius=6.0)
    with cg.Domain(_cfg(cg, p)) as dom:
        dom.set_state(A)
        dom.step(5)
        r1 = dom.reduce()
        r2 = dom.reduce()
        assert np.array_equal(r1, r2)
        S = dom.get_state()
        S[0, 3, 4, 5] = np.nan
        S[1, 7, 8, 9] = -1.0
        S[7, 10, 11, 12] = np.inf   # a Pi component
        S[10, 3, 4, 5] = np.nan     # same cell as the first: still one bad cell
        dom.set_state(S)
        assert dom.reduce()[5] == 3


def test_reduce_counts_saturated_fp16_pi(oracle):
    """FP16-deviation storage holds |Pi| < 16: a larger value saturates to inf on encoding and the
    diagnostics count the cell as non-finite, like a saturated velocity."""
    cg = load_product()
    p = oracle.Params(nx=16, ny=16, nz=16, q=19)
    A = oracle.init_bubble(p, radius=4.0)
    A[9, 2, 3, 4] = 100.0     # Pizz beyond the binary16 range at its scale
    A[3, 5, 6, 7] = 1000.0    # uy beyond the binary16 range at its scale
    with cg.Domain(_cfg(cg, p), storage=cg.STORE_F16) as dom:
        dom.set_state(A)
        assert dom.reduce()[5] == 2


def test_launch_level_reference_layout(oracle):
    """cglbm_launch_rcs on caller-owned arrays in the reference's own NX*NY*NZ ghost-column
    layout (MomentsDevice of memory.cuh), jet case: result equals the compact-layout oracle.
    The caller-owned arrays are the two states of a periodic NX x NY x NZ carrier domain, whose
    compact layout is exactly the reference's x + NX*y + NX*NY*z indexing."""
    cg = load_product()
    p = oracle.Params(nx=14, ny=20, nz=12, q=19, ybc=oracle.Y_JET, jet_radius=3.0)
    NX, NY, NZ = p.nx + 2, p.ny + 2, p.nz + 2
    A0 = oracle.init_jet(p)
    want = oracle.run(p, A0, 4)
    ref = np.zeros((11, NZ, NY, NX), dtype=np.float32)
    ref[:, 1:-1, :, 1:-1] = A0
    cfg = _cfg(cg, p)
    L = cg.Layout(1 + NX * NY, NX, NX * NY)
    with cg.Domain(cg.Config(nx=NX, ny=NY, nz=NZ)) as carrier:
        carrier.set_state(ref)
        bufs = [carrier.moments(0), carrier.moments(1)]
        stream = cg.lib().cglbm_domain_stream(carrier._h)
        for s in range(1, 5):  # A->B->A->B->A: the result is back in the carrier's current buffer
            cg.launch_rcs(bufs[(s - 1) % 2], bufs[s % 2], cfg, L, step=s, kernel=cg.KERNEL_AUTO, stream=stream)
        carrier.sync()
        got = carrier.get_state()
    np.testing.assert_array_equal(got[:, 1:-1, :, 1:-1], want)


# (label, oracle params, initialiser name, keyword arguments)
INIT_CASES = [
    ("bubble_tanh_q19", dict(nx=20, ny=18, nz=16, q=19), "bubble", dict(radius=5.0, width=4.0, sharp=False)),
    ("bubble_sharp_q27", dict(nx=17, ny=20, nz=15, q=27), "bubble", dict(radius=5.0, width=4.0, sharp=True)),
    ("bubble_offcentre_q19", dict(nx=18, ny=16, nz=20, q=19), "bubble",
     dict(radius=4.5, width=3.0, sharp=False, center=(5.0, 9.5, 12.25))),
    ("jet_q27", dict(nx=20, ny=18, nz=16, q=27, ybc=1, jet_radius=4.0), "jet", {}),
    ("rti_walls_q19", dict(nx=16, ny=24, nz=12, q=19, ybc=2), "rti",
     dict(amplitude=2.0, width=4.0, n_random_modes=8, random_amplitude=0.2, seed=5)),
]


@pytest.mark.parametrize("label,kw,kind,args", INIT_CASES, ids=[c[0] for c in INIT_CASES])
def test_launch_level_initialisers_reference_layout(oracle, label, kw, kind, args):
    """cglbm_launch_init_* (launch_InitStaticBubble / launch_InitJet / launch_InitRTI,
    launch.cuh:14-72) on caller arrays in the reference's NX*NY*NZ ghost-column layout: the
    addressed cells equal the oracle's initialisers bit for bit (whose own bits are pinned to the
    reference's initialisers by the golden vectors) and every other cell keeps its previous
    value."""
    cg = load_product()
    p = oracle.Params(**kw)
    NX, NZ = p.nx + 2, p.nz + 2
    NY = p.ny + 2
    if kind == "bubble":
        want = oracle.init_bubble(p, args["radius"], args["width"], args["sharp"], args.get("center"))
    elif kind == "jet":
        want = oracle.init_jet(p)
    else:
        want = oracle.init_rti(p, args["amplitude"], args["width"], args["n_random_modes"], args["random_amplitude"],
                               args["seed"])
    periodic = p.ybc == oracle.Y_PERIODIC
    y0 = 1 if periodic else 0
    L = cg.Layout(1 + NX * NY + (NX if periodic else 0), NX, NX * NY)
    cfg = _cfg(cg, p)
    fill = np.full((11, NZ, NY, NX), 7.5, dtype=np.float32)  # marks cells the call must not touch
    with cg.Domain(cg.Config(nx=NX, ny=NY, nz=NZ)) as carrier:
        carrier.set_state(fill)
        A = carrier.moments(0)
        stream = cg.lib().cglbm_domain_stream(carrier._h)
        print(f"init {label}", flush=True)
        if kind == "bubble":
            cg.launch_init_bubble(A, cfg, L, stream=stream, **args)
        elif kind == "jet":
            cg.launch_init_jet(A, cfg, L, stream=stream)
        else:
            cg.launch_init_rti(A, cfg, L, stream=stream, **args)
        got = carrier.get_state()
    inner = got[:, 1:-1, y0:y0 + p.nyt, 1:-1]
    np.testing.assert_array_equal(inner, want)
    outside = got.copy()
    outside[:, 1:-1, y0:y0 + p.nyt, 1:-1] = 7.5
    assert np.all(outside == 7.5)


SHAPES = [(4, 3, 1), (8, 5, 2), (16, 7, 1), (16, 16, 2), (8, 1, 2), (4, 100, 2), (16, 0, 2), (4, 0, 1)]


@pytest.mark.parametrize("shape", SHAPES, ids=[f"ty{a}L{b}nb{c}" for a, b, c in SHAPES])
@pytest.mark.parametrize("name", ["jet_small_q19", "jet_small_q27", "bubble_q19", "rand55_q27"])
def test_march_shapes_match_reference_golden(oracle, name, shape):
    """Every tile shape of the z-marching kernel (tile rows, chunk length, exchange buffers)
    reproduces the reference's golden vectors bit for bit."""
    cg = load_product()
    z, meta = _golden.load(name)
    p, A = _golden.initial_state(oracle, z, meta)
    ty, L, nbuf = shape
    print(f"case {name} shape {shape}", flush=True)
    with cg.Domain(_cfg(cg, p), kernel=cg.KERNEL_MARCH) as dom:
        dom.set_tile(ty, min(L, p.nz), nbuf)  # L = 0: planned chunks
        dom.set_state(A)
        k = 0
        for s in meta["steps"][:2]:
            dom.step(s - k)
            k = s
            shas = _golden.field_shas(dom.get_state())
            assert shas == list(z[f"sha_{s}"]), f"step {s} differs"


def test_autotune_keeps_state_and_parity(oracle):
    cg = load_product()
    z, meta = _golden.load("bubble_q27")
    p, A = _golden.initial_state(oracle, z, meta)
    with cg.Domain(_cfg(cg, p)) as dom:
        dom.set_state(A)
        ty, L, nb, ms = dom.autotune(2)
        assert ty in (4, 8, 16) and 0 <= L <= p.nz and nb in (1, 2) and ms > 0  # L = 0: planned chunks
        assert not (ty == 16 and nb == 2)  # reports the shape it ran: D3Q27 16-row tiles keep one buffer
        np.testing.assert_array_equal(dom.get_state(), A)  # autotuning does not advance the state
        dom.step(10)
        assert _golden.field_shas(dom.get_state()) == list(z["sha_10"])
