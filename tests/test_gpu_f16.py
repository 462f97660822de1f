"""FP16-deviation storage on the GPU, bit-exact against the oracle's storage model.

A step with FP16 storage is the FP32 update of the decoded state followed by the binary16
rounding of the nine stored deviations (oracle.run_f16). The device must reproduce that exactly,
field for field, for every kernel class (D3Q19/D3Q27, periodic/jet/walls, body force), tile
shape, and in batches mixed with FP32 members."""
import numpy as np
import pytest

from conftest import load_product

pytestmark = pytest.mark.gpu


def _cfg(cg, p):
    return cg.Config(**{k: getattr(p, k) for k in p.__dataclass_fields__})


def _initial(oracle, p):
    if p.ybc == oracle.Y_JET:
        return oracle.init_jet(p)
    if p.ybc == oracle.Y_WALLS:
        return oracle.init_rti(p, 2.0, 4.0, 8, 0.2, 1)
    return oracle.init_bubble(p, radius=min(p.nx, p.ny, p.nz) / 5.0, width=4.0)


def _check(got, want, label):
    if not np.array_equal(got, want):
        diff = np.abs(got.astype(np.float64) - want).reshape(11, -1)
        pytest.fail(f"{label}: cells differing per field {(diff > 0).sum(axis=1)}, max {diff.max(axis=1)}")


def test_copy_roundtrip_is_the_storage_rounding(oracle):
    cg = load_product()
    p = oracle.Params(nx=24, ny=20, nz=12, q=19)
    rng = np.random.default_rng(1)
    A = np.empty((11, p.nz, p.ny, p.nx), dtype=np.float32)
    A[:2] = rng.uniform(0, 1, A[:2].shape)
    A[2:5] = rng.uniform(-0.1, 0.1, A[2:5].shape) * rng.choice([1.0, 1e-4, 1e-7], A[2:5].shape)
    A[5:] = rng.uniform(-1e-2, 1e-2, A[5:].shape) * rng.choice([1.0, 1e-4, 1e-8], A[5:].shape)
    with cg.Domain(_cfg(cg, p), storage=cg.STORE_F16) as d:
        d.set_state(A)
        Q = oracle.quantize_f16(A)
        np.testing.assert_array_equal(d.get_state(), Q)
        # partial copies: rhor only (no staging) and uy only (decode, replace, encode)
        np.testing.assert_array_equal(d.get_state(0x1), Q[:1])
        B = A.copy()
        B[3] = rng.uniform(-0.05, 0.05, B[3].shape)
        d.set_state(B[3:4], mask=1 << 3)
        Q[3] = oracle.quantize_f16(B)[3]
        np.testing.assert_array_equal(d.get_state(), Q)
        with pytest.raises(cg.CglbmError, match="FP32 field views"):
            d.moments()
    # two states of 26 instead of 44 bytes per (padded) cell; the reduce scratch is the same
    fs = (p.nx * p.ny * p.nz + 63) // 64 * 64
    assert cg.domain_bytes(_cfg(cg, p), cg.STORE_F32) - cg.domain_bytes(_cfg(cg, p), cg.STORE_F16) == 2 * 18 * fs


CONFIGS = [
    ("bubble_q19", dict(nx=40, ny=36, nz=32, q=19, tau=0.8, sigma=0.01, beta=0.7), 60),
    ("bubble_q27", dict(nx=36, ny=28, nz=24, q=27, tau=0.7, sigma=0.01, beta=0.7), 30),
    ("jet_noise_q27", dict(nx=24, ny=40, nz=20, q=27, ybc=1, tau=0.55, sigma=0.005, beta=0.7, jet_radius=5.0,
                           jet_noise=0.01, seed=9), 30),
    ("jet_q19", dict(nx=20, ny=36, nz=18, q=19, ybc=1, tau=0.55, sigma=0.005, beta=0.7, jet_radius=4.0), 40),
    ("rti_q19", dict(nx=24, ny=48, nz=20, q=19, ybc=2, tau=0.6, sigma=0.002, beta=0.7, gravity=1e-4, atwood=0.5), 30),
    ("conservative_q19", dict(nx=20, ny=20, nz=20, q=19, conservative_omega2=1, sigma=0.02), 20),
]


@pytest.mark.parametrize("label,kw,steps", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_step_matches_oracle_storage_model(oracle, label, kw, steps):
    cg = load_product()
    p = oracle.Params(**kw)
    A = _initial(oracle, p)
    print(f"f16 config {label} steps {steps}", flush=True)
    want = oracle.run_f16(p, A, steps)
    with cg.Domain(_cfg(cg, p), storage=cg.STORE_F16) as d:
        d.set_state(A)
        d.step(steps)
        _check(d.get_state(), want, label)


def test_initialisers_store_rounded_state(oracle):
    cg = load_product()
    p = oracle.Params(nx=20, ny=18, nz=16, q=27, ybc=oracle.Y_JET, jet_radius=4.0)
    with cg.Domain(_cfg(cg, p), storage=cg.STORE_F16) as d:
        d.init_jet()
        np.testing.assert_array_equal(d.get_state(), oracle.quantize_f16(oracle.init_jet(p)))
    p = oracle.Params(nx=16, ny=24, nz=12, q=19, ybc=oracle.Y_WALLS, gravity=1e-4, atwood=0.5)
    with cg.Domain(_cfg(cg, p), storage=cg.STORE_F16) as d:
        d.init_rti(2.0, 4.0, 8, 0.2, 5)
        np.testing.assert_array_equal(d.get_state(), oracle.quantize_f16(oracle.init_rti(p, 2.0, 4.0, 8, 0.2, 5)))
    p = oracle.Params(nx=20, ny=18, nz=16, q=19)
    with cg.Domain(_cfg(cg, p), storage=cg.STORE_F16) as d:
        d.init_bubble(radius=5.0, width=4.0)
        np.testing.assert_array_equal(d.get_state(), oracle.quantize_f16(oracle.init_bubble(p, 5.0, 4.0)))


@pytest.mark.parametrize("ty,L,nbuf", [(4, 3, 1), (8, 5, 2), (16, 7, 1), (16, 16, 2), (8, 1, 2), (16, 0, 2), (8, 0, 1)])
def test_tile_shapes(oracle, ty, L, nbuf):
    cg = load_product()
    for kw in (dict(nx=40, ny=20, nz=16, q=19), dict(nx=20, ny=30, nz=14, q=27, ybc=1, jet_radius=4.0)):
        p = oracle.Params(**kw)
        A = _initial(oracle, p)
        want = oracle.run_f16(p, A, 7)
        with cg.Domain(_cfg(cg, p), storage=cg.STORE_F16) as d:
            d.set_tile(ty, min(L, p.nz), nbuf)
            d.set_state(A)
            d.step(7)
            _check(d.get_state(), want, f"q{p.q} tile {ty}x{L} nbuf {nbuf}")


def test_reduce_and_autotune(oracle):
    cg = load_product()
    p = oracle.Params(nx=40, ny=36, nz=32, q=19)
    A = oracle.init_bubble(p, radius=8.0)
    A[2:5] = 0.01 * np.random.default_rng(0).standard_normal(A[2:5].shape).astype(np.float32)
    with cg.Domain(_cfg(cg, p), storage=cg.STORE_F16) as d:
        d.set_state(A)
        got = d.reduce()
        np.testing.assert_allclose(got[:5], oracle.reduce(p, oracle.quantize_f16(A)), rtol=1e-12, atol=1e-12)
        assert got[5] == 0
        d.autotune(1)
        np.testing.assert_array_equal(d.get_state(), oracle.quantize_f16(A))
        d.step(3)
        _check(d.get_state(), oracle.run_f16(p, A, 3), "after autotune")


def test_f16_rejects_unsupported_cases():
    cg = load_product()
    with pytest.raises(cg.CglbmError, match="body force"):
        cg.Domain(cg.Config(nx=16, ny=16, nz=16, gravity=1e-5), storage=cg.STORE_F16)
    with cg.Domain(cg.Config(nx=16, ny=16, nz=16), storage=cg.STORE_F16) as d:
        d.set_kernel(cg.KERNEL_GATHER)
        d.init_bubble(radius=4.0)
        with pytest.raises(cg.CglbmError, match="z-marching"):
            d.step(1)


def test_batch_mixes_storage_formats(oracle):
    cg = load_product()
    ps = [oracle.Params(nx=32, ny=24, nz=20, q=19, tau=0.9), oracle.Params(nx=24, ny=24, nz=24, q=27),
          oracle.Params(nx=20, ny=36, nz=16, q=19, ybc=oracle.Y_JET, jet_radius=4.0, jet_noise=0.01, seed=2)]
    doms, want = [], []
    for p in ps:
        A = _initial(oracle, p)
        for storage, model in ((cg.STORE_F16, oracle.run_f16), (cg.STORE_F32, oracle.run)):
            d = cg.Domain(_cfg(cg, p), storage=storage)
            d.set_state(A)
            doms.append(d)
            want.append(model(p, A, 9))
    with cg.Batch() as b:
        for d in doms:
            b.add(d)
        b.set_graph(3)
        b.step(9)
        for i, d in enumerate(doms):
            _check(d.get_state(), want[i], f"member {i}")
        for d in doms:
            b.remove(d)
    for d in doms:
        d.close()
