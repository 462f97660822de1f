"""Planned z-chunks of the marching kernel (tile shapes with L = 0).

Every 64 steps a step records which own waves took the interface form, and a one-work-group
planner cuts each column tile's z range into chunks of about equal modelled cost, one work-group
per CU. The partition only moves work between work-groups, so results must stay bitwise equal to
the oracle (small cases) and to uniform chunks (large cases, across re-plans); the plan itself
must cover every (column tile, plane) exactly once."""
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


def _check_cover(items, nx, ny, nz, ty):
    ntx, nty = -(-nx // 32), -(-ny // ty)
    cover = np.zeros((nty, ntx, nz), dtype=np.int64)
    for tx, tyi, z0, z1 in items:
        if z0 == z1:
            continue
        assert 0 <= tx < ntx and 0 <= tyi < nty and 0 <= z0 < z1 <= nz, (tx, tyi, z0, z1)
        cover[tyi, tx, z0:z1] += 1
    assert (cover == 1).all(), f"planes covered {np.unique(cover)} times"
    return ntx * nty


PLAN_CASES = [
    # (label, params, tile rows, exchange buffers)
    ("drop128_q19", dict(nx=128, ny=128, nz=128, q=19, tau=0.8, sigma=0.01, beta=0.7), 16, 2),
    ("drop96_q27", dict(nx=96, ny=96, nz=96, q=27, tau=0.8, sigma=0.01, beta=0.7), 16, 1),
    ("jet_q19_ty8", dict(nx=48, ny=64, nz=40, q=19, ybc=1, jet_radius=6.0), 8, 2),
    ("rti_q27_ty4", dict(nx=40, ny=48, nz=20, q=27, ybc=2, gravity=1e-4, atwood=0.5), 4, 1),
    # more column tiles (17 x 16 = 272) than CUs: two rounds of work-groups, two bounds per round
    ("wide_q19", dict(nx=512, ny=272, nz=8, q=19), 16, 2),
]


@pytest.mark.parametrize("label,kw,ty,nbuf", PLAN_CASES, ids=[c[0] for c in PLAN_CASES])
def test_plan_covers_every_plane_once(oracle, label, kw, ty, nbuf):
    cg = load_product()
    p = oracle.Params(**kw)
    print(f"plan cover {label}", flush=True)
    with cg.Domain(_cfg(cg, p)) as dom:
        dom.set_tile(ty, 0, nbuf)
        dom.set_state(_initial(oracle, p))
        assert dom.plan()[1].shape[0] == 0  # a new state has no plan until a step records one
        dom.step(2)
        pty, items = dom.plan()
        assert pty == ty
        ncol = _check_cover(items, p.nx, p.ny, p.nz, ty)
        assert items.shape[0] >= ncol and items.shape[0] % 8 == 0
        busy = items[items[:, 2] < items[:, 3]]
        # busy items first, ordered by the middle of their z range (each XCD takes a z band)
        assert (items[: busy.shape[0], 2] < items[: busy.shape[0], 3]).all()
        zmid = busy[:, 2] + busy[:, 3]
        assert (np.diff(zmid) >= 0).all()


def test_plan_follows_the_interface(oracle):
    """Chunks through the drop's interface shell are shorter than chunks in the bulk (the
    greedy cut leaves one short remainder per column, so medians are compared)."""
    cg = load_product()
    p = oracle.Params(nx=128, ny=128, nz=128, q=19, tau=0.8, sigma=0.01, beta=0.7)
    with cg.Domain(_cfg(cg, p)) as dom:
        dom.set_tile(16, 0, 2)
        dom.init_bubble(radius=24.0, width=4.0)
        dom.step(2)
        _, items = dom.plan()
        busy = items[items[:, 2] < items[:, 3]]
        length = busy[:, 3] - busy[:, 2]
        # column tile (1, 3) crosses the drop (x 32..63, y 48..63); tile (0, 0) lies in the bulk
        through = length[(busy[:, 0] == 1) & (busy[:, 1] == 3)]
        bulk = length[(busy[:, 0] == 0) & (busy[:, 1] == 0)]
        assert np.median(through) < np.median(bulk), (through, bulk)
        assert through.sum() == bulk.sum() == p.nz


PLANNED = [(16, 0, 2), (8, 0, 1), (4, 0, 2)]
ORACLE_CASES = [
    ("bubble32_q19", dict(nx=32, ny=32, nz=32, q=19, tau=0.8, sigma=0.01, beta=0.7), 70),
    ("bubble_q27", dict(nx=40, ny=24, nz=20, q=27, tau=0.8, sigma=0.01, beta=0.7), 12),
    ("jet_noise_q27", dict(nx=24, ny=40, nz=20, q=27, ybc=1, tau=0.55, sigma=0.005, beta=0.7, jet_radius=5.0,
                           jet_noise=0.01, seed=9, outlet=1), 12),
    ("rti_q19", dict(nx=24, ny=48, nz=20, q=19, ybc=2, tau=0.6, sigma=0.002, beta=0.7, gravity=1e-4, atwood=0.5), 12),
    # odd extents in every direction (partial tiles in x and y), across a re-plan
    ("odd_q19", dict(nx=45, ny=37, nz=23, q=19, tau=0.7, sigma=0.01, beta=0.7), 70),
    ("odd_jet_q27", dict(nx=35, ny=29, nz=21, q=27, ybc=1, tau=0.6, sigma=0.005, beta=0.7, jet_radius=4.0), 66),
]


@pytest.mark.parametrize("shape", PLANNED, ids=[f"ty{a}nb{c}" for a, _, c in PLANNED])
@pytest.mark.parametrize("label,kw,steps", ORACLE_CASES, ids=[c[0] for c in ORACLE_CASES])
def test_planned_matches_oracle(oracle, label, kw, steps, shape):
    """Bitwise equal to the oracle; the 70-step case crosses a re-plan."""
    cg = load_product()
    p = oracle.Params(**kw)
    A = _initial(oracle, p)
    print(f"planned {label} shape {shape} steps {steps}", flush=True)
    want = oracle.run(p, A, steps)
    with cg.Domain(_cfg(cg, p), kernel=cg.KERNEL_MARCH) as dom:
        dom.set_tile(*shape)
        dom.set_state(A)
        dom.step(steps)
        got = dom.get_state()
        assert dom.plan()[0] == shape[0]
    if not np.array_equal(got, want):
        diff = np.abs(got.astype(np.float64) - want).reshape(11, -1)
        pytest.fail(f"{label}: cells differing per field {(diff > 0).sum(axis=1)}, max {diff.max(axis=1)}")


@pytest.mark.parametrize("q", [19, 27])
def test_planned_equals_uniform_at_128(q):
    """The 128^3 drop over 140 steps (two re-plans): planned chunks (the untuned default) give
    the same bits as uniform 16-plane chunks, in FP32 and in FP16-deviation storage."""
    cg = load_product()
    cfg = cg.Config(nx=128, ny=128, nz=128, q=q, tau=0.8, sigma=0.01, beta=0.7)
    for storage in (cg.STORE_F32, cg.STORE_F16):
        states = []
        for tile in (None, (16, 16, 1)):
            print(f"q{q} storage {storage} tile {tile}", flush=True)
            with cg.Domain(cfg, storage=storage) as dom:
                if tile:
                    dom.set_tile(*tile)
                dom.init_bubble(radius=24.0, width=4.0)
                dom.step(140)
                states.append(dom.get_state())
                pty, items = dom.plan()
                if tile is None:
                    assert pty == 16 and items.shape[0] > 0  # the default plans at this size
        assert np.array_equal(states[0], states[1])


@pytest.mark.parametrize("shape,planned", [((48, 48, 48), False), ((64, 64, 64), True), ((128, 128, 128), True),
                                          ((256, 256, 256), True), ((320, 320, 320), False), ((512, 272, 8), False)])
def test_default_plans_where_it_fits(shape, planned):
    """The untuned default plans its chunks when they average at least 2 planes (not at 48^3), the
    planner's LDS holds the domain's column tiles (not at 320^3, 200 column tiles of 320 planes)
    and one round of work-groups covers them (not for 272 column tiles); otherwise it keeps the
    size-aware uniform shape."""
    cg = load_product()
    n = min(shape)
    cfg = cg.Config(nx=shape[0], ny=shape[1], nz=shape[2], q=19)
    with cg.Domain(cfg) as dom:
        dom.init_bubble(radius=n / 5.0, width=4.0)
        dom.step(3)
        assert (dom.plan()[1].shape[0] > 0) == planned


def test_launch_level_planned_matches_oracle(oracle):
    """cglbm_launch_rcs_planned on caller-owned arrays in the reference's ghost-column layout (jet),
    70 steps (a re-plan included): bitwise the oracle's result."""
    cg = load_product()
    p = oracle.Params(nx=40, ny=36, nz=30, q=19, ybc=oracle.Y_JET, jet_radius=5.0)
    NX, NY, NZ = p.nx + 2, p.ny + 2, p.nz + 2
    A0 = oracle.init_jet(p)
    steps = 70
    want = oracle.run(p, A0, steps)
    ref = np.zeros((11, NZ, NY, NX), dtype=np.float32)
    ref[:, 1:-1, :, 1:-1] = A0
    cfg = _cfg(cg, p)
    L = cg.Layout(1 + NX * NY, NX, NX * NY)
    with cg.Domain(cg.Config(nx=NX, ny=NY, nz=NZ)) as carrier, cg.Plan(cfg) as plan:
        carrier.set_state(ref)
        bufs = [carrier.moments(0), carrier.moments(1)]
        stream = cg.lib().cglbm_domain_stream(carrier._h)
        for s in range(1, steps + 1):
            cg.launch_rcs_planned(bufs[(s - 1) % 2], bufs[s % 2], cfg, plan, L, step=s, stream=stream)
        carrier.sync()
        got = carrier.get_state(0x7FF)
    # an even number of steps leaves the result in the carrier's current buffer
    np.testing.assert_array_equal(got[:, 1:-1, :, 1:-1], want)


def test_launch_level_planned_equals_uniform_at_128():
    """Planned launch-level steps of a 128^3 drop give the bits of uniform launch-level steps."""
    cg = load_product()
    cfg = cg.Config(nx=128, ny=128, nz=128, q=19, tau=0.8, sigma=0.01, beta=0.7)
    states = []
    for planned in (False, True):
        with cg.Domain(cfg) as carrier:
            carrier.init_bubble(radius=24.0, width=4.0)
            bufs = [carrier.moments(0), carrier.moments(1)]
            stream = cg.lib().cglbm_domain_stream(carrier._h)
            plan = cg.Plan(cfg) if planned else None
            for s in range(1, 81):
                if planned:
                    cg.launch_rcs_planned(bufs[(s - 1) % 2], bufs[s % 2], cfg, plan, step=s, stream=stream)
                else:
                    cg.launch_rcs(bufs[(s - 1) % 2], bufs[s % 2], cfg, step=s, kernel=cg.KERNEL_MARCH, stream=stream)
            carrier.sync()
            states.append(carrier.get_state())
            if plan:
                plan.close()
    assert np.array_equal(states[0], states[1])


def test_plan_rejects_another_shape():
    cg = load_product()
    with cg.Domain(cg.Config(nx=32, ny=32, nz=32)) as carrier, cg.Plan(cg.Config(nx=32, ny=32, nz=16)) as plan:
        bufs = [carrier.moments(0), carrier.moments(1)]
        with pytest.raises(cg.CglbmError, match="another domain shape"):
            cg.launch_rcs_planned(bufs[0], bufs[1], cg.Config(nx=32, ny=32, nz=32), plan, step=1)


def _greedy_count(costs, e, T):
    n, acc = 1, e
    for w in costs:
        if acc + w > T:
            n, acc = n + 1, e + w
        else:
            acc += w
    return n


def test_first_plan_is_optimal_under_its_model(oracle):
    """The first cut of a state (before any measured correction) is, within the planner's 0.1-plane
    search tolerance, the best contiguous cut under its cost model: recomputed here from the
    state's interface map (own wave w of a 16-row tile holds rows 2w, 2w+1; a plane costs
    1 + alpha * (most interface waves on one SIMD) / 2 in 1/32 units, a chunk e more; alpha 1.0,
    e 3.6 for D3Q19), the largest modelled chunk equals the smallest bound that 256 greedy chunks
    can meet."""
    cg = load_product()
    n = 128
    p = oracle.Params(nx=n, ny=n, nz=n, q=19, tau=0.8, sigma=0.01, beta=0.7)
    with cg.Domain(_cfg(cg, p)) as dom:
        dom.set_tile(16, 0, 2)
        dom.init_bubble(radius=24.0, width=4.0)
        s = dom.get_state(0x3)
        dom.step(1)  # records the map and makes the first (uncorrected) cut
        ty, items = dom.plan()
    assert ty == 16
    rr, rb = s[0], s[1]
    inv = np.float32(1.0) / (rr + rb)
    on = (np.float32(4.0) * rr * rb * inv * inv) > np.float32(1e-4)          # (z, y, x)
    waves = on.reshape(n, n // 16, 8, 2, n // 32, 32).any(axis=(3, 5))        # (z, tile y, wave, tile x)
    loads = np.stack([waves[:, :, s_].astype(int) + waves[:, :, s_ + 4] for s_ in range(4)], axis=0).max(axis=0)
    cost = np.rint(32.0 * (1.0 + 1.0 * loads / 2.0)) / 32.0                    # (z, tile y, tile x)
    e = 3.6
    cols = {(tx, tyi): cost[:, tyi, tx].astype(np.float64) for tyi in range(n // 16) for tx in range(n // 32)}
    busy = items[items[:, 2] < items[:, 3]]
    worst = max(e + cols[(tx, tyi)][z0:z1].sum() for tx, tyi, z0, z1 in busy)
    lo, hi = e, e + max(c.sum() for c in cols.values()) + 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if sum(_greedy_count(c, e, mid) for c in cols.values()) <= items.shape[0]:
            hi = mid
        else:
            lo = mid
    assert worst <= hi + 0.1 + 1e-3, (worst, hi)
    assert busy.shape[0] <= items.shape[0]
