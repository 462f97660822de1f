"""Batched stepping: every member's state after a batched run is bit-identical to the same domain
stepped alone (and to the oracle), for mixed kernel classes, uneven sizes, graph-captured step
bundles, membership changes and re-initialisation."""
import numpy as np
import pytest

from conftest import load_product

pytestmark = pytest.mark.gpu


def _members(cg):
    """Mixed set: uneven periodic D3Q19 drops, a D3Q27 drop, a D3Q27 jet with inlet noise and a
    forced D3Q19 RTI with walls."""
    specs = []
    rng = np.random.default_rng(4)
    for e in (32, 48, 20, 40):
        specs.append(("bubble", cg.Config(nx=e, ny=e + 8, nz=e - 4, q=19, tau=float(rng.uniform(0.52, 1.0)),
                                          sigma=float(np.exp(rng.uniform(np.log(1e-3), np.log(2e-2)))), beta=0.7)))
    specs.append(("bubble", cg.Config(nx=36, ny=28, nz=24, q=27, tau=0.7, sigma=0.01, beta=0.7)))
    specs.append(("jet", cg.Config(nx=24, ny=40, nz=20, q=27, ybc=cg.Y_JET, tau=0.55, sigma=0.005, beta=0.7,
                                   jet_radius=5.0, jet_noise=0.01, seed=3)))
    specs.append(("rti", cg.Config(nx=24, ny=48, nz=16, q=19, ybc=cg.Y_WALLS, tau=0.6, sigma=0.002, beta=0.7,
                                   gravity=1e-4, atwood=0.5)))
    return specs


def _init(dom, kind):
    c = dom.cfg
    if kind == "jet":
        dom.init_jet()
    elif kind == "rti":
        dom.init_rti(2.0, 4.0, 8, 0.2, 7)
    else:
        dom.init_bubble(radius=min(c.nx, c.ny, c.nz) / 4.0, width=4.0)


def _alone(cg, specs, steps):
    out = []
    for kind, cfg in specs:
        with cg.Domain(cfg) as d:
            _init(d, kind)
            d.step(steps)
            out.append(d.get_state())
    return out


@pytest.mark.parametrize("bundle", [0, 4])
def test_batch_equals_alone(bundle):
    cg = load_product()
    specs = _members(cg)
    steps = 11
    want = _alone(cg, specs, steps)
    doms = [cg.Domain(cfg) for _, cfg in specs]
    for d, (kind, _) in zip(doms, specs):
        _init(d, kind)
    with cg.Batch() as b:
        for d in doms:
            b.add(d)
        b.set_graph(bundle)
        b.step(5)
        b.step(6)
        for i, d in enumerate(doms):
            got = d.get_state()
            assert d.step_count == steps
            assert np.array_equal(got, want[i]), f"member {i} ({specs[i][0]}) differs from its alone run"
        for d in doms:
            b.remove(d)
    for d in doms:
        d.close()


def test_batch_matches_oracle(oracle):
    cg = load_product()
    p1 = oracle.Params(nx=40, ny=32, nz=24, q=19, tau=0.9, sigma=0.015)
    p2 = oracle.Params(nx=20, ny=36, nz=16, q=27, ybc=oracle.Y_JET, jet_radius=4.0, tau=0.55, sigma=0.005)
    A1 = oracle.init_bubble(p1, radius=7.0)
    A2 = oracle.init_jet(p2)
    d1 = cg.Domain(cg.Config(**{k: getattr(p1, k) for k in p1.__dataclass_fields__}))
    d2 = cg.Domain(cg.Config(**{k: getattr(p2, k) for k in p2.__dataclass_fields__}))
    d1.set_state(A1)
    d2.set_state(A2)
    with cg.Batch() as b:
        b.add(d1)
        b.add(d2)
        b.set_graph(3)
        b.step(9)
        np.testing.assert_array_equal(d1.get_state(), oracle.run(p1, A1, 9))
        np.testing.assert_array_equal(d2.get_state(), oracle.run(p2, A2, 9))
        b.remove(d1)
        b.remove(d2)
    d1.close()
    d2.close()


def test_batch_membership_changes_and_reinit(oracle):
    cg = load_product()
    p = oracle.Params(nx=24, ny=24, nz=24, q=19)
    A = oracle.init_bubble(p, radius=5.0)
    cfg = cg.Config(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    d1, d2 = cg.Domain(cfg), cg.Domain(cfg)
    d1.set_state(A)
    d2.set_state(A)
    with cg.Batch() as b:
        b.add(d1)
        b.step(4)
        b.add(d2)  # joins 4 steps behind
        b.step(3)
        np.testing.assert_array_equal(d1.get_state(), oracle.run(p, A, 7))
        np.testing.assert_array_equal(d2.get_state(), oracle.run(p, A, 3))
        with pytest.raises(cg.CglbmError, match="batch member"):
            d1.step(1)
        d1.set_state(A)
        d1.init_bubble(radius=5.0)  # re-initialised member: step counter restarts
        b.remove(d2)
        b.step(2)
        np.testing.assert_array_equal(d1.get_state(), oracle.run(p, oracle.init_bubble(p, radius=5.0), 2))
        np.testing.assert_array_equal(d2.get_state(), oracle.run(p, A, 3))
        b.remove(d1)
    d1.close()
    d2.close()


def test_batch_autotune_and_sequential_baseline_keep_parity(oracle):
    cg = load_product()
    p = oracle.Params(nx=32, ny=32, nz=32, q=19)
    A = oracle.init_bubble(p, radius=6.0)
    cfg = cg.Config(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    doms = [cg.Domain(cfg) for _ in range(3)]
    for d in doms:
        d.set_state(A)
    with cg.Batch() as b:
        for d in doms:
            b.add(d)
        b.autotune(1)
        for d in doms:
            np.testing.assert_array_equal(d.get_state(), A)
        b.time_sequential(2)
        b.time_steps(3)
        for d in doms:
            np.testing.assert_array_equal(d.get_state(), oracle.run(p, A, 5))
        for d in doms:
            b.remove(d)
    for d in doms:
        d.close()
