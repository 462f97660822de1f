"""Continuous-batching server: every result is identical to the same case run alone, whatever
else shares the GPU, under budget-limited admission and slot limits, for every case kind and
both storage formats; plus request validation and a short Poisson run of the serving mix."""
import time

import numpy as np
import pytest

from conftest import load_product

pytestmark = pytest.mark.gpu


def _cases(cg):
    """Small mixed population: (kind, cfg, storage, steps, init kwargs)."""
    F32, F16 = cg.STORE_F32, cg.STORE_F16
    out = []
    rng = np.random.default_rng(11)
    for i, e in enumerate((20, 28, 32, 24)):
        out.append(("bubble", cg.Config(nx=e, ny=e, nz=e, q=19, tau=float(rng.uniform(0.52, 1.0)),
                                        sigma=float(np.exp(rng.uniform(np.log(1e-3), np.log(2e-2)))), beta=0.7),
                    F16 if i % 2 else F32, int(rng.integers(15, 60)), dict(radius=e / 4.0, width=4.0)))
    for i in range(3):
        out.append(("jet", cg.Config(nx=16, ny=32, nz=16, q=27, ybc=cg.Y_JET, outlet=i % 2, tau=0.55, sigma=0.005,
                                     beta=0.7, jet_radius=4.0, jet_velocity=float(rng.uniform(0.02, 0.06)),
                                     jet_noise=0.01, seed=100 + i),
                    F16 if i == 1 else F32, int(rng.integers(15, 60)), {}))
    for i in range(2):
        out.append(("rti", cg.Config(nx=16, ny=48, nz=16, q=19, ybc=cg.Y_WALLS, tau=0.6, sigma=0.002, beta=0.7,
                                     gravity=1e-4, atwood=float(rng.uniform(0.2, 0.6))),
                    F16 if i else F32, int(rng.integers(15, 60)),
                    dict(amplitude=2.0, width=4.0, n_random_modes=8, random_amplitude=0.2, init_seed=7 + i)))
    out.append(("bubble", cg.Config(nx=18, ny=18, nz=18, q=27), F32, 0, dict(radius=4.0, width=4.0)))
    return out


def _alone(cg, kind, cfg, storage, steps, kw):
    with cg.Domain(cfg, storage=storage) as d:
        if kind == "bubble":
            d.init_bubble(radius=kw["radius"], width=kw["width"])
        elif kind == "jet":
            d.init_jet()
        else:
            d.init_rti(kw["amplitude"], kw["width"], kw["n_random_modes"], kw["random_amplitude"], kw["init_seed"])
        d.step(steps)
        return d.get_state(), d.reduce()


def _collect(srv, n, timeout_s=60.0):
    got = []
    t_end = time.monotonic() + timeout_s
    while len(got) < n and time.monotonic() < t_end:
        got += srv.poll(timeout_ms=200)
    assert len(got) == n, f"{len(got)} of {n} results"
    return {r["tag"]: r for r in got}


def _check_against_alone(cg, cases, res):
    for i, (kind, cfg, storage, steps, kw) in enumerate(cases):
        r = res[i]
        assert r["status"] == 0, r["message"]
        assert r["steps"] == steps
        want, diag = _alone(cg, kind, cfg, storage, steps, kw)
        assert np.array_equal(r["fields"], want), f"case {i} ({kind}, storage {storage}) differs from its alone run"
        assert np.array_equal(r["diag"], diag)
        assert r["t_submit"] <= r["t_admit"] <= r["t_start"] <= r["t_done"]


def test_results_equal_alone_in_the_mix():
    cg = load_product()
    cases = _cases(cg)
    with cg.Server(bundle=4) as srv:
        for i, (kind, cfg, storage, steps, kw) in enumerate(cases):
            srv.submit(cfg, kind, steps=steps, storage=storage, result_mask=0x7FF, tag=i, **kw)
            if i % 3 == 2:
                time.sleep(0.02)  # later cases join a batch that is already stepping
        res = _collect(srv, len(cases))
        st = srv.stats()
    assert st["completed"] == len(cases) and st["failed"] == 0 and st["resident"] == 0
    assert st["bytes_reserved"] == 0
    _check_against_alone(cg, cases, res)


def test_budget_and_slots_limit_admission():
    cg = load_product()
    cfg = cg.Config(nx=24, ny=24, nz=24, q=19)
    one = cg.domain_bytes(cfg, cg.STORE_F32)
    cases = [("bubble", cfg, cg.STORE_F32, 20 + 3 * i, dict(radius=5.0, width=4.0)) for i in range(6)]
    for budget, slots, peak in ((int(2.5 * one), 0, 2), (0, 1, 1)):
        with cg.Server(hbm_budget_bytes=budget, max_members=slots, bundle=3) as srv:
            for i, (kind, c, storage, steps, kw) in enumerate(cases):
                srv.submit(c, kind, steps=steps, storage=storage, result_mask=0x7FF, tag=i, **kw)
            srv.drain()
            res = _collect(srv, len(cases))
            st = srv.stats()
        assert st["peak_resident"] <= peak
        if budget:
            assert st["peak_bytes_reserved"] <= budget
        _check_against_alone(cg, cases, res)


def test_state_case_and_validation(oracle):
    cg = load_product()
    p = oracle.Params(nx=20, ny=24, nz=16, q=27, ybc=oracle.Y_JET, jet_radius=4.0)
    A = oracle.init_jet(p)
    cfg = cg.Config(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    with cg.Server(hbm_budget_bytes=64 << 20) as srv:
        srv.submit(cfg, "state", steps=12, initial=A, result_mask=0x7FF, tag=5)
        r = _collect(srv, 1)[5]
        assert r["status"] == 0, r["message"]
        np.testing.assert_array_equal(r["fields"], oracle.run(p, A, 12))
        with pytest.raises(cg.CglbmError, match="budget"):
            srv.submit(cg.Config(nx=256, ny=256, nz=256), "bubble", steps=1)
        with pytest.raises(cg.CglbmError, match="ybc"):
            srv.submit(cg.Config(nx=16, ny=16, nz=16), "jet", steps=1)
        with pytest.raises(cg.CglbmError, match="forced periodic|body force"):
            srv.submit(cg.Config(nx=16, ny=16, nz=16, gravity=1e-5), "bubble", steps=1)
        with pytest.raises(cg.CglbmError, match="negative"):
            srv.submit(cg.Config(nx=16, ny=16, nz=16), "bubble", steps=-1)


def test_destroy_with_cases_in_flight_returns():
    cg = load_product()
    srv = cg.Server(bundle=2, max_members=2)
    for i in range(8):
        srv.submit(cg.Config(nx=32, ny=32, nz=32), "bubble", steps=200, radius=6.0, tag=i)
    time.sleep(0.05)
    t = time.monotonic()
    srv.close()  # resident cases finish, queued ones are failed; must not hang
    assert time.monotonic() - t < 60


def test_poisson_mix_short_run():
    cg = load_product()
    import cglbm.serving as sv
    out = sv.run_poisson(cg, rate=3.0, duration=2.0, seed=1)
    assert out["requests"] >= 1 and out["completed"] == out["requests"] and out["failed"] == 0
    assert out["latency_p50_s"] <= out["latency_p99_s"]
    assert out["aggregate_mlups"] > 0
