"""Restartable snapshots (cglbm_snapshot_write / read / info): a domain restored from a snapshot
steps on to bitwise the state of the uninterrupted run, for the periodic drop, the jet with a
step-keyed inlet perturbation and a zero-gradient outlet, and the forced Rayleigh-Taylor layer in
FP16-deviation storage; the step count travels with the state; a snapshot is refused by a domain
of another configuration or storage, and a damaged file by its checksum."""
import os

import numpy as np
import pytest

from conftest import load_product

pytestmark = pytest.mark.gpu


def _cases(cg):
    return [
        ("drop_q19", cg.Config(nx=24, ny=24, nz=24, q=19, tau=0.8, sigma=0.01, beta=0.7), cg.STORE_F32,
         lambda d: d.init_bubble(radius=7.0, width=4.0)),
        ("jet_q27_noise", cg.Config(nx=16, ny=32, nz=16, q=27, ybc=cg.Y_JET, outlet=cg.OUTLET_ZERO_GRADIENT, tau=0.55,
                                    sigma=0.005, beta=0.7, jet_radius=4.0, jet_velocity=0.05, jet_noise=0.01, seed=3),
         cg.STORE_F32, lambda d: d.init_jet()),
        ("rti_q19_f16", cg.Config(nx=16, ny=32, nz=16, q=19, ybc=cg.Y_WALLS, tau=0.6, sigma=0.002, beta=0.7,
                                  gravity=1e-4, atwood=0.5), cg.STORE_F16, lambda d: d.init_rti(2.0, 4.0, 8, 0.2, 0)),
    ]


@pytest.mark.parametrize("which", [0, 1, 2])
def test_restored_domain_continues_bitwise(tmp_path, which):
    cg = load_product()
    name, cfg, storage, init = _cases(cg)[which]
    path = str(tmp_path / f"{name}.snap")
    with cg.Domain(cfg, storage=storage) as d:
        init(d)
        d.step(7)
        d.save(path)
        d.step(5)
        want = d.get_state()
        assert d.step_count == 12
    info_cfg, info_storage, info_steps = cg.snapshot_info(path)
    assert info_storage == storage and info_steps == 7
    assert info_cfg.nx == cfg.nx and info_cfg.ny == cfg.ny and info_cfg.q == cfg.q and info_cfg.seed == cfg.seed
    with cg.Domain.from_snapshot(path) as r:
        assert r.step_count == 7
        r.step(5)
        got = r.get_state()
        assert r.step_count == 12
    assert np.array_equal(got, want), f"{name}: restored run differs from the uninterrupted one"


def test_snapshot_refused_by_another_domain_and_when_damaged(tmp_path):
    cg = load_product()
    cfg = cg.Config(nx=16, ny=16, nz=16, q=19, tau=0.8, sigma=0.01, beta=0.7)
    path = str(tmp_path / "a.snap")
    with cg.Domain(cfg) as d:
        d.init_bubble(radius=5.0, width=4.0)
        d.step(2)
        d.save(path)
    with cg.Domain(cg.Config(nx=16, ny=16, nz=16, q=19, tau=0.9, sigma=0.01, beta=0.7)) as other:
        with pytest.raises(cg.CglbmError, match="differs"):
            other.load(path)
    with cg.Domain(cfg, storage=cg.STORE_F16) as other:
        with pytest.raises(cg.CglbmError, match="differs"):
            other.load(path)
    raw = bytearray(open(path, "rb").read())
    raw[-5] ^= 0x01  # one bit of the last field
    bad = str(tmp_path / "bad.snap")
    open(bad, "wb").write(bytes(raw))
    with cg.Domain(cfg) as d:
        with pytest.raises(cg.CglbmError, match="checksum"):
            d.load(bad)
    open(bad, "wb").write(bytes(raw[: len(raw) // 2]))
    with cg.Domain(cfg) as d:
        with pytest.raises(cg.CglbmError, match="truncated"):
            d.load(bad)
