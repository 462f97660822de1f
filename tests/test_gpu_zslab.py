"""A domain split into z-slabs (cglbm_launch_rcs_zslab), the part of the multi-GPU decomposition one
GPU can run: each shard holds its planes plus two halo planes on each side, copied from its
neighbours' planes (periodic in z), and updates only its own planes. The shards' updates put
together must be bitwise the oracle's step of the whole domain: the periodic drop, the jet with a
step-keyed inlet perturbation and a zero-gradient outlet (whose disk and key need the global plane
index, z_offset), and the forced Rayleigh-Taylor layer between walls; uneven slabs, slabs of one
and two planes included. Invalid ranges (no room for halos) are refused."""
import numpy as np
import pytest

from conftest import load_product
from test_gpu_fields import DeviceStates

pytestmark = pytest.mark.gpu


def _cases(oracle):
    drop = oracle.Params(nx=20, ny=24, nz=24, q=19, tau=0.8, sigma=0.01, beta=0.7)
    jet = oracle.Params(nx=16, ny=20, nz=24, q=27, ybc=oracle.Y_JET, outlet=oracle.OUTLET_ZERO_GRADIENT, tau=0.55,
                        sigma=0.005, beta=0.7, jet_radius=4.0, jet_velocity=0.05, jet_noise=0.01, seed=5)
    rti = oracle.Params(nx=16, ny=24, nz=24, q=19, ybc=oracle.Y_WALLS, tau=0.6, sigma=0.002, beta=0.7, gravity=1e-4,
                        atwood=0.5)
    odd = oracle.Params(nx=18, ny=13, nz=24, q=27, tau=0.7, sigma=0.01, beta=0.7)  # rows of 18: scalar halo copies
    return {
        "drop_q27_nx18": (odd, lambda: oracle.init_bubble(odd, radius=5.0, width=4.0)),
        "drop_q19": (drop, lambda: oracle.init_bubble(drop, radius=6.0, width=4.0)),
        "jet_q27_noise": (jet, lambda: oracle.init_jet(jet)),
        "rti_q19_forced": (rti, lambda: oracle.init_rti(rti, 2.0, 4.0, 8, 0.2, 0)),
    }


@pytest.mark.parametrize("slabs", [(8, 10, 6), (1, 2, 21)], ids=["uneven", "thin"])
@pytest.mark.parametrize("label", ["drop_q19", "jet_q27_noise", "rti_q19_forced"])
def test_zslab_shards_equal_whole_domain_step(oracle, label, slabs):
    cg = load_product()
    p, init = _cases(oracle)[label]
    S = oracle.run(p, init(), 3)  # a developed state (the jet's perturbed inlet included)
    step = 4
    want = oracle.step(p, S, None, step)
    nz = p.nz
    got = np.zeros_like(want)
    g0 = 0
    for size in slabs:
        g1 = g0 + size
        planes = [(g % nz) for g in range(g0 - 2, g1 + 2)]  # the slab and its halo planes
        local = np.ascontiguousarray(S[:, planes])
        cfg = cg.Config(**{**{k: getattr(p, k) for k in p.__dataclass_fields__}, "nz": len(planes),
                           "jet_z0": p.jet_z0})
        n = local[0].size
        dev = DeviceStates(cg, n, "block")
        try:
            dev.put(0, local)
            cg.launch_rcs_zslab(dev.moments(0), dev.moments(1), cfg, 2, 2 + size, g0 - 2, step=step)
            dev.sync()
            out = dev.get(1, local.shape[1:])
        finally:
            dev.close()
        got[:, g0:g1] = out[:, 2:2 + size]
        g0 = g1
    assert g0 == nz
    assert np.array_equal(got, want), f"{label}: shards differ from the whole-domain step"


@pytest.mark.parametrize("width", [1, 3])
def test_zslab_faces_equal_slab_update(oracle, width):
    """Faces of other widths: interior by cglbm_launch_rcs_zslab, faces by the one-launch form, on
    a 9-plane slab of the forced Rayleigh-Taylor layer; bitwise the whole slab's update."""
    cg = load_product()
    p, init = _cases(oracle)["rti_q19_forced"]
    S = oracle.run(p, init(), 2)
    want = oracle.step(p, S, None, 3)
    g0, size = 7, 9
    planes = [(g % p.nz) for g in range(g0 - 2, g0 + size + 2)]
    local = np.ascontiguousarray(S[:, planes])
    cfg = cg.Config(**{**{k: getattr(p, k) for k in p.__dataclass_fields__}, "nz": len(planes), "jet_z0": p.jet_z0})
    dev = DeviceStates(cg, local[0].size, "block")
    try:
        dev.put(0, local)
        cg.launch_rcs_zslab(dev.moments(0), dev.moments(1), cfg, 2 + width, 2 + size - width, g0 - 2, step=3)
        cg.launch_rcs_zslab_faces(dev.moments(0), dev.moments(1), cfg, 2, 2 + size, width, g0 - 2, step=3)
        dev.sync()
        out = dev.get(1, local.shape[1:])
        with pytest.raises(cg.CglbmError, match="face width"):
            cg.launch_rcs_zslab_faces(dev.moments(0), dev.moments(1), cfg, 2, 2 + size, 5, g0 - 2, step=3)
        with pytest.raises(cg.CglbmError, match="face width"):
            cg.launch_rcs_zslab_faces(dev.moments(0), dev.moments(1), cfg, 2, 2 + size, 0, g0 - 2, step=3)
    finally:
        dev.close()
    assert np.array_equal(out[:, 2:2 + size], want[:, g0:g0 + size])


def test_zslab_refuses_ranges_without_halos():
    cg = load_product()
    cfg = cg.Config(nx=8, ny=8, nz=10, q=19)
    dev = DeviceStates(cg, 8 * 8 * 10, "block")
    try:
        for lo, hi in ((1, 5), (2, 9), (5, 5), (6, 4)):
            with pytest.raises(cg.CglbmError, match="halo"):
                cg.launch_rcs_zslab(dev.moments(0), dev.moments(1), cfg, lo, hi)
    finally:
        dev.close()


@pytest.mark.parametrize("schedule", ["whole", "overlapped"])
@pytest.mark.parametrize("label", ["drop_q19", "jet_q27_noise", "rti_q19_forced", "drop_q27_nx18"])
def test_zslab_exchange_through_pack_and_unpack(oracle, label, schedule):
    """The same decomposition with the halos moved by the library's own exchange kernels: every
    shard holds only its planes (halo planes zero), packs what its neighbours need, receives their
    buffers (periodic in z), unpacks them into its halo planes and updates; the result is the
    whole-domain step again. "overlapped" is the schedule that hides the transfer: each shard
    updates its interior planes (two planes in from each face) before the halos arrive, while its
    halo planes still hold zeros, then unpacks and updates the two planes next to each face in one
    launch (cglbm_launch_rcs_zslab_faces); so the interior provably reads no halo plane. A 1-plane slab cannot feed its neighbours'
    two-plane halos and is refused by the packer."""
    cg = load_product()
    p, init = _cases(oracle)[label]
    S = oracle.run(p, init(), 3)
    step = 4
    want = oracle.step(p, S, None, step)
    slabs = (8, 10, 6)
    starts = [0, 8, 18]
    plane = p.nx * (p.ny if p.ybc == oracle.Y_PERIODIC else p.ny + 2)
    shards = []
    try:
        for g0, size in zip(starts, slabs):
            local = np.zeros((11, size + 4) + S.shape[2:], dtype=np.float32)
            local[:, 2:2 + size] = S[:, g0:g0 + size]  # own planes only
            cfg = cg.Config(**{**{k: getattr(p, k) for k in p.__dataclass_fields__}, "nz": size + 4,
                               "jet_z0": p.jet_z0})
            dev = DeviceStates(cg, local[0].size, "block")
            dev.put(0, local)
            dev.put(1, np.zeros_like(local))
            send = [dev._alloc(4 * 13 * plane), dev._alloc(4 * 13 * plane)]
            shards.append((g0, size, cfg, dev, send, local.shape[1:]))
        for g0, size, cfg, dev, send, _ in shards:
            cg.zslab_pack(dev.moments(0), cfg, 2, 2 + size, send[0], send[1])
        if schedule == "overlapped":
            for g0, size, cfg, dev, send, _ in shards:
                cg.launch_rcs_zslab(dev.moments(0), dev.moments(1), cfg, 4, size, g0 - 2, step=step)
        for s, (g0, size, cfg, dev, send, _) in enumerate(shards):
            below = shards[(s - 1) % len(shards)][4]
            above = shards[(s + 1) % len(shards)][4]
            cg.zslab_unpack(dev.moments(0), cfg, 2, 2 + size, below[1], above[0])
        got = np.zeros_like(want)
        for g0, size, cfg, dev, send, shape in shards:
            if schedule == "overlapped":
                cg.launch_rcs_zslab_faces(dev.moments(0), dev.moments(1), cfg, 2, 2 + size, 2, g0 - 2, step=step)
            else:
                cg.launch_rcs_zslab(dev.moments(0), dev.moments(1), cfg, 2, 2 + size, g0 - 2, step=step)
            dev.sync()
            got[:, g0:g0 + size] = dev.get(1, shape)[:, 2:2 + size]
        assert np.array_equal(got, want), f"{label}: exchanged shards differ from the whole-domain step"
        g0, size, cfg, dev, send, _ = shards[0]
        cfg1 = cg.Config(**{**cfg.__dict__, "nz": 5})
        with pytest.raises(cg.CglbmError, match="at least 2 planes"):
            cg.zslab_pack(dev.moments(0), cfg1, 2, 3, send[0], send[1])
    finally:
        for *_, dev, _send, _shape in shards:
            dev.close()
