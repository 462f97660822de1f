"""Snapshot header reader on the host (no GPU): cglbm_snapshot_info returns the configuration,
storage and step count of a file laid out as documented in csrc/snapshot.cpp (written here by
hand), and refuses a missing file (CGLBM_EIO), a file that is not a snapshot and one written by
another ABI version (CGLBM_EINVAL)."""
import ctypes
import struct

import pytest

from conftest import load_product


def _snapshot_bytes(cg, cfg, storage=0, steps=41, abi=None, magic=b"CGLBMSNP"):
    cc = cfg.c()
    nyt = cfg.ny if cfg.ybc == cg.Y_PERIODIC else cfg.ny + 2
    n = 11 * cfg.nx * nyt * cfg.nz
    abi = cg.lib().cglbm_abi_version() if abi is None else abi
    head = magic + struct.pack("<IIiiqQQ", 1, ctypes.sizeof(cc), abi, storage, steps, n, 0)
    return head + bytes(cc) + bytes(4 * n)


def test_info_reads_the_header(tmp_path):
    cg = load_product()
    cfg = cg.Config(nx=4, ny=5, nz=3, q=27, ybc=cg.Y_WALLS, tau=0.6, sigma=0.002, beta=0.7, gravity=1e-5,
                    atwood=0.5, seed=9)
    p = tmp_path / "s.snap"
    p.write_bytes(_snapshot_bytes(cg, cfg, storage=cg.STORE_F16, steps=41))
    got, storage, steps = cg.snapshot_info(str(p))
    assert (storage, steps) == (cg.STORE_F16, 41)
    assert bytes(got.c()) == bytes(cfg.c())  # the binary32 configuration the library sees


def test_info_refuses_bad_files(tmp_path):
    cg = load_product()
    L = cg.lib()
    cfg = cg.Config(nx=4, ny=4, nz=4, q=19, tau=0.8, sigma=0.01, beta=0.7)
    assert L.cglbm_snapshot_info(str(tmp_path / "missing.snap").encode(), None, None, None) == cg.EIO
    for name, data in (("junk", b"not a snapshot at all, just some bytes......" * 4),
                       ("magic", _snapshot_bytes(cg, cfg, magic=b"CGLBMXXX")),
                       ("abi", _snapshot_bytes(cg, cfg, abi=99)),
                       ("short", _snapshot_bytes(cg, cfg)[:30])):
        p = tmp_path / f"{name}.snap"
        p.write_bytes(data)
        assert L.cglbm_snapshot_info(str(p).encode(), None, None, None) == cg.EINVAL, name
