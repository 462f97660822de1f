"""VTI output parity (io/vtiWriter.cuh:60-164).

The fixtures under tests/golden/vti_*.npz hold what the reference's own writer printed for a
snapshot of its full grid (ghost layers included): the five fields and the sha256 of the file.
CPU tests: ``cglbm_write_vti_host`` reproduces that file byte for byte from the same arrays, the
printed snapshot is the oracle's state, and the error paths. GPU tests: the state a domain reaches
on the device, printed through the same writer, gives the reference's file again, and
``cglbm_write_vti`` of a domain prints exactly its stored extent.
"""
import hashlib
import os

import numpy as np
import pytest

import _golden
from conftest import load_product

CASES = ["vti_rand_q19", "vti_jet_small_q19"]


def _sha_file(path):
    with open(path, "rb") as fh:
        data = fh.read()
    return hashlib.sha256(data).hexdigest(), len(data)


def _parse_vti(path):
    """The five DataArrays of an ASCII VTI as float32 (nz, ny, nx) arrays, plus the extent."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    ext = [int(v) for v in lines[2].split('WholeExtent="')[1].split('"')[0].split()]
    nx, ny, nz = ext[1] + 1, ext[3] + 1, ext[5] + 1
    out, cur = {}, None
    for ln in lines:
        if "<DataArray" in ln:
            cur = ln.split('Name="')[1].split('"')[0]
            out[cur] = []
        elif "</DataArray>" in ln:
            cur = None
        elif cur is not None:
            out[cur].append(np.array(ln.split(), dtype=np.float32))
    return [np.concatenate(out[n]).reshape(nz, ny, nx) for n in ("rhor", "rhob", "ux", "uy", "uz")]


@pytest.mark.parametrize("name", CASES)
def test_host_writer_is_byte_identical_to_reference(name, tmp_path):
    cg = load_product()
    z, meta = _golden.load(name)
    path = str(tmp_path / "out.vti")
    cg.write_vti_host(path, list(z["full5"]))
    digest, size = _sha_file(path)
    assert size == int(z["vti_bytes"])
    assert digest == str(z["vti_sha256"])
    # "%.8e" keeps nine significant digits, enough to read every float32 back exactly
    for got, want in zip(_parse_vti(path), z["full5"]):
        np.testing.assert_array_equal(got, want)


def test_printed_snapshot_is_the_oracle_state(oracle):
    """The fields the reference printed after 2 steps of the random-moment case are the oracle's
    state after 2 steps from the same input (compact region; x/z ghost shells stay zero)."""
    z, meta = _golden.load("vti_rand_q19")
    zr, mr = _golden.load("rand_q19")
    p, A = _golden.initial_state(oracle, zr, mr)
    assert meta["params"] == mr["params"]
    got = oracle.run(p, A, meta["step"])[:5]
    full5 = z["full5"]
    np.testing.assert_array_equal(full5[:, 1:-1, :, 1:-1], got)
    shell = full5.copy()
    shell[:, 1:-1, :, 1:-1] = 0.0
    assert not shell.any()


def test_jet_snapshot_is_the_oracle_state(oracle):
    z, meta = _golden.load("vti_jet_small_q19")
    p = oracle.Params(**meta["params"])
    got = oracle.run(p, oracle.init_jet(p), meta["step"])[:5]
    np.testing.assert_array_equal(z["full5"][:, 1:-1, :, 1:-1], got)


def test_host_writer_creates_folders_and_reports_errors(tmp_path):
    cg = load_product()
    a = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4) * 0.25
    path = tmp_path / "a" / "b" / "jet_00000010.vti"
    cg.write_vti_host(str(path), [a, a + 1, a + 2, a + 3, a + 4])
    got = _parse_vti(str(path))
    for k in range(5):
        np.testing.assert_array_equal(got[k], a + k)
    text = path.read_text().splitlines()
    assert text[2] == '  <ImageData WholeExtent="0 3 0 2 0 1" Origin="0 0 0" Spacing="1 1 1">'
    assert text[6] == "0.00000000e+00 2.50000000e-01 5.00000000e-01 7.50000000e-01 "
    with pytest.raises(cg.CglbmError, match="cannot"):
        cg.write_vti_host(str(tmp_path / "a"), [a] * 5)  # a folder, not a file
    with pytest.raises(ValueError):
        cg.write_vti_host(str(tmp_path / "x.vti"), [a] * 4)
    lib = cg.lib()
    assert lib.cglbm_write_vti_host(str(tmp_path / "y.vti").encode(), 0, 3, 2, *([a.ctypes.data] * 5)) == cg.EINVAL
    assert lib.cglbm_write_vti_host(str(tmp_path / "y.vti").encode(), 4, 3, 2, *([None] * 5)) == cg.EINVAL
    assert not os.path.exists(tmp_path / "y.vti")


def _embed(full5, compact):
    out = np.array(full5, dtype=np.float32, copy=True)
    out[:, 1:-1, :, 1:-1] = compact
    return out


@pytest.mark.gpu
def test_gpu_state_prints_the_reference_file(oracle, tmp_path):
    cg = load_product()
    for name in CASES:
        z, meta = _golden.load(name)
        p = oracle.Params(**meta["params"])
        with cg.Domain(cg.Config(**meta["params"])) as d:
            if meta["mode"] == "jet":
                d.init_jet()
            else:
                zr, mr = _golden.load("rand_q19")
                d.set_state(np.array(zr["full_0"], dtype=np.float32))
            d.step(meta["step"])
            compact = d.get_state(0x1F)
            # the domain's own writer prints its stored extent
            own = str(tmp_path / f"{name}_own.vti")
            d.write_vti(own)
        for got, want in zip(_parse_vti(own), compact):
            np.testing.assert_array_equal(got, want)
        ref_host = str(tmp_path / f"{name}_compact_host.vti")
        cg.write_vti_host(ref_host, list(compact))
        assert _sha_file(own) == _sha_file(ref_host)
        path = str(tmp_path / f"{name}.vti")
        cg.write_vti_host(path, list(_embed(z["full5"], compact)))
        assert _sha_file(path) == (str(z["vti_sha256"]), int(z["vti_bytes"])), name
        del p
