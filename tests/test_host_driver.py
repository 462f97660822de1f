"""The C++ host driver (lightweight-mb-color-gradient-lbm_amd/host/cglbm_run.cpp), the main program
of the reference (main.cu:13-69) rebuilt over the C ABI. On the GPU, a run with the small jet
configuration writes, at step 10, a VTI file byte-identical to the one the reference's own writer
produced for the reference's own update (tests/golden/vti_jet_small_q19.npz)."""
import hashlib
import os
import subprocess

import pytest

import _golden
from conftest import PKG_DIR

EXE = os.path.join(PKG_DIR, "cglbm_run")


def _need_exe():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} missing: build with __graft_entry__.build() or make -C the package")


def test_driver_rejects_bad_options():
    _need_exe()
    r = subprocess.run([EXE, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown option" in r.stderr
    r = subprocess.run([EXE, "--case", "pipe"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


@pytest.mark.gpu
def test_driver_reproduces_reference_output_file(tmp_path):
    _need_exe()
    z, meta = _golden.load("vti_jet_small_q19")
    p = meta["params"]
    out = tmp_path / "vti"
    cmd = [EXE, "--case", "jet", "--q", "19", "--nx", str(p["nx"]), "--ny", str(p["ny"]), "--nz", str(p["nz"]),
           "--nstep", str(meta["step"] + 1), "--noutput", str(meta["step"]), "--tau", str(p["tau"]),
           "--sigma", str(p["sigma"]), "--beta", str(p["beta"]), "--jet-radius", str(p["jet_radius"]),
           "--jet-velocity", str(p["jet_velocity"]), "--out", str(out)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "MLUPS" in r.stdout
    assert (out / "jet_00000000.vti").exists()
    data = (out / f"jet_{meta['step']:08d}.vti").read_bytes()
    assert len(data) == int(z["vti_bytes"])
    assert hashlib.sha256(data).hexdigest() == str(z["vti_sha256"])


@pytest.mark.gpu
def test_driver_other_cases_run(tmp_path):
    _need_exe()
    for extra in (["--case", "bubble", "--q", "27", "--radius", "6"],
                  ["--case", "rti", "--q", "19", "--gravity", "1e-5", "--atwood", "0.5", "--storage", "f16"]):
        cmd = [EXE, "--nx", "24", "--ny", "32", "--nz", "20", "--nstep", "21", "--noutput", "10",
               "--out", str(tmp_path / extra[1])] + extra
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert (tmp_path / extra[1] / "jet_00000020.vti").exists()
        assert "non-finite cells 0" in r.stdout
