"""The C-ABI library loads and exports every entry point include/cglbm.h declares; argument
checking that needs no GPU answers with error codes, never a crash."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, load_product

HEADER = os.path.join(ROOT, "include", "cglbm.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cglbm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    cg = load_product()
    L = cg.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert sorted(cg.EXPORTS) == names, "Python binding list out of sync with the header"


def test_abi_version():
    cg = load_product()
    assert cg.lib().cglbm_abi_version() == 1


@pytest.mark.parametrize("bad", [
    dict(q=21), dict(nx=2), dict(tau=0.5), dict(tau=float("nan")), dict(sigma=-1e-3), dict(beta=1.5),
    dict(ybc=7), dict(outlet=3), dict(jet_noise=-0.1),
])
def test_invalid_configs_are_rejected(bad):
    cg = load_product()
    kw = dict(nx=8, ny=8, nz=8)
    kw.update(bad)
    cfg = cg.Config(**kw)
    with pytest.raises(cg.CglbmError) as e:
        cg.compact_layout(cfg)
    assert "failed (-1)" in str(e.value)
    assert cg.domain_bytes(cfg) == -1


def test_compact_layout_and_bytes():
    cg = load_product()
    cfg = cg.Config(nx=10, ny=12, nz=6, ybc=cg.Y_JET)
    L = cg.compact_layout(cfg)
    assert (L.origin, L.pitch_y, L.pitch_z) == (0, 10, 10 * 14)
    n = 10 * 14 * 6
    stride = (n + 63) // 64 * 64
    assert cg.domain_bytes(cfg) >= 2 * 11 * stride * 4


def test_launch_rejects_bad_layouts_and_aliasing():
    cg = load_product()
    cfg = cg.Config(nx=8, ny=8, nz=8)
    A = cg.Moments(*([ctypes.c_void_p(4096).value] * 11))
    B = cg.Moments(*([ctypes.c_void_p(8192).value] * 11))
    with pytest.raises(cg.CglbmError, match="layout"):
        cg.launch_rcs(A, B, cfg, cg.Layout(0, 4, 64))
    with pytest.raises(cg.CglbmError, match="alias"):
        cg.launch_rcs(A, A, cfg)
    nullB = cg.Moments()
    with pytest.raises(cg.CglbmError, match="null"):
        cg.launch_rcs(A, nullB, cfg)


def test_launch_rejects_overlapping_fields():
    """Any field of B overlapping any field of A, or another field of B, over the addressed cells
    is refused before a device call (the update would read A while writing B). Disjoint fields
    pass the check: every GPU test that launches on caller arrays covers that side."""
    cg = load_product()
    cfg = cg.Config(nx=8, ny=8, nz=8)
    n = 8 * 8 * 8 * 4  # bytes one field spans in the compact layout
    base = 1 << 20
    a = [base + f * n for f in range(11)]
    b = [base + (11 + f) * n for f in range(11)]
    # (A field moved onto, B field, byte shift, first overlapping A field in the check's order)
    for fa, fb, shift, first in ((0, 0, 0, 0), (7, 3, 4, 7), (10, 10, n - 4, 10), (2, 9, -(n - 4), 1)):
        bb = list(b)
        bb[fb] = a[fa] + shift
        with pytest.raises(cg.CglbmError, match=f"A field {first} and B field {fb}"):
            cg.launch_rcs(cg.Moments(*a), cg.Moments(*bb), cfg)
    bb = list(b)
    bb[4] = b[8] + 4
    with pytest.raises(cg.CglbmError, match="B fields 4 and 8 overlap"):
        cg.launch_rcs(cg.Moments(*a), cg.Moments(*bb), cfg)


def test_domain_create_without_gpu_fails_cleanly():
    cg = load_product()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(cg.CglbmError):
        cg.Domain(cg.Config(nx=8, ny=8, nz=8))
    with pytest.raises(cg.CglbmError, match="HIP device"):
        cg.Server()


def test_domain_entry_points_reject_null_domains():
    """Domain calls given no domain return CGLBM_EINVAL with a message, without touching a GPU."""
    import ctypes
    cg = load_product()
    L = cg.lib()
    n, ty = ctypes.c_int32(), ctypes.c_int32()
    assert L.cglbm_domain_plan(None, None, 0, ctypes.byref(n), ctypes.byref(ty)) == cg.EINVAL
    assert "null domain" in L.cglbm_last_error().decode()
    assert L.cglbm_domain_set_tile(None, 16, 0, 2) == cg.EINVAL
    assert L.cglbm_step(None, 1) == cg.EINVAL


def test_server_struct_layout_matches_header():
    """The ctypes mirrors of cglbm_request / cglbm_result / cglbm_server_opts / cglbm_server_stats
    have the sizes the C compiler gives the header's structs (checked with the host compiler)."""
    import ctypes
    import subprocess
    import tempfile
    cg = load_product()
    src = r'''
#include <stdio.h>
#include "cglbm.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu\n", sizeof(cglbm_config), sizeof(cglbm_request), sizeof(cglbm_result),
           sizeof(cglbm_server_opts), sizeof(cglbm_server_stats));
    return 0;
}
'''
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, "sizes.c")
        with open(c, "w") as f:
            f.write(src)
        exe = os.path.join(tmp, "sizes")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(t) for t in (cg._Config, cg._Request, cg._Result, cg._ServerOpts, cg._ServerStats)]
    assert got == want


def test_fp16_state_over_4gib_is_rejected():
    """FP16 states are addressed with 32-bit buffer offsets: a state of 4 GiB or more is refused at
    creation, before any device call (so this runs without a GPU)."""
    cg = load_product()
    with pytest.raises(cg.CglbmError, match="4 GiB"):
        cg.Domain(cg.Config(nx=560, ny=560, nz=560), storage=cg.STORE_F16)
