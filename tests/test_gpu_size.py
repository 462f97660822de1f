"""Full-size parity: the benched and configured workloads at their real sizes, through the
default path of a library-owned domain (no tile set: planned chunks, re-planned as the state
evolves, with the one-off planned-against-uniform timing), checked field by field against the
sha256 of tests/golden/size_*.npz:

* the bench's 128^3 drop (D3Q19 and D3Q27, 100 steps) and the 64^3 static bubble (100 and 1000
  steps): written by the reference's own update built for the host;
* the D3Q27 round jet 128 x 256 x 128 with 1% inlet noise and zero-gradient outflow, and the
  D3Q19 Rayleigh-Taylor case 128 x 512 x 128 with FP32 and FP16-deviation storage (100 steps):
  written by the oracle, which the reference cannot run (noise, zero-gradient outflow and the
  body force are this build's options), and which equals the reference bit for bit wherever
  both apply.

FP16 states are compared through their decoded FP32 values (what cglbm_copy_out returns)."""
import numpy as np
import pytest

import _golden
from conftest import load_product

pytestmark = pytest.mark.gpu

NAMES = _golden.size_names()


@pytest.mark.parametrize("name", NAMES)
def test_full_size_matches_fixture(name):
    cg = load_product()
    z, meta = _golden.load(name)
    cfg = cg.Config(**meta["params"])
    storage = cg.STORE_F16 if meta.get("storage") == "f16" else cg.STORE_F32
    print(f"size case {name}: {cfg.nx} x {cfg.ny} x {cfg.nz} q{cfg.q}, steps {meta['steps']}", flush=True)
    with cg.Domain(cfg, storage=storage) as d:
        if meta["source"] == "reference":
            d.init_bubble(radius=meta["radius"], width=float(meta["width"]))
        elif meta["init"] == "jet":
            d.init_jet()
        else:
            d.init_rti(**meta["init_args"])
        done = 0
        for k in [0] + meta["steps"]:
            d.step(k - done)
            done = k
            got = d.get_state()
            shas = _golden.field_shas(got)
            bad = [f for f in range(11) if shas[f] != z[f"sha_{k}"][f]]
            if bad:
                vals = got.reshape(11, -1)[:, z["idx"]]
                diff = np.abs(vals.astype(np.float64) - z[f"val_{k}"]).max(axis=1)
                pytest.fail(f"{name} step {k}: fields {bad} differ; sampled max |diff| per field {diff}")
            np.testing.assert_array_equal(got.reshape(11, -1)[:, z["idx"]], z[f"val_{k}"])
