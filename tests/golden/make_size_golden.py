"""Write the full-size fixtures (tests/golden/size_*.npz): per-field sha256 of the benched and
configured workloads at their real sizes, for GPU parity through the default (planned) path.

Two sources, recorded in each fixture's ``meta["source"]``:

* ``reference`` -- the reference's OWN update compiled for the host (``oracle/_ref``, built by
  ``make -C oracle ref`` from the reference headers where they lie): the 128^3 drop of the
  bench (R = 24, tanh width 4, tau 0.8, sigma 0.01, beta 0.7; D3Q19 and D3Q27, 100 steps) and the
  64^3 static bubble (R = 12, same physics; 100 and 1000 steps). The oracle is run on the same
  cases here and must give the same sha256 (recorded as ``meta["oracle_matches"]``), which pins
  the oracle at these sizes too.
* ``oracle`` -- the pinned oracle (``oracle/cgref.cpp``) for the cases the reference cannot
  express (zero-gradient outflow with inlet noise; the body force): the D3Q27 round jet
  128 x 256 x 128 (inlet radius 8, u 0.05 with 1% noise, tau 0.55, sigma 0.005) and the D3Q19
  Rayleigh-Taylor case 128 x 512 x 128 (walls, Atwood 0.5, g 1e-5, tau 0.6, sigma 0.002, single
  mode of amplitude 2 plus 8 random modes of 0.2, seed 0) with FP32 and FP16-deviation storage,
  100 steps each.

Stored per fixture: ``sha_<k>`` (11 sha256 strings), ``sum_<k>`` (float64 field sums), ``idx`` and
``val_<k>`` (256 sampled cells), ``meta`` (JSON). Runs here only (the reference tree is read).

Usage: python tests/golden/make_size_golden.py [--ref /root/reference] [--only NAME]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
ORACLE = os.path.join(ROOT, "oracle")
OUT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ORACLE)
sys.path.insert(0, OUT)

from make_golden import REF_FILES, sha  # noqa: E402

BUBBLE = dict(tau=0.8, sigma=0.01, beta=0.7)

# name, binary, width arg, steps, oracle params, drop radius
REF_CASES = [
    ("size_bubble128_q19", "ref_bubble128_q19", 4, [100], dict(nx=128, ny=128, nz=128, q=19, **BUBBLE), 24.0),
    ("size_bubble128_q27", "ref_bubble128_q27", 4, [100], dict(nx=128, ny=128, nz=128, q=27, **BUBBLE), 24.0),
    ("size_bubble64_q19", "ref_bubble64_q19", 4, [100, 1000], dict(nx=64, ny=64, nz=64, q=19, **BUBBLE), 12.0),
]

JET = dict(nx=128, ny=256, nz=128, q=27, ybc=1, outlet=1, tau=0.55, sigma=0.005, beta=0.7, jet_radius=8.0,
           jet_velocity=0.05, jet_noise=0.01, seed=0)
RTI = dict(nx=128, ny=512, nz=128, q=19, ybc=2, tau=0.6, sigma=0.002, beta=0.7, gravity=1e-5, atwood=0.5)
RTI_INIT = dict(amplitude=2.0, width=4.0, n_random_modes=8, random_amplitude=0.2, seed=0)

# name, params, initialiser, storage, steps
ORACLE_CASES = [
    ("size_jet_q27_128x256x128", JET, ("jet", {}), "f32", [100]),
    ("size_rti_q19_128x512x128_f32", RTI, ("rti", RTI_INIT), "f32", [100]),
    ("size_rti_q19_128x512x128_f16", RTI, ("rti", RTI_INIT), "f16", [100]),
]


def record(state: np.ndarray, idx: np.ndarray, k: int, arrays: dict):
    flat = state.reshape(11, -1)
    arrays[f"sha_{k}"] = np.array([sha(flat[f]) for f in range(11)])
    arrays[f"sum_{k}"] = flat.astype(np.float64).sum(axis=1)
    arrays[f"val_{k}"] = flat[:, idx]


def sample(ncell: int) -> np.ndarray:
    rng = np.random.default_rng(12345)
    return np.sort(rng.choice(ncell, size=256, replace=False)).astype(np.int64)


def main() -> int:
    import cgref

    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", default=None)
    args = ap.parse_args()
    subprocess.run(["make", "-C", ORACLE, f"REF={args.ref}", "-j8", "ref"], check=True)
    subprocess.run(["make", "-C", ORACLE], check=True)
    ref_hash = {}
    for f in REF_FILES:
        with open(os.path.join(args.ref, f), "rb") as fh:
            ref_hash[f] = hashlib.sha256(fh.read()).hexdigest()
    with tempfile.TemporaryDirectory() as tmp:
        for name, binary, width, steps, params, radius in REF_CASES:
            if args.only and name != args.only:
                continue
            prefix = os.path.join(tmp, name)
            subprocess.run([os.path.join(ORACLE, "_ref", binary), "bubble_tanh", str(width),
                            ",".join(map(str, steps)), prefix], check=True, stdout=subprocess.DEVNULL)
            p = cgref.Params(**params)
            idx = sample(p.nx * p.nyt * p.nz)
            arrays = {"idx": idx}
            for k in [0] + steps:
                a = np.fromfile(f"{prefix}_step{k}.bin", dtype=np.float32).reshape(p.shape)
                record(a, idx, k, arrays)
                os.remove(f"{prefix}_step{k}.bin")
            # the oracle on the same case, from its own initialiser
            a = cgref.init_bubble(p, radius=radius, width=float(width))
            matches = [list(arrays["sha_0"]) == [sha(f) for f in a.reshape(11, -1)]]
            done = 0
            for k in steps:
                a = cgref.run(p, a, k - done, first_step=done + 1)
                done = k
                matches.append(list(arrays[f"sha_{k}"]) == [sha(f) for f in a.reshape(11, -1)])
            meta = dict(name=name, source="reference", binary=binary, mode="bubble_tanh", width=width, radius=radius,
                        steps=steps, params=params, oracle_matches=all(matches), reference_sha256=ref_hash)
            if not all(matches):
                print(f"{name}: the oracle differs from the reference at snapshots {matches}", file=sys.stderr)
                return 1
            arrays["meta"] = np.array(json.dumps(meta))
            np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
            print("wrote", name, file=sys.stderr, flush=True)
    for name, params, (kind, init_args), storage, steps in ORACLE_CASES:
        if args.only and name != args.only:
            continue
        p = cgref.Params(**params)
        a = cgref.init_jet(p) if kind == "jet" else cgref.init_rti(p, **init_args)
        idx = sample(p.nx * p.nyt * p.nz)
        arrays = {"idx": idx}
        run = cgref.run_f16 if storage == "f16" else cgref.run
        record(cgref.quantize_f16(a) if storage == "f16" else a, idx, 0, arrays)
        done = 0
        for k in steps:
            a = run(p, a, k - done, first_step=done + 1)
            done = k
            record(a, idx, k, arrays)
        meta = dict(name=name, source="oracle", init=kind, init_args=init_args, storage=storage, steps=steps,
                    params=params)
        arrays["meta"] = np.array(json.dumps(meta))
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
        print("wrote", name, file=sys.stderr, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
