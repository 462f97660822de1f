"""Write the golden vectors under tests/golden/ from the reference's OWN update.

Runs in the container that holds the reference tree (never on the GPU box): builds the host
harness ``oracle/ref_harness.cpp`` against the reference headers (``make -C oracle ref``), runs
each case, and stores compact fixtures:

* ``meta``       JSON: case name, harness build, oracle parameters, snapshot steps, sha256 of every
                 reference file compiled in;
* ``sha_<k>``    sha256 of each of the 11 fields (float32 bytes, compact layout) after step k;
* ``sum_<k>``    float64 sum of each field after step k;
* ``idx``        fixed sample of flat cell indices; ``val_<k>`` the 11 fields at those cells;
* ``full_<k>``   the whole state, for the small random-moment cases (their input included).

VTI fixtures (``vti_<case>.npz``): the reference's own writer (io/vtiWriter.cuh:92-145, compiled into
the harness) prints a snapshot over its full grid, ghost layers included. Stored: the five printed
fields over that full extent (``full5``), the sha256 and byte length of the file it wrote.

Usage: python tests/golden/make_golden.py [--ref /root/reference]
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

REF_FILES = [
    "lbm.cuh", "pdf.cuh", "forces.cuh", "collisionOperators.cuh", "stencil_ct.cuh", "constants.cuh",
    "memory.cuh", "utilities/bounds.cuh", "utilities/indexing.cuh", "utilities/mathUtilities.cuh",
    "utilities/constexprFor.cuh", "utilities/types.cuh", "jet/boundary_conditions.cuh", "jet/jetIn.cuh",
    "bubble/bubbleIn.cuh", "rti/rtiIn.cuh", "io/vtiWriter.cuh",
]

SMALL = dict(tau=0.8, sigma=0.01, beta=0.7, jet_velocity=0.05)
JETP = dict(tau=0.55, sigma=0.005, beta=0.7, jet_velocity=0.05)

# name, binary, mode, arg, steps, oracle params, store full arrays
CASES = [
    ("default_q27_jet", "ref_default_q27", "jet", 0, [1, 2],
     dict(nx=126, ny=126, nz=126, q=27, ybc=1, tau=1.0, sigma=0.1, beta=0.9, jet_radius=12.0,
          jet_velocity=0.05), False),
    ("default_q19_jet", "ref_default_q19", "jet", 0, [1, 2],
     dict(nx=126, ny=126, nz=126, q=19, ybc=1, tau=1.0, sigma=0.1, beta=0.9, jet_radius=12.0,
          jet_velocity=0.05), False),
    ("jet_small_q19", "ref_jet_small_q19", "jet", 0, [1, 10, 100],
     dict(nx=16, ny=32, nz=16, q=19, ybc=1, jet_radius=4.0, **SMALL), False),
    ("jet_small_q27", "ref_jet_small_q27", "jet", 0, [1, 10, 100],
     dict(nx=16, ny=32, nz=16, q=27, ybc=1, jet_radius=4.0, **SMALL), False),
    ("rand_q19", "ref_rand_q19", "random", 1, [1, 2, 5],
     dict(nx=10, ny=12, nz=8, q=19, ybc=1, jet_radius=3.0, **SMALL), True),
    ("rand_q27", "ref_rand_q27", "random", 1, [1, 2, 5],
     dict(nx=10, ny=12, nz=8, q=27, ybc=1, jet_radius=3.0, **SMALL), True),
    ("rand55_q19", "ref_rand55_q19", "random", 2, [1, 2, 5],
     dict(nx=10, ny=12, nz=8, q=19, ybc=1, jet_radius=3.0, **JETP), True),
    ("rand55_q27", "ref_rand55_q27", "random", 2, [1, 2, 5],
     dict(nx=10, ny=12, nz=8, q=27, ybc=1, jet_radius=3.0, **JETP), True),
    ("bubble_q19", "ref_bubble_q19", "bubble_tanh", 4, [1, 10, 100],
     dict(nx=24, ny=24, nz=24, q=19, ybc=0, **SMALL), False),
    ("bubble_q27", "ref_bubble_q27", "bubble_tanh", 4, [1, 10, 100],
     dict(nx=24, ny=24, nz=24, q=27, ybc=0, **SMALL), False),
    ("bubble_sharp_q19", "ref_bubble_q19", "bubble_sharp", 0, [1, 10, 50],
     dict(nx=24, ny=24, nz=24, q=19, ybc=0, **SMALL), False),
    ("rti_init", "ref_rti_init", "rti_init", 0, [],
     dict(nx=32, ny=64, nz=32, q=19, ybc=2, **SMALL), True),
]


# name, binary, mode, arg, step, oracle params
VTI_CASES = [
    ("vti_rand_q19", "ref_rand_q19", "random", 1, 2,
     dict(nx=10, ny=12, nz=8, q=19, ybc=1, jet_radius=3.0, **SMALL)),
    ("vti_jet_small_q19", "ref_jet_small_q19", "jet", 0, 10,
     dict(nx=16, ny=32, nz=16, q=19, ybc=1, jet_radius=4.0, **SMALL)),
]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", default=None)
    args = ap.parse_args()
    subprocess.run(["make", "-C", ORACLE, f"REF={args.ref}", "-j8", "ref"], check=True)
    ref_hash = {}
    for f in REF_FILES:
        with open(os.path.join(args.ref, f), "rb") as fh:
            ref_hash[f] = hashlib.sha256(fh.read()).hexdigest()
    with tempfile.TemporaryDirectory() as tmp:
        for name, binary, mode, arg, steps, params, full in CASES:
            if args.only and name != args.only:
                continue
            prefix = os.path.join(tmp, name)
            cmd = [os.path.join(ORACLE, "_ref", binary), mode, str(arg), ",".join(map(str, steps)) or "0", prefix]
            subprocess.run(cmd, check=True)
            nx, ny, nz = params["nx"], params["ny"], params["nz"]
            nyt = ny if params["ybc"] == 0 else ny + 2
            shape = (11, nz, nyt, nx)
            ncell = nx * nyt * nz
            rng = np.random.default_rng(12345)
            idx = np.sort(rng.choice(ncell, size=min(256, ncell), replace=False)).astype(np.int64)
            arrays = {"idx": idx}
            for k in [0] + steps:
                a = np.fromfile(f"{prefix}_step{k}.bin", dtype=np.float32).reshape(shape)
                flat = a.reshape(11, -1)
                arrays[f"sha_{k}"] = np.array([sha(flat[f]) for f in range(11)])
                arrays[f"sum_{k}"] = flat.astype(np.float64).sum(axis=1)
                arrays[f"val_{k}"] = flat[:, idx]
                if full and k in (0, steps[0] if steps else 0):
                    arrays[f"full_{k}"] = a
            meta = dict(name=name, binary=binary, mode=mode, arg=arg, steps=steps, params=params,
                        reference_sha256=ref_hash)
            arrays["meta"] = np.array(json.dumps(meta))
            np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
            print("wrote", name, file=sys.stderr)
        for name, binary, mode, arg, step, params in VTI_CASES:
            if args.only and name != args.only:
                continue
            prefix = os.path.join(tmp, name)
            cmd = [os.path.join(ORACLE, "_ref", binary), mode, str(arg), str(step), prefix]
            subprocess.run(cmd, check=True, env=dict(os.environ, REF_VTI="1"), stdout=subprocess.DEVNULL)
            full = (params["nx"] + 2, params["ny"] + 2, params["nz"] + 2)
            with open(f"{prefix}_step{step}.bin.vti", "rb") as fh:
                text = fh.read()
            f5 = np.fromfile(f"{prefix}_step{step}.bin.full5", dtype=np.float32).reshape(
                5, full[2], full[1], full[0])
            meta = dict(name=name, binary=binary, mode=mode, arg=arg, step=step, params=params,
                        reference_sha256=ref_hash)
            np.savez_compressed(os.path.join(OUT, f"{name}.npz"), full5=f5,
                                vti_sha256=np.array(hashlib.sha256(text).hexdigest()),
                                vti_bytes=np.array(len(text)), meta=np.array(json.dumps(meta)))
            print("wrote", name, file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
