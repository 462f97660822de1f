"""FP16-deviation and FP32 storage against FP64, per workload: writes
profiles/precision_fp16_vs_fp64.json, the per-result error figures that bench.py attaches to its
FP16 results (the bench's FP16 128^3 leg, and the case kinds of the serving mix).

Each workload runs three times from the same initial state with the oracle on the host: FP32
arithmetic with FP16-deviation storage (``run_f16``, the storage model the GPU path equals bit for
bit: tests/test_gpu_f16.py), FP32 arithmetic with FP32 storage (``run``, equal to the GPU path and
to the reference bit for bit), and FP64 (``run`` on float64 arrays). Reported per field: RMS over
the interior of (result - FP64) / RMS of the FP64 field, and the per-phase mass change from the
initial state. The Rayleigh-Taylor case at full size (128 x 512 x 128, 5000 steps) is the study of
tests/precision_rti.py; its summary is carried over from profiles/precision_rti_128x512x128.md.

    OMP_NUM_THREADS=8 python tests/precision_table.py [--only NAME]

Runs on the CPU only (about 15 minutes on 8 threads). The oracle is the checker here.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
OUT = os.path.join(ROOT, "profiles", "precision_fp16_vs_fp64.json")
FIELDS = ("rhor", "rhob", "ux", "uy", "uz", "Pixx", "Pixy", "Piyy", "Piyz", "Pizz", "Pixz")

WORKLOADS = {
    # the bench's FP16 leg: D3Q19 128^3 drop (R 24, width 4)
    "drop128_d3q19": dict(params=dict(nx=128, ny=128, nz=128, q=19, tau=0.8, sigma=0.01, beta=0.7),
                          init=("bubble", dict(radius=24.0, width=4.0)), steps=200),
    # serving mix case kinds at their own parameter ranges (mid-range draws)
    "serving_bubble_d3q19_48": dict(params=dict(nx=48, ny=48, nz=48, q=19, tau=0.76, sigma=0.0045, beta=0.7),
                                    init=("bubble", dict(radius=12.0, width=4.0)), steps=500),
    "serving_jet_d3q27_64x128x64": dict(params=dict(nx=64, ny=128, nz=64, q=27, ybc=1, outlet=1, tau=0.55,
                                                    sigma=0.005, beta=0.7, jet_radius=8.0, jet_velocity=0.04,
                                                    jet_noise=0.01, seed=7),
                                        init=("jet", {}), steps=500),
    "serving_rti_d3q19_64x256x64": dict(params=dict(nx=64, ny=256, nz=64, q=19, ybc=2, tau=0.6, sigma=0.002,
                                                    beta=0.7, gravity=1e-5, atwood=0.4),
                                        init=("rti", dict(amplitude=2.0, width=4.0, n_random_modes=8,
                                                          random_amplitude=0.2, seed=1)), steps=500),
}

RTI_FULL = {
    "source": "profiles/precision_rti_128x512x128.md (tests/precision_rti.py, GPU runs against an FP64 oracle "
              "run; mid-z x-y plane)",
    "steps": 5000,
    "f16": {"rhor": 1.5e-4, "rhob": 1.5e-4, "ux": 1.7e-2, "uy": 1.8e-2, "uz": 5.8e-2, "Pixx": 3.6e-3,
            "Pixy": 1.2e-2, "Piyy": 2.7e-2, "Piyz": 1.7e-1, "Pizz": 3.2e-3, "Pixz": 1.6e-1},
    "f32": {"rhor": 2.0e-4, "rhob": 2.0e-4, "ux": 4.7e-3, "uy": 9.6e-3, "uz": 5.7e-2, "Pixx": 3.5e-3,
            "Pixy": 4.8e-3, "Piyy": 1.3e-2, "Piyz": 6.3e-2, "Pizz": 3.5e-3, "Pixz": 1.1e-1},
}


def rel_rms(p, x, ref):
    import cgref
    x = cgref.interior(p, x).reshape(11, -1).astype(np.float64)
    r = cgref.interior(p, ref).reshape(11, -1)
    return np.sqrt(((x - r) ** 2).mean(axis=1)) / np.sqrt((r ** 2).mean(axis=1))


def mass(p, a):
    import cgref
    s = cgref.interior(p, a).reshape(11, -1)[:2].astype(np.float64).sum(axis=1)
    return s


def drift(m, m0):
    """Relative per-phase mass change; None for a phase that starts empty (the jet's injected one)."""
    return [float(f"{(a - b) / b:.3g}") if b > 0 else None for a, b in zip(m, m0)]


def main() -> int:
    import cgref

    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None)
    args = ap.parse_args()
    out = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            out = json.load(f)
    out["note"] = ("relative RMS error per field against an FP64 run from the same initial state "
                   "(interior cells); mass_drift = per-phase relative mass change from the initial state")
    out["rti_128x512x128_5000"] = RTI_FULL
    for name, w in WORKLOADS.items():
        if args.only and name != args.only:
            continue
        p = cgref.Params(**w["params"])
        kind, kw = w["init"]
        A = (cgref.init_bubble(p, **kw) if kind == "bubble" else cgref.init_jet(p) if kind == "jet"
             else cgref.init_rti(p, **kw))
        t0 = time.time()
        ref = cgref.run(p, A.astype(np.float64), w["steps"])
        res = {"params": w["params"], "init": [kind, kw], "steps": w["steps"]}
        m0 = mass(p, A)
        for label, fn in (("f32", cgref.run), ("f16", cgref.run_f16)):
            a = fn(p, A, w["steps"])
            e = rel_rms(p, a, ref)
            res[label] = {f: float(f"{v:.3g}") for f, v in zip(FIELDS, e)}
            res[label]["mass_drift"] = drift(mass(p, a), m0)
        res["fp64_mass_drift"] = drift(mass(p, ref), m0)
        out[name] = res
        print(f"{name}: {time.time() - t0:.0f} s; f16 {res['f16']}", file=sys.stderr, flush=True)
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
