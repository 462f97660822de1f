"""Storage-precision study on the Rayleigh-Taylor case: FP16-deviation storage and FP32 storage (both
with FP32 arithmetic, on the GPU) against an FP64 run of the oracle on the host.

Case: D3Q19, 128 x 512 x 128 cells with the wall-bounded axis y (512 cells, plus the two wall
planes), heavy phase above light, Atwood 0.5 through density-weighted forcing, gravity 1e-5 along
-y. The interface has a single cosine mode of amplitude 2 cells plus 8 seeded random modes of
amplitude 0.2 (seed 0), width 4. tau = 0.6, sigma = 0.002, beta = 0.7. 5000 steps.

At each checkpoint the script keeps three things. First, all 11 fields on the x-y plane through
the middle of z. Second, the per-phase masses. Third, the interface height of every (x, z)
column: the linear-interpolated y where psi = (rr - rb)/(rr + rb) changes sign, lowest crossing.

    python tests/precision_rti.py gpu   OUT.npz [--nx .. --ny .. --nz .. --steps .. --every ..]
    python tests/precision_rti.py cpu64 OUT.npz [same size options]       (FP64 oracle, slow)
    python tests/precision_rti.py compare GPU.npz CPU64.npz               (markdown table)

The GPU leg needs the built library and a GPU; the cpu64 leg needs oracle/liboracle.so. The
oracle is used only as the checker here.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS = ("rhor", "rhob", "ux", "uy", "uz", "Pixx", "Pixy", "Piyy", "Piyz", "Pizz", "Pixz")


def case(nx, ny, nz):
    return dict(nx=nx, ny=ny, nz=nz, q=19, ybc=2, tau=0.6, sigma=0.002, beta=0.7, gravity=1e-5, atwood=0.5)


INIT = dict(amplitude=2.0, width=4.0, n_random_modes=8, random_amplitude=0.2, seed=0)


def interface_height(rr: np.ndarray, rb: np.ndarray) -> np.ndarray:
    """(nz, nx) lowest y (stored row index, fractional) where psi changes sign along +y."""
    psi = (rr.astype(np.float64) - rb) / (rr.astype(np.float64) + rb)
    s = np.sign(psi)
    ch = (s[:, :-1, :] * s[:, 1:, :]) <= 0  # (nz, ny-1, nx)
    first = np.argmax(ch, axis=1)  # first crossing row
    z, x = np.meshgrid(np.arange(psi.shape[0]), np.arange(psi.shape[2]), indexing="ij")
    p0 = psi[z, first, x]
    p1 = psi[z, first + 1, x]
    frac = np.where(p0 != p1, p0 / (p0 - p1), 0.0)
    h = first + frac
    return np.where(ch.any(axis=1), h, np.nan)


def snapshot(state: np.ndarray, nz: int) -> dict:
    """state: (11, nz, nyt, nx) float32/float64."""
    zc = nz // 2
    return {
        "plane": np.array(state[:, zc, :, :], dtype=np.float64),
        "mass": np.array([state[0][:, 1:-1, :].sum(dtype=np.float64), state[1][:, 1:-1, :].sum(dtype=np.float64)]),
        "height": interface_height(state[0], state[1]),
    }


def save(path, meta, snaps):
    arrays = {"meta": np.array(json.dumps(meta))}
    for label, lst in snaps.items():
        for step, sn in lst:
            for k, v in sn.items():
                arrays[f"{label}_{step}_{k}"] = v.astype(np.float32) if k == "plane" else v
    np.savez_compressed(path, **arrays)


def run_gpu(args):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import load_product
    cg = load_product()
    p = case(args.nx, args.ny, args.nz)
    snaps = {}
    timing = {}
    for label, storage in (("f32", cg.STORE_F32), ("f16", cg.STORE_F16)):
        print(f"precision: GPU {label} {args.nx}x{args.ny}x{args.nz} {args.steps} steps", flush=True)
        with cg.Domain(cg.Config(**p), storage=storage) as d:
            d.init_rti(**INIT)
            lst = []
            done, ms = 0, 0.0
            while done < args.steps:
                n = min(args.every, args.steps - done)
                ms += d.time_steps(n)
                done += n
                st = d.get_state()
                lst.append((done, snapshot(st, args.nz)))
                red = d.reduce()
                print(f"  step {done}: mass {red[0]:.9e} {red[1]:.9e} bad {red[5]:.0f}", flush=True)
            snaps[label] = lst
            timing[label] = ms
    meta = dict(case=p, init=INIT, steps=args.steps, every=args.every, gpu_ms=timing)
    save(args.out, meta, snaps)


def run_cpu64(args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cgref
    p = case(args.nx, args.ny, args.nz)
    P = cgref.Params(**p)
    A = cgref.init_rti(P, **INIT).astype(np.float64)
    B = np.zeros_like(A)
    lst = []
    t0 = time.time()
    for k in range(1, args.steps + 1):
        cgref.step(P, A, B, k)
        A, B = B, A
        if k % args.every == 0 or k == args.steps:
            lst.append((k, snapshot(A, args.nz)))
            print(f"precision: FP64 oracle step {k} ({time.time() - t0:.0f} s)", flush=True)
            save(args.out, dict(case=p, init=INIT, steps=k, every=args.every, threads=os.environ.get("OMP_NUM_THREADS")),
                 {"f64": lst})


def compare(gpu_path, cpu_path):
    g = np.load(gpu_path, allow_pickle=False)
    c = np.load(cpu_path, allow_pickle=False)
    meta = json.loads(str(g["meta"]))
    steps = sorted({int(k.split("_")[1]) for k in c.files if k.startswith("f64_")})
    rows = []
    for step in steps:
        ref = c[f"f64_{step}_plane"]
        mref = c[f"f64_{step}_mass"]
        href = c[f"f64_{step}_height"]
        for label in ("f32", "f16"):
            key = f"{label}_{step}_plane"
            if key not in g.files:
                continue
            pl = g[key].astype(np.float64)
            errs = []
            for f in range(11):
                d = pl[f] - ref[f]
                scale = np.sqrt(np.mean(ref[f] ** 2))
                errs.append((np.sqrt(np.mean(d ** 2)) / scale if scale > 0 else 0.0, np.abs(d).max()))
            mass = g[f"{label}_{step}_mass"]
            dm = (mass - mref) / mref
            h = g[f"{label}_{step}_height"]
            dh = np.nanmax(np.abs(h - href))
            rows.append((step, label, errs, dm, dh, np.nanmax(href) - np.nanmin(href)))
    print(f"Case {meta['case']}, init {meta['init']}")
    print()
    print("| step | storage | " + " | ".join(FIELDS) + " | mass rr, rb (rel) | interface height max abs (cells) |")
    print("|---" * (len(FIELDS) + 5) + "|")
    for step, label, errs, dm, dh, amp in rows:
        cells = " | ".join(f"{e:.1e}" for e, _ in errs)
        print(f"| {step} | {label} | {cells} | {dm[0]:+.1e}, {dm[1]:+.1e} | {dh:.2e} (bubble-spike {amp:.2f}) |")
    print()
    print("Per-field entries: RMS difference from the FP64 run over the mid-z x-y plane, relative to the")
    print("field's RMS there. Max abs differences per field:")
    for step, label, errs, _, _, _ in rows:
        print(f"  step {step} {label}: " + ", ".join(f"{FIELDS[f]} {errs[f][1]:.1e}" for f in range(11)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["gpu", "cpu64", "compare"])
    ap.add_argument("out")
    ap.add_argument("other", nargs="?")
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--every", type=int, default=1000)
    args = ap.parse_args()
    if args.mode == "gpu":
        run_gpu(args)
    elif args.mode == "cpu64":
        run_cpu64(args)
    else:
        compare(args.out, args.other)


if __name__ == "__main__":
    main()
