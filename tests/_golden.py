"""Loader for the golden vectors under tests/golden/ (written from the reference's own update by
tests/golden/make_golden.py). Shared by the oracle tests and the GPU parity tests."""
from __future__ import annotations

import glob
import hashlib
import json
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    """State fixtures of the small cases (the VTI fixtures, vti_*.npz, and the full-size ones,
    size_*.npz, have their own formats and tests)."""
    return sorted(n for n in (os.path.splitext(os.path.basename(p))[0]
                              for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))
                  if not n.startswith(("vti_", "size_")))


def size_names():
    """Full-size fixtures (tests/golden/make_size_golden.py): per-field sha256 of the benched and
    configured workloads at their real sizes."""
    return sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN_DIR, "size_*.npz")))


def load(name: str):
    z = np.load(os.path.join(GOLDEN_DIR, f"{name}.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return z, meta


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


def field_shas(state: np.ndarray):
    flat = state.reshape(11, -1)
    return [sha(flat[f]) for f in range(11)]


def initial_state(cgref, z, meta):
    """The case's input state, rebuilt with the oracle's initialisers (their sha256 is checked
    against the reference's) or taken from the fixture for random inputs."""
    p = cgref.Params(**meta["params"])
    mode = meta["mode"]
    if mode == "jet":
        return p, cgref.init_jet(p)
    if mode == "random":
        return p, np.array(z["full_0"], dtype=np.float32)
    if mode == "bubble_tanh":
        return p, cgref.init_bubble(p, radius=6.0, width=float(meta["arg"]))
    if mode == "bubble_sharp":
        return p, cgref.init_bubble(p, radius=6.0, sharp=True)
    if mode == "rti_init":
        return p, cgref.init_rti(p, amplitude=2.0, width=4.0)
    raise ValueError(mode)
