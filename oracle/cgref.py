"""TEST INFRASTRUCTURE ONLY (oracle).

ctypes binding of liboracle.so, the clean-room CPU restatement of the reference update
(see cgref.cpp for the file:line map). Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module, and only as the checker.

State arrays are numpy float32 (or float64) arrays of shape (11, nz, nyt, nx): fields
rhor rhob ux uy uz Pixx Pixy Piyy Piyz Pizz Pixz, x fastest; ``nyt = ny`` for a periodic y
axis, ``ny + 2`` when the y axis carries the reference's ghost planes (jet inlet/outlet,
RTI walls).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field, fields

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

Y_PERIODIC, Y_JET, Y_WALLS = 0, 1, 2
OUTLET_CONVECTIVE, OUTLET_ZERO_GRADIENT = 0, 1
FIELDS = ("rhor", "rhob", "ux", "uy", "uz", "Pixx", "Pixy", "Piyy", "Piyz", "Pizz", "Pixz")


class _Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("q", ctypes.c_int32), ("ybc", ctypes.c_int32), ("outlet", ctypes.c_int32),
        ("tau", ctypes.c_float), ("sigma", ctypes.c_float), ("beta", ctypes.c_float),
        ("jet_x0", ctypes.c_float), ("jet_z0", ctypes.c_float),
        ("jet_radius", ctypes.c_float), ("jet_velocity", ctypes.c_float),
        ("jet_noise", ctypes.c_float), ("seed", ctypes.c_uint32),
        ("gravity", ctypes.c_float), ("atwood", ctypes.c_float),
        ("conservative_omega2", ctypes.c_int32),
    ]


@dataclass
class Params:
    nx: int
    ny: int
    nz: int
    q: int = 19
    ybc: int = Y_PERIODIC
    outlet: int = OUTLET_CONVECTIVE
    tau: float = 0.8
    sigma: float = 0.01
    beta: float = 0.7
    jet_x0: float = field(default=None)  # type: ignore[assignment]
    jet_z0: float = field(default=None)  # type: ignore[assignment]
    jet_radius: float = 8.0
    jet_velocity: float = 0.05
    jet_noise: float = 0.0
    seed: int = 0
    gravity: float = 0.0
    atwood: float = 0.0
    conservative_omega2: int = 0

    def __post_init__(self):
        if self.jet_x0 is None:
            self.jet_x0 = (self.nx - 1) / 2.0
        if self.jet_z0 is None:
            self.jet_z0 = (self.nz - 1) / 2.0

    @property
    def nyt(self) -> int:
        return self.ny if self.ybc == Y_PERIODIC else self.ny + 2

    @property
    def shape(self):
        return (11, self.nz, self.nyt, self.nx)

    def c(self) -> _Params:
        return _Params(*[getattr(self, f.name) for f in fields(self)])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(_Params)
        fpp = ctypes.POINTER(ctypes.c_void_p)
        for name in ("cgref_step_f32", "cgref_step_f64", "cgref_apply_ybc_f32", "cgref_apply_ybc_f64"):
            getattr(L, name).argtypes = [P, fpp, fpp, ctypes.c_int64]
            getattr(L, name).restype = ctypes.c_int
        L.cgref_init_bubble_f32.argtypes = [P, fpp] + [ctypes.c_double] * 5 + [ctypes.c_int32]
        L.cgref_init_jet_f32.argtypes = [P, fpp]
        L.cgref_init_rti_f32.argtypes = [P, fpp, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                         ctypes.c_double, ctypes.c_uint32]
        L.cgref_reduce_f32.argtypes = [P, fpp, ctypes.POINTER(ctypes.c_double)]
        L.cgref_constants_f32.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
        _lib = L
    return _lib


def _ptrs(a: np.ndarray):
    assert a.flags.c_contiguous and a.shape[0] == 11
    arr = (ctypes.c_void_p * 11)(*[a[f].ctypes.data for f in range(11)])
    return arr


def _check(rc: int, what: str):
    if rc != 0:
        raise ValueError(f"{what} failed with code {rc}")


def zeros(p: Params, dtype=np.float32) -> np.ndarray:
    return np.zeros(p.shape, dtype=dtype)


def step(p: Params, A: np.ndarray, B: np.ndarray | None = None, step_index: int = 1) -> np.ndarray:
    """One update A -> B (returns B). float32 or float64 arrays."""
    if B is None:
        B = np.zeros_like(A)
    pc = p.c()
    fn = lib().cgref_step_f32 if A.dtype == np.float32 else lib().cgref_step_f64
    _check(fn(ctypes.byref(pc), _ptrs(A), _ptrs(B), step_index), "cgref_step")
    return B


def run(p: Params, A: np.ndarray, nsteps: int, first_step: int = 1) -> np.ndarray:
    """nsteps updates with ping-pong buffers; returns the final state (new array)."""
    a = A.copy()
    b = np.zeros_like(a)
    for k in range(nsteps):
        step(p, a, b, first_step + k)
        a, b = b, a
    return a


# FP16-deviation storage: rhor/rhob kept as float32; fields 2..4 (u) stored as binary16 of u * 2^10,
# fields 5..10 (Pi) as binary16 of Pi * 2^12, rounded to nearest even (numpy's float32 -> float16
# cast). Decoding multiplies back by the exact inverse power of two.
F16_SCALE = np.array([1, 1] + [1024] * 3 + [4096] * 6, dtype=np.float32)


def quantize_f16(A: np.ndarray) -> np.ndarray:
    """The float32 state a FP16-deviation store would hold for A (decoded back to float32)."""
    A = np.asarray(A, dtype=np.float32)
    out = A.copy()
    with np.errstate(over="ignore"):  # beyond the stored range a value becomes +-inf, as on the device
        for f in range(2, 11):
            s = F16_SCALE[f]
            out[f] = (A[f] * s).astype(np.float16).astype(np.float32) / s
    return out


def run_f16(p: Params, A: np.ndarray, nsteps: int, first_step: int = 1) -> np.ndarray:
    """nsteps float32 updates, each followed by the FP16-deviation storage rounding."""
    a = quantize_f16(A)
    b = np.zeros_like(a)
    for k in range(nsteps):
        step(p, a, b, first_step + k)
        a, b = quantize_f16(b), a
    return a


def apply_ybc(p: Params, B: np.ndarray, Aold: np.ndarray, step_index: int = 0):
    pc = p.c()
    fn = lib().cgref_apply_ybc_f32 if B.dtype == np.float32 else lib().cgref_apply_ybc_f64
    _check(fn(ctypes.byref(pc), _ptrs(B), _ptrs(Aold), step_index), "cgref_apply_ybc")


def init_bubble(p: Params, radius: float, width: float = 4.0, sharp: bool = False, center=None) -> np.ndarray:
    A = zeros(p)
    if center is None:
        center = ((p.nx - 1) / 2.0, (p.ny - 1) / 2.0, (p.nz - 1) / 2.0)
    pc = p.c()
    _check(lib().cgref_init_bubble_f32(ctypes.byref(pc), _ptrs(A), *center, radius, width, int(sharp)),
           "cgref_init_bubble")
    return A


def init_jet(p: Params) -> np.ndarray:
    A = zeros(p)
    pc = p.c()
    _check(lib().cgref_init_jet_f32(ctypes.byref(pc), _ptrs(A)), "cgref_init_jet")
    return A


def init_rti(p: Params, amplitude=2.0, width=4.0, n_random_modes=0, random_amplitude=0.0, seed=0) -> np.ndarray:
    A = zeros(p)
    pc = p.c()
    _check(lib().cgref_init_rti_f32(ctypes.byref(pc), _ptrs(A), amplitude, width, n_random_modes,
                                    random_amplitude, seed), "cgref_init_rti")
    return A


def reduce(p: Params, A: np.ndarray) -> np.ndarray:
    out = (ctypes.c_double * 5)()
    pc = p.c()
    _check(lib().cgref_reduce_f32(ctypes.byref(pc), _ptrs(A), out), "cgref_reduce")
    return np.array(out[:], dtype=np.float64)


def constants(tau: float, sigma: float) -> np.ndarray:
    out = (ctypes.c_float * 12)()
    lib().cgref_constants_f32(tau, sigma, out)
    return np.array(out[:], dtype=np.float32)


def interior(p: Params, A: np.ndarray) -> np.ndarray:
    """View of the interior cells (drops the y ghost planes when present)."""
    return A if p.ybc == Y_PERIODIC else A[:, :, 1:-1, :]
