"""cglbm — MI355X colour-gradient lattice Boltzmann update (Python side of the C ABI).

Thin ctypes binding of ``libcglbm.so`` (built in this directory from ``csrc/``; declarations in
``include/cglbm.h``). It mirrors the reference's host interface for the solver update:

* :class:`Moments`  — the eleven device arrays of ``MomentsDevice`` (memory.cuh:14-29);
* :func:`launch_rcs` / :func:`launch_ybc` — ``launch_RCS`` / ``launch_JetBoundaryConditions``
  (launch.cuh:78-94);
* :class:`Domain`   — ``LbmDevice`` ownership plus the ``main.cu:38-55`` step loop, the case
  initialisers, ``copy_out_D2H``, ``write_vti_step_device`` and the MLUPS timer.

Errors raise :class:`CglbmError` carrying the library's message. There is no CPU fallback: if
the library is missing or has no GPU, every call fails loudly.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field, fields

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libcglbm.so")
FIELDS = ("rhor", "rhob", "ux", "uy", "uz", "Pixx", "Pixy", "Piyy", "Piyz", "Pizz", "Pixz")

Y_PERIODIC, Y_JET, Y_WALLS = 0, 1, 2
OUTLET_CONVECTIVE, OUTLET_ZERO_GRADIENT = 0, 1
STORE_F32, STORE_F16 = 0, 1
OK, EINVAL, ENOMEM, EHIP, EIO, ESTATE = 0, -1, -2, -3, -4, -5
KERNEL_AUTO, KERNEL_GATHER, KERNEL_MARCH = 0, 1, 2

# Every symbol include/cglbm.h declares (checked by the CPU test suite).
EXPORTS = (
    "cglbm_last_error", "cglbm_abi_version", "cglbm_compact_layout", "cglbm_launch_rcs", "cglbm_launch_ybc",
    "cglbm_domain_create", "cglbm_domain_destroy", "cglbm_domain_moments", "cglbm_domain_config",
    "cglbm_domain_bytes", "cglbm_domain_set_kernel", "cglbm_domain_set_tile", "cglbm_domain_autotune",
    "cglbm_domain_step_count", "cglbm_domain_stream", "cglbm_domain_plan",
    "cglbm_plan_create", "cglbm_plan_destroy", "cglbm_plan_reset", "cglbm_launch_rcs_planned",
    "cglbm_launch_init_bubble", "cglbm_launch_init_jet", "cglbm_launch_init_rti",
    "cglbm_launch_rcs_planned_noalloc", "cglbm_launch_rcs_zslab", "cglbm_launch_rcs_zslab_faces", "cglbm_zslab_pack",
    "cglbm_zslab_unpack",
    "cglbm_init_bubble", "cglbm_init_jet", "cglbm_init_rti", "cglbm_step", "cglbm_time_steps", "cglbm_sync",
    "cglbm_copy_out", "cglbm_copy_in", "cglbm_reduce", "cglbm_write_vti", "cglbm_write_vti_host",
    "cglbm_snapshot_write", "cglbm_snapshot_read", "cglbm_snapshot_info",
    "cglbm_batch_create", "cglbm_batch_destroy", "cglbm_batch_add", "cglbm_batch_remove", "cglbm_batch_size",
    "cglbm_batch_set_tile", "cglbm_batch_set_graph", "cglbm_batch_step", "cglbm_batch_time_steps",
    "cglbm_batch_time_sequential", "cglbm_batch_autotune", "cglbm_batch_sync",
    "cglbm_server_create", "cglbm_server_submit", "cglbm_server_poll", "cglbm_server_drain",
    "cglbm_server_stats_get", "cglbm_server_clock", "cglbm_server_destroy",
)

CASE_BUBBLE, CASE_JET, CASE_RTI, CASE_STATE = 0, 1, 2, 3


class CglbmError(RuntimeError):
    pass


class Moments(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in FIELDS]


class Layout(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_int64), ("pitch_y", ctypes.c_int64), ("pitch_z", ctypes.c_int64)]


class _Config(ctypes.Structure):
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


class _Request(ctypes.Structure):
    _fields_ = [
        ("cfg", _Config), ("storage", ctypes.c_int32), ("case_kind", ctypes.c_int32),
        ("radius", ctypes.c_double), ("width", ctypes.c_double), ("sharp", ctypes.c_int32),
        ("amplitude", ctypes.c_double), ("n_random_modes", ctypes.c_int32),
        ("random_amplitude", ctypes.c_double), ("init_seed", ctypes.c_uint32),
        ("initial", ctypes.c_void_p), ("steps", ctypes.c_int64), ("result_mask", ctypes.c_uint32),
        ("result", ctypes.c_void_p), ("user_tag", ctypes.c_uint64),
    ]


class _Result(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64), ("user_tag", ctypes.c_uint64), ("status", ctypes.c_int32),
        ("storage", ctypes.c_int32), ("steps", ctypes.c_int64), ("cells", ctypes.c_int64),
        ("diag", ctypes.c_double * 6), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
        ("t_start", ctypes.c_double), ("t_done", ctypes.c_double), ("message", ctypes.c_char * 160),
    ]


class _ServerOpts(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_int64), ("max_members", ctypes.c_int32), ("bundle", ctypes.c_int32),
        ("admit_threads", ctypes.c_int32), ("tile_ty", ctypes.c_int32), ("tile_L", ctypes.c_int32),
        ("tile_nbuf", ctypes.c_int32),
    ]


class _ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "submitted", "admitted", "completed", "failed", "queued", "resident", "bytes_reserved",
        "peak_bytes_reserved", "budget_bytes", "peak_resident", "updates", "bundles")] + [
        ("busy_seconds", ctypes.c_double)]


@dataclass
class Config:
    """Runtime form of constants.cuh:9-93 (field meanings in include/cglbm.h)."""
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

    @property
    def interior_cells(self) -> int:
        return self.nx * self.ny * self.nz

    def c(self) -> _Config:
        return _Config(*[getattr(self, f.name) for f in fields(self)])


_lib = None


def lib():
    """Load libcglbm.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CglbmError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or `make -C` the package")
    L = ctypes.CDLL(LIB_PATH)
    C = ctypes.POINTER(_Config)
    M = ctypes.POINTER(Moments)
    LP = ctypes.POINTER(Layout)
    D = ctypes.c_void_p
    sig = {
        "cglbm_last_error": ([], ctypes.c_char_p),
        "cglbm_abi_version": ([], ctypes.c_int),
        "cglbm_compact_layout": ([C, LP], ctypes.c_int),
        "cglbm_launch_rcs": ([M, M, C, LP, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p], ctypes.c_int),
        "cglbm_launch_ybc": ([M, M, C, LP, ctypes.c_int64, ctypes.c_void_p], ctypes.c_int),
        "cglbm_launch_rcs_zslab": ([M, M, C, LP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                    ctypes.c_void_p], ctypes.c_int),
        "cglbm_launch_rcs_zslab_faces": ([M, M, C, LP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int64, ctypes.c_void_p], ctypes.c_int),
        "cglbm_zslab_pack": ([M, C, LP, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p], ctypes.c_int),
        "cglbm_zslab_unpack": ([M, C, LP, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p], ctypes.c_int),
        "cglbm_launch_init_bubble": ([M, C, LP] + [ctypes.c_double] * 5 + [ctypes.c_int32, ctypes.c_void_p],
                                     ctypes.c_int),
        "cglbm_launch_init_jet": ([M, C, LP, ctypes.c_void_p], ctypes.c_int),
        "cglbm_launch_init_rti": ([M, C, LP, ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                                   ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
        "cglbm_plan_create": ([C, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "cglbm_plan_destroy": ([ctypes.c_void_p], ctypes.c_int),
        "cglbm_plan_reset": ([ctypes.c_void_p], ctypes.c_int),
        "cglbm_launch_rcs_planned": ([M, M, C, LP, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "cglbm_launch_rcs_planned_noalloc": ([M, M, C, LP, ctypes.c_int64, ctypes.c_void_p], ctypes.c_int),
        "cglbm_domain_create": ([C, ctypes.c_int32, ctypes.POINTER(D)], ctypes.c_int),
        "cglbm_domain_destroy": ([D], ctypes.c_int),
        "cglbm_domain_moments": ([D, ctypes.c_int32, M], ctypes.c_int),
        "cglbm_domain_config": ([D, C], ctypes.c_int),
        "cglbm_domain_bytes": ([C, ctypes.c_int32], ctypes.c_int64),
        "cglbm_domain_set_kernel": ([D, ctypes.c_int32], ctypes.c_int),
        "cglbm_domain_set_tile": ([D, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32], ctypes.c_int),
        "cglbm_domain_plan": ([D, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                               ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "cglbm_domain_autotune": ([D, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "cglbm_domain_step_count": ([D], ctypes.c_int64),
        "cglbm_domain_stream": ([D], ctypes.c_void_p),
        "cglbm_init_bubble": ([D] + [ctypes.c_double] * 5 + [ctypes.c_int32], ctypes.c_int),
        "cglbm_init_jet": ([D], ctypes.c_int),
        "cglbm_init_rti": ([D, ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_uint32],
                           ctypes.c_int),
        "cglbm_step": ([D, ctypes.c_int32], ctypes.c_int),
        "cglbm_time_steps": ([D, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "cglbm_sync": ([D], ctypes.c_int),
        "cglbm_copy_out": ([D, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
        "cglbm_copy_in": ([D, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
        "cglbm_reduce": ([D, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "cglbm_write_vti": ([D, ctypes.c_char_p], ctypes.c_int),
        "cglbm_snapshot_write": ([D, ctypes.c_char_p], ctypes.c_int),
        "cglbm_snapshot_read": ([D, ctypes.c_char_p], ctypes.c_int),
        "cglbm_snapshot_info": ([ctypes.c_char_p, C, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)],
                                ctypes.c_int),
        "cglbm_write_vti_host": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5,
                                 ctypes.c_int),
        "cglbm_batch_create": ([ctypes.POINTER(D)], ctypes.c_int),
        "cglbm_batch_destroy": ([D], ctypes.c_int),
        "cglbm_batch_add": ([D, D], ctypes.c_int),
        "cglbm_batch_remove": ([D, D], ctypes.c_int),
        "cglbm_batch_size": ([D], ctypes.c_int32),
        "cglbm_batch_set_tile": ([D, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32], ctypes.c_int),
        "cglbm_batch_set_graph": ([D, ctypes.c_int32], ctypes.c_int),
        "cglbm_batch_step": ([D, ctypes.c_int32], ctypes.c_int),
        "cglbm_batch_time_steps": ([D, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "cglbm_batch_time_sequential": ([D, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "cglbm_batch_autotune": ([D, ctypes.c_int32], ctypes.c_int),
        "cglbm_batch_sync": ([D], ctypes.c_int),
        "cglbm_server_create": ([ctypes.POINTER(_ServerOpts), ctypes.POINTER(D)], ctypes.c_int),
        "cglbm_server_submit": ([D, ctypes.POINTER(_Request), ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "cglbm_server_poll": ([D, ctypes.POINTER(_Result), ctypes.c_int32, ctypes.c_int32,
                               ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "cglbm_server_drain": ([D], ctypes.c_int),
        "cglbm_server_stats_get": ([D, ctypes.POINTER(_ServerStats)], ctypes.c_int),
        "cglbm_server_clock": ([D], ctypes.c_double),
        "cglbm_server_destroy": ([D], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().cglbm_last_error().decode(errors="replace")
        raise CglbmError(f"{what} failed ({rc}): {msg}")


def compact_layout(cfg: Config) -> Layout:
    out = Layout()
    cc = cfg.c()
    _check(lib().cglbm_compact_layout(ctypes.byref(cc), ctypes.byref(out)), "cglbm_compact_layout")
    return out


def launch_rcs(A: Moments, B: Moments, cfg: Config, layout: Layout | None = None, step: int = 1,
               kernel: int = KERNEL_AUTO, stream: int | None = None):
    """``launch_RCS`` + the jet boundary kernel (launch.cuh:78-94) on caller-owned arrays."""
    cc = cfg.c()
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_launch_rcs(ctypes.byref(A), ctypes.byref(B), ctypes.byref(cc), ctypes.byref(L), step, kernel,
                                  stream), "cglbm_launch_rcs")


def launch_rcs_zslab(A: Moments, B: Moments, cfg: Config, z_lo: int, z_hi: int, z_offset: int = 0,
                     layout: Layout | None = None, step: int = 1, stream: int | None = None):
    """One z-slab shard's update (cglbm_launch_rcs_zslab): target planes [z_lo, z_hi) of arrays that
    hold the slab and two halo planes on each side; z_offset maps local planes to global ones."""
    cc = cfg.c()
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_launch_rcs_zslab(ctypes.byref(A), ctypes.byref(B), ctypes.byref(cc), ctypes.byref(L), z_lo,
                                        z_hi, z_offset, step, stream), "cglbm_launch_rcs_zslab")


def launch_rcs_zslab_faces(A: Moments, B: Moments, cfg: Config, z_lo: int, z_hi: int, width: int = 2,
                           z_offset: int = 0, layout: Layout | None = None, step: int = 1, stream: int | None = None):
    """Both faces of a z-slab in one launch (cglbm_launch_rcs_zslab_faces): target planes
    [z_lo, z_lo + width) and [z_hi - width, z_hi)."""
    cc = cfg.c()
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_launch_rcs_zslab_faces(ctypes.byref(A), ctypes.byref(B), ctypes.byref(cc), ctypes.byref(L),
                                              z_lo, z_hi, width, z_offset, step, stream),
           "cglbm_launch_rcs_zslab_faces")


def zslab_pack(A: Moments, cfg: Config, z_lo: int, z_hi: int, send_lo: int, send_hi: int,
               layout: Layout | None = None, stream: int | None = None):
    """Pack the 13 plane-fields each neighbour of a z-slab needs into two device buffers."""
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_zslab_pack(ctypes.byref(A), ctypes.byref(cfg.c()), ctypes.byref(L), z_lo, z_hi, send_lo,
                                  send_hi, stream), "cglbm_zslab_pack")


def zslab_unpack(A: Moments, cfg: Config, z_lo: int, z_hi: int, recv_lo: int, recv_hi: int,
                 layout: Layout | None = None, stream: int | None = None):
    """Write two received device buffers into a z-slab's halo planes."""
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_zslab_unpack(ctypes.byref(A), ctypes.byref(cfg.c()), ctypes.byref(L), z_lo, z_hi, recv_lo,
                                    recv_hi, stream), "cglbm_zslab_unpack")


class Plan:
    """A caller-held chunk plan (cglbm_plan_*) for launch-level calls on one pair of states."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        h = ctypes.c_void_p()
        _check(lib().cglbm_plan_create(ctypes.byref(cfg.c()), ctypes.byref(h)), "cglbm_plan_create")
        self._h = h

    def reset(self):
        _check(lib().cglbm_plan_reset(self._h), "cglbm_plan_reset")

    def close(self):
        if getattr(self, "_h", None):
            lib().cglbm_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def launch_rcs_planned(A: Moments, B: Moments, cfg: Config, plan: Plan, layout: Layout | None = None, step: int = 1,
                       stream: int | None = None):
    """``cglbm_launch_rcs`` with the chunk plan `plan` (balanced z-chunks, same results)."""
    cc = cfg.c()
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_launch_rcs_planned(ctypes.byref(A), ctypes.byref(B), ctypes.byref(cc), ctypes.byref(L), step,
                                          plan._h, stream), "cglbm_launch_rcs_planned")


def launch_rcs_planned_noalloc(A: Moments, B: Moments, cfg: Config, layout: Layout | None = None, step: int = 1,
                               stream: int | None = None):
    """``cglbm_launch_rcs_planned`` with the plan in the library's per-GPU scratch array (no device
    allocation); calls that overlap on one GPU race."""
    cc = cfg.c()
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_launch_rcs_planned_noalloc(ctypes.byref(A), ctypes.byref(B), ctypes.byref(cc), ctypes.byref(L),
                                                  step, stream), "cglbm_launch_rcs_planned_noalloc")


def launch_ybc(B: Moments, Aold: Moments, cfg: Config, layout: Layout | None = None, step: int = 0,
               stream: int | None = None):
    cc = cfg.c()
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_launch_ybc(ctypes.byref(B), ctypes.byref(Aold), ctypes.byref(cc), ctypes.byref(L), step,
                                  stream), "cglbm_launch_ybc")


def launch_init_bubble(A: Moments, cfg: Config, layout: Layout | None = None, radius: float = 8.0,
                       width: float = 4.0, sharp: bool = False, center=None, stream: int | None = None):
    """``launch_InitStaticBubble`` (launch.cuh:14-20) on caller-owned arrays (sharp = the
    reference's sphere, else the diffuse tanh drop); centre defaults to the domain middle."""
    cc = cfg.c()
    L = layout if layout is not None else compact_layout(cfg)
    if center is None:
        center = ((cfg.nx - 1) / 2.0, (cfg.ny - 1) / 2.0, (cfg.nz - 1) / 2.0)
    _check(lib().cglbm_launch_init_bubble(ctypes.byref(A), ctypes.byref(cc), ctypes.byref(L), *center, radius, width,
                                          int(sharp), stream), "cglbm_launch_init_bubble")


def launch_init_jet(A: Moments, cfg: Config, layout: Layout | None = None, stream: int | None = None):
    """``launch_InitJet`` (launch.cuh:52-72) plus main.cu:26's boundary call, on caller-owned arrays."""
    cc = cfg.c()
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_launch_init_jet(ctypes.byref(A), ctypes.byref(cc), ctypes.byref(L), stream),
           "cglbm_launch_init_jet")


def launch_init_rti(A: Moments, cfg: Config, layout: Layout | None = None, amplitude: float = 2.0,
                    width: float = 4.0, n_random_modes: int = 0, random_amplitude: float = 0.0, seed: int = 0,
                    stream: int | None = None):
    """``launch_InitRTI`` (launch.cuh:42-46) on caller-owned arrays."""
    cc = cfg.c()
    L = layout if layout is not None else compact_layout(cfg)
    _check(lib().cglbm_launch_init_rti(ctypes.byref(A), ctypes.byref(cc), ctypes.byref(L), amplitude, width,
                                       n_random_modes, random_amplitude, seed, stream), "cglbm_launch_init_rti")


def write_vti_host(path: str, fields) -> None:
    """The reference's ``write_vti(filename, LbmHost)`` (io/vtiWriter.cuh:92-145): ``fields`` holds
    rhor, rhob, ux, uy, uz as float32 arrays of one shape (nz, ny, nx), x fastest. Needs no GPU."""
    arrs = [np.ascontiguousarray(f, dtype=np.float32) for f in fields]
    if len(arrs) != 5 or any(a.ndim != 3 or a.shape != arrs[0].shape for a in arrs):
        raise ValueError("write_vti_host needs five float32 arrays of one (nz, ny, nx) shape")
    nz, ny, nx = arrs[0].shape
    _check(lib().cglbm_write_vti_host(path.encode(), nx, ny, nz, *[a.ctypes.data for a in arrs]),
           "cglbm_write_vti_host")


def domain_bytes(cfg: Config, storage: int = STORE_F32) -> int:
    cc = cfg.c()
    return int(lib().cglbm_domain_bytes(ctypes.byref(cc), storage))


def snapshot_info(path: str) -> tuple["Config", int, int]:
    """(configuration, storage, step count) of a snapshot file."""
    cc = _Config()
    st = ctypes.c_int32()
    steps = ctypes.c_int64()
    _check(lib().cglbm_snapshot_info(path.encode(), ctypes.byref(cc), ctypes.byref(st), ctypes.byref(steps)),
           "cglbm_snapshot_info")
    cfg = Config(**{f.name: getattr(cc, f.name) for f in fields(Config)})
    return cfg, int(st.value), int(steps.value)


class Domain:
    """A library-owned domain (ping-pong moment state on the GPU)."""

    def __init__(self, cfg: Config, storage: int = STORE_F32, kernel: int = KERNEL_AUTO):
        self.cfg = cfg
        self.storage = storage
        h = ctypes.c_void_p()
        cc = cfg.c()
        _check(lib().cglbm_domain_create(ctypes.byref(cc), storage, ctypes.byref(h)), "cglbm_domain_create")
        self._h = h
        if kernel != KERNEL_AUTO:
            self.set_kernel(kernel)

    # -- lifecycle ---------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cglbm_domain_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration ------------------------------------------------------------------------
    def set_kernel(self, kernel: int):
        _check(lib().cglbm_domain_set_kernel(self._h, kernel), "cglbm_domain_set_kernel")

    def set_tile(self, ty: int, L: int = 0, nbuf: int = 2):
        _check(lib().cglbm_domain_set_tile(self._h, ty, L, nbuf), "cglbm_domain_set_tile")

    def plan(self):
        """(tile rows, items) of the current chunk plan; items is an (n, 4) int32 array of
        {tile x, tile y, z0, z1} (z0 == z1: idle), empty when no plan exists."""
        n, ty = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().cglbm_domain_plan(self._h, None, 0, ctypes.byref(n), ctypes.byref(ty)), "cglbm_domain_plan")
        items = np.zeros((max(n.value, 0), 4), dtype=np.int32)
        if n.value > 0:
            _check(lib().cglbm_domain_plan(self._h, items.ctypes.data, n.value, ctypes.byref(n), ctypes.byref(ty)),
                   "cglbm_domain_plan")
        return ty.value, items

    def autotune(self, steps_per_candidate: int = 3):
        """Returns (tile rows, chunk length, exchange buffers, ms per step) of the fastest shape."""
        ty, L, nb, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        _check(lib().cglbm_domain_autotune(self._h, steps_per_candidate, ctypes.byref(ty), ctypes.byref(L),
                                           ctypes.byref(nb), ctypes.byref(ms)), "cglbm_domain_autotune")
        return ty.value, L.value, nb.value, ms.value

    @property
    def step_count(self) -> int:
        return int(lib().cglbm_domain_step_count(self._h))

    def moments(self, which: int = 0) -> Moments:
        m = Moments()
        _check(lib().cglbm_domain_moments(self._h, which, ctypes.byref(m)), "cglbm_domain_moments")
        return m

    # -- initial conditions -------------------------------------------------------------------
    def init_bubble(self, radius: float, width: float = 4.0, sharp: bool = False, center=None):
        c = self.cfg
        if center is None:
            center = ((c.nx - 1) / 2.0, (c.ny - 1) / 2.0, (c.nz - 1) / 2.0)
        _check(lib().cglbm_init_bubble(self._h, *center, radius, width, int(sharp)), "cglbm_init_bubble")

    def init_jet(self):
        _check(lib().cglbm_init_jet(self._h), "cglbm_init_jet")

    def init_rti(self, amplitude=2.0, width=4.0, n_random_modes=0, random_amplitude=0.0, seed=0):
        _check(lib().cglbm_init_rti(self._h, amplitude, width, n_random_modes, random_amplitude, seed),
               "cglbm_init_rti")

    # -- stepping -----------------------------------------------------------------------------
    def step(self, nsteps: int = 1):
        _check(lib().cglbm_step(self._h, nsteps), "cglbm_step")

    def time_steps(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        _check(lib().cglbm_time_steps(self._h, nsteps, ctypes.byref(ms)), "cglbm_time_steps")
        return ms.value

    def sync(self):
        _check(lib().cglbm_sync(self._h), "cglbm_sync")

    # -- data ---------------------------------------------------------------------------------
    def get_state(self, mask: int = 0x7FF) -> np.ndarray:
        nf = bin(mask).count("1")
        out = np.empty((nf,) + self.cfg.shape[1:], dtype=np.float32)
        _check(lib().cglbm_copy_out(self._h, mask, out.ctypes.data), "cglbm_copy_out")
        return out

    def set_state(self, state: np.ndarray, mask: int = 0x7FF):
        a = np.ascontiguousarray(state, dtype=np.float32)
        nf = bin(mask).count("1")
        if a.shape != (nf,) + self.cfg.shape[1:]:
            raise CglbmError(f"state shape {a.shape} != {(nf,) + self.cfg.shape[1:]}")
        _check(lib().cglbm_copy_in(self._h, mask, a.ctypes.data), "cglbm_copy_in")

    def reduce(self) -> np.ndarray:
        out = (ctypes.c_double * 6)()
        _check(lib().cglbm_reduce(self._h, out), "cglbm_reduce")
        return np.array(out[:], dtype=np.float64)

    def write_vti(self, path: str):
        _check(lib().cglbm_write_vti(self._h, path.encode()), "cglbm_write_vti")

    def save(self, path: str):
        """Restartable snapshot of the configuration, storage, step count and full state."""
        _check(lib().cglbm_snapshot_write(self._h, path.encode()), "cglbm_snapshot_write")

    def load(self, path: str):
        """Restore a snapshot written from a domain of the same configuration and storage."""
        _check(lib().cglbm_snapshot_read(self._h, path.encode()), "cglbm_snapshot_read")

    @classmethod
    def from_snapshot(cls, path: str, kernel: int = KERNEL_AUTO) -> "Domain":
        """A new domain restored from a snapshot."""
        cfg, storage, _ = snapshot_info(path)
        d = cls(cfg, storage=storage, kernel=kernel)
        try:
            d.load(path)
        except Exception:
            d.close()
            raise
        return d


class Batch:
    """Many domains stepped together (one launch per kernel class per step)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().cglbm_batch_create(ctypes.byref(h)), "cglbm_batch_create")
        self._h = h
        self.members = []

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cglbm_batch_destroy(self._h)
            self._h = None
            self.members = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, dom: "Domain"):
        _check(lib().cglbm_batch_add(self._h, dom._h), "cglbm_batch_add")
        self.members.append(dom)

    def remove(self, dom: "Domain"):
        _check(lib().cglbm_batch_remove(self._h, dom._h), "cglbm_batch_remove")
        self.members.remove(dom)

    def __len__(self):
        return int(lib().cglbm_batch_size(self._h))

    def set_tile(self, ty: int, L: int = 16, nbuf: int = 2):
        _check(lib().cglbm_batch_set_tile(self._h, ty, L, nbuf), "cglbm_batch_set_tile")

    def set_graph(self, bundle: int):
        _check(lib().cglbm_batch_set_graph(self._h, bundle), "cglbm_batch_set_graph")

    def step(self, nsteps: int = 1):
        _check(lib().cglbm_batch_step(self._h, nsteps), "cglbm_batch_step")

    def time_steps(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        _check(lib().cglbm_batch_time_steps(self._h, nsteps, ctypes.byref(ms)), "cglbm_batch_time_steps")
        return ms.value

    def time_sequential(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        _check(lib().cglbm_batch_time_sequential(self._h, nsteps, ctypes.byref(ms)), "cglbm_batch_time_sequential")
        return ms.value

    def autotune(self, steps_per_candidate: int = 2):
        _check(lib().cglbm_batch_autotune(self._h, steps_per_candidate), "cglbm_batch_autotune")

    def sync(self):
        _check(lib().cglbm_batch_sync(self._h), "cglbm_batch_sync")


class Server:
    """Continuous-batching server (include/cglbm.h, cglbm_server_*): submit cases at any time;
    they are admitted by HBM budget, stepped together while resident, and returned as results."""

    _CASES = {"bubble": CASE_BUBBLE, "jet": CASE_JET, "rti": CASE_RTI, "state": CASE_STATE}

    def __init__(self, hbm_budget_bytes: int = 0, max_members: int = 0, bundle: int = 0, admit_threads: int = 0,
                 tile=None):
        o = _ServerOpts(hbm_budget_bytes, max_members, bundle, admit_threads, *(tile or (0, 0, 0)))
        h = ctypes.c_void_p()
        _check(lib().cglbm_server_create(ctypes.byref(o), ctypes.byref(h)), "cglbm_server_create")
        self._h = h
        self._keep = {}  # id -> (result buffer, initial state) kept alive until polled

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cglbm_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, cfg: Config, case: str = "bubble", steps: int = 100, storage: int = STORE_F32,
               radius: float = 8.0, width: float = 4.0, sharp: bool = False, amplitude: float = 2.0,
               n_random_modes: int = 0, random_amplitude: float = 0.0, init_seed: int = 0, initial=None,
               result_mask: int = 0, tag: int = 0) -> int:
        buf = None
        if result_mask:
            buf = np.empty((bin(result_mask).count("1"),) + cfg.shape[1:], dtype=np.float32)
        init = None
        if initial is not None:
            init = np.ascontiguousarray(initial, dtype=np.float32)
            if init.shape != cfg.shape:
                raise CglbmError(f"initial state shape {init.shape} != {cfg.shape}")
        r = _Request(cfg.c(), storage, self._CASES[case], radius, width, int(sharp), amplitude, n_random_modes,
                     random_amplitude, init_seed, init.ctypes.data if init is not None else None, steps,
                     result_mask, buf.ctypes.data if buf is not None else None, tag)
        rid = ctypes.c_uint64()
        _check(lib().cglbm_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)), "cglbm_server_submit")
        self._keep[rid.value] = buf
        return rid.value

    def poll(self, max_results: int = 256, timeout_ms: int = 0):
        """Finished cases as dicts (``fields`` holds the requested result fields)."""
        arr = (_Result * max_results)()
        n = ctypes.c_int32()
        _check(lib().cglbm_server_poll(self._h, arr, max_results, timeout_ms, ctypes.byref(n)), "cglbm_server_poll")
        out = []
        for k in range(n.value):
            r = arr[k]
            out.append(dict(id=r.id, tag=r.user_tag, status=r.status, storage=r.storage, steps=r.steps,
                            cells=r.cells, diag=np.array(r.diag[:]), t_submit=r.t_submit, t_admit=r.t_admit,
                            t_start=r.t_start, t_done=r.t_done, message=r.message.decode(errors="replace"),
                            fields=self._keep.pop(r.id, None)))
        return out

    def drain(self):
        _check(lib().cglbm_server_drain(self._h), "cglbm_server_drain")

    def stats(self) -> dict:
        st = _ServerStats()
        _check(lib().cglbm_server_stats_get(self._h, ctypes.byref(st)), "cglbm_server_stats_get")
        return {name: getattr(st, name) for name, _ in _ServerStats._fields_}

    def clock(self) -> float:
        return float(lib().cglbm_server_clock(self._h))
