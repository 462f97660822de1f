"""Benchmark: fused colour-gradient step on one MI355X.

Workload (the metric's configuration): D3Q19 periodic 128^3 domain holding a diffuse drop
(R = 24, tanh profile of width 4, blue inside), tau = 0.8, sigma = 0.01, beta = 0.7, FP32 moment
storage, inputs resident in HBM. A step is one fused update of all 2,097,152 cells. The D3Q27
128^3 figure of the same workload is reported alongside.

value = MLUPS = cells * steps / seconds / 1e6 over exactly --steps timed steps (HIP events on
the domain's stream, bracketed by device synchronisation). Tiles are autotuned on the live state
and reported as [tile rows, chunk planes, exchange buffers]; chunk planes 0 = planned chunks (the
library re-cuts the z-chunks every 64 steps from where the interface is, the planner's launches
included in the timed steps). roofline.achieved = 88 B per cell
update (11 FP32 fields read + 11 written) * cells / average step duration. cpu_baseline = the
reference's own update (compiled for the host from the reference sources, oracle/_ref) on a
bounded sample of the same workload, on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--kernel auto|gather|march]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE_F32 = 88
# FP16-deviation storage: rhor, rhob binary32 (2 x 4 B), nine binary16 deviations (9 x 2 B), read + write
BYTES_PER_UPDATE_F16 = 52


def _load():
    import __graft_entry__ as ge
    return ge._import_package()


def measure(cg, q: int, n: int, steps: int, warmup: int, kernel: int, tile=None, storage=None):
    cfg = cg.Config(nx=n, ny=n, nz=n, q=q, tau=0.8, sigma=0.01, beta=0.7)
    storage = cg.STORE_F32 if storage is None else storage
    with cg.Domain(cfg, storage=storage, kernel=kernel) as dom:
        dom.init_bubble(radius=24.0 * n / 128.0, width=4.0)
        m0 = dom.reduce()
        shape = None
        if kernel != cg.KERNEL_GATHER:
            if tile:
                dom.set_tile(*tile)
                shape = tuple(tile)
            else:
                shape = dom.autotune(3)[:3]  # tile rows, chunk length, exchange buffers
        dom.step(warmup)
        dom.sync()
        ms = dom.time_steps(steps)
        m1 = dom.reduce()
    cells = cfg.interior_cells
    sec = ms / 1e3
    mlups = cells * steps / sec / 1e6
    per_update = BYTES_PER_UPDATE_F16 if storage == cg.STORE_F16 else BYTES_PER_UPDATE_F32
    achieved = per_update * cells / (sec / steps) / 1e9
    return dict(
        mlups=mlups,
        ms_per_step=ms / steps,
        achieved_gbs=achieved,
        mass_drift_rel=[float((m1[0] - m0[0]) / m0[0]), float((m1[1] - m0[1]) / m0[1])],
        bad_cells=int(m1[5]),
        tile=shape,
    )


def calibrate_copy(cg, mib: int):
    """One hipMemcpy device-to-device of `mib` MiB through the HIP runtime the library uses: a
    kernel of known traffic (mib MiB read + mib MiB written) for scaling FETCH_SIZE/WRITE_SIZE."""
    import ctypes
    cg.lib()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                path = line.split()[-1]
                break
    hip = ctypes.CDLL(path)
    n = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n)) == 0
    hip.hipMemset(src, 1, ctypes.c_size_t(n))
    hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3)
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"bench: calibration copy of {mib} MiB done", file=sys.stderr, flush=True)


def _hip_runtime(cg):
    """The HIP runtime the library loaded (ctypes), for events and copies around raw launches."""
    import ctypes
    cg.lib()
    with open("/proc/self/maps") as f:
        path = next(line.split()[-1] for line in f if "libamdhip64" in line)
    hip = ctypes.CDLL(path)
    hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    return hip


def measure_shard(cg, case: str = "rti", steps: int = 200, ranks: int = 8):
    """One rank's share of a domain split into `ranks` z-slabs (the decomposition a multi-GPU run
    would use; one GPU runs one shard), the fourth rank's slab held with two halo planes on each
    side: its update (cglbm_launch_rcs_zslab) and the device side of its halo exchange, packing
    the 13 plane-fields of each face into send buffers and unpacking received ones into the halo
    planes (cglbm_zslab_pack / unpack). case "rti": the 128 x 512 x 128 Rayleigh-Taylor layer
    (walls and gravity along y, so z-slabs of 16 planes with 128 x 514 faces); "periodic": a
    periodic 128 x 128 x 512 column with a drop in each slab (64-plane slabs, 128 x 128 faces). The
    exchange itself (RCCL send/receive over xGMI) is not run: one GPU."""
    import ctypes
    hip = _hip_runtime(cg)
    if case == "rti":
        nz_full, ny, nx = 128, 512, 128
        own = nz_full // ranks
        cfg = cg.Config(nx=nx, ny=ny, nz=own + 4, q=19, ybc=cg.Y_WALLS, tau=0.6, sigma=0.002, beta=0.7,
                        gravity=1e-5, atwood=0.5)
        plane = nx * (ny + 2)  # cells of one stored z-plane
        label = "RTI 128 x 512 x 128 (D3Q19, walls, forced), z-slabs"
    elif case == "periodic":
        nz_full, ny, nx = 512, 128, 128
        own = nz_full // ranks
        cfg = cg.Config(nx=nx, ny=ny, nz=own + 4, q=19, tau=0.8, sigma=0.01, beta=0.7)
        plane = nx * ny
        label = "periodic 128 x 128 x 512 column (D3Q19, a drop of radius 24 per slab), z-slabs"
    else:
        raise ValueError(case)
    with cg.Domain(cfg) as d:
        if case == "rti":
            d.init_rti(2.0, 4.0, 8, 0.2, 0)
        else:
            d.init_bubble(radius=24.0, width=4.0)
        d.sync()
        s = ctypes.c_void_p(cg.lib().cglbm_domain_stream(d._h))
        views = [d.moments(0), d.moments(1)]
        ev = [ctypes.c_void_p() for _ in range(12)]
        for e in ev:
            assert hip.hipEventCreate(ctypes.byref(e)) == 0
        off = 3 * own - 2  # the fourth rank's slab: global planes 3*own .. 4*own - 1
        for k in range(10):  # warm-up
            cg.launch_rcs_zslab(views[k & 1], views[(k + 1) & 1], cfg, 2, 2 + own, off, step=k + 1, stream=s.value)
        hip.hipEventRecord(ev[0], s)
        for k in range(steps):
            cg.launch_rcs_zslab(views[k & 1], views[(k + 1) & 1], cfg, 2, 2 + own, off, step=k + 11, stream=s.value)
        hip.hipEventRecord(ev[1], s)
        # halo exchange, device side: pack both faces (13 plane-fields each) into send buffers, and
        # unpack two buffers into the halo planes (the same buffers stand in for received ones)
        bufs = [ctypes.c_void_p(), ctypes.c_void_p()]
        for b in bufs:
            assert hip.hipMalloc(ctypes.byref(b), 13 * plane * 4) == 0
        for k in range(5):
            cg.zslab_pack(views[0], cfg, 2, 2 + own, bufs[0].value, bufs[1].value, stream=s.value)
        hip.hipEventRecord(ev[2], s)
        for k in range(steps):
            cg.zslab_pack(views[0], cfg, 2, 2 + own, bufs[0].value, bufs[1].value, stream=s.value)
        hip.hipEventRecord(ev[3], s)
        hip.hipEventRecord(ev[4], s)
        for k in range(steps):
            cg.zslab_unpack(views[0], cfg, 2, 2 + own, bufs[0].value, bufs[1].value, stream=s.value)
        hip.hipEventRecord(ev[5], s)
        # the schedule that hides the transfer: pack, update the interior (two planes in from each
        # face; the transfer would run meanwhile), unpack, update the two planes next to each face
        # (both faces in one launch)
        hip.hipEventRecord(ev[6], s)
        for k in range(steps):
            A, B = views[k & 1], views[(k + 1) & 1]
            cg.zslab_pack(A, cfg, 2, 2 + own, bufs[0].value, bufs[1].value, stream=s.value)
            cg.launch_rcs_zslab(A, B, cfg, 4, own, off, step=k + 11, stream=s.value)
            cg.zslab_unpack(A, cfg, 2, 2 + own, bufs[0].value, bufs[1].value, stream=s.value)
            cg.launch_rcs_zslab_faces(A, B, cfg, 2, 2 + own, 2, off, step=k + 11, stream=s.value)
        hip.hipEventRecord(ev[7], s)
        # the schedule's two update parts alone
        hip.hipEventRecord(ev[8], s)
        for k in range(steps):
            cg.launch_rcs_zslab(views[k & 1], views[(k + 1) & 1], cfg, 4, own, off, step=k + 11, stream=s.value)
        hip.hipEventRecord(ev[9], s)
        hip.hipEventRecord(ev[10], s)
        for k in range(steps):
            cg.launch_rcs_zslab_faces(views[k & 1], views[(k + 1) & 1], cfg, 2, 2 + own, 2, off, step=k + 11,
                                      stream=s.value)
        hip.hipEventRecord(ev[11], s)
        assert hip.hipEventSynchronize(ev[11]) == 0
        ms_inner, ms_faces = ctypes.c_float(), ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms_inner), ev[8], ev[9])
        hip.hipEventElapsedTime(ctypes.byref(ms_faces), ev[10], ev[11])
        ms_step, ms_pack, ms_unpack, ms_sched = ctypes.c_float(), ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms_sched), ev[6], ev[7])
        hip.hipEventElapsedTime(ctypes.byref(ms_step), ev[0], ev[1])
        hip.hipEventElapsedTime(ctypes.byref(ms_pack), ev[2], ev[3])
        hip.hipEventElapsedTime(ctypes.byref(ms_unpack), ev[4], ev[5])
        for b in bufs:
            hip.hipFree(b)
        for e in ev:
            hip.hipEventDestroy(e)
    us_step = ms_step.value * 1e3 / steps
    cells = nx * ny * own
    return {"domain": label, "ranks": ranks,
            "slab": [nx, ny, own], "stored_planes": own + 4, "steps": steps,
            "us_per_step": round(us_step, 2), "value": round(cells / (us_step * 1e-6) / 1e6, 1), "unit": "MLUPS",
            "halo_face_bytes": 13 * plane * 4,
            "us_pack_both_faces": round(ms_pack.value * 1e3 / steps, 2),
            "us_unpack_both_faces": round(ms_unpack.value * 1e3 / steps, 2),
            "us_per_step_overlap_schedule": round(ms_sched.value * 1e3 / steps, 2),
            "us_interior_update": round(ms_inner.value * 1e3 / steps, 2),
            "us_face_updates": round(ms_faces.value * 1e3 / steps, 2),
            "note": "update of one rank's slab with its halo planes in place (cglbm_launch_rcs_zslab), and the "
                    "device side of its exchange: packing both faces' 13 plane-fields into send buffers and "
                    "unpacking two received buffers into the halo planes (cglbm_zslab_pack / unpack); "
                    "us_per_step_overlap_schedule: pack, interior update, unpack and the two face updates in "
                    "sequence on one stream, the device time of a rank's step whose transfer hides behind the "
                    "interior update; the transfer between GPUs is not run"}


def batched_workload(cg, seed: int = 0, count: int = 256):
    """Config 4: `count` independent periodic D3Q19 cubes, edges uniform in {32, 48, 64, 80, 96},
    drop radius uniform in [4, min(24, e/3)], tau uniform in [0.52, 1.0], sigma log-uniform in
    [1e-3, 2e-2], beta 0.7 (seeded)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        e = int(rng.choice([32, 48, 64, 80, 96]))
        r = float(rng.uniform(4.0, min(24.0, e / 3.0)))
        tau = float(rng.uniform(0.52, 1.0))
        sigma = float(np.exp(rng.uniform(np.log(1e-3), np.log(2e-2))))
        out.append((cg.Config(nx=e, ny=e, nz=e, q=19, tau=tau, sigma=sigma, beta=0.7), r))
    return out


def measure_batched(cg, steps: int, seq_steps: int, bundle: int = 10):
    work = batched_workload(cg)
    cells = sum(c.interior_cells for c, _ in work)
    doms = []
    for cfg, r in work:
        d = cg.Domain(cfg)
        d.init_bubble(radius=r, width=4.0)
        doms.append(d)
    b = cg.Batch()
    for d in doms:
        b.add(d)
    b.autotune(2)
    b.set_graph(bundle)
    b.step(2 * bundle)
    b.sync()
    ms = b.time_steps(steps)
    ms_seq = b.time_sequential(seq_steps)
    # the reference's launch style: one thread per cell (the gather kernel), member after member
    for d in doms:
        d.set_kernel(cg.KERNEL_GATHER)
    ms_ref = b.time_sequential(seq_steps)
    for d in doms:
        d.set_kernel(cg.KERNEL_AUTO)
    for d in doms:
        b.remove(d)
        d.close()
    b.close()
    mlups = cells * steps / (ms / 1e3) / 1e6
    mlups_seq = cells * seq_steps / (ms_seq / 1e3) / 1e6
    mlups_ref = cells * seq_steps / (ms_ref / 1e3) / 1e6
    return {"domains": len(work), "cells": cells, "steps": steps, "ms_per_step": round(ms / steps, 4),
            "value": round(mlups, 1), "unit": "MLUPS", "graph_bundle": bundle,
            "sequential_plain_launches": {"kernel": "z-marching, size-aware default tile", "steps": seq_steps,
                                          "ms_per_step": round(ms_seq / seq_steps, 4), "value": round(mlups_seq, 1)},
            "sequential_reference_shaped_launches": {"kernel": "one thread per cell (gather)", "steps": seq_steps,
                                                     "ms_per_step": round(ms_ref / seq_steps, 4),
                                                     "value": round(mlups_ref, 1)},
            "speedup_vs_sequential": round(mlups / mlups_seq, 2),
            "speedup_vs_reference_shaped": round(mlups / mlups_ref, 2),
            "achieved_gbs": round(BYTES_PER_UPDATE_F32 * cells / (ms / steps / 1e3) / 1e9, 1)}


def measure_case(cg, cfg, init, steps: int, storage=None, warmup: int = 5):
    """One configuration run for `steps` timed steps from its initial state, twice: on the default
    path (no tile set: planned chunks where the library plans them, timed once against its uniform
    default, which it keeps if faster) and with the tile the autotuner picks on the initial state.
    `value` is the default path's; the autotuned run is reported beside it."""
    storage = cg.STORE_F32 if storage is None else storage
    runs = []
    for how in ("default", "autotuned at step 0"):
        with cg.Domain(cfg, storage=storage) as d:
            init(d)
            m0 = d.reduce()
            tile = None if how == "default" else tuple(d.autotune(2)[:3])
            d.step(warmup)
            d.sync()
            ms = d.time_steps(steps)
            m1 = d.reduce()
        runs.append((how, ms, tile, m0, m1))
    cells = cfg.interior_cells
    how, ms, tile, m0, m1 = runs[0]
    # relative per-phase mass change (None where a phase starts empty, as the jet's injected one)
    drift = [float((m1[k] - m0[k]) / m0[k]) if m0[k] > 0 else None for k in (0, 1)]
    return {"value": round(cells * steps / (ms / 1e3) / 1e6, 1), "unit": "MLUPS", "steps": steps,
            "ms_total": round(ms, 2), "how": how, "bad_cells": int(m1[5]), "mass_drift_rel": drift,
            "runs": [{"how": h, "tile": list(t) if t else None, "ms_total": round(r, 2),
                      "value": round(cells * steps / (r / 1e3) / 1e6, 1)} for h, r, t, _, _ in runs]}


def measure_configs(cg):
    """The other workloads of the reference's configuration set, each at its stated size and length:
    the 64^3 D3Q19 bubble (1000 steps), the D3Q27 round jet (128 x 256 x 128 with the jet axis y,
    inlet radius 8, u 0.05 with 1% noise, zero-gradient outflow, 2000 steps) and the D3Q19
    Rayleigh-Taylor case (128 x 512 x 128, walls at the ends of y, Atwood 0.5, g 1e-5, 5000 steps)
    with FP16-deviation and FP32 storage."""
    out = {}
    c1 = cg.Config(nx=64, ny=64, nz=64, q=19, tau=0.8, sigma=0.01, beta=0.7)
    out["bubble64_d3q19"] = measure_case(cg, c1, lambda d: d.init_bubble(radius=12.0, width=4.0), 1000)
    c2 = cg.Config(nx=128, ny=256, nz=128, q=27, ybc=cg.Y_JET, outlet=cg.OUTLET_ZERO_GRADIENT, tau=0.55,
                   sigma=0.005, beta=0.7, jet_radius=8.0, jet_velocity=0.05, jet_noise=0.01, seed=0)
    out["jet_d3q27_128x256x128"] = measure_case(cg, c2, lambda d: d.init_jet(), 2000)
    c3 = cg.Config(nx=128, ny=512, nz=128, q=19, ybc=cg.Y_WALLS, tau=0.6, sigma=0.002, beta=0.7, gravity=1e-5,
                   atwood=0.5)
    rti = lambda d: d.init_rti(2.0, 4.0, 8, 0.2, 0)
    out["rti_d3q19_128x512x128_f16"] = measure_case(cg, c3, rti, 5000, storage=cg.STORE_F16)
    out["rti_d3q19_128x512x128_f16"]["error_vs_fp64"] = precision_table().get("rti_128x512x128_5000")
    out["rti_d3q19_128x512x128_f32"] = measure_case(cg, c3, rti, 5000)
    out["bubble64_d3q19"]["mass_modes"] = mass_modes(cg)
    return out


def mass_modes(cg):
    """Per-phase relative mass change of the 64^3 static bubble after 100 and 1000 steps, in the
    reference's mode and in the conservative variant (no Omega2 where |F| <= 1e-4), from the FP64
    diagnostics. Binary32 arithmetic grows the mass by about the binary32 weight sum's excess over
    1 per step in both modes; the variant removes only the reference's Omega2 leak."""
    out = {}
    for cons in (0, 1):
        cfg = cg.Config(nx=64, ny=64, nz=64, q=19, tau=0.8, sigma=0.01, beta=0.7, conservative_omega2=cons)
        with cg.Domain(cfg) as d:
            d.init_bubble(radius=12.0, width=4.0)
            m0 = d.reduce()
            d.step(100)
            m100 = d.reduce()
            d.step(900)
            m1000 = d.reduce()
        key = "conservative_omega2" if cons else "reference_mode"
        out[key] = {"steps_100": [float((m100[k] - m0[k]) / m0[k]) for k in (0, 1)],
                    "steps_1000": [float((m1000[k] - m0[k]) / m0[k]) for k in (0, 1)]}
    return out


# binary32 VALU issue rate of one MI355X without packed math: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12


def profile_counters(workload: str, leg: str):
    """Per-launch counters of the leg's dominant kernel from the committed rocprof summary
    (profiles/counters_<workload>.json: calibrated HBM bytes per launch and VALU lane-ops per cell
    update, from SQ_INSTS_VALU x 64 / cells)."""
    path = os.path.join(ROOT, "profiles", f"counters_{workload}.json")
    try:
        with open(path) as f:
            return json.load(f).get(leg)
    except (OSError, ValueError):
        return None


def roofline(bytes_per_update: int, cells: int, ms_per_step: float, counters) -> dict:
    """Both roofs of the step kernel. HBM: algorithmic bytes per update x updates / step time
    against 8 TB/s. VALU: lane-ops per update (profile) x updates / step time against the binary32
    VALU issue rate. `bound` names the roof with the larger fraction; the top-level achieved, peak,
    unit and frac are that roof's (both are kept in `hbm` and `valu`)."""
    ups = cells / (ms_per_step / 1e3)
    gbs = bytes_per_update * ups / 1e9
    hbm = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
           "bytes_per_update": bytes_per_update}
    out = {"bound": "hbm", **{k: hbm[k] for k in ("achieved", "peak", "unit", "frac")},
           "traffic": counters.get("hbm_bytes_per_launch") if counters else None, "hbm": hbm, "valu": None}
    if counters and counters.get("valu_lane_ops_per_update"):
        lops = counters["valu_lane_ops_per_update"]
        rate = lops * ups / 1e12
        valu = {"lane_ops_per_update": lops, "achieved": round(rate, 2), "peak": round(VALU_PEAK_TOPS, 2),
                "unit": "T lane-ops/s", "frac": round(rate / VALU_PEAK_TOPS, 4), "source": counters.get("source")}
        out["valu"] = valu
        if valu["frac"] > hbm["frac"]:
            out.update(bound="valu", achieved=valu["achieved"], peak=valu["peak"], unit=valu["unit"], frac=valu["frac"])
    return out


def precision_table():
    """Per-result error of FP16-deviation (and FP32) storage against FP64 runs, per workload
    (profiles/precision_fp16_vs_fp64.json, written by tests/precision_table.py)."""
    try:
        with open(os.path.join(ROOT, "profiles", "precision_fp16_vs_fp64.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _cpu_threads() -> int:
    """Host threads for the CPU baseline: the OMP_NUM_THREADS share this process was given, else
    every CPU it may run on."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if env > 0:
        return env
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_reference(exe: str, mode: str, arg: str, threads: int, budget_s: float, fixed_steps: int = 0):
    """Seconds per step of the reference's own update (oracle/_ref host build) on `threads` host
    threads: one probe step (after one untimed step), then as many timed steps as fit `budget_s`."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    k = fixed_steps
    if not k:
        probe = subprocess.run([exe, mode, arg, "1", "-"], env=env, capture_output=True, text=True, check=True,
                               timeout=600)
        sps = float(probe.stdout.split()[1])
        k = max(2, min(1000, int(budget_s / max(sps, 1e-4))))
    run = subprocess.run([exe, mode, arg, str(k), "-"], env=env, capture_output=True, text=True, check=True,
                         timeout=900)
    words = run.stdout.split()
    return float(words[1]), int(words[3]), k


def cpu_baseline(budget_s: float = 4.0):
    """The reference's own update (compiled for the host from the reference sources, oracle/_ref)
    timed on this host's cores: the 128^3 D3Q19 drop of the metric (the value), and the other
    cases of the CPU-baseline table (64^3 static bubble for its full 1000 steps, 128^3 D3Q27 drop,
    the D3Q27 jet at 128 x 256 x 128, the 128 x 512 x 128 Rayleigh-Taylor domain, and the 128^3
    drop on one thread), each a bounded sample. Falls back to the oracle port if the reference
    builds are absent."""
    threads = _cpu_threads()
    ref = os.path.join(ROOT, "oracle", "_ref")
    exe = os.path.join(ref, "ref_bubble128_q19")
    info = {"cpu_model": _cpu_model(), "host_cpus": os.cpu_count(), "threads": threads}
    if os.path.exists(exe):
        sps, cells, k = _time_reference(exe, "time_bubble_tanh", "4", threads, 5.0)
        table = []
        cases = [
            ("static bubble 64^3", 19, "ref_bubble64_q19", "time_bubble_tanh", "4", threads, 1000),
            ("drop 128^3", 27, "ref_bubble128_q27", "time_bubble_tanh", "4", threads, 0),
            ("round jet 128 x 256 x 128 (reference convective outlet)", 27, "ref_jet256_q27", "time_jet", "0", threads, 0),
            ("Rayleigh-Taylor 128 x 512 x 128 (no body force in the reference)", 19, "ref_rti512_q19", "time_rti", "0",
             threads, 0),
            ("drop 128^3, one thread", 19, "ref_bubble128_q19", "time_bubble_tanh", "4", 1, 0),
        ]
        for label, q, binary, mode, arg, th, fixed in cases:
            path = os.path.join(ref, binary)
            if not os.path.exists(path):
                continue
            t, n, kk = _time_reference(path, mode, arg, th, budget_s, fixed)
            table.append({"case": label, "q": q, "threads": th, "timed_steps": kk, "cells": n,
                          "value": round(n / t / 1e6, 2), "unit": "MLUPS"})
            print(f"bench: CPU baseline {label} D3Q{q}: {n / t / 1e6:.1f} MLUPS on {th} threads", file=sys.stderr,
                  flush=True)
        return {"value": round(cells / sps / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "reference",
                "sample": f"{k} timed steps (after 1 warm-up) of the same D3Q19 128^3 drop; reference lbm.cuh RCS "
                          f"built for the host (g++ -O2, OpenMP over z), y wrapped through its own wrapy",
                **info, "table": table}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cgref
    p = cgref.Params(nx=128, ny=128, nz=128, q=19, tau=0.8, sigma=0.01, beta=0.7)
    A = cgref.init_bubble(p, radius=24.0, width=4.0)
    B = A.copy()
    cgref.step(p, A, B, 1)
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < budget_s or k < 2:
        cgref.step(p, B, A, k + 2)
        A, B = B, A
        k += 1
    dt = (time.perf_counter() - t0) / k
    return {"value": 2097152 / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{k} steps of the same D3Q19 128^3 drop with the oracle", **info}


SPLIT_CASES = (
    # (label, q, storage, state, tile rows, chunk planes (0 = planned), exchange buffers)
    ("flat_q19_ty16_L128", 19, 0, "flat", 16, 128, 2),
    ("flat_q19_ty16_L16", 19, 0, "flat", 16, 16, 2),
    ("flat_q19_ty8_L128", 19, 0, "flat", 8, 128, 2),
    ("flat_q19_ty4_L128", 19, 0, "flat", 4, 128, 2),
    ("drop_q19_ty16_L128", 19, 0, "drop", 16, 128, 2),
    ("drop_q19_ty16_Lplan", 19, 0, "drop", 16, 0, 2),
    ("flat_q27_ty16_L128", 27, 0, "flat", 16, 128, 1),
    ("flat_q27_ty16_L16", 27, 0, "flat", 16, 16, 1),
    ("drop_q27_ty16_L128", 27, 0, "drop", 16, 128, 1),
    ("drop_q27_ty16_Lplan", 27, 0, "drop", 16, 0, 1),
    ("flat_f16_q19_ty16_L128", 19, 1, "flat", 16, 128, 2),
    ("drop_f16_q19_ty16_Lplan", 19, 1, "drop", 16, 0, 2),
)


def valu_split(cg, n: int = 128):
    """Launches for an instruction-counter profile of the marching kernel (run under rocprofv3
    --pmc SQ_INSTS_VALU ...): the same kernel on a state with no interface (rho = 1, one phase) and
    on the 128^3 drop, with one chunk per column tile (no chunk edges) or 16-plane or planned
    chunks, and 16-, 8- and 4-row tiles, so that per-update VALU work splits into own-wave bulk
    work, halo (ring) work, the interface form and chunk edges. A few steps each, in the order of
    SPLIT_CASES; the dispatch order identifies the cases."""
    import numpy as np
    for label, q, st, state, ty, L, nb in SPLIT_CASES:
        cfg = cg.Config(nx=n, ny=n, nz=n, q=q, tau=0.8, sigma=0.01, beta=0.7)
        with cg.Domain(cfg, storage=cg.STORE_F16 if st else cg.STORE_F32) as d:
            if state == "flat":
                S = np.zeros(cfg.shape, np.float32)
                S[0] = 1.0
                d.set_state(S)
            else:
                d.init_bubble(radius=24.0 * n / 128.0, width=4.0)
            d.set_tile(ty, L, nb)
            d.step(8 if L == 0 else 3)
            d.sync()
        print(f"bench: split case {label} done", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", default="auto", choices=["auto", "gather", "march"])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-domain batched measurement")
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--no-serve", action="store_true", help="skip the Poisson serving measurement")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations (bubble64, jet, RTI)")
    ap.add_argument("--serve-rates", default="6,22",
                    help="serving arrival rates (per second), comma-separated: one Poisson window each")
    ap.add_argument("--serve-durations", default="40,30", help="arrival window of each rate (s), comma-separated")
    ap.add_argument("--tile19", default=None, help="fixed D3Q19 tile 'rows,chunk,buffers' instead of autotuning")
    ap.add_argument("--tile27", default=None, help="fixed D3Q27 tile 'rows,chunk,buffers' instead of autotuning")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--valu-split", action="store_true",
                    help="only the launches of the VALU split profile (profiles/march_bubble128.md), no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the update of one domain "
              f"runs on one GPU; the `shard` leg times one rank's share of a domain split into z-slabs, "
              f"and the exchange between GPUs is not run)", file=sys.stderr)
        return 2
    cg = _load()
    if args.valu_split:
        valu_split(cg, args.n)
        return 0
    kern = {"auto": cg.KERNEL_AUTO, "gather": cg.KERNEL_GATHER, "march": cg.KERNEL_MARCH}[args.kernel]
    print(f"bench: D3Q19 {args.n}^3 warmup {args.warmup} steps {args.steps}", file=sys.stderr, flush=True)
    t19 = [int(v) for v in args.tile19.split(",")] if args.tile19 else None
    t27 = [int(v) for v in args.tile27.split(",")] if args.tile27 else None
    r19 = measure(cg, 19, args.n, args.steps, args.warmup, kern, t19)
    print(f"bench: D3Q19 {r19['mlups']:.0f} MLUPS; D3Q27 next", file=sys.stderr, flush=True)
    r27 = measure(cg, 27, args.n, args.steps, args.warmup, kern, t27)
    print(f"bench: D3Q27 {r27['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    print("bench: D3Q19 FP16-deviation storage", file=sys.stderr, flush=True)
    r16 = measure(cg, 19, args.n, args.steps, args.warmup, cg.KERNEL_AUTO, None, storage=cg.STORE_F16)
    print(f"bench: D3Q19 FP16 storage {r16['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    batched = None
    if not args.no_batched:
        print("bench: 256 batched domains", file=sys.stderr, flush=True)
        batched = measure_batched(cg, args.batched_steps, 10)
        print(f"bench: batched {batched['value']:.0f} MLUPS, {batched['speedup_vs_sequential']}x sequential, "
              f"{batched['speedup_vs_reference_shaped']}x reference-shaped sequential",
              file=sys.stderr, flush=True)
    configs = None
    if not args.no_configs:
        print("bench: other configurations (64^3 bubble, D3Q27 jet, RTI FP16/FP32)", file=sys.stderr, flush=True)
        configs = measure_configs(cg)
        print("bench: " + ", ".join(f"{k} {v['value']:.0f}" for k, v in configs.items()), file=sys.stderr, flush=True)
    shard = None
    if not args.no_configs:
        shard = {}
        for key, case in (("rti_128x512x128", "rti"), ("periodic_128x128x512", "periodic")):
            print(f"bench: one z-slab shard of the {key} domain (8-way split)", file=sys.stderr, flush=True)
            r = shard[key] = measure_shard(cg, case)
            print(f"bench: shard {r['us_per_step']:.1f} us/step, halo pack {r['us_pack_both_faces']:.1f} us, "
                  f"unpack {r['us_unpack_both_faces']:.1f} us, overlap schedule "
                  f"{r['us_per_step_overlap_schedule']:.1f} us/step", file=sys.stderr, flush=True)
    serving = None
    if not args.no_serve:
        import cglbm.serving as sv
        rates = [float(v) for v in args.serve_rates.split(",")]
        durs = [float(v) for v in args.serve_durations.split(",")]
        runs = []
        for rate, dur in zip(rates, durs):
            print(f"bench: serving mix, Poisson {rate}/s over {dur} s", file=sys.stderr, flush=True)
            r = sv.run_poisson(cg, rate, dur, seed=0, log=lambda m: print(m, file=sys.stderr, flush=True))
            # offered load: the updates the window's requests ask for, per second of arrivals
            r["offered_mlups"] = r["updates"] / dur / 1e6 if r["completed"] == r["requests"] else None
            r["saturated"] = bool(r["offered_mlups"] and r["aggregate_mlups"] < 0.95 * r["offered_mlups"])
            print(f"bench: serving {rate}/s: p50 {r['latency_p50_s']:.2f} s p99 {r['latency_p99_s']:.2f} s, "
                  f"{r['aggregate_mlups']:.0f} MLUPS aggregate (offered {r['offered_mlups']:.0f})", file=sys.stderr,
                  flush=True)
            runs.append(r)
        sat = [r for r in runs if r["saturated"]]
        serving = {"runs": runs, "capacity_mlups": max(r["aggregate_mlups"] for r in sat) if sat else None,
                   "capacity_note": "aggregate MLUPS of the windows whose offered load exceeds what the server "
                                    "completes (arrivals faster than service); None if no window saturated"}
    if args.calibrate_copy:
        calibrate_copy(cg, args.calibrate_copy)
    workload = f"bubble{args.n}_d3q19"
    prof = f"bubble{args.n}"
    cells = args.n ** 3
    roof19 = roofline(BYTES_PER_UPDATE_F32, cells, r19["ms_per_step"], profile_counters(prof, "d3q19"))
    roof27 = roofline(BYTES_PER_UPDATE_F32, cells, r27["ms_per_step"], profile_counters(prof, "d3q27"))
    roof16 = roofline(BYTES_PER_UPDATE_F16, cells, r16["ms_per_step"], profile_counters(prof, "d3q19_f16"))
    prec = precision_table()
    cpu = None
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline()
    line = {
        "metric": "MLUPS (D3Q19 128^3 fused colour-gradient step, % of HBM roofline)",
        "value": round(r19["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r19["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: periodic 128^3 domain, diffuse drop R=24 (tanh, w=4), rho=1, u=Pi=0 at t=0",
        "config": {"workload": workload, "grid": [args.n] * 3, "q": 19, "tau": 0.8, "sigma": 0.01, "beta": 0.7,
                   "drop_radius": 24.0 * args.n / 128.0, "storage": "f32", "kernel": args.kernel,
                   "tile_d3q19": r19["tile"], "tile_d3q27": r27["tile"],
                   "timing": "HIP events on the domain stream around exactly `steps` steps; with planned chunks "
                             "(chunk planes 0) a re-plan (recording step plus planner, ~50 us) falls every 64 "
                             "steps, so fewer timed steps than that include at most one"},
        "hbm_frac": roof19["hbm"]["frac"],
        "roofline": roof19,
        "cpu_baseline": cpu,
        "d3q27": {"value": round(r27["mlups"], 1), "unit": "MLUPS", "ms_per_step": round(r27["ms_per_step"], 5),
                  "roofline": roof27},
        "d3q19_f16_storage": {"value": round(r16["mlups"], 1), "unit": "MLUPS", "ms_per_step": round(r16["ms_per_step"], 5),
                              "bytes_per_update": BYTES_PER_UPDATE_F16, "tile": r16["tile"], "roofline": roof16,
                              "error_vs_fp64": prec.get(f"drop{args.n}_d3q19"),
                              "error_note": "relative RMS per field against an FP64 run of the same drop "
                                            "(profiles/precision_fp16_vs_fp64.json; the FP16 storage model there is "
                                            "this path's bit for bit)"},
        "mass_drift_rel": {"d3q19": r19["mass_drift_rel"], "d3q27": r27["mass_drift_rel"],
                           "d3q19_f16_storage": r16["mass_drift_rel"]},
        "batched": batched,
        "serving": None if serving is None else {
            **serving,
            "fp16_error_vs_fp64_by_kind": {k: prec.get(n, {}).get("f16") for k, n in (
                ("bubble", "serving_bubble_d3q19_48"), ("jet", "serving_jet_d3q27_64x128x64"),
                ("rti", "serving_rti_d3q19_64x256x64"))},
            "fp16_error_note": "relative RMS per field after 500 steps against FP64, per case kind at a mid-range draw "
                               "(profiles/precision_fp16_vs_fp64.json)"},
        "other_configs": configs,
        "shard": shard,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
