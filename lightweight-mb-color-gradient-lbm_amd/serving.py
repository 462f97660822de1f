"""Serving workload: Poisson arrivals of a mixed case population into a continuous-batching Server.

The population (seeded):
  60 % bubbles  D3Q19, periodic cube, edge uniform in [32, 96], drop radius uniform in
                [4, min(24, edge/3)], tanh width 4, tau uniform in [0.52, 1.0], sigma log-uniform in
                [1e-3, 2e-2], beta 0.7;
  30 % jets     D3Q27, jet axis y, nx = nz = 64 s, ny = 128 s with s uniform in [1, 2] (cells rounded
                to multiples of 8: 64 x 128 x 64 up to 128 x 256 x 128), inlet radius 8, inlet speed
                uniform in [0.02, 0.06] with 1 % uniform noise, zero-gradient outflow, tau 0.55,
                sigma 0.005, beta 0.7;
  10 % RTI      D3Q19, 64 x 256 x 64 with walls at the ends of y, Atwood uniform in [0.2, 0.6],
                gravity 1e-5, single cosine mode of amplitude 2 plus 8 random modes of 0.2, tau 0.6,
                sigma 0.002, beta 0.7;
each running a uniform 500..3000 steps, with FP16-deviation or FP32 storage (even odds).

``run_poisson`` submits the population at exponentially distributed gaps over ``duration``
seconds, polls results as they finish, drains, and reports p50/p99 time-to-result (submission to
results on the host) and the aggregate update rate.
"""
from __future__ import annotations

import math
import time

import numpy as np


def draw_case(cg, rng) -> dict:
    u = rng.random()
    steps = int(rng.integers(500, 3001))
    storage = cg.STORE_F16 if rng.random() < 0.5 else cg.STORE_F32
    if u < 0.6:
        e = int(rng.integers(32, 97))
        r = float(rng.uniform(4.0, min(24.0, e / 3.0)))
        cfg = cg.Config(nx=e, ny=e, nz=e, q=19, tau=float(rng.uniform(0.52, 1.0)),
                        sigma=float(math.exp(rng.uniform(math.log(1e-3), math.log(2e-2)))), beta=0.7)
        return dict(kind="bubble", cfg=cfg, steps=steps, storage=storage, radius=r, width=4.0)
    if u < 0.9:
        sc = float(rng.uniform(1.0, 2.0))
        n = int(round(64 * sc / 8.0)) * 8
        cfg = cg.Config(nx=n, ny=2 * n, nz=n, q=27, ybc=cg.Y_JET, outlet=cg.OUTLET_ZERO_GRADIENT, tau=0.55,
                        sigma=0.005, beta=0.7, jet_radius=8.0, jet_velocity=float(rng.uniform(0.02, 0.06)),
                        jet_noise=0.01, seed=int(rng.integers(0, 2**31)))
        return dict(kind="jet", cfg=cfg, steps=steps, storage=storage)
    cfg = cg.Config(nx=64, ny=256, nz=64, q=19, ybc=cg.Y_WALLS, tau=0.6, sigma=0.002, beta=0.7, gravity=1e-5,
                    atwood=float(rng.uniform(0.2, 0.6)))
    return dict(kind="rti", cfg=cfg, steps=steps, storage=storage, amplitude=2.0, width=4.0, n_random_modes=8,
                random_amplitude=0.2, init_seed=int(rng.integers(0, 2**31)))


def submit_case(server, cg, case: dict, tag: int = 0, result_mask: int = 0) -> int:
    kw = {k: v for k, v in case.items() if k not in ("kind", "cfg", "steps", "storage")}
    return server.submit(case["cfg"], case["kind"], steps=case["steps"], storage=case["storage"],
                         result_mask=result_mask, tag=tag, **kw)


def percentile(values, q):
    if not values:
        return None
    return float(np.percentile(np.asarray(values, dtype=np.float64), q))


def run_poisson(cg, rate: float, duration: float, seed: int = 0, server=None, log=None, **server_kw) -> dict:
    """Poisson arrivals at `rate` cases/s for `duration` s; returns the latency/throughput summary."""
    rng = np.random.default_rng(seed)
    arrivals, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            break
        arrivals.append(t)
    cases = [draw_case(cg, rng) for _ in arrivals]
    own = server is None
    srv = server if server is not None else cg.Server(**server_kw)
    results = []
    try:
        t0 = time.monotonic()
        c0 = srv.clock()
        for i, (ta, case) in enumerate(zip(arrivals, cases)):
            while True:
                wait = t0 + ta - time.monotonic()
                if wait <= 0:
                    break
                results += srv.poll(timeout_ms=max(1, int(wait * 1000)))
            submit_case(srv, cg, case, tag=i)
            if log and (i + 1) % 50 == 0:
                st = srv.stats()
                log(f"serving: {i + 1}/{len(cases)} submitted, {st['completed']} done, {st['resident']} resident, "
                    f"{st['queued']} queued")
        srv.drain()
        results += srv.poll(max_results=len(cases) + 1)
        stats = srv.stats()
    finally:
        if own:
            srv.close()
    ok = [r for r in results if r["status"] == 0]
    lat = [r["t_done"] - r["t_submit"] for r in ok]
    span = max(r["t_done"] for r in ok) - c0 if ok else float("nan")
    updates = sum(r["cells"] * r["steps"] for r in ok)
    kinds = {}
    for c in cases:
        kinds[c["kind"]] = kinds.get(c["kind"], 0) + 1
    return dict(
        rate=rate, duration_s=duration, seed=seed, requests=len(cases), completed=len(ok),
        failed=len(results) - len(ok), mix=kinds,
        fp16=sum(1 for c in cases if c["storage"] == cg.STORE_F16),
        latency_p50_s=percentile(lat, 50), latency_p99_s=percentile(lat, 99),
        latency_max_s=max(lat) if lat else None,
        queue_wait_p99_s=percentile([r["t_admit"] - r["t_submit"] for r in ok], 99),
        aggregate_mlups=updates / span / 1e6 if ok else None, wall_s=span, updates=updates,
        peak_resident=stats["peak_resident"], peak_bytes_reserved=stats["peak_bytes_reserved"],
        budget_bytes=stats["budget_bytes"], bundles=stats["bundles"], busy_s=stats["busy_seconds"],
    )
