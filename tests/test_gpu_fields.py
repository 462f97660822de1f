"""Caller-owned states whose 11 fields are not at one stride from one base, as with the reference's
separately allocated arrays (one device array per moment, memory.cuh:59-81 `allocate_device`).
The z-marching kernel takes them through its pointer form (one 64-bit base per field): results
must be bitwise the oracle's, and equal to the same states held as one block per state, at about
the same rate.

Device arrays come straight from the HIP runtime the library uses (hipMalloc through ctypes; the
library sees plain pointers). In the "pool" arrangement the fields of a state sit at uneven gaps
in reverse order inside one allocation, so the eleven bases are never at one stride and the
pointer form is certain to run; in the "separate" arrangement every field is its own hipMalloc,
exactly as the reference allocates them."""
import ctypes

import numpy as np
import pytest

from conftest import load_product

pytestmark = pytest.mark.gpu


def _cfg(cg, p):
    return cg.Config(**{k: getattr(p, k) for k in p.__dataclass_fields__})


def _hip(cg):
    """The HIP runtime already loaded by the library (same process-wide instance)."""
    cg.lib()
    with open("/proc/self/maps") as f:
        path = next(line.split()[-1] for line in f if "libamdhip64" in line)
    hip = ctypes.CDLL(path)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
    return hip


class DeviceStates:
    """Two caller-owned states of `n` floats per field. arrangement: "pool" (one allocation per
    state, fields at uneven gaps in reverse order), "separate" (one allocation per field) or
    "block" (one allocation per state, fields at one stride)."""

    def __init__(self, cg, n, arrangement):
        self.hip, self.n, self.cg = _hip(cg), n, cg
        self.allocs, self.ptrs = [], []
        for s in range(2):
            if arrangement == "separate":
                ptrs = [self._alloc(4 * (n + 64 * f)) for f in range(11)]
            elif arrangement == "pool":
                gaps = [64 * (1 + (f * 7 + s * 3) % 11) for f in range(11)]  # distinct, 64-float aligned
                base = self._alloc(4 * (11 * n + sum(gaps)))
                ptrs, o = [0] * 11, 0
                for f in reversed(range(11)):
                    ptrs[f] = base + 4 * o
                    o += n + gaps[f]
                assert len({ptrs[f + 1] - ptrs[f] for f in range(10)}) > 1
            else:
                stride = (n + 63) // 64 * 64
                base = self._alloc(4 * 11 * stride)
                ptrs = [base + 4 * stride * f for f in range(11)]
            self.ptrs.append(ptrs)

    def _alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        assert self.hip.hipMemset(p, 0, nbytes) == 0
        self.allocs.append(p.value)
        return p.value

    def moments(self, s):
        m = self.cg.Moments()
        for f, name in enumerate(self.cg.FIELDS):
            setattr(m, name, self.ptrs[s][f])
        return m

    def put(self, s, state):
        a = np.ascontiguousarray(state, dtype=np.float32).reshape(11, -1)
        assert a.shape[1] == self.n
        for f in range(11):
            assert self.hip.hipMemcpy(self.ptrs[s][f], a[f].ctypes.data, 4 * self.n, 1) == 0

    def get(self, s, shape):
        out = np.empty((11, self.n), dtype=np.float32)
        for f in range(11):
            assert self.hip.hipMemcpy(out[f].ctypes.data, self.ptrs[s][f], 4 * self.n, 2) == 0
        return out.reshape((11,) + shape)

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def close(self):
        self.sync()
        for p in self.allocs:
            self.hip.hipFree(p)
        self.allocs = []


CASES = {
    "jet_q19": dict(nx=40, ny=36, nz=30, q=19, jet_radius=5.0),
    "jet_q27": dict(nx=34, ny=21, nz=17, q=27, jet_radius=4.0),
}


@pytest.mark.parametrize("arrangement", ["pool", "separate"])
@pytest.mark.parametrize("planned", [False, True], ids=["uniform", "planned"])
@pytest.mark.parametrize("label", list(CASES))
def test_separate_fields_reference_layout_match_oracle(oracle, label, planned, arrangement):
    """Jet case in the reference's NX*NY*NZ ghost-column layout with fields apart: the marching
    kernel (forced) and the planned launch give the oracle's bits."""
    cg = load_product()
    p = oracle.Params(ybc=oracle.Y_JET, **CASES[label])
    NX, NY, NZ = p.nx + 2, p.ny + 2, p.nz + 2
    A0 = oracle.init_jet(p)
    steps = 70 if planned else 6
    want = oracle.run(p, A0, steps)
    ref = np.zeros((11, NZ, NY, NX), dtype=np.float32)
    ref[:, 1:-1, :, 1:-1] = A0
    cfg = _cfg(cg, p)
    L = cg.Layout(1 + NX * NY, NX, NX * NY)
    dev = DeviceStates(cg, NX * NY * NZ, arrangement)
    try:
        dev.put(0, ref)
        mom = [dev.moments(0), dev.moments(1)]
        print(f"{label} planned={planned} {arrangement} steps={steps}", flush=True)
        plan = cg.Plan(cfg) if planned else None
        for s in range(1, steps + 1):
            if planned:
                cg.launch_rcs_planned(mom[(s - 1) % 2], mom[s % 2], cfg, plan, L, step=s)
            else:
                cg.launch_rcs(mom[(s - 1) % 2], mom[s % 2], cfg, L, step=s, kernel=cg.KERNEL_MARCH)
        dev.sync()
        if plan:
            plan.close()
        got = dev.get(steps % 2, (NZ, NY, NX))
    finally:
        dev.close()
    np.testing.assert_array_equal(got[:, 1:-1, :, 1:-1], want)


@pytest.mark.parametrize("q", [19, 27])
def test_separate_fields_equal_one_block_at_128(q):
    """128^3 drop, 40 launch-level steps on the marching kernel: fields apart (pointer form) give
    the bits of one-block states (buffer form), at no less than 85% of the rate (reported)."""
    cg = load_product()
    cfg = cg.Config(nx=128, ny=128, nz=128, q=q, tau=0.8, sigma=0.01, beta=0.7)
    n = 128 ** 3
    with cg.Domain(cfg) as d:
        d.init_bubble(radius=24.0, width=4.0)
        init = d.get_state()
    hip = _hip(cg)
    out, ms = {}, {}
    for kind in ("block", "pool"):
        dev = DeviceStates(cg, n, kind)
        try:
            dev.put(0, init)
            mom = [dev.moments(0), dev.moments(1)]

            def run(s0, k):
                for s in range(s0, s0 + k):
                    cg.launch_rcs(mom[(s - 1) % 2], mom[s % 2], cfg, step=s, kernel=cg.KERNEL_MARCH)

            run(1, 10)
            dev.sync()
            ev = [ctypes.c_void_p(), ctypes.c_void_p()]
            for e in ev:
                assert hip.hipEventCreate(ctypes.byref(e)) == 0
            hip.hipEventRecord(ev[0], None)
            run(11, 30)
            hip.hipEventRecord(ev[1], None)
            assert hip.hipEventSynchronize(ev[1]) == 0
            t = ctypes.c_float()
            assert hip.hipEventElapsedTime(ctypes.byref(t), ev[0], ev[1]) == 0
            for e in ev:
                hip.hipEventDestroy(e)
            ms[kind] = t.value / 30
            out[kind] = dev.get(0, (128, 128, 128))
        finally:
            dev.close()
    print(f"q{q} ms/step {ms} MLUPS " + str({k: round(n / v / 1e3) for k, v in ms.items()}), flush=True)
    assert np.array_equal(out["block"], out["pool"])
    assert ms["pool"] < ms["block"] / 0.85
