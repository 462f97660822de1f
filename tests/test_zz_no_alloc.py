"""cglbm_launch_rcs_planned_noalloc (the planned launch with its plan in the library's per-GPU
scratch array, no device allocation) against cglbm_launch_rcs_planned with a cglbm_plan (the
entry point it is a variant of): bit for bit, over enough steps to cross the recording steps and
a re-plan, for two calls in a row of different domains, the second smaller (a D3Q19 96^3 drop in
the compact layout, then a D3Q27 jet in the reference's ghost-column layout)."""
import numpy as np
import pytest

from conftest import load_product

pytestmark = pytest.mark.gpu

STEPS = 70  # two recording steps, planned steps, a re-plan at 64


def _run(cg, cfg, layout, carrier_cfg, init, planned_with):
    """STEPS launch-level steps on the two states of a carrier domain; returns the final carrier
    state."""
    with cg.Domain(carrier_cfg) as carrier:
        carrier.set_state(init)
        mom = [carrier.moments(0), carrier.moments(1)]
        stream = cg.lib().cglbm_domain_stream(carrier._h)
        plan = cg.Plan(cfg) if planned_with == "plan" else None
        try:
            for s in range(1, STEPS + 1):
                if plan is not None:
                    cg.launch_rcs_planned(mom[(s - 1) % 2], mom[s % 2], cfg, plan, layout, step=s, stream=stream)
                else:
                    cg.launch_rcs_planned_noalloc(mom[(s - 1) % 2], mom[s % 2], cfg, layout, step=s, stream=stream)
            carrier.sync()
        finally:
            if plan is not None:
                plan.close()
        # STEPS is even: the result is back in state 0, the carrier's current state
        return carrier.get_state()


def test_noalloc_planned_equals_planned_two_calls(oracle):
    cg = load_product()
    assert STEPS % 2 == 0
    # first call: D3Q19 96^3 drop, compact layout
    cfg = cg.Config(nx=96, ny=96, nz=96, q=19, tau=0.8, sigma=0.01, beta=0.7)
    p = oracle.Params(nx=96, ny=96, nz=96, q=19, tau=0.8, sigma=0.01, beta=0.7)
    A0 = oracle.init_bubble(p, radius=18.0, width=4.0)
    print("no-alloc call 1: D3Q19 96^3 drop", flush=True)
    want = _run(cg, cfg, cg.compact_layout(cfg), cfg, A0, "plan")
    got = _run(cg, cfg, cg.compact_layout(cfg), cfg, A0, "noalloc")
    np.testing.assert_array_equal(got, want)
    # second call, smaller: D3Q27 jet 40 x 36 x 30 in the reference's NX*NY*NZ ghost-column layout
    p = oracle.Params(nx=40, ny=36, nz=30, q=27, ybc=oracle.Y_JET, jet_radius=5.0)
    cfg = cg.Config(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    NX, NY, NZ = p.nx + 2, p.ny + 2, p.nz + 2
    ref = np.zeros((11, NZ, NY, NX), dtype=np.float32)
    ref[:, 1:-1, :, 1:-1] = oracle.init_jet(p)
    L = cg.Layout(1 + NX * NY, NX, NX * NY)
    carrier = cg.Config(nx=NX, ny=NY, nz=NZ)
    print("no-alloc call 2: D3Q27 jet 40 x 36 x 30, reference layout", flush=True)
    want = _run(cg, cfg, L, carrier, ref, "plan")
    got = _run(cg, cfg, L, carrier, ref, "noalloc")
    np.testing.assert_array_equal(got, want)
    # and the planned result is the update itself: the oracle's for the small case
    np.testing.assert_array_equal(got[:, 1:-1, :, 1:-1], oracle.run(p, oracle.init_jet(p), STEPS))


def test_noalloc_rejects_a_plan_beyond_its_scratch():
    """A single-round plan whose interface map exceeds the 384 KiB array is refused with the bytes
    it needs (checked before any device work)."""
    cg = load_product()
    big = cg.Config(nx=64, ny=64, nz=8192, q=19)  # 2 x 4 column tiles x 8192 planes x 8 waves: 512 KiB map
    # the states are real (right-sized) so that nothing could run out of bounds even if the check
    # did not fire
    with cg.Domain(big) as carrier:
        A, B = carrier.moments(0), carrier.moments(1)
        with pytest.raises(cg.CglbmError, match="bytes of scratch"):
            cg.launch_rcs_planned_noalloc(A, B, big)
