"""lsamd_sweep_static, the sweep that allocates nothing, through the C ABI.

It runs lsamd_sweep's kernels on the library's static device scratch instead
of a caller's workspace, so its periodograms and peaks must equal
lsamd_sweep's bit for bit.  Two calls in a row per case, the second on a
smaller input (the scratch still holds the first call's contents), over the
matrix-core and the VALU split32 tiles, the FP64 split form, both variants
and both schedules; and a sweep whose workspace exceeds the scratch is
refused before anything is enqueued.
"""
from ctypes import byref, c_void_p

import numpy as np
import pytest

from gpu_lomb_scargle_amd import _native, synth

pytestmark = pytest.mark.gpu
F64 = np.float64
SCRATCH = 64 << 20  # LSAMD_STATIC_SCRATCH_BYTES


def _plan(torch, variant, form, K, nf, schedule, seed, tile=None):
    from gpu_lomb_scargle_amd import device
    gen = variant == _native.GENERALIZED
    b = synth.lognormal_batch(K, 140, 0.6, 20, 1000, 1800, gen, seed)
    t, y, dy, off = device.numpy_to_device(b["obj_id"], b["t"], b["y"], b["dy"], F64)
    return device.SweepPlan(t, y, dy, off, min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX, num_freqs=nf,
                            variant=variant, form=form, top_k=3, schedule=schedule, tile=tile)


def _outputs(torch, plan):
    torch.cuda.synchronize()
    return [x.clone() for x in (plan.pgram, plan.peak_index, plan.peak_power, plan.peak_period)]


def _clear(torch, plan):
    for x in (plan.pgram, plan.peak_power, plan.peak_period):
        x.fill_(-7.0)
    plan.peak_index.fill_(-1)
    torch.cuda.synchronize()


CASES = [  # (form, tile, schedule): first call, then a smaller second one
    (_native.FORM_SPLIT32, None, _native.SCHEDULE_GRID),
    (_native.FORM_SPLIT32, (256, 4), _native.SCHEDULE_PERSISTENT),
    (_native.FORM_SPLIT, None, _native.SCHEDULE_PERSISTENT),
]


@pytest.mark.parametrize("variant", [_native.STANDARD, _native.GENERALIZED], ids=["standard", "generalized"])
@pytest.mark.parametrize("case", range(len(CASES)), ids=["split32_mfma_grid", "split32_valu_persistent",
                                                          "split_persistent"])
def test_static_scratch_sweep_equals_workspace_sweep(variant, case):
    torch = pytest.importorskip("torch")
    L = _native.core()
    form, tile, schedule = CASES[case]
    if tile is not None and variant == _native.GENERALIZED:
        tile = (256, 4)
    for K, nf, seed in ((240, 60000, 901 + case), (17, 7001, 902 + case)):
        plan = _plan(torch, variant, form, K, nf, schedule, seed, tile)
        need = int(L.lsamd_sweep_workspace_size(byref(plan.args)))
        print(f"variant {variant} form {form} tile {tile} K {K} nf {nf}: workspace {need} bytes", flush=True)
        assert need <= SCRATCH
        plan.run()
        ref = _outputs(torch, plan)
        _clear(torch, plan)
        rc = L.lsamd_sweep_static(byref(plan.args), c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == _native.OK, L.lsamd_last_error()
        got = _outputs(torch, plan)
        for g, r in zip(got, ref):
            assert torch.equal(g, r)
        assert (ref[1] >= 0).all()


def test_static_scratch_refuses_a_larger_sweep():
    """1000 objects x 1e6 bins need ~180 MB of workspace: refused with a
    message naming the limit, before any launch (the device pointers are
    never dereferenced)."""
    pytest.importorskip("torch")
    L = _native.core()
    a = _native.SweepArgs()
    a.variant, a.dtype, a.form = _native.STANDARD, _native.F64, _native.FORM_SPLIT32
    a.num_objects, a.num_freqs, a.num_obs, a.top_k = 1000, 1000000, 163652, 1
    a.min_freq, a.max_freq = synth.OMEGA_MIN, synth.OMEGA_MAX
    a.offsets = a.t = a.y = a.peak_index = 1
    assert L.lsamd_sweep_workspace_size(byref(a)) > SCRATCH
    assert L.lsamd_sweep_static(byref(a), None) == _native.ERR_INVALID
    assert b"static scratch" in L.lsamd_last_error()
