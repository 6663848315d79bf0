"""Benchmark of the Lomb-Scargle hot path on one MI355X.

Workload: the standard L-S batch of 1000 objects with log-normal observation
counts (median 140, sigma 0.6, clipped to 20..1000; ~165k rows) over 10^6
angular frequencies 1.005..150.796 rad/day, inputs resident in HBM (FP64
arrays), periodogram written to HBM and the per-object best period fused on
the device.  A step is one full sweep (statistics, records, sweep kernel,
peak reduction).

Reported (one JSON line on stdout):
  value           freq*obs terms/s of the FP32-split form: split32 (FP64 phase
                  reduction, FP32 hardware sin/cos seeds; sums on the FP16
                  matrix cores as row x column factor products with
                  head/tail splitting and FP32 accumulation folded into
                  FP64) when it
                  meets max relative power error <= 1e-5 against the FP64 form
                  with identical best periods, else the split form (FP64
                  rotation and sums); config.form names the one reported
  split, split32  both forms: time, kernel time, full-size accuracy
  fp64_plain      the same workload in the reference's FP64 arithmetic
  objects_per_s   objects per second of the reported form
  roofline        the reported form's sweep kernel: algorithmic work per
                  launch / its duration (HIP events on the launch stream)
                  against the peak of the units it runs on: for the
                  matrix-core tile 16 flops per term against the dense FP16
                  matrix peak (2.5 PFLOP/s); for a VALU tile the canonical
                  20 flop-eq per term against the FP32 vector peak
                  (157.3 TFLOP/s); traffic = HBM bytes per launch
                  from the committed rocprofv3 counter summary
  cpu_baseline    the reference's own CPU batch path (mode 4, OpenMP) timed on
                  the host on a bounded object sample of the same workload
  fp32_batch, generalized_batch, uneven_batch, refinement_batch, single_object
                  the same batch with FP32 inputs and periodogram, the
                  generalized 1000-object batch, the 100000-object top-3
                  batch (persistent, graph, autotuned), the 4096-object grid
                  search + Newton refinement + gradient workload, and the
                  3555-observation single object beside the reference's CPU
                  MODE 5
  end_to_end_pcie the 1000-object batch through LombScarglePy with host
                  arrays (PCIe-inclusive; never the headline value)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# VALU issue model, canonical work per freq*obs term (lane-cycles, FP32,
# non-packed): standard = 1 phase multiply + sin + cos + 5 FMA into XC, XS,
# CC, SS, CS; generalized = the same + 2 (w*c and the weighted C sum) as the
# minimal single-pass form.  Transcendental cost measured on gfx950 by
# gpu_lomb_scargle_amd/lib/valu_rates: v_sin_f32 / v_cos_f32 issue at half the
# v_mul_f32 rate, so each costs 2 lane-cycles.
TRANS_CYCLES = 2.0
CYCLES_PER_TERM = 1.0 + 2.0 * TRANS_CYCLES + 5.0          # 10 (standard)
CYCLES_PER_TERM_GEN = CYCLES_PER_TERM + 2.0               # 12 (generalized)
# Peaks.  MI355X's published FP32 vector peak is 157.3 TFLOP/s.  The 4 SIMD x
# 16 lane x 2.4 GHz issue model (one wave64 VALU instruction per SIMD per 4
# cycles) gives 3.93e13 lane-cycles/s = 78.6 TFLOP/s in non-packed FMAs; the
# rocprofv3 counters show the split32 sweep issuing more VALU lane-
# instructions per CU-cycle than that model allows, so the published peak is
# the ceiling roofline.frac is taken against, and the 78.6 figure is reported
# beside it.
PEAK_LANE_CYCLES = 256 * 4 * 16 * 2.4e9
PEAK_TFLOPS_NONPACKED = 2.0 * PEAK_LANE_CYCLES / 1e12   # 78.6
PEAK_TFLOPS = 157.3
# The standard split32 sweep's default tile runs on the matrix cores: each
# term is two complex multiply-accumulates (Z1 = sum y e^{i w t}, Z2 = sum
# e^{2 i w t}) = 16 flops, executed as FP16 products with head/tail splitting
# (3 FP16 products per real product, 48 flops per term).  MI355X's published
# dense FP16 matrix peak is 2.5 PFLOP/s.
MFMA_TILE = (64, 16)
VALU_TILE = (256, 8)   # the fastest split32 tile on the vector ALUs
MFMA_FLOPS_PER_TERM = 16.0
MFMA_EXEC_FLOPS_PER_TERM = 48.0
PEAK_TFLOPS_MFMA_F16 = 2500.0
MAX_REL_ERR = 1e-5
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_sweep_hbm_traffic_1000x1e6.txt")


def measured_traffic(kernel_prefix):
    """HBM bytes per launch of a sweep kernel from the rocprofv3 counter
    summary (calibrated FETCH_SIZE + WRITE_SIZE), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("TRAFFIC_JSON "):
                    table = json.loads(line[len("TRAFFIC_JSON "):])
                    for name, rec in table.items():
                        if name.startswith(kernel_prefix):
                            return rec
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(data, nf, sample_objects, threads):
    """Reference CPU batch (MODE 4) on the first `sample_objects` objects."""
    import oracle
    counts = data["counts"][:sample_objects]
    rows = int(counts.sum())
    o, t, y = data["obj_id"][:rows], data["t"][:rows], data["y"][:rows]
    kind = "reference" if oracle.ref_available(False, np.float64, True, threads) else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        res = oracle.ref_batch(o, t, y, None, 1.005, 150.796, nf, False, np.float64, threads=threads)
    else:
        res = oracle.batch(o, t, y, None, 1.005, 150.796, nf, False, np.float64, nthreads=threads)
    dt = time.perf_counter() - t0
    terms = float(rows) * nf
    return {"value": terms / dt, "unit": "terms/s", "cores": threads, "kind": kind,
            "sample": f"{sample_objects} of the 1000 objects ({rows} rows) x {nf} freqs, FP64, "
                      f"reference MODE 4 (OpenMP dynamic over objects), {dt:.1f} s",
            "objects_per_s": sample_objects / dt}, np.asarray(res[1], np.float64)


def periods_vs_cpu(gpu_periods, cpu_periods):
    """Best periods of the sampled objects: each GPU form against the
    reference CPU batch mode on the same objects."""
    out = {}
    for name, p in gpu_periods.items():
        p = np.asarray(p, np.float64)[:cpu_periods.size]
        out[name] = {"identical": f"{int((p == cpu_periods).sum())}/{cpu_periods.size}",
                     "max_rel_diff": float(np.max(np.abs(p - cpu_periods) / cpu_periods))}
    return out


def full_accuracy(torch, res, ref, K):
    """Split-form output vs the FP64 form over every bin."""
    a, b = res["pgram"], ref["pgram"]
    pmax = b.max(dim=1, keepdim=True).values
    mask = b >= 1e-3 * pmax
    rel = ((a - b).abs() / b.abs()).where(mask, torch.zeros_like(a))
    out = {"max_rel_err_masked": float(rel.max()),
           "median_rel_err": float(((a - b).abs() / b.abs()).flatten()[::97].median()),
           "max_abs_err_over_pmax": float(((a - b).abs() / pmax).max()),
           "argmax_identical": f"{int((res['index'] == ref['index']).sum())}/{K}",
           "sum_periods": float(res["period"].sum()),
           "sum_periods_fp64": float(ref["period"].sum()),
           "mask": "P >= 1e-3 * P_max per object"}
    del rel, mask
    return out


def timed_plan(torch, L, plan, steps, warmup, stream=None):
    for _ in range(warmup):
        plan.run(stream)
    torch.cuda.synchronize()
    L.lsamd_set_timing(1)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run(stream)
        torch.cuda.synchronize()
        kms.append(L.lsamd_last_sweep_kernel_ms())
    t1 = time.perf_counter()
    L.lsamd_set_timing(0)
    return (t1 - t0) * 1e3 / steps, float(np.mean(kms))


def fp32_batch(torch, L, dev, data, nf, steps):
    """The same 1000-object batch with FP32 inputs and FP32 periodogram (the
    reference's float build): plain FP32 (the reference's float arithmetic,
    sin/cos/atan2 of float arguments correctly rounded) and split32, the
    latter checked against FP64 on the float-rounded inputs."""
    from gpu_lomb_scargle_amd import _native, device, synth
    K = int(data["counts"].size)
    terms = float(data["counts"].sum()) * nf
    out = {"workload": "the 1000-object batch with FP32 inputs and an FP32 periodogram", "terms_per_step": terms}
    t32, y32, _, off = device.numpy_to_device(data["obj_id"], data["t"], data["y"], None, np.float32, dev)
    mk = lambda t, y, form: device.SweepPlan(t, y, None, off, min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX,
                                             num_freqs=nf, variant=_native.STANDARD, form=form, top_k=1)
    p = mk(t32, y32, _native.FORM_PLAIN)
    ms, kms = timed_plan(torch, L, p, 1, 1)
    out["fp32_plain"] = {"value": terms / (ms / 1e3), "ms_per_step": round(ms, 3), "kernel_ms": round(kms, 3)}
    del p
    p = mk(t32, y32, _native.FORM_SPLIT32)
    ms, kms = timed_plan(torch, L, p, steps, 1)
    ref = mk(t32.double(), y32.double(), _native.FORM_PLAIN)
    ref.run()
    torch.cuda.synchronize()
    res = p.results()
    res = {k: (v.double() if v.is_floating_point() else v) for k, v in res.items()}
    out["split32"] = {"value": terms / (ms / 1e3), "ms_per_step": round(ms, 3), "kernel_ms": round(kms, 3),
                      "accuracy": full_accuracy(torch, res, ref.results(), K),
                      "note": "against FP64 on the float-rounded inputs"}
    del p, ref, res
    torch.cuda.empty_cache()
    return out


def generalized_batch(torch, L, dev, steps):
    """1000 objects over 2000 d, un-centred magnitudes (15..20) with dmag:
    split forms against the FP64 generalized form, 1e6 bins."""
    from gpu_lomb_scargle_amd import _native, device, synth
    data = synth.generalized_batch_1000()
    K = int(data["counts"].size)
    nf = 1000000
    terms = float(data["counts"].sum()) * nf
    t, y, dy, off = device.numpy_to_device(data["obj_id"], data["t"], data["y"], data["dy"],
                                           np.float64, dev)
    mk = lambda form: device.SweepPlan(t, y, dy, off, min_freq=synth.OMEGA_MIN,
                                       max_freq=synth.OMEGA_MAX, num_freqs=nf,
                                       variant=_native.GENERALIZED, form=form, top_k=1)
    out = {"workload": "generalized L-S batch: 1000 objects (median 140 obs), 2000 d span, "
                       "magnitudes 15..20 not subtracted, dmag logU(0.01, 0.2), 1e6 bins",
           "terms_per_step": terms}
    plain = mk(_native.FORM_PLAIN)
    ms, kms = timed_plan(torch, L, plain, 1, 1)
    out["fp64_plain"] = {"value": terms / (ms / 1e3), "ms_per_step": round(ms, 3)}
    ref = plain.results()
    for name, form in (("split", _native.FORM_SPLIT), ("split32", _native.FORM_SPLIT32)):
        p = mk(form)
        ms, kms = timed_plan(torch, L, p, steps, 1)
        out[name] = {"value": terms / (ms / 1e3), "ms_per_step": round(ms, 3),
                     "kernel_ms": round(kms, 3),
                     "valu_roofline_frac": round(terms / (kms / 1e3) * 2.0 * CYCLES_PER_TERM_GEN / (PEAK_TFLOPS * 1e12), 4),
                     "accuracy": full_accuracy(torch, p.results(), ref, K)}
        del p
    del plain, ref
    torch.cuda.empty_cache()
    return out


def uneven_batch(torch, L, dev, steps):
    """100000 objects, N log-normal (median 60, sigma 1, 5..5000), 2e5 bins,
    top-3 peaks, no periodogram: persistent + graph-captured + autotuned vs
    one workgroup per item, vs a per-object launch loop."""
    from ctypes import byref, c_void_p
    from gpu_lomb_scargle_amd import _native, device, synth
    out = {"workload": "uneven batch: 100000 objects, log-normal N (median 60, sigma 1.0, 5..5000), "
                       "2e5 angular freqs 1.005..150.796, top-3 peaks per object, no periodogram"}
    nf = 200000
    for vname, variant in (("standard", _native.STANDARD), ("generalized", _native.GENERALIZED)):
        data = synth.uneven_batch_100000(with_error=variant == _native.GENERALIZED)
        K = int(data["counts"].size)
        rows = int(data["counts"].sum())
        terms = float(rows) * nf
        t, y, dy, off = device.numpy_to_device(data["obj_id"], data["t"], data["y"], data["dy"],
                                               np.float64, dev)
        res = {"objects": K, "rows": rows, "terms_per_step": terms}
        form = _native.FORM_SPLIT32
        kw = dict(min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX, num_freqs=nf, variant=variant,
                  form=form, top_k=3, pgram=False)
        pers = device.SweepPlan(t, y, dy, off, schedule=_native.SCHEDULE_PERSISTENT, **kw)
        T, M = pers.autotune()
        stream = torch.cuda.Stream()
        pers.capture_graph(stream)
        ms, kms = timed_plan(torch, L, pers, steps, 1, stream)
        cyc = CYCLES_PER_TERM_GEN if variant == _native.GENERALIZED else CYCLES_PER_TERM
        res["persistent_graph"] = {"value": terms / (ms / 1e3), "objects_per_s": K / (ms / 1e3),
                                   "ms_per_step": round(ms, 3), "tile": f"{T}x{M}",
                                   "valu_roofline_frac": round(terms / (ms / 1e3) * 2.0 * cyc / (PEAK_TFLOPS * 1e12), 4)}
        grid = device.SweepPlan(t, y, dy, off, schedule=_native.SCHEDULE_GRID, **kw)
        ms_g, _ = timed_plan(torch, L, grid, steps, 1)
        res["grid"] = {"value": terms / (ms_g / 1e3), "ms_per_step": round(ms_g, 3)}
        same = bool(torch.equal(pers.results()["index"], grid.results()["index"]))
        # the per-object launch loop: one lsamd_sweep (4 kernels) per object
        a = _native.SweepArgs()
        ctypes_mod = __import__("ctypes")
        ctypes_mod.memmove(byref(a), byref(pers.args), ctypes_mod.sizeof(a))
        # the same kernels as the batch, one object per launch (no observation
        # slices: a loop over objects is what the batch schedules replace)
        a.num_objects, a.order, a.schedule, a.obs_slices = 1, None, _native.SCHEDULE_GRID, 1
        ws1 = torch.empty(int(L.lsamd_sweep_workspace_size(byref(a))), dtype=torch.uint8, device=dev)
        idx = torch.zeros(K * 3, dtype=torch.int32, device=dev)
        pw = torch.zeros(K * 3, dtype=torch.float64, device=dev)
        pe = torch.zeros(K * 3, dtype=torch.float64, device=dev)
        s0 = torch.cuda.current_stream().cuda_stream
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        base_off, base_i, base_p, base_e = off.data_ptr(), idx.data_ptr(), pw.data_ptr(), pe.data_ptr()
        for k in range(K):
            a.offsets = base_off + 8 * k
            a.peak_index = base_i + 12 * k
            a.peak_power = base_p + 24 * k
            a.peak_period = base_e + 24 * k
            rc = L.lsamd_sweep(byref(a), c_void_p(ws1.data_ptr()), ws1.numel(), c_void_p(s0))
            if rc:
                raise RuntimeError(L.lsamd_last_error())
        torch.cuda.synchronize()
        ms_loop = (time.perf_counter() - t0) * 1e3
        loop_same = bool(torch.equal(idx.view(K, 3), grid.results()["index"]))
        res["per_object_loop"] = {"value": terms / (ms_loop / 1e3), "ms": round(ms_loop, 1),
                                  "launches": K}
        res["persistent_speedup_vs_loop"] = round(ms_loop / ms, 1)
        res["peaks_identical_persistent_grid_loop"] = same and loop_same
        out[vname] = res
        del pers, grid, ws1
        torch.cuda.empty_cache()
    return out


def refinement_batch(torch, L, dev):
    """4096 objects (median 200 obs, sigma 0.5), true P log-uniform in
    [0.05, 10] d: FP32-split grid search over 5e5 bins with the device argmax,
    then Newton refinement (lsamd_refine_peaks) and one forward+backward
    evaluation of the power (lsamd_power_grad: dP/domega, dP/dy, dP/d(dy))."""
    from gpu_lomb_scargle_amd import _native, device, grad, synth
    nf = 500000
    dw = (synth.OMEGA_MAX - synth.OMEGA_MIN) / nf
    out = {"workload": "refinement: 4096 objects, log-normal N (median 200, sigma 0.5), 2000 d span, "
                       "split32 grid search over 5e5 angular freqs (top-1 on device), Newton steps on "
                       "omega to |step| < 1e-9 rad/day, power gradient w.r.t. omega, y, dy"}
    for vname, variant in (("standard", _native.STANDARD), ("generalized", _native.GENERALIZED)):
        data = synth.refinement_batch_4096(with_error=variant == _native.GENERALIZED)
        K = int(data["counts"].size)
        rows = int(data["counts"].sum())
        t, y, dy, off = device.numpy_to_device(data["obj_id"], data["t"], data["y"], data["dy"],
                                               np.float64, dev)
        plan = device.SweepPlan(t, y, dy, off, min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX,
                                num_freqs=nf, variant=variant, form=_native.FORM_SPLIT32, pgram=False,
                                top_k=1)
        ms_sweep, _ = timed_plan(torch, L, plan, 3, 1)
        idx = plan.results()["index"][:, 0].to(torch.float64)
        om0 = synth.OMEGA_MIN + idx * dw
        for _ in range(2):
            grad.refine_peaks(om0, t, y, dy, off, variant, max_step=dw, tol=1e-9, max_iter=30)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        om, P, it = grad.refine_peaks(om0, t, y, dy, off, variant, max_step=dw, tol=1e-9, max_iter=30)
        torch.cuda.synchronize()
        ms_ref = (time.perf_counter() - t0) * 1e3
        # one forward + backward through the autograd Function
        yy = y.clone().requires_grad_(True)
        dd = dy.clone().requires_grad_(True) if dy is not None else None
        omr = om.clone().requires_grad_(True)
        for _ in range(2):
            grad.power(omr, t, yy, dd, off, variant).sum().backward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        grad.power(omr, t, yy, dd, off, variant).sum().backward()
        torch.cuda.synchronize()
        ms_fb = (time.perf_counter() - t0) * 1e3
        P0, _, _ = grad.power_and_derivatives(om0, t, y, dy, off, variant)
        om0n, omn, itn = om0.cpu().numpy(), om.cpu().numpy(), it.cpu().numpy()
        Ptrue = data["period"]
        wt = 2 * np.pi / Ptrue
        hit = np.abs(om0n - wt) < 2 * dw
        e0 = np.abs(2 * np.pi / om0n - Ptrue)[hit] / Ptrue[hit]
        e1 = np.abs(2 * np.pi / omn - Ptrue)[hit] / Ptrue[hit]
        out[vname] = {
            "objects": K, "rows": rows, "grid_search_ms": round(ms_sweep, 3),
            "grid_search_terms_per_s": rows * float(nf) / (ms_sweep / 1e3),
            "refine_ms": round(ms_ref, 3), "refine_objects_per_s": K / (ms_ref / 1e3),
            "newton_iterations_median": float(np.median(np.abs(itn))),
            "converged": f"{int((itn > 0).sum())}/{K}",
            "power_never_lower": bool((P >= P0 * (1 - 1e-12)).all().item()),
            "forward_backward_ms": round(ms_fb, 3),
            "grid_peak_at_true_frequency": f"{int(hit.sum())}/{K}",
            "period_rel_err_grid": {"median": float(np.median(e0)), "p90": float(np.percentile(e0, 90))},
            "period_rel_err_refined": {"median": float(np.median(e1)), "p90": float(np.percentile(e1, 90))},
            "note": "errors over objects whose grid peak lies within 2 bins of the true frequency",
        }
        del plan, yy, dd, omr
        torch.cuda.empty_cache()
    return out


def single_object(torch, L, dev, cpu_threads):
    """One object, N = 3555 over 1800 nights, 1e5 angular freqs 1.005..150.796:
    GPU plain FP64, split and split32 against the reference CPU single-object
    path (MODE 5, OpenMP over frequencies, FP64), full run."""
    import oracle
    from gpu_lomb_scargle_amd import _native, device, synth
    data = synth.single_object_3555()
    nf = 100000
    n = data["t"].size
    terms = float(n) * nf
    t, y, dy, off = device.numpy_to_device(None, data["t"], data["y"], None, np.float64, dev)
    out = {"workload": "single object: N = 3555 over 1800 nights, 1e5 angular freqs 1.005..150.796, "
                       "normalised standard power, periodogram + best period",
           "terms_per_step": terms}
    res = {}
    for name, form in (("fp64_plain", _native.FORM_PLAIN), ("split", _native.FORM_SPLIT),
                       ("split32", _native.FORM_SPLIT32)):
        p = device.SweepPlan(t, y, None, off, min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX,
                             num_freqs=nf, variant=_native.STANDARD, form=form, top_k=1)
        ms, kms = timed_plan(torch, L, p, 5, 2)
        res[name] = p.results()
        out[name] = {"value": terms / (ms / 1e3), "ms_per_step": round(ms, 4), "kernel_ms": round(kms, 4),
                     "period": float(res[name]["period"][0, 0])}
    ref = res["fp64_plain"]["pgram"]
    for name in ("split", "split32"):
        a = res[name]["pgram"]
        m = ref >= 1e-3 * ref.max()
        out[name]["max_rel_err_masked"] = float(((a - ref).abs() / ref.abs()).where(m, torch.zeros_like(a)).max())
        out[name]["argmax_identical"] = bool(int(res[name]["index"][0, 0]) == int(res["fp64_plain"]["index"][0, 0]))
    kind = "reference" if oracle.ref_available(False, np.float64, True, cpu_threads) else "port"
    yy = data["y"].copy()
    t0 = time.perf_counter()
    if kind == "reference":
        pg = oracle.ref_single(data["t"], yy, None, synth.OMEGA_MIN, synth.OMEGA_MAX, nf, False, np.float64,
                               threads=cpu_threads)
    else:
        pg = oracle.single(data["t"], yy, None, synth.OMEGA_MIN, synth.OMEGA_MAX, nf, False, np.float64,
                           nthreads=cpu_threads)
    dt = time.perf_counter() - t0
    pg = np.asarray(pg[0] if isinstance(pg, tuple) else pg).reshape(-1)
    gpu = res["fp64_plain"]["pgram"].cpu().numpy().reshape(-1)
    out["cpu_mode5"] = {"value": terms / dt, "seconds": round(dt, 3), "cores": cpu_threads, "kind": kind,
                        "fp64_plain_max_abs_diff_over_pmax": float(np.abs(gpu - pg).max() / pg.max()),
                        "argmax_identical": bool(int(np.argmax(pg)) == int(res["fp64_plain"]["index"][0, 0]))}
    out["true_period"] = float(data["period"][0])
    return out


def end_to_end(torch, nf):
    """The 1000-object batch through the drop-in entry point LombScarglePy
    (libpylsnoerrordouble.so, MODE 1, LSAMD_FORM=split32): host arrays in,
    the 8 GB periodogram copied back into pageable host memory.  PCIe-inclusive;
    not the headline value."""
    import ctypes as ct
    from gpu_lomb_scargle_amd import _native, synth
    data = synth.standard_batch_1000()
    lib = _native.load(_native.DROPIN_LIBS[(False, "double")])
    P = ct.POINTER
    lib.LombScarglePy.argtypes = [P(ct.c_uint), P(ct.c_double), P(ct.c_double), P(ct.c_double), ct.c_uint,
                                  ct.c_double, ct.c_double, ct.c_uint, ct.c_int, P(ct.c_double)]
    lib.LombScarglePy.restype = None
    o = np.ascontiguousarray(data["obj_id"], np.uint32)
    t = np.ascontiguousarray(data["t"], np.float64)
    y = np.ascontiguousarray(data["y"], np.float64)
    dy = np.ones_like(t)
    K = int(data["counts"].size)
    pgram = np.empty(K * nf, np.float64)
    pgram[::4096] = 0.0
    old_form = os.environ.get("LSAMD_FORM")
    os.environ["LSAMD_FORM"] = "split32"
    os.environ["LSAMD_GPU_STATS"] = "0"
    cp = lambda a, ty: a.ctypes.data_as(P(ty))
    saved = os.dup(1)
    times = []
    try:
        os.dup2(2, 1)  # the entry point prints progress lines on stdout, as the reference does
        for _ in range(3):
            t0 = time.perf_counter()
            lib.LombScarglePy(cp(o, ct.c_uint), cp(t, ct.c_double), cp(y, ct.c_double), cp(dy, ct.c_double),
                              o.size, synth.OMEGA_MIN, synth.OMEGA_MAX, nf, 1, cp(pgram, ct.c_double))
            times.append(time.perf_counter() - t0)
    finally:
        ct.CDLL(None).fflush(None)  # the C library's stdout buffer goes to stderr too
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)
        if old_form is None:
            del os.environ["LSAMD_FORM"]
        else:
            os.environ["LSAMD_FORM"] = old_form
    dt = min(times[1:])
    terms = float(t.size) * nf
    return {"workload": "1000-object batch through LombScarglePy (MODE 1, double library, LSAMD_FORM=split32): "
                        "host arrays in, 8 GB periodogram copied to pageable host memory",
            "value": terms / dt, "unit": "terms/s", "seconds": round(dt, 3),
            "host_copy_GBps": round(K * nf * 8 / dt / 1e9, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-objects", type=int, default=32)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nf", type=int, default=1000000)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the generalized and uneven-batch workloads")
    ap.add_argument("--profile-once", action="store_true",
                    help="for rocprofv3 counter runs: one 1 GiB device copy (calibration of the "
                         "HBM byte counters), then one sweep of each form; no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (a batch would shard "
              "by objects, one slice per GPU, with no exchange)", file=sys.stderr)
        sys.exit(2)

    import torch

    from gpu_lomb_scargle_amd import _native, device, synth

    dev = torch.device("cuda:0")
    L = _native.core()
    data = synth.standard_batch_1000()
    K = int(data["counts"].size)
    rows = int(data["counts"].sum())
    nf = args.nf
    terms = float(rows) * nf
    log(f"workload: {K} objects, {rows} rows, {nf} freqs, {terms:.3e} terms/step")
    t, y, dy, off = device.numpy_to_device(data["obj_id"], data["t"], data["y"], None, np.float64, dev)

    def make(form):
        return device.SweepPlan(t, y, None, off, min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX,
                                num_freqs=nf, variant=_native.STANDARD, form=form, normalize=True,
                                top_k=1)

    def timed(plan, steps, warmup):
        for _ in range(warmup):
            plan.run()
        torch.cuda.synchronize()
        L.lsamd_set_timing(1)
        kms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.run()
            torch.cuda.synchronize()  # one step = one sweep; sync so the kernel time is read per step
            kms.append(L.lsamd_last_sweep_kernel_ms())
        t1 = time.perf_counter()
        L.lsamd_set_timing(0)
        return (t1 - t0) * 1e3 / steps, float(np.mean(kms))

    if args.profile_once:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)  # 1 GiB
        dst = src.clone()
        torch.cuda.synchronize()
        for form, tile in ((_native.FORM_SPLIT32, None), (_native.FORM_SPLIT32, VALU_TILE),
                           (_native.FORM_SPLIT, None), (_native.FORM_PLAIN, None)):
            if tile:
                L.lsamd_set_tile(form, _native.STANDARD, *tile)
            plan = make(form)
            plan.run()
            torch.cuda.synchronize()
            if tile:
                L.lsamd_set_tile(form, _native.STANDARD, 0, 0)
            del plan
        log("profile-once: 1 GiB copy, one split32 sweep with each tile (matrix-core, VALU 256x8), "
            "one split and one FP64 plain sweep")
        del src, dst
        return

    def tile_of(form):
        T, M = ctypes.c_int(), ctypes.c_int()
        L.lsamd_get_tile(form, _native.STANDARD, _native.F64, ctypes.byref(T), ctypes.byref(M))
        return T.value, M.value

    runs = {}
    # split32 with its default (matrix-core) tile and with the VALU tile
    for name, form, tile in (("split32", _native.FORM_SPLIT32, None),
                             ("split32_valu", _native.FORM_SPLIT32, VALU_TILE),
                             ("split", _native.FORM_SPLIT, None)):
        if tile:
            L.lsamd_set_tile(form, _native.STANDARD, *tile)
        plan = make(form)
        ms, kms = timed(plan, args.steps, args.warmup)
        T, M = tile_of(form)
        if tile:
            L.lsamd_set_tile(form, _native.STANDARD, 0, 0)
        runs[name] = {"plan": plan, "ms": ms, "kms": kms, "tile": (T, M)}
        log(f"{name}: {ms:.3f} ms/step, sweep kernel {kms:.3f} ms, tile {T}x{M}")

    plain = make(_native.FORM_PLAIN)
    ms_plain, kms_plain = timed(plain, args.fp64_steps, 1)
    log(f"fp64 plain: {ms_plain:.3f} ms/step, sweep kernel {kms_plain:.3f} ms")

    # full-size accuracy of the split forms against the FP64 form (all bins)
    pp = plain.results()
    for name, r in runs.items():
        r["acc"] = full_accuracy(torch, r["plan"].results(), pp, K)
        log(f"accuracy {name}: {r['acc']}")
    # best periods of the objects the CPU baseline also runs (MODE 4 check)
    sample_periods = {"fp64_plain": pp["period"][:args.cpu_objects, 0].cpu().numpy()}
    for name, r in runs.items():
        sample_periods[name] = r["plan"].results()["period"][:args.cpu_objects, 0].cpu().numpy()

    def meets(acc):
        return acc["max_rel_err_masked"] <= MAX_REL_ERR and acc["argmax_identical"] == f"{K}/{K}"

    chosen = "split32" if meets(runs["split32"]["acc"]) else "split"
    r = runs[chosen]
    T, M = r["tile"]
    value = terms / (r["ms"] / 1e3)
    kernel_terms_s = terms / (r["kms"] / 1e3)
    valu_achieved = kernel_terms_s * 2.0 * CYCLES_PER_TERM / 1e12
    mfma = chosen == "split32" and (T, M) == MFMA_TILE
    if mfma:
        kname = "k_sweep_mfma<double>"
        achieved = kernel_terms_s * MFMA_FLOPS_PER_TERM / 1e12
        bound = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_TFLOPS_MFMA_F16,
                 "unit": "TFLOP/s", "frac": round(achieved / PEAK_TFLOPS_MFMA_F16, 4),
                 "work_per_term": f"{MFMA_FLOPS_PER_TERM:g} flops (two complex multiply-accumulates) on "
                                  f"FP16 matrix cores; executed {MFMA_EXEC_FLOPS_PER_TERM:g} (FP16 head/tail "
                                  f"split, 3 products per real product)",
                 "executed_tflops": round(kernel_terms_s * MFMA_EXEC_FLOPS_PER_TERM / 1e12, 3),
                 "frac_executed": round(kernel_terms_s * MFMA_EXEC_FLOPS_PER_TERM / 1e12
                                        / PEAK_TFLOPS_MFMA_F16, 4),
                 "roofline_terms_per_s": PEAK_TFLOPS_MFMA_F16 * 1e12 / MFMA_EXEC_FLOPS_PER_TERM,
                 "valu_canonical": {"achieved": round(valu_achieved, 3), "peak": round(PEAK_TFLOPS, 2),
                                    "frac": round(valu_achieved / PEAK_TFLOPS, 4),
                                    "note": "the VALU-form roofline (20 flop-eq per term against the "
                                            "FP32 vector peak), which the matrix-core tile exceeds"}}
    else:
        kname = "k_sweep_%s<false, double, %d, %d>" % ("split32" if chosen == "split32" else "split64", T, M)
        achieved = valu_achieved
        bound = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(PEAK_TFLOPS, 2),
                 "unit": "TFLOP/s", "frac": round(achieved / PEAK_TFLOPS, 4),
                 "peak_nonpacked_issue_model": round(PEAK_TFLOPS_NONPACKED, 2),
                 "frac_nonpacked_issue_model": round(achieved / PEAK_TFLOPS_NONPACKED, 4),
                 "work_per_term": f"{CYCLES_PER_TERM:g} VALU lane-cycles = {2 * CYCLES_PER_TERM:g} flop-eq "
                                  f"(mul + sin + cos at 1/{TRANS_CYCLES:g} rate + 5 FMA), canonical",
                 "roofline_terms_per_s": PEAK_TFLOPS * 1e12 / (2.0 * CYCLES_PER_TERM)}
    tr = measured_traffic(kname)
    roofline = dict(bound)
    roofline.update({
                "traffic": tr["bytes"] if tr else None,
                "kernel": kname, "kernel_ms": round(r["kms"], 4),
                "algorithmic_hbm_bytes": K * nf * 8 + rows * 16,
                "pgram_write_GBps": round(K * nf * 8 / (r["kms"] / 1e3) / 1e9, 1),
                "traffic_source": ("rocprofv3 FETCH_SIZE x%.1f + WRITE_SIZE x%.1f (copy-calibrated), %s"
                                   % (tr["fetch_factor"], tr["write_factor"], os.path.relpath(TRAFFIC_FILE, ROOT))
                                   if tr else None)})

    def form_report(name):
        q = runs[name]
        return {"value": terms / (q["ms"] / 1e3), "ms_per_step": round(q["ms"], 3),
                "kernel_ms": round(q["kms"], 3), "tile": "%dx%d" % q["tile"], "accuracy": q["acc"],
                "meets_1e-5_and_argmax": meets(q["acc"])}

    forms = {"split32": "FP64 phase reduction, FP32 hardware sin/cos seeds; matrix-core tile: per 32x32 bins the "
                        "sums as complex products of a row and a column factor on FP16 matrix cores (head/tail "
                        "split, FP32 accumulation folded every 256 observations), FP64 power",
             "split32_valu": "FP64 phase reduction, FP32 hardware sin/cos seeds, FP32 rotation or three-term "
                             "recurrence across bins, 32-observation FP32 partial sums folded into FP64, FP64 power",
             "split": "FP64 phase reduction, FP32 hardware sin/cos seeds, FP64 rotation recurrence and sums, "
                      "FP64 power"}
    out = {
        "metric": "freq*obs terms/s and objects/s, FP32-split vs FP64 at <=1e-5 power error",
        "value": value, "unit": "terms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f16x2+f32+f64" if mfma else "f32+f64", "data": "synthetic",
        "config": {"workload": "standard L-S batch: 1000 objects, log-normal N (median 140, sigma 0.6, "
                               "20..1000), 1e6 angular freqs 1.005..150.796 rad/d, FP64 inputs in HBM, "
                               "f64 periodogram written + fused argmax",
                   "objects": K, "rows": rows, "freqs": nf, "terms_per_step": terms,
                   "form": chosen, "form_detail": forms[chosen], "tile": f"{T}x{M}"},
        "objects_per_s": K / (r["ms"] / 1e3),
        "accuracy": r["acc"],
        "split32": form_report("split32"),
        "split32_valu_tile": dict(form_report("split32_valu"), form_detail=forms["split32_valu"]),
        "split": form_report("split"),
        "fp64_plain": {"value": terms / (ms_plain / 1e3), "ms_per_step": round(ms_plain, 3),
                       "kernel_ms": round(kms_plain, 3),
                       "speedup_of_reported_form": round(ms_plain / r["ms"], 2)},
        "roofline": roofline,
    }
    if not args.no_extra:
        runs.clear()
        del plain, pp
        torch.cuda.empty_cache()
        out["fp32_batch"] = fp32_batch(torch, L, dev, data, nf, max(2, args.steps // 2))
        log(f"fp32 batch: {out['fp32_batch']}")
        out["generalized_batch"] = generalized_batch(torch, L, dev, max(2, args.steps // 2))
        log(f"generalized batch: {out['generalized_batch']}")
        out["uneven_batch"] = uneven_batch(torch, L, dev, max(2, args.steps // 2))
        log(f"uneven batch: {out['uneven_batch']}")
        out["refinement_batch"] = refinement_batch(torch, L, dev)
        log(f"refinement batch: {out['refinement_batch']}")
        out["single_object"] = single_object(torch, L, dev, args.cpu_threads)
        log(f"single object: {out['single_object']}")
        out["end_to_end_pcie"] = end_to_end(torch, nf)
        log(f"end to end: {out['end_to_end_pcie']}")
    if not args.no_cpu:
        out["cpu_baseline"], cpu_periods = cpu_baseline(data, nf, args.cpu_objects, args.cpu_threads)
        out["cpu_baseline"]["best_periods_vs_gpu"] = periods_vs_cpu(sample_periods, cpu_periods)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
