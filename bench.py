"""Benchmark of the Lomb-Scargle hot path on one MI355X.

Workload: the standard L-S batch of 1000 objects with log-normal observation
counts (median 140, sigma 0.6, clipped to 20..1000; ~165k rows) over 10^6
angular frequencies 1.005..150.796 rad/day, inputs resident in HBM (FP64
arrays), periodogram written to HBM and the per-object best period fused on
the device.  A step is one full sweep (statistics, records, sweep kernel,
peak reduction).

Output: the LAST stdout line is one compact JSON object (a few KB):
  value           freq*obs terms/s of the FP32-split form: split32 (FP64 phase
                  reduction, FP32 hardware sin/cos; sums on the FP16 matrix
                  cores as row x column factor products with head/tail
                  splitting and FP32 accumulation) when it meets max relative
                  power error <= 1e-5 against the FP64 form with identical
                  best periods, else the split form (FP64 rotation and sums);
                  config.form names the one reported
  accuracy        that form against the FP64 form over all 10^9 bins
  roofline        the reported form's sweep kernel: algorithmic work per
                  launch / its duration (HIP events on the launch stream)
                  against the peak of the units it runs on: for the
                  matrix-core tile 16 flops per term against the dense FP16
                  matrix peak (2.5 PFLOP/s); for a VALU tile the canonical
                  20 flop-eq per term against the FP32 vector peak
                  (157.3 TFLOP/s); traffic = HBM bytes per launch from the
                  committed rocprofv3 counter summary
  cpu_baseline    the reference's own CPU batch path (mode 4, OpenMP) timed on
                  the host's physical cores (as many as this process may use)
                  on a bounded object sample of the same workload, and with
                  the reference Python build's 16 threads
  configs         one line per other workload: the generalized 1000-object
                  batch, the 100000-object top-3 batch (persistent, graph,
                  autotuned), the 4096-object grid search + refinement, the
                  3555-observation single object, FP32 inputs, the
                  PCIe-inclusive drop-in call
The full nested report goes to stderr (and to --report PATH when given).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# VALU issue model, canonical work per freq*obs term (lane-cycles, FP32,
# non-packed): standard = 1 phase multiply + sin + cos + 5 FMA into XC, XS,
# CC, SS, CS.  Transcendental cost measured on gfx950 (profiles/
# valu_rates_gfx950.txt): v_sin_f32 / v_cos_f32 issue at half the v_mul_f32
# rate, so each costs 2 lane-cycles.  MI355X's published FP32 vector peak is
# 157.3 TFLOP/s.
TRANS_CYCLES = 2.0
CYCLES_PER_TERM = 1.0 + 2.0 * TRANS_CYCLES + 5.0          # 10 (standard)
PEAK_TFLOPS = 157.3
# The split32 default tiles run on the matrix cores: each standard term is two
# complex multiply-accumulates (Z1 = sum y e^{i w t}, Z2 = sum e^{2 i w t}) =
# 16 flops, each generalized term three (Zw, Zy, Z2) = 24 flops, executed as
# FP16 products with head/tail splitting (3 FP16 products per real product).
# MI355X's published dense FP16 matrix peak is 2.5 PFLOP/s.
MFMA_TILE = (64, 16)
VALU_TILE = (256, 8)   # the fastest split32 standard tile on the vector ALUs
MFMA_FLOPS_PER_TERM = {"standard": 16.0, "generalized": 24.0}
MFMA_EXEC_FACTOR = 3.0
PEAK_TFLOPS_MFMA_F16 = 2500.0
MAX_REL_ERR = 1e-5
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_sweep_hbm_traffic_1000x1e6.txt")
REF_BUILD_THREADS = 16   # NTHREADSCPU of the reference's Python build (source/Makefile:49)


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


def host_cpu():
    """Physical cores of the host (lscpu), the CPUs this process may run on
    (affinity mask, cgroup quota), and the thread count used: the physical
    cores, capped by what the process is allowed to use."""
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                             timeout=10).stdout
        pairs = {ln.strip() for ln in out.splitlines() if ln and not ln.startswith("#")}
        phys = len(pairs) or None
    except (OSError, subprocess.SubprocessError):
        pass
    logical = os.cpu_count() or 1
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else logical
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    threads = min(x for x in (phys or logical, affinity, quota or logical) if x)
    return {"physical_cores": phys, "logical_cpus": logical, "affinity_cpus": affinity,
            "cgroup_cpu_quota": quota, "threads": threads}


def cpu_run(data, nf, objects, threads):
    """Reference CPU batch (MODE 4) on the first `objects` objects."""
    import oracle
    counts = data["counts"][:objects]
    rows = int(counts.sum())
    o, t, y = data["obj_id"][:rows], data["t"][:rows], data["y"][:rows]
    kind = "reference" if oracle.ref_available(False, np.float64, True, threads) else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        res = oracle.ref_batch(o, t, y, None, 1.005, 150.796, nf, False, np.float64, threads=threads)
    else:
        res = oracle.batch(o, t, y, None, 1.005, 150.796, nf, False, np.float64, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": float(rows) * nf / dt, "objects_per_s": objects / dt, "seconds": round(dt, 2),
            "objects": objects, "rows": rows, "threads": threads, "kind": kind}, np.asarray(res[1], np.float64)


def cpu_baseline(data, nf, threads, target_s):
    """The reference's CPU batch path on the host's physical cores, with a
    sample sized to about target_s seconds (about 4e7 terms/s per core), and
    with the reference Python build's 16 threads on 32 objects."""
    mean_n = float(data["counts"].mean())
    objs = int(np.clip(round(target_s * 4e7 * threads / (mean_n * nf)), 32, data["counts"].size))
    full, periods = cpu_run(data, nf, objs, threads)
    ref16, _ = cpu_run(data, nf, 32, REF_BUILD_THREADS)
    return full, ref16, periods


def periods_vs_cpu(gpu_periods, cpu_periods):
    """Best periods of the sampled objects: each GPU form against the
    reference CPU batch mode on the same objects."""
    out = {}
    for name, p in gpu_periods.items():
        p = np.asarray(p, np.float64)[:cpu_periods.size]
        out[name] = {"identical": f"{int((p == cpu_periods).sum())}/{cpu_periods.size}",
                     "max_rel_diff": float(np.max(np.abs(p - cpu_periods) / cpu_periods))}
    return out


def full_accuracy(torch, res, ref, K, worst=5):
    """Split-form output vs the FP64 form over every bin, with the (object,
    bin) of the worst masked errors."""
    a, b = res["pgram"], ref["pgram"]
    pmax = b.max(dim=1, keepdim=True).values
    mask = b >= 1e-3 * pmax
    rel = ((a - b).abs() / b.abs()).where(mask, torch.zeros_like(a))
    top = torch.topk(rel.flatten(), worst)
    nf = a.shape[1]
    out = {"max_rel_err_masked": float(rel.max()),
           "median_rel_err": float(((a - b).abs() / b.abs()).flatten()[::97].median()),
           "max_abs_err_over_pmax": float(((a - b).abs() / pmax).max()),
           "argmax_identical": f"{int((res['index'] == ref['index']).sum())}/{K}",
           "sum_periods": float(res["period"].sum()),
           "sum_periods_fp64": float(ref["period"].sum()),
           "mask": "P >= 1e-3 * P_max per object",
           "worst": [{"object": int(i) // nf, "bin": int(i) % nf, "rel_err": float(v),
                      "p_over_pmax": float(b.flatten()[i] / pmax.flatten()[int(i) // nf])}
                     for v, i in zip(top.values.tolist(), top.indices.tolist())]}
    del rel, mask
    return out


def mfma_roofline(kernel_terms_s, variant):
    """Matrix-core tile: algorithmic and executed FP16 matrix flops per
    second against the dense FP16 matrix peak."""
    fpt = MFMA_FLOPS_PER_TERM[variant]
    ach = kernel_terms_s * fpt / 1e12
    return {"bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_TFLOPS_MFMA_F16, "unit": "TFLOP/s",
            "frac": round(ach / PEAK_TFLOPS_MFMA_F16, 4),
            "frac_executed": round(ach * MFMA_EXEC_FACTOR / PEAK_TFLOPS_MFMA_F16, 4)}


def timed_plan(torch, L, plan, steps, warmup, stream=None):
    for _ in range(warmup):
        plan.run(stream)
    torch.cuda.synchronize()
    L.lsamd_set_timing(1)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run(stream)
        torch.cuda.synchronize()  # one step = one sweep; sync so the kernel time is read per step
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
    split forms against the FP64 generalized form, 1e6 bins; then the split32
    accuracy with a dmag drawn per observation (non-uniform weights)."""
    from gpu_lomb_scargle_amd import _native, device, synth
    nf = 1000000
    out = {"workload": "generalized L-S batch: 1000 objects (median 140 obs), 2000 d span, "
                       "magnitudes 15..20 not subtracted, dmag logU(0.01, 0.2) per object, 1e6 bins"}
    for dname, per_obs in (("per_object_dmag", False), ("per_observation_dmag", True)):
        data = synth.generalized_batch_1000(dmag_per_obs=per_obs)
        K = int(data["counts"].size)
        terms = float(data["counts"].sum()) * nf
        t, y, dy, off = device.numpy_to_device(data["obj_id"], data["t"], data["y"], data["dy"],
                                               np.float64, dev)
        mk = lambda form: device.SweepPlan(t, y, dy, off, min_freq=synth.OMEGA_MIN,
                                           max_freq=synth.OMEGA_MAX, num_freqs=nf,
                                           variant=_native.GENERALIZED, form=form, top_k=1)
        plain = mk(_native.FORM_PLAIN)
        ms, kms = timed_plan(torch, L, plain, 1, 1)
        ref = plain.results()
        res = {"terms_per_step": terms, "fp64_plain": {"value": terms / (ms / 1e3), "ms_per_step": round(ms, 3)}}
        forms = (("split", _native.FORM_SPLIT), ("split32", _native.FORM_SPLIT32)) if not per_obs \
            else (("split32", _native.FORM_SPLIT32),)
        for name, form in forms:
            p = mk(form)
            ms, kms = timed_plan(torch, L, p, steps, 1)
            res[name] = {"value": terms / (ms / 1e3), "ms_per_step": round(ms, 3), "kernel_ms": round(kms, 3),
                         "accuracy": full_accuracy(torch, p.results(), ref, K)}
            if form == _native.FORM_SPLIT32:
                res[name]["roofline"] = mfma_roofline(terms / (kms / 1e3), "generalized")
            del p
        out[dname] = res
        del plain, ref
        torch.cuda.empty_cache()
    return out


def uneven_batch(torch, L, dev, steps):
    """100000 objects, N log-normal (median 60, sigma 1, 5..5000), 2e5 bins,
    top-3 peaks, no periodogram: persistent + graph-captured + autotuned vs
    one workgroup per item, vs a per-object launch loop; then the accuracy of
    the same path against the FP64 form on the first 2000 objects."""
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
        res["persistent_graph"] = {"value": terms / (ms / 1e3), "objects_per_s": K / (ms / 1e3),
                                   "ms_per_step": round(ms, 3), "tile": f"{T}x{M}"}
        if (T, M) == MFMA_TILE:
            res["persistent_graph"]["roofline"] = mfma_roofline(terms / (ms / 1e3), vname)
        grid = device.SweepPlan(t, y, dy, off, schedule=_native.SCHEDULE_GRID, tile=(T, M), **kw)
        ms_g, _ = timed_plan(torch, L, grid, steps, 1)
        res["grid"] = {"value": terms / (ms_g / 1e3), "ms_per_step": round(ms_g, 3)}
        same = bool(torch.equal(pers.results()["index"], grid.results()["index"]))
        # the per-object launch loop: one lsamd_sweep (4 kernels) per object
        a = _native.SweepArgs()
        ctypes.memmove(byref(a), byref(pers.args), ctypes.sizeof(a))
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
        del pers, grid, ws1
        torch.cuda.empty_cache()
        res["accuracy_first_2000"] = uneven_accuracy(torch, data, variant, (T, M), 2000, nf, dev)
        out[vname] = res
        torch.cuda.empty_cache()
    return out


def uneven_accuracy(torch, data, variant, tile, objects, nf, dev):
    """The uneven batch's persistent top-3 path on its first `objects`
    objects (the same kernels, with the periodogram kept) against the FP64
    form: powers over every bin, top-1 agreement, and top-3 agreement
    wherever the FP64 peaks are separated by more than the tolerance."""
    from gpu_lomb_scargle_amd import _native, device, synth
    rows = int(data["counts"][:objects].sum())
    dy = data["dy"][:rows] if data["dy"] is not None else None
    t, y, dd, off = device.numpy_to_device(data["obj_id"][:rows], data["t"][:rows], data["y"][:rows], dy,
                                           np.float64, dev)
    kw = dict(min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX, num_freqs=nf, variant=variant, top_k=3)
    p = device.SweepPlan(t, y, dd, off, form=_native.FORM_SPLIT32, schedule=_native.SCHEDULE_PERSISTENT,
                         tile=tile, **kw)
    r = device.SweepPlan(t, y, dd, off, form=_native.FORM_PLAIN, **kw)
    p.run()
    r.run()
    torch.cuda.synchronize()
    acc = full_accuracy(torch, p.results(), r.results(), objects)
    del acc["argmax_identical"]  # counts all three peaks; top1_identical below is per object
    gi, ri = p.results()["index"], r.results()["index"]
    rp = r.results()["power"]
    sep = ((rp[:, :2] - rp[:, 1:]) > 1e-4 * rp[:, :1]).all(dim=1)  # FP64 top-3 powers 1e-4 apart
    acc["top3_identical_where_separated"] = f"{int((gi == ri).all(dim=1)[sep].sum())}/{int(sep.sum())}"
    acc["top1_identical"] = f"{int((gi[:, 0] == ri[:, 0]).sum())}/{objects}"
    n = data["counts"][:objects]
    small = torch.as_tensor(n <= 20, device=dev)
    if bool(small.any()):
        a, b = p.results()["pgram"][small], r.results()["pgram"][small]
        pm = b.max(dim=1, keepdim=True).values
        m = b >= 1e-3 * pm
        acc["max_rel_err_masked_n_le_20"] = float(((a - b).abs() / b.abs()).where(m, torch.zeros_like(a)).max())
    del p, r
    return acc


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


def compact(out):
    """The driver-facing line: headline, accuracy, roofline, CPU baseline
    and one short entry per other workload."""
    acc = out["accuracy"]
    line = {k: out[k] for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                                "higher_is_better", "scaling", "vs_baseline", "dtype", "data")}
    c = out["config"]
    line["config"] = {"workload": c["workload"], "form": c["form"], "tile": c["tile"],
                      "terms_per_step": c["terms_per_step"]}
    line["objects_per_s"] = round(out["objects_per_s"], 1)
    line["accuracy"] = {"max_rel_err": acc["max_rel_err_masked"], "median_rel_err": acc["median_rel_err"],
                        "argmax_identical": acc["argmax_identical"]}
    line["roofline"] = {k: out["roofline"][k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic",
                                                         "algorithmic_bytes", "kernel", "kernel_ms")
                        if k in out["roofline"]}
    if "frac_executed" in out["roofline"]:
        line["roofline"]["frac_executed"] = out["roofline"]["frac_executed"]
    if "cpu_baseline" in out:
        cb = out["cpu_baseline"]
        line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample",
                                                   "value_16_threads")}
    line["fp64_plain_terms_per_s"] = out["fp64_plain"]["value"]
    line["speedup_vs_fp64_plain"] = out["fp64_plain"]["speedup_of_reported_form"]
    cfg = {}
    g = out.get("generalized_batch")
    if g:
        for d in ("per_object_dmag", "per_observation_dmag"):
            s = g[d]["split32"]
            cfg["generalized_split32_" + d] = {"terms_per_s": s["value"], "ms": s["ms_per_step"],
                                               "max_rel_err": s["accuracy"]["max_rel_err_masked"],
                                               "argmax": s["accuracy"]["argmax_identical"]}
    u = out.get("uneven_batch")
    if u:
        for v in ("standard", "generalized"):
            s = u[v]
            a = s["accuracy_first_2000"]
            cfg["uneven_" + v] = {"terms_per_s": s["persistent_graph"]["value"],
                                  "objects_per_s": round(s["persistent_graph"]["objects_per_s"]),
                                  "ms": s["persistent_graph"]["ms_per_step"],
                                  "speedup_vs_loop": s["persistent_speedup_vs_loop"],
                                  "max_rel_err": a["max_rel_err_masked"],
                                  "top3": a["top3_identical_where_separated"]}
    r = out.get("refinement_batch")
    if r:
        for v in ("standard", "generalized"):
            cfg["refine_" + v] = {"grid_ms": r[v]["grid_search_ms"], "refine_ms": r[v]["refine_ms"],
                                  "p_err_median": [r[v]["period_rel_err_grid"]["median"],
                                                   r[v]["period_rel_err_refined"]["median"]]}
    s = out.get("single_object")
    if s:
        cfg["single_3555x1e5"] = {"split32_ms": s["split32"]["ms_per_step"], "fp64_ms": s["fp64_plain"]["ms_per_step"],
                                  "cpu_mode5_s": s["cpu_mode5"]["seconds"]}
    e = out.get("end_to_end_pcie")
    if e:
        cfg["dropin_pcie_terms_per_s"] = e["value"]
    line["configs"] = cfg
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target duration of the all-cores CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the host's physical cores")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nf", type=int, default=1000000)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workloads (generalized, uneven, refinement, single, FP32, PCIe)")
    ap.add_argument("--report", default=None, help="also write the full nested report to this file")
    ap.add_argument("--profile-once", action="store_true",
                    help="for rocprofv3 counter runs: one 1 GiB device copy (calibration of the "
                         "HBM byte counters), then one sweep of each form; no JSON")
    ap.add_argument("--profile-repeat", type=int, default=1,
                    help="with --profile-once: sweeps of the split32 plans (standard and generalized "
                         "batches), for steady-state kernel traces")
    ap.add_argument("--profile-uneven", type=int, default=0, metavar="K",
                    help="for rocprofv3 runs: K eager persistent split32 sweeps (matrix-core tile, top-3, "
                         "no periodogram) of the 100000-object uneven batch per variant; no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (a batch would shard "
              "by objects, one slice per GPU, with no exchange)", file=sys.stderr)
        sys.exit(2)

    import torch

    from gpu_lomb_scargle_amd import _native, device, synth

    dev = torch.device("cuda:0")
    L = _native.core()
    if args.profile_uneven:
        for variant in (_native.STANDARD, _native.GENERALIZED):
            u = synth.uneven_batch_100000(with_error=variant == _native.GENERALIZED)
            ut, uy, udy, uoff = device.numpy_to_device(u["obj_id"], u["t"], u["y"], u["dy"], np.float64, dev)
            plan = device.SweepPlan(ut, uy, udy, uoff, min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX,
                                    num_freqs=200000, variant=variant, form=_native.FORM_SPLIT32, top_k=3,
                                    pgram=False, schedule=_native.SCHEDULE_PERSISTENT, tile=MFMA_TILE)
            for _ in range(args.profile_uneven):
                plan.run()
            torch.cuda.synchronize()
            log(f"profile-uneven: {args.profile_uneven} persistent split32 sweeps, variant {variant}, "
                f"{int(u['counts'].sum())} rows x 200000 freqs")
            del plan, ut, uy, udy, uoff
            torch.cuda.empty_cache()
        return
    cpu = host_cpu()
    cpu_threads = args.cpu_threads or cpu["threads"]
    data = synth.standard_batch_1000()
    K = int(data["counts"].size)
    rows = int(data["counts"].sum())
    nf = args.nf
    terms = float(rows) * nf
    log(f"workload: {K} objects, {rows} rows, {nf} freqs, {terms:.3e} terms/step; host {cpu}")
    t, y, dy, off = device.numpy_to_device(data["obj_id"], data["t"], data["y"], None, np.float64, dev)

    def make(form, tile=None):
        return device.SweepPlan(t, y, None, off, min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX,
                                num_freqs=nf, variant=_native.STANDARD, form=form, normalize=True,
                                top_k=1, tile=tile)

    if args.profile_once:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)  # 1 GiB
        dst = src.clone()
        torch.cuda.synchronize()
        for form, tile in ((_native.FORM_SPLIT32, None), (_native.FORM_SPLIT32, VALU_TILE),
                           (_native.FORM_SPLIT, None), (_native.FORM_PLAIN, None)):
            plan = make(form, tile)
            for _ in range(args.profile_repeat if form == _native.FORM_SPLIT32 else 1):
                plan.run()
            torch.cuda.synchronize()
            del plan
        gdata = synth.generalized_batch_1000()
        gt, gy, gdy, goff = device.numpy_to_device(gdata["obj_id"], gdata["t"], gdata["y"], gdata["dy"],
                                                   np.float64, dev)
        plan = device.SweepPlan(gt, gy, gdy, goff, min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX,
                                num_freqs=nf, variant=_native.GENERALIZED, form=_native.FORM_SPLIT32, top_k=1)
        for _ in range(args.profile_repeat):
            plan.run()
        torch.cuda.synchronize()
        log("profile-once: 1 GiB copy, one standard split32 sweep with each tile (matrix-core, VALU 256x8), "
            "one split and one FP64 plain standard sweep, one generalized split32 sweep")
        del src, dst, plan
        return

    runs = {}
    # split32 with its default (matrix-core) tile and with the VALU tile
    for name, form, tile in (("split32", _native.FORM_SPLIT32, None),
                             ("split32_valu", _native.FORM_SPLIT32, VALU_TILE),
                             ("split", _native.FORM_SPLIT, None)):
        plan = make(form, tile)
        ms, kms = timed_plan(torch, L, plan, args.steps, args.warmup)
        runs[name] = {"plan": plan, "ms": ms, "kms": kms, "tile": plan.tile()}
        log(f"{name}: {ms:.3f} ms/step, sweep kernel {kms:.3f} ms, tile {runs[name]['tile']}")

    plain = make(_native.FORM_PLAIN)
    ms_plain, kms_plain = timed_plan(torch, L, plain, args.fp64_steps, 1)
    log(f"fp64 plain: {ms_plain:.3f} ms/step, sweep kernel {kms_plain:.3f} ms")

    # full-size accuracy of the split forms against the FP64 form (all bins)
    pp = plain.results()
    for name, r in runs.items():
        r["acc"] = full_accuracy(torch, r["plan"].results(), pp, K)
        log(f"accuracy {name}: {r['acc']}")

    def meets(acc):
        return acc["max_rel_err_masked"] <= MAX_REL_ERR and acc["argmax_identical"] == f"{K}/{K}"

    chosen = "split32" if meets(runs["split32"]["acc"]) else "split"
    r = runs[chosen]
    T, M = r["tile"]
    value = terms / (r["ms"] / 1e3)
    kernel_terms_s = terms / (r["kms"] / 1e3)
    mfma = chosen == "split32" and (T, M) == MFMA_TILE
    if mfma:
        kname = "k_sweep_mfma<double>"
        bound = mfma_roofline(kernel_terms_s, "standard")
        bound["work_per_term"] = ("16 flops (two complex multiply-accumulates) on the FP16 matrix cores; "
                                  "48 executed (FP16 head/tail split, 3 products per real product)")
    else:
        kname = "k_sweep_%s<false, double, %d, %d>" % ("split32" if chosen == "split32" else "split64", T, M)
        achieved = kernel_terms_s * 2.0 * CYCLES_PER_TERM / 1e12
        bound = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_TFLOPS,
                 "unit": "TFLOP/s", "frac": round(achieved / PEAK_TFLOPS, 4),
                 "work_per_term": f"{CYCLES_PER_TERM:g} VALU lane-cycles = {2 * CYCLES_PER_TERM:g} flop-eq "
                                  f"(mul + sin + cos at 1/{TRANS_CYCLES:g} rate + 5 FMA), canonical"}
    tr = measured_traffic(kname)
    roofline = dict(bound)
    roofline.update({
        "traffic": tr["bytes"] if tr else None,
        "kernel": kname, "kernel_ms": round(r["kms"], 4),
        "algorithmic_bytes": K * nf * 8 + rows * 16,
        "pgram_write_GBps": round(K * nf * 8 / (r["kms"] / 1e3) / 1e9, 1),
        "traffic_source": ("rocprofv3 FETCH_SIZE x%.1f + WRITE_SIZE x%.1f (copy-calibrated), %s"
                           % (tr["fetch_factor"], tr["write_factor"], os.path.relpath(TRAFFIC_FILE, ROOT))
                           if tr else None)})

    def form_report(name):
        q = runs[name]
        return {"value": terms / (q["ms"] / 1e3), "ms_per_step": round(q["ms"], 3),
                "kernel_ms": round(q["kms"], 3), "tile": "%dx%d" % q["tile"], "accuracy": q["acc"],
                "meets_1e-5_and_argmax": meets(q["acc"])}

    forms = {"split32": "FP64 phase reduction, FP32 hardware sin/cos; matrix-core tile: per 32x32 bins the "
                        "sums as complex products of a row and a column factor on FP16 matrix cores (head/tail "
                        "split, FP32 accumulation), FP64 power",
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
        "host": cpu,
    }
    # best periods of the objects the CPU baseline also runs (MODE 4 check)
    sample_periods = {"fp64_plain": pp["period"][:, 0].cpu().numpy()}
    for name, q in runs.items():
        sample_periods[name] = q["plan"].results()["period"][:, 0].cpu().numpy()
    if not args.no_extra:
        runs.clear()
        del plain, pp
        torch.cuda.empty_cache()
        for key, fn in (("fp32_batch", lambda: fp32_batch(torch, L, dev, data, nf, max(2, args.steps // 2))),
                        ("generalized_batch", lambda: generalized_batch(torch, L, dev, max(2, args.steps // 2))),
                        ("uneven_batch", lambda: uneven_batch(torch, L, dev, max(2, args.steps // 2))),
                        ("refinement_batch", lambda: refinement_batch(torch, L, dev)),
                        ("single_object", lambda: single_object(torch, L, dev, cpu_threads)),
                        ("end_to_end_pcie", lambda: end_to_end(torch, nf))):
            out[key] = fn()
            log(f"{key}: {json.dumps(out[key])}")
    if not args.no_cpu:
        full, ref16, cpu_periods = cpu_baseline(data, nf, cpu_threads, args.cpu_seconds)
        out["cpu_baseline"] = {
            "value": full["value"], "unit": "terms/s", "cores": cpu_threads, "kind": full["kind"],
            "objects_per_s": full["objects_per_s"],
            "sample": (f"first {full['objects']} of the 1000 objects ({full['rows']} rows) x {nf} freqs, FP64, "
                       f"reference MODE 4 (OpenMP dynamic over objects), {full['seconds']} s on {cpu_threads} "
                       f"threads (physical cores {cpu['physical_cores']}, usable CPUs "
                       f"{min(cpu['affinity_cpus'], cpu['cgroup_cpu_quota'] or cpu['affinity_cpus'])})"),
            "value_16_threads": ref16["value"],
            "sample_16_threads": f"first 32 objects, {ref16['seconds']} s (NTHREADSCPU=16, source/Makefile:49)",
            "best_periods_vs_gpu": periods_vs_cpu(sample_periods, cpu_periods)}
        log(f"cpu baseline: {out['cpu_baseline']}")
    full_report = json.dumps(out)
    log("full report: " + full_report)
    if args.report:
        with open(args.report, "w") as f:
            f.write(full_report + "\n")
    line = json.dumps(compact(out))
    # the reference's CPU code prints progress text without a final newline
    # into the C library's stdout buffer: flush it, then start a fresh line
    import ctypes as ct
    ct.CDLL(None).fflush(None)
    sys.stdout.write("\n")
    print(line, flush=True)


if __name__ == "__main__":
    main()
