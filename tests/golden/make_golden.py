"""Write the golden fixtures of tests/golden/ from the reference's own CPU code.

Runs the reference's CPU path (source/main.cu modes 4/5, compiled from the
reference's sources into oracle/_ref/ by oracle/build_ref.sh with
NTHREADSCPU=1, i.e. every reduction in index order) on:
  * the reference's data files (copied into tests/golden/data/),
  * small seeded synthetic batches (inputs stored in the fixture),
and stores, per case and build: the periodogram itself when it is small,
otherwise per-object argmax / period / power, samples around each peak and
at fixed bins, and a SHA-256 over the full periodogram's bytes.

Usage (in the build container, where the reference tree exists):
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gpu_lomb_scargle_amd import synth  # noqa: E402

DATA = os.path.join(HERE, "data")
TWO_PI = 2.0 * np.pi


def load(name):
    d = np.loadtxt(os.path.join(DATA, name), delimiter=",")
    return d[:, 0].astype(np.uint32), d[:, 1], d[:, 2], (d[:, 3] if d.shape[1] > 3 else None)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def peaks_and_samples(pg, fixed):
    K, nf = pg.shape
    am = np.zeros(K, np.int64)
    for k in range(K):  # computePeriod: first strict maximum from 0
        best, idx = 0.0, 0
        row = pg[k]
        m = row.max()
        if m > 0:
            idx = int(np.flatnonzero(row == m)[0])
        am[k] = idx
    win = np.clip(am[:, None] + np.arange(-8, 9)[None, :], 0, nf - 1)
    return am, np.take_along_axis(pg, win, 1), win, pg[:, fixed]


def main():
    if not oracle.ref_available(False, np.float64):
        raise SystemExit("oracle/_ref is not built (run oracle/build_ref.sh)")
    out = {}

    # 8205 single object, f = 0.1..10 cycles/day, 1e4 bins: full pgrams
    o, t, y, dy = load("8205_normalized_with_error.txt")
    fmin, fmax, nf = TWO_PI * 0.1, TWO_PI * 10.0, 10000
    for dt, name in ((np.float64, "double"), (np.float32, "float")):
        pg, per, pw = oracle.ref_single(t, y, None, fmin, fmax, nf, False, dt, normalize=True)
        out[f"s8205_std_{name}"] = pg
        out[f"s8205_std_{name}_period"] = np.array([per], dt)
        pg, per, pw = oracle.ref_single(t, y, dy, fmin, fmax, nf, True, dt)
        out[f"s8205_gen_{name}"] = pg
        out[f"s8205_gen_{name}_period"] = np.array([per], dt)
    pg, _, _ = oracle.ref_single(t, y, None, fmin, fmax, nf, False, np.float64, normalize=False)
    out["s8205_std_double_raw"] = pg

    # 243 ZTF (t ~ 2.4587e6 JD), omega 3.142..150.796, 1e4 bins
    o, t, y, _ = load("243_normalized_ztf_filter2.txt")
    for dt, name in ((np.float64, "double"), (np.float32, "float")):
        pg, per, pw = oracle.ref_single(t, y, None, 3.142, 150.796, 10000, False, dt, normalize=True)
        out[f"z243_std_{name}"] = pg
    pg, _, _ = oracle.ref_single(t, y, None, 3.142, 150.796, 10000, False, np.float64, normalize=False)
    out["z243_std_double_raw"] = pg

    # SDSS stripe 82 batch (136 objects), f = 0.1..10 cycles/day, 2e4 bins
    o, t, y, dy = load("SDSS_stripe82_band_z.txt")
    nf = 20000
    fixed = np.linspace(0, nf - 1, 64).astype(np.int64)
    for err in (False, True):
        for dt, name in ((np.float64, "double"), (np.float32, "float")):
            pg, per, pw = oracle.ref_batch(o, t, y, dy if err else None, fmin, fmax, nf, err, dt)
            am, win_v, win_i, fix_v = peaks_and_samples(pg, fixed)
            key = f"sdss_{'gen' if err else 'std'}_{name}"
            out[key + "_sha256"] = np.array(sha(pg))
            out[key + "_argmax"] = am
            out[key + "_period"] = per
            out[key + "_power"] = pw
            out[key + "_win_values"] = win_v
            out[key + "_win_index"] = win_i
            out[key + "_fixed_values"] = fix_v
    out["sdss_fixed_index"] = fixed

    # seeded synthetic batches: 32 objects each, 2e4 bins over the
    # benchmark's angular range
    nf = 20000
    fixed = np.linspace(0, nf - 1, 64).astype(np.int64)
    for err, seed in ((False, 7), (True, 8)):
        b = synth.lognormal_batch(32, 140, 0.6, 20, 1000, 2000 if err else 1800, err, seed)
        tag = "gen" if err else "std"
        out[f"syn_{tag}_obj"] = b["obj_id"]
        out[f"syn_{tag}_t"] = b["t"]
        out[f"syn_{tag}_y"] = b["y"]
        if err:
            out[f"syn_{tag}_dy"] = b["dy"]
        for dt, name in ((np.float64, "double"), (np.float32, "float")):
            pg, per, pw = oracle.ref_batch(b["obj_id"], b["t"], b["y"], b["dy"], synth.OMEGA_MIN,
                                           synth.OMEGA_MAX, nf, err, dt)
            am, win_v, win_i, fix_v = peaks_and_samples(pg, fixed)
            key = f"syn_{tag}_{name}"
            out[key + "_sha256"] = np.array(sha(pg))
            out[key + "_argmax"] = am
            out[key + "_period"] = per
            out[key + "_power"] = pw
            out[key + "_win_values"] = win_v
            out[key + "_win_index"] = win_i
            out[key + "_fixed_values"] = fix_v
    out["syn_fixed_index"] = fixed
    out["syn_nf"] = np.array(nf)

    path = os.path.join(HERE, "reference_cpu_fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
