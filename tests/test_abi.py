"""The C ABI on the CPU: every library loads and exports what
include/lombscargle_amd.h declares; the drop-in LombScarglePy reproduces the
reference's CPU modes 4/5 bit for bit; argument checking and error
reporting work without a GPU.  No GPU compute is attempted here."""
import os
import re
import subprocess
import sys
from ctypes import byref, c_void_p

import numpy as np
import pytest

import oracle
from conftest import ROOT, load_data
from gpu_lomb_scargle_amd import _native
from gpu_lomb_scargle_amd import lombscarglegpu as ls

TWO_PI = 2.0 * np.pi
HEADER = os.path.join(ROOT, "include", "lombscargle_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(?:void|int|size_t|double|uint32_t|const char \*)\s*\**\s*(\w+)\s*\(", src)
    return sorted(set(names))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_header_declares_expected_api():
    names = declared_functions()
    assert "LombScarglePy" in names
    assert set(_native.EXPORTED) <= set(names)


@pytest.mark.parametrize("lib", [_native.CORE_LIB] + sorted(_native.DROPIN_LIBS.values()))
def test_library_loads_and_exports_every_declared_symbol(lib):
    path = os.path.join(_native.LIB_DIR, lib)
    assert os.path.exists(path), f"{path} not built"
    _native.load(lib)
    syms = exported(path)
    want = set(declared_functions())
    if lib == _native.CORE_LIB:
        want.discard("LombScarglePy")
    missing = want - syms
    assert not missing, f"{lib} lacks {sorted(missing)}"


def test_no_device_is_reported_not_crashed():
    L = _native.core()
    if L.lsamd_device_count() > 0:
        pytest.skip("a GPU is visible")
    t = np.linspace(0, 10, 20)
    with pytest.raises(_native.LsamdError) as ei:
        _native.batch_host(None, t, np.sin(t), None, 1.0, 2.0, 10, _native.STANDARD)
    assert ei.value.code == _native.ERR_NODEVICE


def test_sweep_argument_checks():
    L = _native.core()
    a = _native.SweepArgs()
    a.variant, a.dtype, a.form = 0, 0, 0
    a.num_objects, a.num_freqs, a.num_obs, a.top_k = 0, 10, 10, 1
    a.min_freq, a.max_freq = 1.0, 2.0
    assert L.lsamd_sweep(byref(a), None, 0, None) == _native.ERR_INVALID
    assert b"empty" in L.lsamd_last_error()
    a.num_objects = 1
    a.offsets = a.t = a.y = 1  # never dereferenced: rejected before any launch
    a.top_k = 2
    assert L.lsamd_sweep(byref(a), None, 0, None) == _native.ERR_INVALID
    a.top_k = 1
    a.max_freq = 0.5
    assert L.lsamd_sweep(byref(a), None, 0, None) == _native.ERR_INVALID
    a.max_freq = 2.0
    need = L.lsamd_sweep_workspace_size(byref(a))
    assert need > 0
    assert L.lsamd_sweep(byref(a), c_void_p(4096), need - 1, None) == _native.ERR_INVALID
    assert b"workspace" in L.lsamd_last_error()
    a.variant = 1  # generalized without dy
    assert L.lsamd_sweep(byref(a), c_void_p(4096), need, None) == _native.ERR_INVALID


def test_observation_slices_workspace():
    """A split-form sweep of one long object reserves per-slice FP64 sums;
    with obs_slices = 1, or for a batch, it does not."""
    L = _native.core()
    a = _native.SweepArgs()
    a.variant, a.dtype, a.form = _native.STANDARD, _native.F64, _native.FORM_SPLIT32
    a.num_objects, a.num_freqs, a.num_obs, a.top_k = 1, 100000, 3555, 1
    a.min_freq, a.max_freq = 1.005, 150.796
    sliced = L.lsamd_sweep_workspace_size(byref(a))
    a.obs_slices = 1
    whole = L.lsamd_sweep_workspace_size(byref(a))
    # chunks of T x M bins, slices of <= T observations (one LDS stage), 4 sums
    from ctypes import c_int
    T, M = c_int(), c_int()
    L.lsamd_get_tile(a.form, a.variant, a.dtype, byref(T), byref(M))
    chunks = -(-100000 // (T.value * M.value))
    slices = min(-(-2048 // chunks), -(-3555 // T.value))
    assert slices >= 2
    assert sliced - whole >= chunks * slices * 4 * M.value * T.value * 8
    a.obs_slices, a.num_objects = 0, 2
    assert L.lsamd_sweep_workspace_size(byref(a)) < sliced
    a.form, a.num_objects = _native.FORM_PLAIN, 1
    assert L.lsamd_sweep_workspace_size(byref(a)) < whole
    a.obs_slices = 2
    a.offsets = a.t = a.y = 1  # never dereferenced: rejected before any launch
    assert L.lsamd_sweep(byref(a), c_void_p(4096), 1 << 30, None) == _native.ERR_INVALID
    assert b"obs_slices" in L.lsamd_last_error()


def test_per_call_tile_overrides_the_default():
    """A tile in the args plans the workspace exactly like the same process
    default set by lsamd_set_tile, leaves that default alone, and an
    uncompiled per-call tile is rejected before anything is launched."""
    from ctypes import c_int
    L = _native.core()
    a = _native.SweepArgs()
    a.variant, a.dtype, a.form = _native.STANDARD, _native.F64, _native.FORM_SPLIT32
    a.num_objects, a.num_freqs, a.num_obs, a.top_k, a.obs_slices = 1000, 1000000, 163652, 1, 1
    a.min_freq, a.max_freq = 1.005, 150.796
    try:
        assert L.lsamd_set_tile(a.form, a.variant, 256, 8) == _native.OK
        by_default = L.lsamd_sweep_workspace_size(byref(a))
    finally:
        L.lsamd_set_tile(a.form, a.variant, 0, 0)
    builtin = L.lsamd_sweep_workspace_size(byref(a))
    a.tile_threads, a.tile_fpt = 256, 8
    assert L.lsamd_sweep_workspace_size(byref(a)) == by_default != builtin
    T, M = c_int(), c_int()
    L.lsamd_get_tile(a.form, a.variant, a.dtype, byref(T), byref(M))
    assert (T.value, M.value) == (64, 16)
    a.tile_threads, a.tile_fpt = 96, 3
    a.offsets = a.t = a.y = 1  # never dereferenced: rejected before any launch
    a.peak_index = 1
    assert L.lsamd_sweep(byref(a), c_void_p(4096), 1 << 40, None) == _native.ERR_INVALID
    assert b"not compiled" in L.lsamd_last_error()


def test_matrix_core_tile_workspace():
    """The split32 standard matrix-core tile (64 x 16) reserves its column-
    factor fragments: 8 fragments x 64 lanes x 16 B per 16-observation block,
    in slots floor(beg/16) + k .. + ceil(N/16) per object, which never overlap
    and end below n/16 + K + 1."""
    from ctypes import c_int
    L = _native.core()
    a = _native.SweepArgs()
    a.variant, a.dtype, a.form = _native.STANDARD, _native.F64, _native.FORM_SPLIT32
    a.num_objects, a.num_freqs, a.num_obs, a.top_k, a.obs_slices = 1000, 1000000, 163652, 1, 1
    a.min_freq, a.max_freq = 1.005, 150.796
    T, M = c_int(), c_int()
    L.lsamd_get_tile(a.form, a.variant, a.dtype, byref(T), byref(M))
    assert (T.value, M.value) == (64, 16)   # the built-in split32 standard tile
    try:
        mfma = L.lsamd_sweep_workspace_size(byref(a))
        assert L.lsamd_set_tile(a.form, a.variant, 256, 8) == _native.OK
        valu = L.lsamd_sweep_workspace_size(byref(a))
    finally:
        L.lsamd_set_tile(a.form, a.variant, 0, 0)
    colfac = 8 * 64 * 16 * (163652 // 16 + 1000 + 1)
    assert mfma - valu >= colfac   # (its 1024-bin chunks also need more peak partials)
    # the generalized tile keeps 12 fragments per block (the weights ride on the columns)
    a.variant = _native.GENERALIZED
    try:
        gen = L.lsamd_sweep_workspace_size(byref(a))
        assert L.lsamd_set_tile(a.form, a.variant, 256, 4) == _native.OK
        gen_valu = L.lsamd_sweep_workspace_size(byref(a))
    finally:
        L.lsamd_set_tile(a.form, a.variant, 0, 0)
    assert gen - gen_valu >= colfac * 12 // 8
    assert gen - mfma >= colfac // 2
    # the slot rule on random object sizes (including empty and 1-observation objects)
    rng = np.random.default_rng(5)
    for _ in range(20):
        sizes = rng.integers(0, 90, rng.integers(1, 60))
        beg = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        first = beg // 16 + np.arange(sizes.size)
        last = first + (sizes + 15) // 16          # one past the object's last slot
        assert np.all(last[:-1] <= first[1:])
        assert last.max() <= sizes.sum() // 16 + sizes.size + 1


def test_point_argument_checks():
    """lsamd_power_grad / lsamd_refine_peaks validate before any launch."""
    L = _native.core()
    assert L.lsamd_power_grad(None, None) == _native.ERR_INVALID
    a = _native.PointArgs()
    assert L.lsamd_power_grad(byref(a), None) == _native.ERR_INVALID
    assert b"empty" in L.lsamd_last_error()
    a.num_objects, a.num_obs = 1, 10
    a.offsets = a.t = a.y = a.omega = 1  # never dereferenced: rejected before any launch
    assert L.lsamd_power_grad(byref(a), None) == _native.ERR_INVALID
    assert b"nothing to compute" in L.lsamd_last_error()
    a.variant = 1  # generalized without dy
    a.power = 1
    assert L.lsamd_power_grad(byref(a), None) == _native.ERR_INVALID
    assert b"dy" in L.lsamd_last_error()
    a.variant = 0
    a.max_step, a.tol, a.max_iter = 0.0, 1e-9, 10
    assert L.lsamd_refine_peaks(byref(a), None) == _native.ERR_INVALID
    assert b"max_step" in L.lsamd_last_error()
    a.dtype = 7
    assert L.lsamd_power_grad(byref(a), None) == _native.ERR_INVALID


def test_tile_registry():
    L = _native.core()
    assert L.lsamd_set_tile(1, 0, 256, 4) == _native.OK
    from ctypes import c_int
    T, M = c_int(), c_int()
    L.lsamd_get_tile(1, 0, 0, byref(T), byref(M))
    assert (T.value, M.value) == (256, 4)
    assert L.lsamd_set_tile(1, 0, 0, 0) == _native.OK
    assert L.lsamd_set_tile(1, 0, 96, 3) == _native.ERR_INVALID


def test_segmentation_helpers_match_oracle():
    L = _native.core()
    ids = np.array([3, 3, 1, 1, 1, 3, 8, 8, 2], np.uint32)
    K = L.lsamd_count_objects(_native.ptr(ids), ids.size)
    off = np.zeros(K + 1, np.uint64)
    got_ids = np.zeros(K, np.uint32)
    assert L.lsamd_object_offsets(_native.ptr(ids), ids.size, _native.ptr(off), _native.ptr(got_ids)) == K
    lo, hi, oids = oracle.object_ranges(ids)
    assert off[:-1].tolist() == lo.tolist() and (off[1:] - 1).tolist() == hi.tolist()
    assert got_ids.tolist() == oids.tolist()


# --------------------------------------------------------------------------
# drop-in LombScarglePy, CPU modes, against the oracle (bit for bit)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", ["float", "double"])
@pytest.mark.parametrize("error", [False, True])
def test_dropin_cpu_batch_mode4_bitwise(dtype, error):
    o, t, y, dy = load_data("SDSS_stripe82_band_z.txt")
    sl = slice(0, 1500)
    # a float64 magY is pre-centred in place by the error libraries: pass a copy
    ids, per, pg = ls.lombscargle(o[sl], t[sl], y[sl].copy(), 0.1, 10.0, error, "CPU",
                                  dy[sl] if error else None, 2500, dtype, False)
    npt = np.float32 if dtype == "float" else np.float64
    ref = oracle.batch(o[sl], t[sl], y[sl], dy[sl] if error else None, TWO_PI * 0.1, TWO_PI * 10.0,
                       2500, error, npt)
    assert pg.dtype == npt and pg.tobytes() == ref[0].tobytes()
    assert ids.tolist() == oracle.object_ranges(o[sl])[2].tolist()
    exp = 1.0 / (0.1 + (10.0 - 0.1) / 2500 * np.argmax(ref[0], axis=1))
    assert np.array_equal(per, exp)


@pytest.mark.parametrize("dtype", ["float", "double"])
@pytest.mark.parametrize("error", [False, True])
def test_dropin_cpu_single_mode5_bitwise(dtype, error):
    o, t, y, dy = load_data("8205_normalized_with_error.txt")
    ids, per, pg = ls.lombscargle(o, t, y.copy(), 0.1, 10.0, error, "CPU", dy if error else None, 1500,
                                  dtype, False)
    npt = np.float32 if dtype == "float" else np.float64
    ref = oracle.single(t, y, dy if error else None, TWO_PI * 0.1, TWO_PI * 10.0, 1500, error, npt)
    assert pg.shape == (1, 1500) and pg[0].tobytes() == ref[0].tobytes()


def test_dropin_precentres_callers_float64_array_in_place():
    """Reference side effect (main.cu:1631-1662 on the caller's buffer)."""
    o, t, y, dy = load_data("8205_normalized_with_error.txt")
    y64 = np.ascontiguousarray(y[:200], np.float64)
    before = y64.copy()
    ls.lombscargle(o[:200], t[:200], y64, 0.1, 10.0, True, "CPU", dy[:200], 50, "double", False)
    assert np.array_equal(y64, oracle.precenter(before, dy[:200]))
    y32 = np.ascontiguousarray(y[:200], np.float32)  # float input is converted, caller untouched
    before32 = y32.copy()
    ls.lombscargle(o[:200], t[:200], y32, 0.1, 10.0, True, "CPU", dy[:200], 50, "float", False)
    assert np.array_equal(y32, before32)


def test_dropin_unknown_mode_does_nothing():
    lib = _native.load(_native.DROPIN_LIBS[(False, "double")])
    from ctypes import c_double, c_int, c_uint
    t = np.linspace(0, 10, 16)
    y = np.sin(t)
    ids = np.zeros(16, np.uint32)
    pg = np.full(8, -7.0)
    lib.LombScarglePy.argtypes = None
    lib.LombScarglePy(_native.ptr(ids), _native.ptr(t), _native.ptr(y), _native.ptr(np.ones(16)),
                      c_uint(16), c_double(1.0), c_double(2.0), c_uint(8), c_int(3), _native.ptr(pg))
    assert np.all(pg == -7.0)


def test_python_error_without_magdy_exits():
    t = np.linspace(0, 10, 16)
    with pytest.raises(SystemExit):
        ls.lombscargle(np.zeros(16), t, np.sin(t), 0.1, 1.0, True, "CPU", None, 8, "double", False)


def test_python_helpers():
    start, end, uniq = ls.computeIndexRangesForEachObject(np.array(["a", "a", "b", "a", "a"]))
    assert start.tolist() == [0, 2, 3] and end.tolist() == [1, 2, 4]
    assert uniq.tolist() == ["a", "b", "a"]
    assert ls.enumerateObjects(start, end).tolist() == [0, 0, 1, 2, 2]
    a = np.arange(6, dtype=np.float32)[::2]
    b = ls.convert_type(a, np.float32)
    assert b.flags["C_CONTIGUOUS"] and b.dtype == np.float32


def test_gpu_mode_without_gpu_raises_cleanly():
    if _native.core().lsamd_device_count() > 0:
        pytest.skip("a GPU is visible")
    t = np.linspace(0, 10, 16)
    with pytest.raises(RuntimeError):
        ls.lombscargle(np.zeros(16), t, np.sin(t), 0.1, 1.0, False, "GPU", None, 8, "double", False)
