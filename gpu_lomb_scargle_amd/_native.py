"""ctypes bindings of the C ABI in include/lombscargle_amd.h.

The libraries are built in-tree into ``gpu_lomb_scargle_amd/lib/`` by
``__graft_entry__.build()`` (``make -C gpu_lomb_scargle_amd/csrc``).  There is
no fallback: if a library is missing, loading raises ``NativeLibraryError``.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_int, c_int32, c_size_t, c_uint32,
                    c_uint64, c_void_p)

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.environ.get("LSAMD_LIB_DIR", os.path.join(PKG_DIR, "lib"))
CORE_LIB = "libgpulsamd.so"
DROPIN_LIBS = {
    (False, "float"): "libpylsnoerrorfloat.so",
    (False, "double"): "libpylsnoerrordouble.so",
    (True, "float"): "libpylserrorfloat.so",
    (True, "double"): "libpylserrordouble.so",
}

STANDARD, GENERALIZED = 0, 1
F32, F64 = 0, 1
FORM_PLAIN, FORM_SPLIT, FORM_SPLIT32 = 0, 1, 2
OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_NODEVICE = 0, 1, 2, 3, 4
SCHEDULE_GRID, SCHEDULE_PERSISTENT = 0, 1

# Every function include/lombscargle_amd.h declares (LombScarglePy only in the
# drop-in libraries).
EXPORTED = [
    "lsamd_sweep_workspace_size", "lsamd_sweep", "lsamd_batch_host", "lsamd_cpu_batch",
    "lsamd_cpu_single", "lsamd_count_objects", "lsamd_object_offsets", "lsamd_last_error_code",
    "lsamd_last_error", "lsamd_clear_error", "lsamd_device_count", "lsamd_set_timing",
    "lsamd_last_sweep_kernel_ms", "lsamd_set_tile", "lsamd_get_tile", "lsamd_version",
    "lsamd_graph_create", "lsamd_graph_launch", "lsamd_graph_destroy", "lsamd_autotune",
    "lsamd_autotune_workspace_size", "lsamd_power_grad", "lsamd_refine_peaks", "lsamd_read_table",
    "lsamd_write_periods", "lsamd_segment_workspace_size", "lsamd_segment", "lsamd_sweep_static",
]


class NativeLibraryError(RuntimeError):
    pass


class LsamdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lsamd error {code}: {msg}")
        self.code = code


class SweepArgs(Structure):
    """Mirror of ``lsamd_sweep_args``."""
    _fields_ = [
        ("variant", c_int32), ("dtype", c_int32), ("form", c_int32), ("normalize", c_int32),
        ("precenter", c_int32), ("num_objects", c_uint32), ("num_freqs", c_uint32),
        ("top_k", c_uint32), ("num_obs", c_uint64), ("min_freq", c_double), ("max_freq", c_double),
        ("offsets", c_void_p), ("order", c_void_p), ("t", c_void_p), ("y", c_void_p),
        ("dy", c_void_p), ("pgram", c_void_p), ("peak_index", c_void_p), ("peak_power", c_void_p),
        ("peak_period", c_void_p), ("schedule", c_int32), ("obs_slices", c_int32),
        ("tile_threads", c_int32), ("tile_fpt", c_int32),
    ]


class PointArgs(Structure):
    """Mirror of ``lsamd_point_args``."""
    _fields_ = [
        ("variant", c_int32), ("dtype", c_int32), ("normalize", c_int32), ("num_objects", c_uint32),
        ("num_obs", c_uint64), ("offsets", c_void_p), ("t", c_void_p), ("y", c_void_p),
        ("dy", c_void_p), ("omega", c_void_p), ("power", c_void_p), ("dpower", c_void_p),
        ("d2power", c_void_p), ("grad_scale", c_void_p), ("grad_y", c_void_p), ("grad_dy", c_void_p),
        ("max_step", c_double), ("tol", c_double), ("max_iter", c_int32), ("iterations", c_void_p),
    ]


def _declare(L):
    L.lsamd_sweep_workspace_size.argtypes = [POINTER(SweepArgs)]
    L.lsamd_sweep_workspace_size.restype = c_size_t
    L.lsamd_sweep.argtypes = [POINTER(SweepArgs), c_void_p, c_size_t, c_void_p]
    L.lsamd_sweep.restype = c_int
    L.lsamd_sweep_static.argtypes = [POINTER(SweepArgs), c_void_p]
    L.lsamd_sweep_static.restype = c_int
    L.lsamd_batch_host.argtypes = [c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_uint64, c_double, c_double, c_uint32, c_void_p,
                                   c_uint32, c_void_p, c_void_p, c_void_p, POINTER(c_uint32)]
    L.lsamd_batch_host.restype = c_int
    L.lsamd_cpu_batch.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_uint64, c_double, c_double, c_uint32, c_void_p, c_int]
    L.lsamd_cpu_batch.restype = c_int
    L.lsamd_cpu_single.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_uint64,
                                   c_double, c_double, c_uint32, c_void_p, c_int]
    L.lsamd_cpu_single.restype = c_int
    L.lsamd_count_objects.argtypes = [c_void_p, c_uint64]
    L.lsamd_count_objects.restype = c_uint32
    L.lsamd_object_offsets.argtypes = [c_void_p, c_uint64, c_void_p, c_void_p]
    L.lsamd_object_offsets.restype = c_uint32
    L.lsamd_last_error_code.restype = c_int
    L.lsamd_last_error.restype = c_char_p
    L.lsamd_device_count.restype = c_int
    L.lsamd_set_timing.argtypes = [c_int]
    L.lsamd_last_sweep_kernel_ms.restype = c_double
    L.lsamd_set_tile.argtypes = [c_int, c_int, c_int, c_int]
    L.lsamd_set_tile.restype = c_int
    L.lsamd_get_tile.argtypes = [c_int, c_int, c_int, POINTER(c_int), POINTER(c_int)]
    L.lsamd_version.restype = c_int
    L.lsamd_graph_create.argtypes = [POINTER(SweepArgs), c_void_p, c_size_t, c_void_p, POINTER(c_void_p)]
    L.lsamd_graph_create.restype = c_int
    L.lsamd_graph_launch.argtypes = [c_void_p, c_void_p]
    L.lsamd_graph_launch.restype = c_int
    L.lsamd_graph_destroy.argtypes = [c_void_p]
    L.lsamd_graph_destroy.restype = None
    L.lsamd_autotune.argtypes = [POINTER(SweepArgs), c_void_p, c_size_t, c_void_p, POINTER(c_int),
                                 POINTER(c_int)]
    L.lsamd_autotune.restype = c_int
    L.lsamd_autotune_workspace_size.argtypes = [POINTER(SweepArgs)]
    L.lsamd_autotune_workspace_size.restype = c_size_t
    L.lsamd_power_grad.argtypes = [POINTER(PointArgs), c_void_p]
    L.lsamd_power_grad.restype = c_int
    L.lsamd_refine_peaks.argtypes = [POINTER(PointArgs), c_void_p]
    L.lsamd_refine_peaks.restype = c_int
    L.lsamd_read_table.argtypes = [c_char_p, c_int, c_int, c_uint64, c_void_p, c_void_p, c_void_p,
                                   c_void_p, POINTER(c_uint64)]
    L.lsamd_read_table.restype = c_int
    L.lsamd_write_periods.argtypes = [c_char_p, c_void_p, c_uint32, c_uint32, c_int, c_void_p, c_void_p]
    L.lsamd_write_periods.restype = c_int
    L.lsamd_segment_workspace_size.argtypes = [c_uint64, c_uint64]
    L.lsamd_segment_workspace_size.restype = c_size_t
    L.lsamd_segment.argtypes = [c_void_p, c_uint64, c_void_p, c_uint64, c_void_p, c_void_p, c_void_p,
                                c_size_t, c_void_p]
    L.lsamd_segment.restype = c_int
    return L


_cache = {}


def _share_torch_runtime():
    """torch-ROCm bundles its own libamdhip64.so.7 (same soname as ROCm's).
    Whichever copy is loaded first serves the whole process, and torch only
    initialises against its own, so when torch is installed it is imported
    before this library is loaded; both then share torch's HIP runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(name=CORE_LIB, lib_dir=None):
    """Load one of the built libraries (raises NativeLibraryError if absent)."""
    path = os.path.join(lib_dir or LIB_DIR, name)
    if path not in _cache:
        _share_torch_runtime()
        if not os.path.exists(path):
            raise NativeLibraryError(
                f"{path} is not built; run __graft_entry__.build() or make -C gpu_lomb_scargle_amd/csrc")
        L = ctypes.CDLL(path)
        _declare(L)
        _cache[path] = L
    return _cache[path]


def core():
    return load(CORE_LIB)


def check(L, rc):
    if rc != OK:
        msg = L.lsamd_last_error()
        raise LsamdError(rc, msg.decode() if msg else "")
    return rc


def dtype_code(dtype):
    dt = np.dtype(dtype)
    if dt == np.float32:
        return F32
    if dt == np.float64:
        return F64
    raise ValueError(f"unsupported dtype {dt}")


def ptr(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def batch_host(obj_id, t, y, dy, fmin, fmax, nf, variant, dtype=np.float64, form=FORM_PLAIN,
               normalize=True, precenter=True, pgram=True, top_k=1, lib=None):
    """Run lsamd_batch_host on host arrays.  obj_id None = one object.
    Returns dict(pgram[K,nf] or None, index[K,top_k], power, period, y) where
    y is the (possibly pre-centred) magnitude array the library saw."""
    L = lib or core()
    npt = np.dtype(dtype).type
    t = np.ascontiguousarray(t, npt)
    y = np.array(y, npt)
    dyv = None if dy is None else np.ascontiguousarray(dy, npt)
    o = None if obj_id is None else np.ascontiguousarray(obj_id, np.uint32)
    n = t.size
    K = 1 if o is None else int(L.lsamd_count_objects(ptr(o), n))
    pg = np.zeros((K, nf), npt) if pgram else None
    tk = top_k
    idx = np.zeros((K, max(tk, 1)), np.uint32)
    pw = np.zeros((K, max(tk, 1)), npt)
    per = np.zeros((K, max(tk, 1)), npt)
    kout = c_uint32(0)
    rc = L.lsamd_batch_host(variant, dtype_code(dtype), form, int(normalize), int(precenter), ptr(o),
                            ptr(t), ptr(y), ptr(dyv), n, float(fmin), float(fmax), nf, ptr(pg), tk,
                            ptr(idx) if tk else None, ptr(pw) if tk else None,
                            ptr(per) if tk else None, ctypes.byref(kout))
    check(L, rc)
    return {"pgram": pg, "index": idx[:, :tk], "power": pw[:, :tk], "period": per[:, :tk], "y": y,
            "K": int(kout.value)}


def read_table(path, dtype=np.float64, with_dy=False, lib=None):
    """Observation table in the reference's input format (lsamd_read_table):
    returns (objectId uint32, t, y, dy or None) arrays of dtype."""
    L = lib or core()
    code = dtype_code(dtype)
    n = c_uint64(0)
    check(L, L.lsamd_read_table(os.fsencode(path), code, int(with_dy), 0, None, None, None, None,
                                ctypes.byref(n)))
    rows = n.value
    ids = np.zeros(rows, np.uint32)
    t, y = np.zeros(rows, dtype), np.zeros(rows, dtype)
    dy = np.zeros(rows, dtype) if with_dy else None
    check(L, L.lsamd_read_table(os.fsencode(path), code, int(with_dy), rows, ptr(ids), ptr(t), ptr(y),
                                ptr(dy), ctypes.byref(n)))
    return ids, t, y, dy


def write_periods(path, object_ids, periods, powers, lib=None):
    """Best periods in the reference's file format (lsamd_write_periods):
    periods/powers of shape [K] (bestperiods.txt) or [K, 3]
    (bestperiods_top_three.txt)."""
    L = lib or core()
    per = np.ascontiguousarray(periods)
    pw = np.ascontiguousarray(powers, per.dtype)
    ids = np.ascontiguousarray(object_ids, np.uint32)
    top_k = 1 if per.ndim == 1 else per.shape[1]
    check(L, L.lsamd_write_periods(os.fsencode(path), ptr(ids), ids.size, top_k, dtype_code(per.dtype),
                                   ptr(per), ptr(pw)))
