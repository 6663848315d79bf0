"""CPU oracle for the Lomb-Scargle hot path (test infrastructure only).

Python bindings over ``liblsoracle.so`` (the plain-C restatement of the
reference's CPU path, ``ls_oracle.c``) and, when it has been built, over the
reference's own CPU functions compiled into ``oracle/_ref/``.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only to check GPU results against: the
GPU library never routes through it.

Reference citations (``source/main.cu`` of mgowanlock/gpu_lomb_scargle):
  * ``batch``   -> lombscargleCPUBatch[Error], main.cu:1401-1543 (MODE 4)
  * ``single``  -> lombscargleCPUOneObject[Error], main.cu:1358-1398 (MODE 5)
  * ``std_power`` -> lombscarglecpuinnerloop, main.cu:1566-1599
  * ``gen_power`` -> lombscarglecpuinnerloopAstroPy, main.cu:1701-1777
  * ``precenter`` -> updateYerrorfactor, main.cu:1631-1662
  * ``stddev_sq`` -> computeStandardDevSquared, main.cu:1295-1317
  * ``object_ranges`` -> computeObjectRanges, main.cu:468-510
  * ``top3`` -> PRINTPERIODS==3 block, main.cu:1165-1219
  * ``period_of_index`` -> main.cu:1136-1153
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_uint, c_uint64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsoracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build() -> None:
    """Compile liblsoracle.so (and _ref/ when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liblsoracle.so"], check=True)
    subprocess.run([os.path.join(HERE, "build_ref.sh")], check=True)


def _np_ptr(a, ctype):
    return a.ctypes.data_as(POINTER(ctype))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "liblsoracle.so"], check=True)
        L = ctypes.CDLL(LIB_PATH)
        for sfx, R in (("d", c_double), ("f", c_float)):
            PR = POINTER(R)
            getattr(L, f"lso_std_power_{sfx}").argtypes = [R, PR, PR, c_uint]
            getattr(L, f"lso_std_power_{sfx}").restype = R
            getattr(L, f"lso_gen_power_{sfx}").argtypes = [R, PR, PR, PR, c_uint]
            getattr(L, f"lso_gen_power_{sfx}").restype = R
            getattr(L, f"lso_precenter_{sfx}").argtypes = [PR, PR, c_uint]
            getattr(L, f"lso_stddev_sq_{sfx}").argtypes = [PR, c_uint64]
            getattr(L, f"lso_stddev_sq_{sfx}").restype = c_double
            getattr(L, f"lso_normalize_{sfx}").argtypes = [PR, PR, c_uint64, c_uint64]
            getattr(L, f"lso_compute_period_{sfx}").argtypes = [PR, c_uint, R, R, PR, PR, POINTER(c_uint)]
            getattr(L, f"lso_freq_step_{sfx}").argtypes = [R, R, c_uint]
            getattr(L, f"lso_freq_step_{sfx}").restype = R
            getattr(L, f"lso_batch_{sfx}").argtypes = [c_int, c_int, POINTER(c_uint), PR, PR, PR, c_uint,
                                                       c_double, c_double, c_uint, PR, PR, PR,
                                                       POINTER(c_uint), c_int]
            getattr(L, f"lso_single_{sfx}").argtypes = [c_int, c_int, PR, PR, PR, c_uint, c_double,
                                                        c_double, c_uint, PR, PR, PR, POINTER(c_uint), c_int]
            getattr(L, f"lso_period_of_index_{sfx}").argtypes = [c_double, c_double, c_uint, c_uint64]
            getattr(L, f"lso_period_of_index_{sfx}").restype = R
            getattr(L, f"lso_top3_{sfx}").argtypes = [PR, c_uint, c_uint, c_double, c_double, PR, PR,
                                                      POINTER(c_uint)]
        L.lso_object_ranges.argtypes = [POINTER(c_uint), c_uint, POINTER(c_uint), POINTER(c_uint),
                                        POINTER(c_uint)]
        L.lso_object_ranges.restype = c_uint
        L.lso_set_float_libm_cr.argtypes = [c_int]
        _lib = L
    return _lib


def _sfx(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "d", c_double, np.float64
    if dtype == np.float32:
        return "f", c_float, np.float32
    raise ValueError(f"unsupported dtype {dtype}")


def set_float_libm_cr(on):
    """Float oracle: glibc sinf/cosf/atan2f (False, the reference's libm) or
    correctly rounded versions (True)."""
    lib().lso_set_float_libm_cr(int(bool(on)))


class float_libm_cr:
    """Context manager: run the float oracle with correctly rounded sin/cos/atan2."""

    def __enter__(self):
        set_float_libm_cr(True)

    def __exit__(self, *exc):
        set_float_libm_cr(False)
        return False


def object_ranges(obj_id):
    """(idx_min, idx_max, ids) of the contiguous runs of obj_id."""
    o = np.ascontiguousarray(obj_id, dtype=np.uint32)
    n = o.size
    K = lib().lso_object_ranges(_np_ptr(o, c_uint), n, None, None, None)
    lo = np.zeros(K, np.uint32)
    hi = np.zeros(K, np.uint32)
    ids = np.zeros(K, np.uint32)
    lib().lso_object_ranges(_np_ptr(o, c_uint), n, _np_ptr(lo, c_uint), _np_ptr(hi, c_uint),
                            _np_ptr(ids, c_uint))
    return lo, hi, ids


def std_power(freq, x, y, dtype=np.float64):
    s, R, npt = _sfx(dtype)
    x = np.ascontiguousarray(x, npt)
    y = np.ascontiguousarray(y, npt)
    return getattr(lib(), f"lso_std_power_{s}")(R(freq), _np_ptr(x, R), _np_ptr(y, R), x.size)


def gen_power(freq, x, y, dy, dtype=np.float64):
    """Generalized power of already pre-centred y (the reference sweeps
    pre-centred magnitudes)."""
    s, R, npt = _sfx(dtype)
    x = np.ascontiguousarray(x, npt)
    y = np.ascontiguousarray(y, npt)
    dy = np.ascontiguousarray(dy, npt)
    return getattr(lib(), f"lso_gen_power_{s}")(R(freq), _np_ptr(x, R), _np_ptr(y, R),
                                                _np_ptr(dy, R), x.size)


def precenter(y, dy, dtype=np.float64):
    s, R, npt = _sfx(dtype)
    y = np.array(y, npt)
    dy = np.ascontiguousarray(dy, npt)
    getattr(lib(), f"lso_precenter_{s}")(_np_ptr(y, R), _np_ptr(dy, R), y.size)
    return y


def stddev_sq(y, dtype=np.float64):
    s, R, npt = _sfx(dtype)
    y = np.ascontiguousarray(y, npt)
    return getattr(lib(), f"lso_stddev_sq_{s}")(_np_ptr(y, R), y.size)


def freq_step(fmin, fmax, nf, dtype=np.float64):
    s, R, npt = _sfx(dtype)
    return getattr(lib(), f"lso_freq_step_{s}")(R(fmin), R(fmax), nf)


def freq_grid(fmin, fmax, nf, dtype=np.float64):
    """Angular frequencies of the CPU path: min + step*i in DTYPE
    (main.cu:1371, 1611)."""
    npt = np.dtype(dtype).type
    step = npt(freq_step(fmin, fmax, nf, dtype))
    i = np.arange(nf, dtype=np.uint32).astype(npt)
    return (npt(fmin) + step * i).astype(npt)


def batch(obj_id, t, y, dy, fmin, fmax, nf, error, dtype=np.float64, normalize=True, nthreads=1):
    """Reference MODE 4 over contiguous runs of obj_id.  fmin/fmax are angular
    (rad/day).  Returns (pgram[K, nf], periods[K], powers[K], argmax[K]) and
    leaves the caller's y untouched (the reference pre-centres in place)."""
    s, R, npt = _sfx(dtype)
    o = np.ascontiguousarray(obj_id, np.uint32)
    t = np.ascontiguousarray(t, npt)
    y = np.array(y, npt)
    dy = np.ascontiguousarray(np.ones_like(t) if dy is None else dy, npt)
    K = lib().lso_object_ranges(_np_ptr(o, c_uint), o.size, None, None, None)
    pg = np.zeros((K, nf), npt)
    per = np.zeros(K, npt)
    pw = np.zeros(K, npt)
    ix = np.zeros(K, np.uint32)
    getattr(lib(), f"lso_batch_{s}")(int(bool(error)), int(bool(normalize)), _np_ptr(o, c_uint),
                                     _np_ptr(t, R), _np_ptr(y, R), _np_ptr(dy, R), t.size,
                                     float(fmin), float(fmax), nf, _np_ptr(pg, R), _np_ptr(per, R),
                                     _np_ptr(pw, R), _np_ptr(ix, c_uint), int(nthreads))
    return pg, per, pw, ix


def single(t, y, dy, fmin, fmax, nf, error, dtype=np.float64, normalize=True, nthreads=1):
    """Reference MODE 5: the whole input as one object."""
    s, R, npt = _sfx(dtype)
    t = np.ascontiguousarray(t, npt)
    y = np.array(y, npt)
    dy = np.ascontiguousarray(np.ones_like(t) if dy is None else dy, npt)
    pg = np.zeros(nf, npt)
    per = R()
    pw = R()
    ix = c_uint()
    getattr(lib(), f"lso_single_{s}")(int(bool(error)), int(bool(normalize)), _np_ptr(t, R),
                                      _np_ptr(y, R), _np_ptr(dy, R), t.size, float(fmin), float(fmax),
                                      nf, _np_ptr(pg, R), ctypes.byref(per), ctypes.byref(pw),
                                      ctypes.byref(ix), int(nthreads))
    return pg, per.value, pw.value, ix.value


def period_of_index(fmin, fmax, nf, idx, dtype=np.float64):
    s, R, npt = _sfx(dtype)
    return getattr(lib(), f"lso_period_of_index_{s}")(float(fmin), float(fmax), nf, int(idx))


def top3(pgram, fmin, fmax, dtype=np.float64):
    """Top three distinct bins per object -> (periods[K,3], powers[K,3], idx[K,3])."""
    s, R, npt = _sfx(dtype)
    pg = np.ascontiguousarray(pgram, npt)
    K, nf = pg.shape
    per = np.zeros((K, 3), npt)
    pw = np.zeros((K, 3), npt)
    ix = np.zeros((K, 3), np.uint32)
    getattr(lib(), f"lso_top3_{s}")(_np_ptr(pg, R), K, nf, float(fmin), float(fmax), _np_ptr(per, R),
                                    _np_ptr(pw, R), _np_ptr(ix, c_uint))
    return per, pw, ix


# --------------------------------------------------------------------------
# The reference's own CPU functions (oracle/_ref/, built by build_ref.sh).
# --------------------------------------------------------------------------
_ref_cache = {}


def ref_path(error, dtype, normalize=True, threads=1):
    name = "libref_%s%s%s_t%d.so" % ("error" if error else "noerror",
                                    "double" if np.dtype(dtype) == np.float64 else "float",
                                    "" if (normalize or error) else "_raw", threads)
    return os.path.join(REF_DIR, name)


def ref_available(error=False, dtype=np.float64, normalize=True, threads=1):
    return os.path.exists(ref_path(error, dtype, normalize, threads))


def ref_lib(error, dtype, normalize=True, threads=1):
    key = (bool(error), np.dtype(dtype).str, bool(normalize), threads)
    if key not in _ref_cache:
        path = ref_path(error, dtype, normalize, threads)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        _, R, _ = _sfx(dtype)
        PR = POINTER(R)
        L.ref_batch.argtypes = [POINTER(c_uint), PR, PR, PR, c_uint, c_double, c_double, c_uint, PR, PR, PR]
        L.ref_single.argtypes = [PR, PR, PR, c_uint, c_double, c_double, c_uint, PR, PR, PR]
        L.ref_power.argtypes = [R, PR, PR, PR, c_uint]
        L.ref_power.restype = R
        L.ref_precenter.argtypes = [PR, PR, c_uint]
        L.ref_stddev_sq.argtypes = [PR, c_uint]
        L.ref_stddev_sq.restype = c_double
        L.ref_object_ranges.argtypes = [POINTER(c_uint), c_uint, POINTER(c_uint), POINTER(c_uint),
                                        POINTER(c_uint)]
        L.ref_object_ranges.restype = c_uint
        _ref_cache[key] = L
    return _ref_cache[key]


def ref_batch(obj_id, t, y, dy, fmin, fmax, nf, error, dtype=np.float64, normalize=True, threads=1):
    """Reference MODE 4 (its own code).  Same return convention as batch()
    minus the argmax index."""
    _, R, npt = _sfx(dtype)
    L = ref_lib(error, dtype, normalize, threads)
    o = np.ascontiguousarray(obj_id, np.uint32)
    t = np.ascontiguousarray(t, npt)
    y = np.array(y, npt)
    dy = np.ascontiguousarray(np.ones_like(t) if dy is None else dy, npt)
    K = lib().lso_object_ranges(_np_ptr(o, c_uint), o.size, None, None, None)
    pg = np.zeros((K, nf), npt)
    per = np.zeros(K, npt)
    pw = np.zeros(K, npt)
    L.ref_batch(_np_ptr(o, c_uint), _np_ptr(t, R), _np_ptr(y, R), _np_ptr(dy, R), t.size, float(fmin),
                float(fmax), nf, _np_ptr(pg, R), _np_ptr(per, R), _np_ptr(pw, R))
    return pg, per, pw


def ref_single(t, y, dy, fmin, fmax, nf, error, dtype=np.float64, normalize=True, threads=1):
    _, R, npt = _sfx(dtype)
    L = ref_lib(error, dtype, normalize, threads)
    t = np.ascontiguousarray(t, npt)
    y = np.array(y, npt)
    dy = np.ascontiguousarray(np.ones_like(t) if dy is None else dy, npt)
    pg = np.zeros(nf, npt)
    per = R()
    pw = R()
    L.ref_single(_np_ptr(t, R), _np_ptr(y, R), _np_ptr(dy, R), t.size, float(fmin), float(fmax), nf,
                 _np_ptr(pg, R), ctypes.byref(per), ctypes.byref(pw))
    return pg, per.value, pw.value


def ref_power(freq, x, y, dy, error, dtype=np.float64):
    _, R, npt = _sfx(dtype)
    L = ref_lib(error, dtype)
    x = np.ascontiguousarray(x, npt)
    y = np.ascontiguousarray(y, npt)
    dy = np.ascontiguousarray(np.ones_like(x) if dy is None else dy, npt)
    return L.ref_power(R(freq), _np_ptr(x, R), _np_ptr(y, R), _np_ptr(dy, R), x.size)


def ref_precenter(y, dy, dtype=np.float64):
    _, R, npt = _sfx(dtype)
    L = ref_lib(True, dtype)
    y = np.array(y, npt)
    dy = np.ascontiguousarray(dy, npt)
    L.ref_precenter(_np_ptr(y, R), _np_ptr(dy, R), y.size)
    return y


def ref_stddev_sq(y, dtype=np.float64):
    _, R, npt = _sfx(dtype)
    L = ref_lib(False, dtype)
    y = np.ascontiguousarray(y, npt)
    return L.ref_stddev_sq(_np_ptr(y, R), y.size)


def ref_object_ranges(obj_id):
    L = ref_lib(False, np.float64)
    o = np.ascontiguousarray(obj_id, np.uint32)
    lo = np.zeros(o.size, np.uint32)
    hi = np.zeros(o.size, np.uint32)
    ids = np.zeros(o.size, np.uint32)
    K = L.ref_object_ranges(_np_ptr(o, c_uint), o.size, _np_ptr(lo, c_uint), _np_ptr(hi, c_uint),
                            _np_ptr(ids, c_uint))
    return lo[:K], hi[:K], ids[:K]
