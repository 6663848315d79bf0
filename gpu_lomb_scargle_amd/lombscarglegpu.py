"""Python interface of the MI355X-native Lomb-Scargle engine.

Drop-in counterpart of the reference's ``source/lombscarglegpu.py``: the same
``lombscargle(objId, timeX, magY, minFreq, maxFreq, error, mode, magDY=None,
freqToTest="auto", dtype="float", verbose=False)`` call returning
``(uniqueObjectIds, periods, pgram[K, Nf])``, loading the same four library
names (``libpyls{noerror,error}{float,double}.so``) and calling the same C
entry point ``LombScarglePy`` through ctypes (reference ``py:220-395``).

Behaviour kept from the reference:
  * minFreq/maxFreq are ordinary frequencies (cycles/day), converted to
    angular before the C call (py:227-228); ``freqToTest="auto"`` picks
    int((fmax-fmin)/(0.1/max_window)) bins (Richards et al. 2011, py:179-206);
  * objects are contiguous runs of objId, enumerated 0..K-1 in run order
    (py:120-175); MODE is 1/2 (GPU) or 4/5 (CPU) by object count (py:249-257);
  * periods are 1/(fmin + df*argmax) with numpy's argmax over the returned,
    normalised periodogram (py:378-383);
  * error=True without magDY prints a message and exits (py:261-263);
  * with error=True the library pre-centres the magnitude array it is given
    in place; with dtype="double" and a contiguous float64 magY that is the
    caller's own array, as in the reference.

Differences (all additive): the libraries are loaded from this package's
``lib/`` directory (or ``$LSAMD_LIB_DIR``) instead of the working directory;
a failed library call raises ``RuntimeError`` instead of killing the
process; an optional ``form="split"`` (or ``"split32"``) keyword selects
the fast GPU arithmetic (default: the reference's arithmetic).
"""
from __future__ import annotations

import csv
import os
import sys
from contextlib import contextmanager
from ctypes import c_double, c_float, c_int, c_uint

import numpy as np
import numpy.ctypeslib as npct

from . import _native


def getColumn(filename, column):
    """One CSV column as a list of strings (reference py:10-13)."""
    with open(filename) as fh:
        return [row[column] for row in csv.reader(fh, delimiter=",")]


def convert_type(in_array, new_dtype):
    """Return a C-contiguous array of new_dtype, reusing in_array when it
    already is one (reference py:17-30)."""
    ret = in_array
    if not isinstance(in_array, np.ndarray):
        ret = np.array(in_array, dtype=new_dtype)
    elif in_array.dtype != new_dtype:
        ret = np.array(ret, dtype=new_dtype)
    if not ret.flags["C_CONTIGUOUS"]:
        ret = np.ascontiguousarray(ret)
    return ret


class HideOutput:
    """Silence file descriptor 1 (the C library's stdout) for a scope."""

    def __enter__(self):
        sys.stdout.flush()
        self._saved = os.dup(1)
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)
        os.close(devnull)
        return self

    def __exit__(self, exc_type, exc, tb):
        sys.stdout.flush()
        os.dup2(self._saved, 1)
        os.close(self._saved)
        return False


def computeIndexRangesForEachObject(objId):
    """(start, end, ids) of the contiguous runs of objId (reference py:120-143)."""
    ids = np.asarray(objId)
    n = ids.shape[0]
    if n == 0:
        return np.zeros(0, int), np.zeros(0, int), ids[:0]
    change = np.nonzero(ids[1:] != ids[:-1])[0] + 1
    start = np.concatenate(([0], change)).astype(int)
    end = np.concatenate((change - 1, [n - 1])).astype(int)
    return start, end, ids[start]


def enumerateObjects(start_index_arr, end_index_arr):
    """Run-order object numbers 0..K-1 per row (reference py:165-175)."""
    counts = np.asarray(end_index_arr) - np.asarray(start_index_arr) + 1
    return np.repeat(np.arange(counts.size), counts).astype(int)


def computeNumFreqAuto(objId, timeX, fmin, fmax):
    """Number of frequency bins from the longest observing window:
    int((fmax - fmin) / (0.1 / window)) (reference py:179-206)."""
    start, end, _ = computeIndexRangesForEachObject(objId)
    t = np.asarray(timeX, dtype=float)
    window = np.max(t[end] - t[start])
    deltaf = 0.1 / window
    num_freqs = int((fmax - fmin) / deltaf)
    print("*********************")
    print("Automatically generating the number of frequencies based on maximum observing window:")
    print("Max. Observing Window: %f, Delta f: %f" % (window, deltaf))
    print("Number of frequencies: ", num_freqs)
    print("*********************")
    return num_freqs


@contextmanager
def _form_env(form):
    if form is None:
        yield
        return
    if form not in ("plain", "split", "split32"):
        raise ValueError("form must be 'plain', 'split' or 'split32'")
    old = os.environ.get("LSAMD_FORM")
    os.environ["LSAMD_FORM"] = form
    try:
        yield
    finally:
        if old is None:
            os.environ.pop("LSAMD_FORM", None)
        else:
            os.environ["LSAMD_FORM"] = old


def lombscargle(objId, timeX, magY, minFreq, maxFreq, error, mode, magDY=None, freqToTest="auto",
                dtype="float", verbose=False, form=None):
    """Reference-compatible entry point (py:210-217)."""
    if not verbose:
        with HideOutput():
            return lombscarglemain(objId, timeX, magY, minFreq, maxFreq, error, mode, magDY,
                                   freqToTest, dtype, form=form)
    return lombscarglemain(objId, timeX, magY, minFreq, maxFreq, error, mode, magDY, freqToTest,
                           dtype, verbose, form=form)


def lombscarglemain(objId, timeX, magY, minFreq, maxFreq, error, mode, magDY=None,
                    freqToTest="auto", dtype="float", verbose=False, form=None):
    """Reference py:220-395."""
    minFreqStandard, maxFreqStandard = minFreq, maxFreq
    minFreq = 2.0 * np.pi * minFreq
    maxFreq = 2.0 * np.pi * maxFreq

    if isinstance(freqToTest, str) and freqToTest == "auto":
        freqToTest = computeNumFreqAuto(objId, timeX, minFreqStandard, maxFreqStandard)

    numObjects = np.size(np.unique(objId))
    if mode == "GPU" and numObjects > 1:
        setmode = 1
    elif mode == "GPU" and numObjects == 1:
        setmode = 2
    elif mode == "CPU" and numObjects > 1:
        setmode = 4
    elif mode == "CPU" and numObjects == 1:
        setmode = 5
    else:
        raise ValueError(f"mode must be 'GPU' or 'CPU' (got {mode!r})")

    if error is True and magDY is None:
        print("[Python] Error: No input error array, but the error mode is True. Set error mode to False.")
        sys.exit(0)

    start, end, uniqueIdsOrdered = computeIndexRangesForEachObject(objId)
    objId = enumerateObjects(start, end)

    if dtype not in ("float", "double"):
        raise ValueError("dtype must be 'float' or 'double'")
    lib = _native.load(_native.DROPIN_LIBS[(bool(error), dtype)])
    ctype = c_float if dtype == "float" else c_double
    arr = npct.ndpointer(dtype=ctype, ndim=1, flags="CONTIGUOUS")
    lib.LombScarglePy.argtypes = [npct.ndpointer(dtype=c_uint, ndim=1, flags="CONTIGUOUS"), arr, arr,
                                  arr, c_uint, c_double, c_double, c_uint, c_int, arr]
    lib.LombScarglePy.restype = None

    sizeData = len(objId)
    print("[Python] Number of rows in file: %d" % sizeData)

    objId = np.asarray(objId, dtype=int)
    # np.asfarray semantics of the reference (py:298-302): float64, no copy
    # when the input already is a float64 array
    timeX = np.asarray(timeX, dtype=np.float64)
    magY = np.asarray(magY, dtype=np.float64)
    if error:
        magDY = np.asarray(magDY, dtype=np.float64)
    else:
        magDY = np.full(objId.size, 1.0)

    c_objId = convert_type(objId, c_uint)
    c_timeX = convert_type(timeX, ctype)
    c_magY = convert_type(magY, ctype)
    c_magDY = convert_type(magDY, ctype)

    dfstandard = (maxFreqStandard - minFreqStandard) / freqToTest * 1.0
    uniqueObjects = np.size(np.unique(objId))
    print("[Python] Unique objects: %d" % uniqueObjects)
    ret_pgram = np.zeros(uniqueObjects * freqToTest, dtype=ctype)
    print("[Python] Size of pgram in elems: %d (%f GiB)" %
          (ret_pgram.size, ret_pgram.nbytes / (1024.0 ** 3)))

    lib.lsamd_clear_error()
    with _form_env(form):
        lib.LombScarglePy(c_objId, c_timeX, c_magY, c_magDY, c_uint(sizeData), c_double(minFreq),
                          c_double(maxFreq), c_uint(freqToTest), c_int(setmode), ret_pgram)
    rc = lib.lsamd_last_error_code()
    if rc != _native.OK:
        raise RuntimeError("LombScarglePy failed (%d): %s" % (rc, lib.lsamd_last_error().decode()))

    ret_pgram = ret_pgram.reshape([uniqueObjects, freqToTest])
    ret_periods = np.zeros(uniqueObjects)
    for x in range(uniqueObjects):
        ret_periods[x] = 1.0 / (minFreqStandard + (dfstandard * np.argmax(ret_pgram[x])))

    if uniqueObjects > 1:
        print("[Python] Sum of all periods: %f" % np.sum(ret_periods))
    else:
        print("[Python] Period for object: %f" % ret_periods[0])
    return uniqueIdsOrdered, ret_periods, ret_pgram
