"""The Lomb-Scargle sweep on device-resident torch tensors.

PyTorch-ROCm only supplies device memory and the stream here; the work is
done by ``lsamd_sweep`` of the C ABI (include/lombscargle_amd.h), called
through ctypes with raw device pointers.  Nothing falls back to a CPU path:
without a GPU or without the built library the calls raise.

``SweepPlan`` allocates every buffer once (outputs and workspace) so that
``run()`` only enqueues kernels: that is what the benchmark times.
"""
from __future__ import annotations

from ctypes import byref, c_void_p

import numpy as np
import torch

from . import _native

_TORCH_DTYPE = {_native.F32: torch.float32, _native.F64: torch.float64}


def _dtype_code(tdtype):
    if tdtype == torch.float32:
        return _native.F32
    if tdtype == torch.float64:
        return _native.F64
    raise ValueError(f"unsupported dtype {tdtype}")


def offsets_from_counts(counts, device):
    c = torch.as_tensor(counts, dtype=torch.int64)
    off = torch.zeros(c.numel() + 1, dtype=torch.int64)
    off[1:] = torch.cumsum(c, 0)
    return off.to(device)


class SweepPlan:
    """Pre-allocated sweep over K objects held as contiguous row ranges.

    t, y, dy: 1-D float32/float64 CUDA tensors (dy only for the generalized
    variant); offsets: int64 CUDA tensor of K+1 row offsets.  order (optional
    int32/uint32 CUDA tensor of K object numbers) is the processing order;
    by default objects are processed longest first.  obs_slices=1 keeps a
    single-object split-form sweep from spreading its observations over
    workgroups (0, the default, lets it).  tile=(threads, freqs_per_thread)
    picks a compiled tile shape for this plan only (None: the process
    default); autotune() sets it to the fastest.
    """

    def __init__(self, t, y, dy, offsets, *, min_freq, max_freq, num_freqs, variant,
                 form=_native.FORM_SPLIT, normalize=True, precenter=True, pgram=True, top_k=1,
                 order="longest_first", schedule=_native.SCHEDULE_GRID, obs_slices=0, tile=None,
                 lib=None):
        if not t.is_cuda:
            raise RuntimeError("SweepPlan needs CUDA (HIP) tensors")
        self.lib = lib or _native.core()
        self.dtype = _dtype_code(t.dtype)
        dev = t.device
        self.t, self.y = t.contiguous(), y.contiguous()
        self.dy = dy.contiguous() if dy is not None else None
        self.offsets = offsets.to(device=dev, dtype=torch.int64).contiguous()
        K = self.offsets.numel() - 1
        self.K, self.nf = K, int(num_freqs)
        if isinstance(order, str) and order == "longest_first":
            lens = (self.offsets[1:] - self.offsets[:-1]).cpu()
            order = torch.argsort(-lens, stable=True).to(torch.int32).to(dev)
        self.order = order.to(device=dev, dtype=torch.int32).contiguous() if order is not None else None
        td = _TORCH_DTYPE[self.dtype]
        self.pgram = torch.empty(K * self.nf, dtype=td, device=dev) if pgram else None
        self.top_k = int(top_k)
        kk = max(self.top_k, 1)
        self.peak_index = torch.zeros(K * kk, dtype=torch.int32, device=dev)
        self.peak_power = torch.zeros(K * kk, dtype=td, device=dev)
        self.peak_period = torch.zeros(K * kk, dtype=td, device=dev)
        a = _native.SweepArgs()
        a.variant, a.dtype, a.form = int(variant), self.dtype, int(form)
        a.normalize, a.precenter = int(bool(normalize)), int(bool(precenter))
        a.num_objects, a.num_freqs, a.top_k = K, self.nf, self.top_k
        a.num_obs = self.t.numel()
        a.min_freq, a.max_freq = float(min_freq), float(max_freq)
        a.offsets = self.offsets.data_ptr()
        a.order = self.order.data_ptr() if self.order is not None else None
        a.t, a.y = self.t.data_ptr(), self.y.data_ptr()
        a.dy = self.dy.data_ptr() if self.dy is not None else None
        a.pgram = self.pgram.data_ptr() if self.pgram is not None else None
        if self.top_k:
            a.peak_index = self.peak_index.data_ptr()
            a.peak_power = self.peak_power.data_ptr()
            a.peak_period = self.peak_period.data_ptr()
        a.schedule = int(schedule)
        a.obs_slices = int(obs_slices)
        if tile is not None:
            a.tile_threads, a.tile_fpt = int(tile[0]), int(tile[1])
        self.args = a
        ws = int(self.lib.lsamd_sweep_workspace_size(byref(a)))
        self.workspace = torch.empty(ws, dtype=torch.uint8, device=dev)
        self.graph = None
        self._graph_stream = None

    def _stream(self, stream):
        return stream if stream is not None else torch.cuda.current_stream(self.t.device)

    def tile(self):
        """(threads, freqs_per_thread) of this plan's sweep kernel."""
        from ctypes import c_int
        a = self.args
        if a.tile_threads and a.tile_fpt:
            return a.tile_threads, a.tile_fpt
        T, M = c_int(), c_int()
        self.lib.lsamd_get_tile(a.form, a.variant, a.dtype, byref(T), byref(M))
        return T.value, M.value

    def autotune(self, stream=None):
        """Pick the fastest compiled tile shape for this plan's form/variant
        (lsamd_autotune), keep it for this plan and resize the workspace for
        it; returns (T, M)."""
        from ctypes import c_int
        need = int(self.lib.lsamd_autotune_workspace_size(byref(self.args)))
        big = torch.empty(max(need, 256), dtype=torch.uint8, device=self.t.device)
        T, M = c_int(), c_int()
        rc = self.lib.lsamd_autotune(byref(self.args), c_void_p(big.data_ptr()), big.numel(),
                                     c_void_p(self._stream(stream).cuda_stream), byref(T), byref(M))
        _native.check(self.lib, rc)
        del big
        self.args.tile_threads, self.args.tile_fpt = T.value, M.value
        ws = int(self.lib.lsamd_sweep_workspace_size(byref(self.args)))
        self.workspace = torch.empty(ws, dtype=torch.uint8, device=self.t.device)
        self.destroy_graph()
        return T.value, M.value

    def capture_graph(self, stream):
        """Capture this sweep into a HIP graph on `stream` (a non-default
        torch.cuda.Stream); later run() calls replay it."""
        self.destroy_graph()
        g = c_void_p()
        rc = self.lib.lsamd_graph_create(byref(self.args), c_void_p(self.workspace.data_ptr()),
                                         self.workspace.numel(), c_void_p(stream.cuda_stream), byref(g))
        _native.check(self.lib, rc)
        self.graph = g
        self._graph_stream = stream

    def destroy_graph(self):
        if self.graph is not None:
            self.lib.lsamd_graph_destroy(self.graph)
            self.graph = None

    def __del__(self):
        try:
            self.destroy_graph()
        except Exception:
            pass

    def run(self, stream=None):
        """Enqueue the sweep on `stream` (default: torch's current stream),
        replaying the captured graph when there is one."""
        s = self._stream(stream)
        if self.graph is not None:
            rc = self.lib.lsamd_graph_launch(self.graph, c_void_p(s.cuda_stream))
        else:
            rc = self.lib.lsamd_sweep(byref(self.args), c_void_p(self.workspace.data_ptr()),
                                      self.workspace.numel(), c_void_p(s.cuda_stream))
        _native.check(self.lib, rc)

    def results(self):
        kk = max(self.top_k, 1)
        out = {
            "index": self.peak_index.view(self.K, kk)[:, :self.top_k],
            "power": self.peak_power.view(self.K, kk)[:, :self.top_k],
            "period": self.peak_period.view(self.K, kk)[:, :self.top_k],
        }
        if self.pgram is not None:
            out["pgram"] = self.pgram.view(self.K, self.nf)
        return out


def sweep(t, y, dy, offsets, **kw):
    """One-shot sweep; returns the result dict of SweepPlan.results()."""
    plan = SweepPlan(t, y, dy, offsets, **kw)
    plan.run()
    return plan.results()


def segment(obj_id, max_objects=None, stream=None):
    """Device segmentation (lsamd_segment) of a CUDA uint32/int32 tensor of
    object ids: returns (offsets int64 [K+1], order int32 [K]), objects being
    maximal runs of equal consecutive ids and the order longest class first.
    One synchronisation (to read K)."""
    if not obj_id.is_cuda:
        raise RuntimeError("segment needs a CUDA (HIP) tensor")
    lib = _native.core()
    ids = obj_id.to(torch.int32).contiguous()
    n = ids.numel()
    cap = n if max_objects is None else int(max_objects)
    dev = ids.device
    off = torch.empty(cap + 1, dtype=torch.int64, device=dev)
    order = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
    k = torch.empty(1, dtype=torch.int32, device=dev)
    ws = torch.empty(int(lib.lsamd_segment_workspace_size(n, cap)), dtype=torch.uint8, device=dev)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    _native.check(lib, lib.lsamd_segment(c_void_p(ids.data_ptr()), n, c_void_p(off.data_ptr()), cap,
                                         c_void_p(order.data_ptr()), c_void_p(k.data_ptr()),
                                         c_void_p(ws.data_ptr()), ws.numel(), c_void_p(s.cuda_stream)))
    K = int(k.item()) & 0xFFFFFFFF
    if K > cap:
        raise ValueError(f"{K} objects exceed max_objects={cap}")
    return off[:K + 1], order[:K]


def numpy_to_device(obj_id, t, y, dy, dtype=np.float64, device="cuda"):
    """Host observation arrays -> (t, y, dy, offsets) CUDA tensors, objects
    as runs of obj_id (None = one object)."""
    npt = np.dtype(dtype)
    t = np.ascontiguousarray(t, npt)
    n = t.size
    if obj_id is None:
        off = np.array([0, n], np.int64)
    else:
        o = np.ascontiguousarray(obj_id, np.uint32)
        K = int(_native.core().lsamd_count_objects(_native.ptr(o), n))
        off = np.zeros(K + 1, np.uint64)
        _native.core().lsamd_object_offsets(_native.ptr(o), n, _native.ptr(off), None)
        off = off.astype(np.int64)
    tt = torch.from_numpy(t).to(device)
    yy = torch.from_numpy(np.ascontiguousarray(y, npt)).to(device)
    dd = None if dy is None else torch.from_numpy(np.ascontiguousarray(dy, npt)).to(device)
    return tt, yy, dd, torch.from_numpy(off).to(device)
