"""Batched multi-element evaluation: one persistent launch replayed from a hipGraph.

    plan = BatchPlan(elements, offsets, K, q, MUON_MASS, precision="f64")
    plan.launch()                 # replays the captured graph on the current stream
    brems, pair, photo, ion = plan.outputs

Element e owns pairs [offsets[e], offsets[e+1]) of K, q and the outputs.
The reference has no batched form: a caller loops dcs::vmap over elements
(src/noa/pms/dcs.hh:35-48).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

_PREC = {"f64": _lib.PREC_F64, "f32mixed": _lib.PREC_F32MIXED}


class BatchPlan:
    def __init__(self, elements, offsets, K: torch.Tensor, q: torch.Tensor, mass: float,
                 precision: str = "f64", outputs=None, tile: int = 0, waves_per_cu: int = 0,
                 use_graph: bool = True, tune_cache: str | None = None, unroll: int = 0):
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {list(_PREC)}")
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        if offsets.ndim != 1 or len(offsets) != len(elements) + 1:
            raise ValueError("offsets must have len(elements) + 1 entries")
        for name, t in (("K", K), ("q", q)):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous float64 GPU tensor")
        if K.numel() != q.numel() or K.numel() < int(offsets[-1]):
            raise ValueError("K and q must hold offsets[-1] pairs")
        odt = torch.float64 if precision == "f64" else torch.float32
        if outputs is None:
            outputs = tuple(torch.zeros(K.numel(), dtype=odt, device=K.device) for _ in range(4))
        for o in outputs:
            if o is not None and (o.dtype != odt or o.device != K.device or not o.is_contiguous()
                                  or o.numel() < int(offsets[-1])):
                raise ValueError(f"outputs must be contiguous {odt} tensors on {K.device} with offsets[-1] elements")
        self.K, self.q, self.outputs = K, q, tuple(outputs)
        self.elements = list(elements)
        self.offsets = offsets
        self.device = K.device
        els = (_lib.Element * len(elements))(*[_lib.Element(float(e.A), float(e.I), int(e.Z)) for e in elements])
        offs = (ctypes.c_int64 * len(offsets))(*[int(x) for x in offsets])
        cfg = _lib.BatchConfig(_PREC[precision], tile, waves_per_cu, int(use_graph),
                               tune_cache.encode() if tune_cache else None, unroll)
        handle = ctypes.c_void_p()
        ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())  # noqa: E731
        with torch.cuda.device(self.device):
            stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
            _lib.check(_lib.lib().noa_dcs_batch_create(els, offs, len(elements), float(mass), ptr(K), ptr(q),
                                                       *(ptr(o) for o in self.outputs), ctypes.byref(cfg), stream,
                                                       ctypes.byref(handle)))
        self._handle = handle

    def launch(self) -> None:
        if self._handle is None:
            raise RuntimeError("plan was closed")
        with torch.cuda.device(self.device):
            stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
            _lib.check(_lib.lib().noa_dcs_batch_launch(self._handle, stream))

    def info(self) -> dict:
        t, w, u, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _lib.check(_lib.lib().noa_dcs_batch_info(self._handle, ctypes.byref(t), ctypes.byref(w), ctypes.byref(u),
                                                 ctypes.byref(n)))
        return {"tile": t.value, "waves_per_cu": w.value, "unroll": u.value, "n_items": n.value}

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None:
            _lib.lib().noa_dcs_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
