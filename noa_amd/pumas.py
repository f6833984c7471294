"""PUMAS v1.2.1 DCS tabulation on MI355X.

The physics tables that PUMAS (vendored by the reference under
src/noa/3rdparty/_pumas) builds from its DCS models: compute_dcs_table
(pumas.c:11693-11787) and dcs_tabulate_row (pumas.c:11422-11507) for the
default models, 0 SSR bremsstrahlung, 1 SSR pair production, 2 DRSS
photonuclear (pumas.c:190-198).

    table, x, envelope = pumas.dcs_table(Z, A, MUON_MASS, 0.05, K)

Z and A are sequences of floats (or host arrays), K a float64 GPU tensor; table is [3, n_el, n_K, n_x, 2] float32
(ln DCS and PCHIP slope per node, PUMAS's table_DCS layout), x the n_x nodes
ln(q / K), envelope [3, n_el, n_K, 2] (the exponent; entry 1, the envelope
maximum, is 0: not computed).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(t: torch.Tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _check(name: str, t: torch.Tensor, device=None) -> None:
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float64 GPU tensor")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")


def n_table_dcs(cutoff: float) -> int:
    """Grid nodes of a relative cutoff (pumas.c:1548-1554)."""
    n = _lib.lib().noa_pumas_n_table_dcs(float(cutoff))
    if n < 0:
        _lib.check(n)
    return n


def dcs(process: int, Z: float, A: float, mass: float, K: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
    """Model DCS (m^2/kg) of process 0, 1 or 2 over (K, q) pairs
    (dcs_get(process), pumas.c:11018-11112)."""
    _check("K", K)
    _check("q", q, K.device)
    if q.numel() != K.numel():
        raise ValueError("K and q must have equal numel")
    out = torch.empty_like(K)
    with torch.cuda.device(K.device):
        _lib.check(_lib.lib().noa_pumas_dcs_f64(int(process), float(Z), float(A), float(mass), _ptr(K), _ptr(q),
                                                K.numel(), _ptr(out), _stream(K)))
    return out


def dcs_table(Z, A, mass: float, cutoff: float, K: torch.Tensor, fp64: bool = False, csn=None):
    """compute_dcs_table for every element (Z[e], A[e]) at the kinetic
    energies K: (table [3, n_el, n_K, n_x, 2], x [n_x], envelope [3, n_el, n_K,
    2]: exponent and maximum), plus the FP64 table when fp64. csn (a device
    tensor [3, n_el, n_K] of doubles, or None: every row positive) is the
    cross-section normalisation that gates rows and envelopes."""
    _check("K", K)
    if csn is not None:
        _check("csn", csn)
        if csn.numel() != 3 * len(np.atleast_1d(Z)) * K.numel():
            raise ValueError("csn must hold 3 x n_elements x n_K values")
    Zh = np.ascontiguousarray(np.asarray(Z, dtype=np.float64).ravel())
    Ah = np.ascontiguousarray(np.asarray(A, dtype=np.float64).ravel())
    if Zh.size != Ah.size:
        raise ValueError("Z and A must have equal length")
    n = n_table_dcs(cutoff)
    n_el, n_K = Zh.size, K.numel()
    table = torch.empty(3, n_el, n_K, n, 2, dtype=torch.float32, device=K.device)
    x = torch.empty(n, dtype=torch.float32, device=K.device)
    env = torch.empty(3, n_el, n_K, 2, dtype=torch.float32, device=K.device)
    t64 = torch.empty(3, n_el, n_K, n, 2, dtype=torch.float64, device=K.device) if fp64 else None
    with torch.cuda.device(K.device):
        _lib.check(_lib.lib().noa_pumas_dcs_table(Zh.ctypes.data_as(ctypes.c_void_p), Ah.ctypes.data_as(ctypes.c_void_p),
                                                  n_el, float(mass), float(cutoff), _ptr(K), n_K,
                                                  _ptr(table), _ptr(x), _ptr(env), _ptr(t64),
                                                  _ptr(csn) if csn is not None else None, _stream(K)))
    return (table, x, env, t64) if fp64 else (table, x, env)
