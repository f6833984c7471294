"""Seeded synthetic (K, q) workloads, on the host (numpy) and on the device.

u(seed, stream, i) is a 53-bit uniform from a SplitMix64 output function
applied to a (seed, stream, i) counter; the device generator in
csrc/dcs_eval.hip computes the same integers, so the uniforms agree bit for
bit and K, q agree to the last ulp of the two pow implementations.

Standard-rock workload: K = 10^(-3 + 9 u0) GeV (log-uniform in [1e-3, 1e6]),
nu = 10^(-6 + 6 u1) (log-uniform in [1e-6, 1)), q = nu K; u0 from stream 2s,
u1 from stream 2s + 1.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from .physics import AtomicElement

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
STREAM_MUL = 0xD1B54A32D192ED03

LOG_K = (-3.0, 6.0)
LOG_NU = (-6.0, 0.0)


def _mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform53(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    key0 = (seed * GOLDEN + (stream + 1) * STREAM_MUL) & M64
    key = _mix64(np.array([key0], dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        z = _mix64(key + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(GOLDEN))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def pairs(n: int, seed: int = 0, stream: int = 0, offset: int = 0,
          log_k=LOG_K, log_nu=LOG_NU) -> tuple[np.ndarray, np.ndarray]:
    """Host (K, q) of the standard workload, indices [offset, offset + n)."""
    idx = np.arange(offset, offset + n, dtype=np.uint64)
    u0 = uniform53(seed, 2 * stream, idx)
    u1 = uniform53(seed, 2 * stream + 1, idx)
    K = np.power(10.0, log_k[0] + (log_k[1] - log_k[0]) * u0)
    nu = np.power(10.0, log_nu[0] + (log_nu[1] - log_nu[0]) * u1)
    return K, nu * K


def fill_device(K, q, seed: int = 0, stream: int = 0, log_k=LOG_K, log_nu=LOG_NU) -> None:
    """Fills two contiguous float64 GPU tensors with the same workload."""
    import torch

    from . import _lib
    if K.dtype != torch.float64 or q.dtype != torch.float64 or K.numel() != q.numel():
        raise ValueError("K and q must be float64 tensors of equal size")
    if not (K.is_cuda and q.is_cuda and K.is_contiguous() and q.is_contiguous()):
        raise ValueError("K and q must be contiguous GPU tensors")
    s = ctypes.c_void_p(torch.cuda.current_stream(K.device).cuda_stream)
    with torch.cuda.device(K.device):
        _lib.check(_lib.lib().noa_dcs_fill_workload(ctypes.c_void_p(K.data_ptr()), ctypes.c_void_p(q.data_ptr()),
                                                    K.numel(), seed, stream, log_k[0], log_k[1], log_nu[0],
                                                    log_nu[1], s))


def mean_excitation(Z: int) -> float:
    """Mean excitation energy rule for synthetic elements: I = 16 eV Z^0.9 (GeV)."""
    return 16e-9 * Z ** 0.9


def element_batch(n_elements: int = 48, seed: int = 0, log2_min: float = 10.0,
                  log2_max: float = 20.0) -> tuple[list[AtomicElement], np.ndarray]:
    """Uneven multi-element batch: Z uniform in 1..92, A = Z U(2.0, 2.6) g/mol,
    per-element pair counts floor(2^U(log2_min, log2_max)). Returns the
    elements and the offsets (n_elements + 1) of their pair ranges."""
    e = np.arange(n_elements, dtype=np.uint64)
    uz = uniform53(seed, 1000, e)
    ua = uniform53(seed, 1001, e)
    un = uniform53(seed, 1002, e)
    elements, counts = [], []
    for i in range(n_elements):
        Z = 1 + int(math.floor(92 * uz[i]))
        A = Z * (2.0 + 0.6 * ua[i])
        elements.append(AtomicElement(A=float(A), I=mean_excitation(Z), Z=Z))
        counts.append(int(math.floor(2.0 ** (log2_min + (log2_max - log2_min) * un[i]))))
    offsets = np.zeros(n_elements + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)
    return elements, offsets


def batch_pairs(offsets: np.ndarray, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Host (K, q) of a batch: element e draws from stream 1 + e."""
    Ks, qs = [], []
    for e in range(len(offsets) - 1):
        k, q = pairs(int(offsets[e + 1] - offsets[e]), seed=seed, stream=1 + e)
        Ks.append(k)
        qs.append(q)
    return np.concatenate(Ks), np.concatenate(qs)


def fill_device_batch(K, q, offsets: np.ndarray, seed: int = 0) -> None:
    """Device twin of batch_pairs."""
    for e in range(len(offsets) - 1):
        a, b = int(offsets[e]), int(offsets[e + 1])
        fill_device(K[a:b], q[a:b], seed=seed, stream=1 + e)
