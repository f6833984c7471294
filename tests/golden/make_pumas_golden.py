"""Writes tests/golden/pumas_golden.npz: golden vectors of the PUMAS v1.2.1
DCS tabulation, computed by the reference's vendored PUMAS
(src/noa/3rdparty/_pumas/pumas.c) compiled in place into
oracle/_ref/libpumas_ref.so (`make -C oracle ref`).

    python tests/golden/make_pumas_golden.py

Contents:
  table_<set>_{Z,A,K,mass,cutoff}   the inputs of a table set
  table_<set>_dcs                   compute_dcs_table's rows [3, n_el, n_K, n_x, 2] (float32)
  table_<set>_x                     its nodes table_DCS_x [n_x] (float32)
  table_<set>_alpha                 the envelope exponents [3, n_el, n_K] (float32)
  table_<set>_envmax                the envelope maxima [3, n_el, n_K] (float32; the cross-section
                                    normalisation set to 1, so every row runs the search)
  point_{Z,A,K,q}, point_dcs        the model DCS (m^2/kg) of processes 0, 1, 2 at (K, q)
                                    pairs of four elements, muon: [3, n] (float64)
Set "muon": 6 elements, 61 kinetic energies 1e-3..1e7 GeV, cutoff 0.05;
set "tau": 2 elements, 20 kinetic energies 1e-1..1e8 GeV, cutoff 1e-3.
The pairs are log-uniform plus each model's kinematic edges approached from
both sides.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

MUON, TAU = 0.10565839, 1.77682
ME = 0.510998910E-03
M_N = 0.5 * (0.938272 + 0.939565)
MPI = 0.13957018

SETS = {
    "muon": dict(Z=[1.0, 6.0, 11.0, 26.0, 82.0, 92.0], A=[1.00794, 12.0107, 22.0, 55.845, 207.2, 238.02891],
                 K=10.0 ** np.linspace(-3, 7, 61), mass=MUON, cutoff=0.05),
    "tau": dict(Z=[11.0, 82.0], A=[22.0, 207.2], K=10.0 ** np.linspace(-1, 8, 20), mass=TAU, cutoff=1e-3),
}


def edge_pairs(Z: float, mass: float, rng) -> tuple[np.ndarray, np.ndarray]:
    """(K, q) at the models' kinematic edges, from both sides."""
    K = 10.0 ** rng.uniform(-2, 7, 60)
    eps = np.array([-1e-9, -1e-12, 0.0, 1e-12, 1e-9])
    qs, Ks = [], []
    sqrte = 1.6487212707
    for k in K:
        qmax_pair = k + mass * (1.0 - 0.75 * sqrte * Z ** (1.0 / 3.0))
        qmin_ph = MPI + 0.5 * MPI * MPI / M_N
        qmax_ph = k + mass - 0.5 * (M_N + mass * mass / M_N)
        for edge in (4 * ME, qmax_pair, qmin_ph, qmax_ph):
            for e in eps:
                qs.append(edge * (1 + e))
                Ks.append(k)
    return np.array(Ks), np.array(qs)


def main() -> None:
    ref = oracle.load_pumas_ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libpumas_ref.so is not built (make -C oracle ref)")
    out = {}
    for name, s in SETS.items():
        dcs, x, env = oracle.ref_pumas_dcs_table(s["Z"], s["A"], s["mass"], s["cutoff"], s["K"])
        out.update({f"table_{name}_Z": np.array(s["Z"]), f"table_{name}_A": np.array(s["A"]),
                    f"table_{name}_K": s["K"], f"table_{name}_mass": np.array(s["mass"]),
                    f"table_{name}_cutoff": np.array(s["cutoff"]), f"table_{name}_dcs": dcs,
                    f"table_{name}_x": x, f"table_{name}_alpha": env[..., 0],
                    f"table_{name}_envmax": env[..., 1]})
    rng = np.random.default_rng(2024)
    Zs, As, Ks, qs = [], [], [], []
    for Z, A in ((1.0, 1.00794), (11.0, 22.0), (26.0, 55.845), (92.0, 238.02891)):
        K = 10.0 ** rng.uniform(-3, 7, 700)
        q = K * 10.0 ** rng.uniform(-6, 0, 700)
        Ke, qe = edge_pairs(Z, MUON, rng)
        K, q = np.concatenate([K, Ke]), np.concatenate([q, qe])
        Zs.append(np.full(K.size, Z))
        As.append(np.full(K.size, A))
        Ks.append(K)
        qs.append(q)
    Z, A, K, q = (np.concatenate(v) for v in (Zs, As, Ks, qs))
    dcs = np.zeros((3, K.size))
    for p in range(3):
        for z in np.unique(Z):
            sel = Z == z
            dcs[p, sel] = oracle.ref_pumas_dcs(p, z, A[sel][0], MUON, K[sel], q[sel])
    out.update(point_Z=Z, point_A=A, point_K=K, point_q=q, point_dcs=dcs)
    path = os.path.join(HERE, "pumas_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes;", {p: int((dcs[p] > 0).sum()) for p in range(3)}, "positive of", K.size)


if __name__ == "__main__":
    main()
