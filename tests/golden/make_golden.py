"""Writes tests/golden/dcs_golden.npz and tests/golden/coulomb_golden.npz
from the reference implementation itself.

    make -C oracle ref && python tests/golden/make_golden.py [dcs|coulomb|all]

oracle/_ref/libnoa_ref.so is the reference's own DCS headers
(src/noa/pms/dcs.hh, physics.hh, utils/numerics.hh, utils/common.hh) compiled
in place by oracle/Makefile; this script only calls its dcs::vmap adaptor.
The fixture holds, per element, the inputs (K, q) and the four reference
outputs (bremsstrahlung, pair production, photonuclear, ionisation) for:
  - 2048 pairs of the standard workload (K log-uniform in [1e-3, 1e6] GeV,
    nu log-uniform in [1e-6, 1), seed 11),
  - edge cases: every threshold of the four processes approached from both
    sides (q = 4 m_e, q = K + m(1 - 0.75 sqrt(e) Z^1/3), q = 1 GeV,
    q = 2e-3 K, q = K, q = Wmax, q = 0.62 I, K = (m - m_e)^2 / 2 m_e, the
    zeta switch gamma = 35, q -> K), and the notebook inputs
    (docs/pms/muon_dcs_calc.ipynb: K = linspace(1e-3, 1e6, 10000) in float32
    then float64, q = 0.0505 K; first 64 values).
Elements: standard rock (Z=11), hydrogen (Z=1 branches), iron, uranium; muon
mass, plus a tau-mass block for standard rock.

coulomb_golden.npz holds the Coulomb-scattering helpers (dcs.hh:499-950) on a
K grid from 1e-9 GeV (the K cutoff and the zeta^2 > 1 series) to 1e7 GeV:
dcs::coulomb_data, dcs::coulomb_transport at mu = 1 (as in
test/unit/test-dcs-calc.cc:134-166) and at per-K mu from 1e-14 to 1 (both
transport branches), dcs::hard_scattering for one element and for a
three-element mixture, and dcs::soft_scattering.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from noa_amd import workload  # noqa: E402
from noa_amd.physics import ELECTRON_MASS, MUON_MASS, STANDARD_ROCK, TAU_MASS, AtomicElement  # noqa: E402

ELEMENTS = {
    "rock": STANDARD_ROCK,
    "hydrogen": AtomicElement(A=1.00794, I=19.2e-9, Z=1),
    "iron": AtomicElement(A=55.845, I=286e-9, Z=26),
    "uranium": AtomicElement(A=238.02891, I=890e-9, Z=92),
}


def edge_inputs(el: AtomicElement, mass: float) -> tuple[np.ndarray, np.ndarray]:
    me = ELECTRON_MASS
    Ks, qs = [], []
    rel = np.array([-1e-9, -1e-12, 0.0, 1e-12, 1e-9])
    for K in 10.0 ** np.arange(-3, 7, 0.5):
        E = K + mass
        P2 = K * (K + 2 * mass)
        wmax = 2 * me * P2 / (mass * mass + me * (me + 2 * E))
        cut_pair = K + mass * (1 - 0.75 * 1.6487212707 * el.Z ** (1 / 3))
        for q0 in (4 * me, cut_pair, 1.0, 2e-3 * K, K, wmax, 0.62 * el.I, 0.5 * K, 0.99 * K):
            if q0 <= 0:
                continue
            for r in rel:
                Ks.append(K)
                qs.append(q0 * (1 + r))
    # radiative-correction switch of ionisation and the zeta switch of pair production
    kthr = 0.5 * (mass - me) ** 2 / me
    for K in (kthr * (1 - 1e-12), kthr, kthr * (1 + 1e-12), 34 * mass * (1 - 1e-9), 34 * mass, 34 * mass * (1 + 1e-9)):
        for nu in (1e-3, 1e-2, 0.1, 0.5):
            Ks.append(K)
            qs.append(nu * K)
    # notebook inputs (float32 linspace, as in docs/pms/muon_dcs_calc.ipynb)
    import torch
    kn = torch.linspace(1e-3, 1e6, 10000).double().numpy()[:64]
    Ks += list(kn)
    qs += list(0.0505 * kn)
    return np.array(Ks), np.array(qs)


def block(el: AtomicElement, mass: float, seed: int) -> dict:
    K0, q0 = workload.pairs(2048, seed=seed)
    K1, q1 = edge_inputs(el, mass)
    K = np.concatenate([K0, K1])
    q = np.concatenate([q0, q1])
    out = np.stack([oracle.ref_map(p, K, q, el, mass) for p in range(4)])
    return {"K": K, "q": q, "dcs": out}


COULOMB_K = np.concatenate([[1e-9, 1e-7, 1e-5, 1e-4], 10.0 ** np.linspace(-3, 6, 91), [1e7]])


def coulomb_main() -> None:
    K = COULOMB_K
    n = K.size
    mu_k = 10.0 ** np.linspace(-14, 0, n)
    arrays = {"K": K, "mu_k": mu_k}
    blocks = {"rock_muon": (STANDARD_ROCK, MUON_MASS), "hydrogen_muon": (ELEMENTS["hydrogen"], MUON_MASS),
              "uranium_muon": (ELEMENTS["uranium"], MUON_MASS), "rock_tau": (STANDARD_ROCK, TAU_MASS)}
    per = {}
    for name, (el, mass) in blocks.items():
        fCM, scr, fs, il = oracle.coulomb_data(K, el, mass, use_ref=True)
        G1 = oracle.coulomb_transport(scr, fs, 1.0, use_ref=True)
        Gk = oracle.coulomb_transport(scr, fs, mu_k, use_ref=True)
        mu0, lbh = oracle.hard_scattering(G1, fCM, scr, il, fs, use_ref=True)
        soft = oracle.soft_scattering(K, el, mass, use_ref=True)
        per[name] = (G1, fCM, scr, il, fs)
        for k, v in (("fCM", fCM), ("screening", scr), ("fspin", fs), ("invlambda", il), ("G_mu1", G1),
                     ("G_muk", Gk), ("mu0", mu0), ("lb_h", lbh), ("soft", soft)):
            arrays[f"{name}_{k}"] = v
        arrays[f"{name}_element"] = np.array([el.A, el.I, el.Z, mass], dtype=np.float64)
    mix = ("rock_muon", "hydrogen_muon", "uranium_muon")
    stacked = [np.stack([per[m][j] for m in mix]) for j in range(5)]
    mu0, lbh = oracle.hard_scattering(*stacked, use_ref=True)
    arrays["mixture_mu0"] = mu0
    arrays["mixture_lb_h"] = lbh
    np.savez_compressed(os.path.join(HERE, "coulomb_golden.npz"), **arrays)
    print({k: v.shape for k, v in arrays.items()})


def main() -> None:
    if oracle.load_ref() is None:
        raise SystemExit("oracle/_ref/libnoa_ref.so is missing: run `make -C oracle ref` first")
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which in ("coulomb", "all"):
        coulomb_main()
    if which not in ("dcs", "all"):
        return
    arrays = {}
    for name, el in ELEMENTS.items():
        b = block(el, MUON_MASS, seed=11)
        for k, v in b.items():
            arrays[f"{name}_muon_{k}"] = v
        arrays[f"{name}_element"] = np.array([el.A, el.I, el.Z], dtype=np.float64)
    b = block(STANDARD_ROCK, TAU_MASS, seed=12)
    for k, v in b.items():
        arrays[f"rock_tau_{k}"] = v
    # DEL / CEL recoil integrals (dcs::vmap_integral(dcs::recoil_integral(...)),
    # dcs.hh:89-130, 955-1001): xlow = X_FRACTION, min_points 180 as in
    # test/unit/test-dcs-calc.cc:22-132, plus 7 and 61; K over the whole
    # range and around the ionisation analytic switch
    m1 = MUON_MASS - ELECTRON_MASS
    kthr = 0.5 * m1 * m1 / ELECTRON_MASS
    K = np.concatenate([10.0 ** np.linspace(-3, 6, 97), kthr * (1 + np.array([-1e-9, 0.0, 1e-9]))])
    for name in ("rock", "hydrogen"):
        el = ELEMENTS[name]
        for mp in (180, 7, 61):
            out = np.stack([[oracle.ref_recoil(p, cel, K, 0.05, el, MUON_MASS, mp) for p in range(4)]
                            for cel in (0, 1)])
            arrays[f"recoil_{name}_mp{mp}"] = out
    arrays["recoil_K"] = K
    np.savez_compressed(os.path.join(HERE, "dcs_golden.npz"), **arrays)
    print({k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
