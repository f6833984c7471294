"""DEL / CEL recoil integrals on the GPU (row f1) against the oracle, which is
bit-identical to the reference build on the golden fixture."""
import numpy as np
import pytest
import torch

import oracle
from noa_amd import _lib, dcs
from noa_amd.physics import ELECTRON_MASS, MUON_MASS, STANDARD_ROCK, TAU_MASS, AtomicElement
from tests.parity import compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("el,mass", [(STANDARD_ROCK, MUON_MASS), (AtomicElement(1.00794, 19.2e-9, 1), MUON_MASS),
                                     (AtomicElement(207.2, 823e-9, 82), TAU_MASS)], ids=["rock", "H", "Pb-tau"])
@pytest.mark.parametrize("min_points", [180, 7])
def test_recoil_integrals_parity(device, el, mass, min_points):
    m1 = mass - ELECTRON_MASS
    kthr = 0.5 * m1 * m1 / ELECTRON_MASS
    Kh = np.concatenate([10.0 ** np.linspace(-3, 6, 301), kthr * (1 + np.array([-1e-9, 0.0, 1e-9]))])
    K = torch.from_numpy(Kh).to(device)
    dels = tuple(torch.full_like(K, np.nan) for _ in range(4))
    cels = tuple(torch.full_like(K, np.nan) for _ in range(4))
    dcs.recoil_integrals(K, 0.05, el, mass, min_points, dels=dels, cels=cels)
    for cel, outs in ((0, dels), (1, cels)):
        for p in range(4):
            exp = oracle.recoil(p, cel, Kh, 0.05, el, mass, min_points, threads=0)
            r = compare(outs[p].cpu().numpy(), exp)
            assert r["finite"] and r["zero_mismatch"] == 0 and r["max_rel"] <= 1e-10, (cel, p, r)


def test_reference_call_shape_and_chunking(device):
    # more K values than one pass (65536) exercises the chunk loop
    Kh = 10.0 ** np.random.default_rng(2).uniform(-3, 6, 70000)
    K = torch.from_numpy(Kh).to(device)
    res = torch.empty_like(K)
    dcs.vmap_integral(dcs.recoil_integral(dcs.pair_production, dcs.del_integrand))(
        res, K, 0.05, STANDARD_ROCK, MUON_MASS, 180)
    exp = oracle.recoil(1, 0, Kh, 0.05, STANDARD_ROCK, MUON_MASS, 180, threads=0)
    r = compare(res.cpu().numpy(), exp)
    assert r["zero_mismatch"] == 0 and r["max_rel"] <= 1e-10, r
    res2 = torch.empty_like(K)
    dcs.vmap_integral(dcs.recoil_integral(dcs.ionisation, dcs.cel_integrand))(
        res2, K, 0.05, STANDARD_ROCK, MUON_MASS, 180)
    exp2 = oracle.recoil(3, 1, Kh, 0.05, STANDARD_ROCK, MUON_MASS, 180, threads=0)
    r = compare(res2.cpu().numpy(), exp2)
    assert r["zero_mismatch"] == 0 and r["max_rel"] <= 1e-10, r
    with pytest.raises(_lib.NoaDcsError, match="xlow"):
        dcs.recoil_integrals(K, 1.5, STANDARD_ROCK, MUON_MASS, dels=(res, None, None, None))
