"""Per-point cosmic-ray ionisation rate (pdrchem_rates_device_zeta).

The reference's rate routine takes one global zeta (calc_reac_rates.F90:24);
the per-point entry point replaces it point by point in the cosmic-ray
(:247), cosmic-ray-photon (:275-276) and cosmic-ray desorption (:334-340)
rates.  Checked against the oracle's rate evaluation run point by point with
that point's zeta (the oracle's rates are pinned to the reference Fortran's
own output in test_oracle.py), and for linearity in zeta of the cosmic-ray
reactions.
"""
import numpy as np
import pytest

import oracle
from pdrchem import workloads


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["reduced", "full"])
def test_per_point_zeta_matches_oracle(name):
    import torch
    from pdrchem import lib
    orc = oracle.Oracle(name)
    n = 48
    P = workloads.random_points(n, seed=314, columns=(name != "reduced"))
    rng = np.random.default_rng(2718)
    zeta = 10.0 ** rng.uniform(-1.0, 3.0, n)          # units of 1.3e-17 s^-1
    dev = torch.device("cuda:0")
    t = lambda v: torch.tensor(np.ascontiguousarray(v), device=dev)
    nid = lib.network_id(name)
    got, neg = lib.rates_device(nid, t(P.T), t(P.Td), t(P.G), t(P.Av), t(P.cols), zeta=t(zeta))
    base, _ = lib.rates_device(nid, t(P.T), t(P.Td), t(P.G), t(P.Av), t(P.cols))
    torch.cuda.synchronize()
    got, base = got.cpu().numpy(), base.cpu().numpy()
    assert int(neg.item()) == 0
    for i in range(n):
        ph = oracle.default_phys(zeta=float(zeta[i]))
        ref = orc.rates(P.T[i:i + 1], P.Td[i:i + 1], P.G[i:i + 1], P.Av[i:i + 1], P.cols[i:i + 1], phys=ph)[0]
        rel = np.where((got[i] == 0) & (ref == 0), 0.0, np.abs(got[i] - ref) / np.maximum(np.abs(ref), 1e-300))
        assert rel.max() < 1e-11, (name, i, rel.max())
    # zeta = NULL keeps the global value; reactions that depend on zeta scale with it
    ref0 = orc.rates(P.T, P.Td, P.G, P.Av, P.cols)
    np.testing.assert_allclose(base, ref0, rtol=1e-11, atol=0)
    z0 = oracle.default_phys().zeta
    dep = np.any(np.abs(got - base) > 1e-12 * np.abs(base), axis=0)
    assert dep.sum() > 0
    ratio = (got[:, dep] / np.where(base[:, dep] == 0, 1.0, base[:, dep]))
    lin = (base[:, dep] > 1e-90) & (got[:, dep] < 1.0) & (base[:, dep] < 1.0)   # clamps k <= 1 aside
    np.testing.assert_allclose(ratio[lin], np.broadcast_to((zeta / z0)[:, None], ratio.shape)[lin], rtol=1e-12)
