"""GPU parity of the fit with the final propagation to the PCA (MKF_FIT_PROP_TO_PCA)."""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from _common import par_rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_pca_parity(gpu_device, form):
    import torch
    geo = mk.geometry(1)
    rl, xi = mk.material_grid(geo)
    ev = mk.generate(geo, mk.default_toy_params(seed=8), 30000, geo.n_barrel)
    ctx = mk.Context(bfield=3.8, rl_grid=rl, xi_grid=xi)
    ff = mk.FIT_CHI2 | mk.FIT_PROP_TO_PCA
    b = mk.FitBatch.from_numpy(ev, gpu_device, pflags=mk.PF_APPLY_MATERIAL, fit_flags=ff, out_of_place=True,
                               pca_pflags=0)
    (ctx.fit_plain if form == "plain" else ctx.fit_batched)(b)
    torch.cuda.synchronize()
    O = oracle.Oracle(bfield=3.8, rl_grid=rl, xi_grid=xi)
    par, err, chg, chi2 = O.fit(ev, pflags=2, fit_flags=1 | 8, n_threads=8, pca_pflags=0)
    pr = par_rel(b.par.cpu().numpy(), par)
    assert np.mean(pr.max(axis=0) <= 1e-5) > 0.98, np.mean(pr.max(axis=0) <= 1e-5)
    assert np.percentile(pr, 99.9) < 1e-3
