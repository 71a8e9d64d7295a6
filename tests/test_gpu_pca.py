"""GPU parity of the fit with the final propagation to the PCA (MKF_FIT_PROP_TO_PCA)."""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from _common import TOL_COV, TOL_PAR, assert_parity, par_rel

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
    p64, e64, _, _ = O.fit(ev, pflags=2, fit_flags=1 | 8, dtype=np.float64, n_threads=8, pca_pflags=0)
    # The step to the PCA extrapolates the fitted state over up to a metre, so FP32 rounding of the
    # angles (~1e-7 rad) moves the position there by ~1e-5 cm: the FP32 oracle itself is that far from
    # FP64.  The GPU must be as close to FP32 and FP64 as the FP32 oracle is to FP64.
    floor = float(np.mean(par_rel(par, p64) <= TOL_PAR))
    assert_parity("pca", b.par.cpu().numpy(), par, p64, "par", TOL_PAR, frac=floor - 0.05)
    assert_parity("pca", b.err.cpu().numpy(), err, e64, "cov", TOL_COV, frac=0.95)
