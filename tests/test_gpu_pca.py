"""GPU parity of the fit with the final propagation to the PCA (MKF_FIT_PROP_TO_PCA)."""
import numpy as np
import pytest

import mkfit_amd as mk
from oracle.parity import fit_reference
from _common import assert_track_parity

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
    # The step to the PCA extrapolates the fitted state over up to a metre, so FP32 rounding of the
    # angles (~1e-7 rad) moves the position there by ~1e-5 cm: most tracks are rounding-sensitive, and
    # the per-track criterion holds each to its own measured spread.
    o, o64, chk = fit_reference(3.8, rl, xi, ev, mk.PF_APPLY_MATERIAL, ff, pca_pflags=0)
    g = (b.par.cpu().numpy(), b.err.cpu().numpy(), b.chi2.cpu().numpy())
    assert_track_parity("pca", g, (o[0], o[1], o[3]), (o64[0], o64[1], o64[3]), chk)
