"""GPU parity of the backward fit of found candidates (mkf_backward_fit) against the CPU oracle.

The candidates are the oracle clone engine's best candidate per seed, handed to the GPU and to the oracle
alike, so the comparison isolates the backward fit.  Every candidate is compared on its own
(oracle/parity.py): unflagged within 1e-5 (parameters) and 1e-4 (covariance, chi2) of the FP32 oracle,
flagged ones (rounding alone moves the oracle by more than a third of that, or a threshold tie) within their
bound.  Charges must agree where the parameters are unflagged; scores must follow the chi2.  A last test runs
the GPU clone engine and the GPU backward fit back to back on device arrays.
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from _common import assert_track_parity
from oracle.parity import bkfit_reference, flag_mask
import test_oracle_bkfit as obk

pytestmark = pytest.mark.gpu

PF_PARAM_B, PF_MATERIAL = 1, 2


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _gpu(ev, descs, cand, dev, layer_base=None, rl=None, xi=None, **kw):
    import torch
    ctx = mk.Context(bfield=3.8, rl_grid=rl, xi_grid=xi)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], dev))
    b = mk.BkFitBatch(cand, device=dev, layer_base=layer_base, **kw)
    ctx.backward_fit(eoh, b)
    torch.cuda.synchronize()
    return b.numpy()


def _check(name, g, ev, descs, cand, layer_base=None, rl=None, xi=None, **kw):
    okw = dict(kw)
    okw["include_pca"] = okw.pop("include_pca", True)
    c32, c64, chk = bkfit_reference(ev, descs, cand, layer_base=layer_base, rl_grid=rl, xi_grid=xi, bfield=3.8,
                                    n_threads=8, **okw)
    r = assert_track_parity(name, (g["par"], g["err"], g["chi2"]), (c32[0], c32[1], c32[3]),
                            (c64[0], c64[1], c64[3]), chk)
    unflagged = ~flag_mask(chk)[0]
    assert np.array_equal(g["chg"][unflagged], c32[2][unflagged]), name
    # getScoreCand: the same counts and pT class on both sides, so the scores differ by the chi2 difference
    pt_g, pt_c = np.abs(1 / g["par"][3]), np.abs(1 / c32[0][3])
    same_class = ((pt_g < 0.9) == (pt_c < 0.9)) & np.isfinite(c32[3]) & np.isfinite(g["chi2"])
    ds = np.abs((g["score"] + g["chi2"]) - (c32[4] + c32[3]))[same_class]
    assert ds.max() <= 1e-4 * max(1.0, np.abs(c32[4]).max()), (name, ds.max())
    return r


@pytest.fixture(scope="module")
def barrel():
    return obk._barrel(3000, noise=2.0, seed=71)


@pytest.mark.parametrize("group", [1, 16, 64])
def test_backward_fit_matches_oracle(gpu_device, barrel, group):
    geo, ev, descs, cand = barrel
    kw = dict(group=group, pflags=PF_PARAM_B, pca_pflags=PF_PARAM_B)
    g = _gpu(ev, descs, cand, gpu_device, **kw)
    _check("bkfit_g%d" % group, g, ev, descs, cand, **kw)


def test_backward_fit_without_pca_and_from_layer_3(gpu_device, barrel):
    geo, ev, descs, cand = barrel
    kw = dict(group=16, pflags=PF_PARAM_B, pca_pflags=PF_PARAM_B, include_pca=False, last_layer=3)
    g = _gpu(ev, descs, cand, gpu_device, **kw)
    _check("bkfit_nopca_last3", g, ev, descs, cand, **kw)


@pytest.mark.parametrize("fake_material", [False, True])
def test_backward_fit_with_material(gpu_device, barrel, fake_material):
    """Material on every propagation to a hit; fake hits take it only in the best-hit driver's form."""
    geo, ev, descs, cand = barrel
    rl, xi = mk.material_grid(geo)
    kw = dict(group=16, pflags=PF_PARAM_B | PF_MATERIAL, pca_pflags=PF_PARAM_B, fake_material=fake_material)
    g = _gpu(ev, descs, cand, gpu_device, rl=rl, xi=xi, **kw)
    _check("bkfit_material_fake%d" % fake_material, g, ev, descs, cand, rl=rl, xi=xi, **kw)


def test_backward_fit_mixed_events(gpu_device):
    """A barrel and a disk event in one event of hits, candidates interleaved: blocks of 16 hold both layer
    kinds, each lane fits on its own event's layers (layer_base); candidates whose plan would run past the
    last layer are left untouched."""
    cat, descs, cand, base, L = obk._mixed()
    base = base.copy()
    base[::11] = len(descs)
    kw = dict(group=16, pflags=PF_PARAM_B, pca_pflags=PF_PARAM_B)
    g = _gpu(cat, descs, cand, gpu_device, layer_base=base, **kw)
    bad = base == len(descs)
    for k in ("par", "err", "chg", "score"):
        assert np.array_equal(g[k][..., bad], np.asarray(cand[k])[..., bad]), k
    _check("bkfit_mixed", g, cat, descs, cand, layer_base=base, **kw)


def test_clone_engine_then_backward_fit_on_device(gpu_device):
    """The GPU clone engine's slot 0 fed to the GPU backward fit without leaving the device, against the
    oracle's clone engine and backward fit, on the seeds whose best candidates carry the same hits."""
    import torch
    geo, ev, descs, _ = obk._barrel(2000, noise=2.0, seed=83)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
    ce = mk.CEBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=30.0, max_cands=4)
    ctx.find_tracks_clone_engine(eoh, ce)
    cand_dev = {"par": ce.par[0], "err": ce.err[0], "chg": ce.chg[0], "hot_idx": ce.hot_idx[0], "score": ce.score[0]}
    b = mk.BkFitBatch(cand_dev, device=gpu_device, group=16, pflags=PF_PARAM_B, pca_pflags=PF_PARAM_B)
    ctx.backward_fit(eoh, b)
    torch.cuda.synchronize()
    g = b.numpy()
    gce = ce.numpy()
    cce = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=4, n_threads=8)
    same = (gce["n_cands"] > 0) & (cce["n_cands"] > 0) & np.all(gce["hot_idx"][0] == cce["hot_idx"][0], axis=0)
    assert same.mean() > 0.95, same.mean()
    cand = {k: np.ascontiguousarray(cce[k][0]) for k in ("par", "err", "chg", "hot_idx", "score")}
    c = oracle.backward_fit(ev, descs, cand, group=16, pflags=PF_PARAM_B, pca_pflags=PF_PARAM_B, bfield=3.8,
                            n_threads=8)
    # the inputs already differ by the clone engine's own rounding: hold the fitted states to the input
    # differences' scale, and the hit counts exactly
    from oracle.parity import par_rel
    d_in = par_rel(gce["par"][0][:, same], cce["par"][0][:, same]).max(axis=0)
    d_out = par_rel(g["par"][:, same], c["par"][:, same]).max(axis=0)
    assert np.quantile(d_out, 0.99) < max(1e-4, 10 * np.quantile(d_in, 0.99)), (np.quantile(d_out, 0.99),
                                                                               np.quantile(d_in, 0.99))
    assert np.median(np.abs(g["chi2"][same] - c["chi2"][same]) / np.maximum(1, c["chi2"][same])) < 1e-4
