"""GPU parity of standard combinatorial finding (mkf_find_tracks_standard) against the CPU oracle.

As for the clone engine (test_gpu_ce.py): per seed the candidate lists (count, order, hit index per layer of
every candidate) must match the oracle's exactly, except for seeds where the oracle itself took a decision
within float32 rounding (a hit at a window or bin edge, a chi2 at the cut, two scores tied), each explained
(tests/_find_common.py::explain_clone_engine); candidates of equal seeds agree to the contract's tolerances
seed by seed (oracle/parity.py).
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from _find_common import assert_ce_parity, clone_engine_reference

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _event(seed, n, noise, pt_min=0.5):
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=seed, simulate_material=0, pt_min=pt_min, pt_max=10.0)
    return mk.generate_find_event(geo, p, n, noise_per_cm2=noise), mk.barrel_layer_descs(geo)


def _run(ev, descs, dev, K, first_layer=3, chi2_cut=30.0, prop_to_hit=True, intra_pflags=0, layer_base=None,
         n_layers=None):
    import torch
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], dev))
    b = mk.CEBatch(ev, n_layers or len(descs), first_layer, device=dev, chi2_cut=chi2_cut, max_cands=K,
                   prop_to_hit=prop_to_hit, intra_pflags=intra_pflags, layer_base=layer_base)
    ctx.find_tracks_standard(eoh, b)
    torch.cuda.synchronize()
    return b.numpy()


def _compare(g, ev, descs, name, first_layer=3, **kw):
    c, spread = clone_engine_reference(ev, descs, first_layer, n_threads=8, standard=True, **kw)
    assert_ce_parity(name, g, c, spread, first_layer)
    return c


@pytest.mark.parametrize("K,noise", [(6, 2.0), (1, 2.0), (8, 3.0), (4, 8.0)])
def test_standard_matches_oracle(gpu_device, K, noise):
    ev, descs = _event(200 + K, 4000, noise)
    g = _run(ev, descs, gpu_device, K)
    _compare(g, ev, descs, "std_K%d" % K, chi2_cut=30.0, max_cands=K)


@pytest.mark.parametrize("K,prop,intra", [(6, False, 0), (6, True, 1)])
def test_standard_propagation_settings(gpu_device, K, prop, intra):
    """Without the propagation to each hit (the toy geometry's setting, Geoms/CylCowWLids.cc:225), and with
    the parametrised field on the propagation to the hit only (intra-layer flags)."""
    ev, descs = _event(240 + intra, 3000, 2.0)
    g = _run(ev, descs, gpu_device, K, prop_to_hit=prop, intra_pflags=intra)
    _compare(g, ev, descs, "std_prop%d_intra%d" % (prop, intra), chi2_cut=30.0, max_cands=K, prop_to_hit=prop,
             intra_pflags=intra)


def test_standard_short_layers_and_disks(gpu_device):
    """Barrel layers cut to |z| <= r / 2 + 2 cm (window edges, candidates leaving through the layer ends, held
    back and stopped replicas), then an endcap-disk event."""
    ev, descs = _event(251, 3000, 2.0)
    descs = [dict(d, zmin=-(0.5 * d["propagate_to"] + 2), zmax=0.5 * d["propagate_to"] + 2) if i >= 3 else d
             for i, d in enumerate(descs)]
    g = _run(ev, descs, gpu_device, 6)
    _compare(g, ev, descs, "std_short_barrel", chi2_cut=30.0, max_cands=6)
    geo = mk.geometry(4)
    p = mk.default_toy_params(seed=16, simulate_material=0, eta_max=2.5, pt_min=1.0)
    ev = mk.generate_find_event(geo, p, 4000, noise_per_cm2=1.0)
    descs = mk.disk_layer_descs(geo)
    g = _run(ev, descs, gpu_device, 6)
    _compare(g, ev, descs, "std_disks", chi2_cut=30.0, max_cands=6)


def test_standard_differs_from_clone_engine_as_the_oracle_does(gpu_device):
    """The GPU's two algorithms differ where the oracle's do: same best-candidate hits on almost every seed,
    states apart (the clone engine updates from the previous layer, the standard from the layer)."""
    import torch
    ev, descs = _event(261, 3000, 2.0)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
    out = []
    for f in (ctx.find_tracks_clone_engine, ctx.find_tracks_standard):
        b = mk.CEBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=30.0, max_cands=6)
        f(eoh, b)
        torch.cuda.synchronize()
        out.append(b.numpy())
    same = np.all(out[0]["hot_idx"][0] == out[1]["hot_idx"][0], axis=0)
    assert same.mean() > 0.98
    assert not np.array_equal(out[0]["par"][0][:, same], out[1]["par"][0][:, same])
