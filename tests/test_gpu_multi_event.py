"""Several events in one event of hits and one finder launch (batch field layer_base): every event's
results equal those of finding it alone, and the oracle's."""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from _find_common import assert_find_parity, best_hit_reference

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.fixture(scope="module")
def events():
    geo = mk.geometry(1)
    descs = mk.barrel_layer_descs(geo)
    evs = [mk.generate_find_event(geo, mk.default_toy_params(seed=40 + e, simulate_material=0), 1500 + 700 * e,
                                  noise_per_cm2=1.0 + e) for e in range(3)]
    return evs, descs


def _slice(out, k, lo, hi):
    return out[k][..., lo:hi]


def test_best_hit_events_in_one_launch(gpu_device, events):
    import torch
    evs, descs = events
    L = len(descs)
    cat = mk.concat_find_events(evs, L)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs * len(evs))
    eoh.load(cat["hit_offset"], _t(cat["hits_flat"], gpu_device))
    b = mk.FindBatch(cat, L, 3, device=gpu_device, chi2_cut=15.0, layer_base=cat["layer_base"])
    ctx.find_tracks_best_hit(eoh, b)
    torch.cuda.synchronize()
    got = b.numpy()
    off = cat["event_seed_offset"]
    for e, ev in enumerate(evs):
        one = mk.EventOfHits(ctx, descs)
        one.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
        b1 = mk.FindBatch(ev, L, 3, device=gpu_device, chi2_cut=15.0)
        ctx.find_tracks_best_hit(one, b1)
        torch.cuda.synchronize()
        ref = b1.numpy()
        for k in ref:
            assert np.array_equal(_slice(got, k, off[e], off[e + 1]), ref[k]), (e, k)
        cpu, chk = best_hit_reference(ev, descs, 3, chi2_cut=15.0, bfield=3.8, n_threads=8)
        assert_find_parity("event %d" % e, ref, cpu, chk, 3)


def test_clone_engine_events_in_one_launch(gpu_device, events):
    import torch
    evs, descs = events
    L = len(descs)
    cat = mk.concat_find_events(evs, L)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs * len(evs))
    eoh.load(cat["hit_offset"], _t(cat["hits_flat"], gpu_device))
    b = mk.CEBatch(cat, L, 3, device=gpu_device, chi2_cut=30.0, max_cands=6, layer_base=cat["layer_base"])
    ctx.find_tracks_clone_engine(eoh, b)
    torch.cuda.synchronize()
    got = b.numpy()
    off = cat["event_seed_offset"]
    for e, ev in enumerate(evs):
        one = mk.EventOfHits(ctx, descs)
        one.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
        b1 = mk.CEBatch(ev, L, 3, device=gpu_device, chi2_cut=30.0, max_cands=6)
        ctx.find_tracks_clone_engine(one, b1)
        torch.cuda.synchronize()
        ref = b1.numpy()
        for k in ref:
            assert np.array_equal(_slice(got, k, off[e], off[e + 1]), ref[k]), (e, k)


def _first_layers(ev, n_layers):
    """The event restricted to its first n_layers layers (the seeds' hits lie in the first three)."""
    out = dict(ev)
    n = int(ev["hit_offset"][n_layers])
    out["hit_offset"] = ev["hit_offset"][:n_layers + 1]
    out["hits_flat"] = ev["hits_flat"][:n]
    out["hit_mc"] = ev["hit_mc"][:n]
    return out


def test_clone_engine_mixed_layer_kinds_interleaved(gpu_device):
    """A barrel event and an endcap-disk event in one launch, their seeds interleaved so that every
    wavefront holds candidates of both: a lane that evaluates another candidate's window hits must use
    that candidate's layer (kind and hit offset).  Each event's results equal those of finding it alone."""
    import torch
    gb, gd = mk.geometry(1), mk.geometry(4)
    dd = mk.disk_layer_descs(gd)
    L = len(dd)
    db = mk.barrel_layer_descs(gb)[:L]
    eb = _first_layers(mk.generate_find_event(gb, mk.default_toy_params(seed=61, simulate_material=0), 1500,
                                              noise_per_cm2=2.0), L)
    ed = _first_layers(mk.generate_find_event(gd, mk.default_toy_params(seed=62, simulate_material=0, eta_max=2.5,
                                                                        pt_min=1.0), 4000, noise_per_cm2=1.0), L)
    evs, descs = [eb, ed], [db, dd]
    cat = mk.concat_find_events(evs, L)
    perm = np.random.default_rng(7).permutation(cat["n_seeds"])    # new position p holds seed perm[p]
    for k in ("seed_par", "seed_err", "seed_hits"):
        cat[k] = np.ascontiguousarray(cat[k][..., perm])
    for k in ("seed_chg", "seed_mc", "layer_base"):
        cat[k] = np.ascontiguousarray(cat[k][perm])
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, db + dd)
    eoh.load(cat["hit_offset"], _t(cat["hits_flat"], gpu_device))
    b = mk.CEBatch(cat, L, 3, device=gpu_device, chi2_cut=30.0, max_cands=6, layer_base=cat["layer_base"])
    ctx.find_tracks_clone_engine(eoh, b)
    torch.cuda.synchronize()
    got = b.numpy()
    off = cat["event_seed_offset"]
    where = np.argsort(perm)
    for e, ev in enumerate(evs):
        one = mk.EventOfHits(ctx, descs[e])
        one.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
        b1 = mk.CEBatch(ev, L, 3, device=gpu_device, chi2_cut=30.0, max_cands=6)
        ctx.find_tracks_clone_engine(one, b1)
        torch.cuda.synchronize()
        ref = b1.numpy()
        assert ev["n_seeds"] > 100 and (ref["n_found"][0] > 3).mean() > 0.5, e
        for k in ref:
            assert np.array_equal(got[k][..., where[off[e]:off[e + 1]]], ref[k]), (e, k)


def test_layer_base_out_of_range_leaves_seeds_untouched(gpu_device, events):
    import torch
    evs, descs = events
    L = len(descs)
    cat = mk.concat_find_events(evs[:2], L)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs * 2)
    eoh.load(cat["hit_offset"], _t(cat["hits_flat"], gpu_device))
    base = cat["layer_base"].copy()
    base[::5] = 2 * L          # one past the last event
    base[1::7] = -L
    b = mk.FindBatch(cat, L, 3, device=gpu_device, layer_base=base)
    hot0 = b.hot_idx.clone()
    ctx.find_tracks_best_hit(eoh, b)
    torch.cuda.synchronize()
    bad = (base < 0) | (base > L)
    g = b.numpy()
    assert np.array_equal(g["par"][:, bad], cat["seed_par"][:, bad])
    assert np.array_equal(g["hot_idx"][:, bad], hot0.cpu().numpy()[:, bad])
    assert (g["hot_idx"][3:, ~bad] >= 0).mean() > 0.5          # the others were searched
    with pytest.raises(mk.MkfError, match="layer_base"):
        ctx.find_tracks_best_hit(eoh, mk.FindBatch(cat, L, 3, device=gpu_device))


def test_clone_engine_layer_base_out_of_range_gives_no_candidates(gpu_device, events):
    import torch
    evs, descs = events
    L = len(descs)
    cat = mk.concat_find_events(evs[:2], L)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs * 2)
    eoh.load(cat["hit_offset"], _t(cat["hits_flat"], gpu_device))
    base = cat["layer_base"].copy()
    base[::5] = 2 * L
    base[1::7] = -L
    b = mk.CEBatch(cat, L, 3, device=gpu_device, max_cands=6, layer_base=base)
    b.n_cands.fill_(77)
    b.hot_idx.fill_(123)
    b.score.fill_(5.0)
    b.n_found.fill_(9)
    ctx.find_tracks_clone_engine(eoh, b)
    torch.cuda.synchronize()
    bad = (base < 0) | (base > L)
    g = b.numpy()
    assert np.all(g["n_cands"][bad] == 0)
    assert np.all(g["hot_idx"][:, :, bad] == -9)
    assert np.all(g["score"][:, bad] == np.float32(-1e16)) and np.all(g["n_found"][:, bad] == 0)
    assert np.all(g["n_cands"][~bad] >= 1)
