"""An event read back from a binary event file drives the GPU finders to the same result as the
event it was written from, and that result is the oracle's on the read-back event (best hit and clone
engine, seed by seed with the decision-margin explanations of tests/_find_common.py)."""
import numpy as np
import pytest

import mkfit_amd as mk
from _find_common import assert_ce_parity, assert_find_parity, best_hit_reference, clone_engine_reference
from mkfit_amd import event_file as ef

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _find(ev, descs, dev):
    import torch
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], dev))
    b = mk.FindBatch(ev, len(descs), 3, device=dev, chi2_cut=15.0)
    ctx.find_tracks_best_hit(eoh, b)
    c = mk.CEBatch(ev, len(descs), 3, device=dev, chi2_cut=30.0, max_cands=6)
    ctx.find_tracks_clone_engine(eoh, c)
    torch.cuda.synchronize()
    return b.numpy(), c.numpy()


def test_file_event_finds_the_same_tracks(gpu_device, tmp_path):
    geo = mk.geometry(1)
    descs = mk.barrel_layer_descs(geo)
    evs = [mk.generate_find_event(geo, mk.default_toy_params(seed=s, simulate_material=0), 2000, noise_per_cm2=2.0)
           for s in (5, 6)]
    path = tmp_path / "events.bin"
    with ef.EventFileWriter(path, geo.n_barrel, extra_sections=ef.SEEDS) as w:
        for ev in evs:
            w.write_event(ef.find_event_to_file_event(ev))
    with ef.EventFileReader(path) as f:
        back = [ef.file_event_to_find_event(e) for e in f]
    assert len(back) == 2
    for ev, bk in zip(evs, back):
        (b0, c0), (b1, c1) = _find(ev, descs, gpu_device), _find(bk, descs, gpu_device)
        for k in b0:
            assert np.array_equal(b0[k], b1[k]), k
        for k in c0:
            assert np.array_equal(c0[k], c1[k]), k
        cpu, spread = best_hit_reference(bk, descs, 3, chi2_cut=15.0, bfield=3.8, n_threads=8)
        assert_find_parity("file event, best hit", b1, cpu, spread, 3)
        cpu, spread = clone_engine_reference(bk, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=8)
        assert_ce_parity("file event, clone engine", c1, cpu, spread, 3)
