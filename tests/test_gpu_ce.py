"""GPU parity of clone-engine finding (mkf_find_tracks_clone_engine) against the CPU oracle.

Per seed the candidate lists (count, order, hit index per layer of every candidate) must match the
oracle's exactly, except for seeds where the oracle itself took a decision within float32 rounding (a
hit at a window or bin edge, a chi2 at the cut, two scores tied): every differing seed must be explained
by one (tests/_find_common.py::explain_clone_engine).  Candidates of equal seeds must agree to the
contract's tolerances, seed by seed (oracle/parity.py).
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from _common import RawDeviceBuffer
from _find_common import assert_ce_parity, clone_engine_reference, pad_seed_hits

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _event(seed, n, noise, pt_min=0.5):
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=seed, simulate_material=0, pt_min=pt_min, pt_max=10.0)
    return mk.generate_find_event(geo, p, n, noise_per_cm2=noise), mk.barrel_layer_descs(geo)


def _run(ev, descs, dev, K, first_layer=3, chi2_cut=30.0, prop_to_hit=True):
    import torch
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    hits = _t(ev["hits_flat"], dev) if len(ev["hits_flat"]) else torch.zeros((1, 9), device=dev)
    eoh.load(ev["hit_offset"], hits)
    b = mk.CEBatch(ev, len(descs), first_layer, device=dev, chi2_cut=chi2_cut, max_cands=K, prop_to_hit=prop_to_hit)
    ctx.find_tracks_clone_engine(eoh, b)
    torch.cuda.synchronize()
    return b.numpy()


def _dump(name, **arrays):
    """Keeps the arrays of a failed comparison when MKF_TEST_DUMP names a directory."""
    import os
    d = os.environ.get("MKF_TEST_DUMP")
    if d:
        os.makedirs(d, exist_ok=True)
        np.savez_compressed(os.path.join(d, name + ".npz"), **arrays)


def _compare(g, ev, descs, name="ce", first_layer=3, **kw):
    """The GPU's candidates against the oracle's, seed by seed (assert_ce_parity); returns the oracle's."""
    c, spread = clone_engine_reference(ev, descs, first_layer, n_threads=8, **kw)
    try:
        assert_ce_parity(name, g, c, spread, first_layer)
    except AssertionError:
        _dump(name, **{"g_" + k: v for k, v in g.items()}, **{"c_" + k: v for k, v in c.items()})
        raise
    return c


def _true_fraction(hot, ev, first_layer):
    h = hot[first_layer:]; off = ev["hit_offset"][first_layer:-1]
    ok = h >= 0
    truth = np.where(ok, ev["hit_mc"][off[:, None] + np.maximum(h, 0)], -5)
    return (truth == ev["seed_mc"][None, :]).sum() / h.size


# K = 1 and the 8/cm^2 noise give windows with more accepted hits than K: the kernel keeps
# only each candidate's best K records, which must not change the seed's selection.
@pytest.mark.parametrize("K,noise", [(6, 2.0), (1, 2.0), (8, 3.0), (4, 8.0)])
def test_ce_matches_oracle(gpu_device, K, noise):
    ev, descs = _event(100 + K, 4000, noise)
    g = _run(ev, descs, gpu_device, K)
    c = _compare(g, ev, descs, name="ce_K%d" % K, chi2_cut=30.0, max_cands=K)
    assert abs(_true_fraction(g["hot_idx"][0], ev, 3) - _true_fraction(c["hot_idx"][0], ev, 3)) < 3e-3


@pytest.mark.parametrize("K", [1, 6])
def test_ce_without_propagation_to_hit(gpu_device, K):
    """The toy geometry's setting (finding_requires_propagation_to_hit_pos = false, Geoms/CylCowWLids.cc:225):
    chi2 and the update use the state at the layer, not one propagated to each hit."""
    ev, descs = _event(140 + K, 3000, 2.0)
    g = _run(ev, descs, gpu_device, K, prop_to_hit=False)
    _compare(g, ev, descs, name="ce_noprop_K%d" % K, chi2_cut=30.0, max_cands=K, prop_to_hit=False)


def test_ce_noise_free(gpu_device):
    ev, descs = _event(7, 3000, 0.0, pt_min=1.0)
    g = _run(ev, descs, gpu_device, 6)
    _compare(g, ev, descs, name="ce_noise_free", chi2_cut=30.0, max_cands=6)
    assert _true_fraction(g["hot_idx"][0], ev, 3) > 0.985


def test_ce_on_endcap_disks(gpu_device):
    import torch
    geo = mk.geometry(4)
    p = mk.default_toy_params(seed=15, simulate_material=0, eta_max=2.5, pt_min=1.0)
    ev = mk.generate_find_event(geo, p, 4000, noise_per_cm2=1.0)
    descs = mk.disk_layer_descs(geo)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
    b = mk.CEBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=30.0, max_cands=6)
    ctx.find_tracks_clone_engine(eoh, b)
    torch.cuda.synchronize()
    _compare(b.numpy(), ev, descs, name="ce_disks", chi2_cut=30.0, max_cands=6)


def test_ce_edge_cases(gpu_device):
    ev, descs = _event(8, 500, 2.0)
    L = len(descs)
    evL = pad_seed_hits(ev, L)
    g = _run(evL, descs, gpu_device, 6, first_layer=L)         # nothing to search: the seed itself
    c = oracle.find_clone_engine(evL, descs, L, chi2_cut=30.0, max_cands=6, n_threads=4)
    assert np.array_equal(g["n_cands"], np.ones(ev["n_seeds"], np.int32))
    assert np.array_equal(g["hot_idx"], c["hot_idx"])
    off = ev["hit_offset"]
    keep = np.ones(off[-1], bool); keep[off[5]:off[6]] = False
    cnt = np.diff(off).copy(); cnt[5] = 0
    ev2 = dict(ev, hits_flat=ev["hits_flat"][keep], hit_offset=np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32),
               hit_mc=ev["hit_mc"][keep])
    g = _run(ev2, descs, gpu_device, 6)
    _compare(g, ev2, descs, name="ce_empty_layer", chi2_cut=30.0, max_cands=6)
    ev3 = dict(ev, hits_flat=np.zeros((0, 9), np.float32), hit_offset=np.zeros(L + 1, np.int32),
               hit_mc=np.zeros(0, np.int32))
    g = _run(ev3, descs, gpu_device, 6)
    c = oracle.find_clone_engine(ev3, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=4)
    for k in ("n_cands", "hot_idx", "n_found"):
        assert np.array_equal(g[k], c[k]), k


def test_ce_rejects_bad_arguments(gpu_device):
    import torch
    ev, descs = _event(9, 50, 1.0)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
    b = mk.CEBatch(ev, len(descs), 3, device=gpu_device, max_cands=6)
    b.K = 9
    with pytest.raises(mk.MkfError, match="max_cands"):
        ctx.find_tracks_clone_engine(eoh, b)
    b = mk.CEBatch(ev, len(descs), 3, device=gpu_device, max_cands=6)
    b.first_layer = 0
    with pytest.raises(mk.MkfError, match="first_layer"):
        ctx.find_tracks_clone_engine(eoh, b)
    b = mk.CEBatch(ev, len(descs), 3, device=gpu_device, max_cands=6)
    with RawDeviceBuffer(6 * 2 * b.n * 4) as short:               # room for 2 of the 17 layers
        b.hot_idx = short
        with pytest.raises(mk.MkfError, match="hot_idx"):
            ctx.find_tracks_clone_engine(eoh, b)


def test_every_find_shape_gives_the_same_candidates(gpu_device):
    """The finders' launch shapes (6 or 8 candidate slots per seed in a wavefront, registers for one or two
    waves per SIMD) give the same candidates bit for bit, for the clone engine and standard finding; the
    autotuner picks one of them and keeps it; a shape with fewer slots than max_cands is refused."""
    import torch
    ev, descs = _event(33, 3000, 2.0)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))

    def run(standard, K=6):
        b = mk.CEBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=30.0, max_cands=K)
        (ctx.find_tracks_standard if standard else ctx.find_tracks_clone_engine)(eoh, b)
        torch.cuda.synchronize()
        return b.numpy()

    def same(a, b):
        assert np.array_equal(a["n_cands"], b["n_cands"])
        for k in ("hot_idx", "score", "n_found"):
            assert np.array_equal(a[k], b[k]), k
        K = a["par"].shape[0]
        valid = np.arange(K)[:, None] < a["n_cands"][None, :]
        for k in ("chg", "chi2"):
            assert np.array_equal(a[k][valid], b[k][valid]), k
        for k in ("par", "err"):
            assert np.array_equal(np.moveaxis(a[k], 1, 2)[valid], np.moveaxis(b[k], 1, 2)[valid]), k

    for standard in (False, True):
        assert ctx.get_find_shape(standard) == (0, 2)
        ref = run(standard)
        for slots in (6, 8):
            for waves in (2, 1):
                ctx.set_find_shape(standard, slots, waves)
                same(run(standard), ref)
        b = mk.CEBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=30.0, max_cands=6)
        shape, ms = ctx.autotune_find(eoh, b, standard=standard)
        torch.cuda.synchronize()
        assert shape in {(6, 2), (6, 1), (8, 2), (8, 1)} and ms > 0
        assert ctx.get_find_shape(standard) == shape
        same(b.numpy(), ref)
        same(run(standard), ref)
    ctx.set_find_shape(False, 6, 2)
    with pytest.raises(mk.MkfError):
        run(False, K=8)
    ctx.close()


@pytest.mark.parametrize("standard", [False, True])
def test_dead_regions_match_oracle(gpu_device, standard):
    """Dead modules (mkf_event_of_hits_set_dead_regions): 60 regions of 0.15 rad x 16 cm on the finding layers,
    some wrapping across phi = +-pi, with a fifth of the hits masked so that windows in them often hold no hit:
    the candidates (with their -7 records) match the oracle run with the same regions, seed by seed; clearing
    the regions gives the run without them back bit for bit."""
    import torch
    ev, descs = _event(34, 3000, 2.0)
    rng = np.random.default_rng(34)
    n = 60
    phi1 = rng.uniform(-np.pi, np.pi - 0.15, n)
    phi1[:6] = rng.uniform(np.pi - 0.1, np.pi, 6)                 # wrapping across +-pi
    phi2 = phi1 + 0.15
    phi2[:6] -= 2 * np.pi
    zc = rng.uniform(-60, 60, n)
    regions = np.stack([rng.integers(3, len(descs), n), phi1, phi2, zc - 8, zc + 8], 1).astype(np.float32)
    total = int(ev["hit_offset"][-1])
    mask = (rng.random(total) < 0.2).astype(np.uint8)
    mask[:ev["hit_offset"][3]] = 0
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
    eoh.set_hit_mask(_t(mask, gpu_device))

    def run():
        b = mk.CEBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=30.0, max_cands=6)
        (ctx.find_tracks_standard if standard else ctx.find_tracks_clone_engine)(eoh, b)
        torch.cuda.synchronize()
        return b.numpy()
    without = run()
    eoh.set_dead_regions(regions)
    g = run()
    assert (g["hot_idx"] == -7).sum() >= 10, (g["hot_idx"] == -7).sum()
    evd = dict(ev, hit_mask=mask, dead_regions=regions)
    cpu, spread = clone_engine_reference(evd, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=8, standard=standard)
    assert (cpu["hot_idx"] == -7).sum() >= 10
    assert_ce_parity("dead regions %s" % ("standard" if standard else "clone engine"), g, cpu, spread, 3)
    eoh.set_dead_regions(None)
    again = run()
    for k in ("n_cands", "hot_idx", "score"):
        assert np.array_equal(again[k], without[k]), k
    ctx.close()


@pytest.mark.parametrize("standard", [False, True])
def test_strip_charge_matches_oracle(gpu_device, standard):
    """Strip cluster charge (mkf_event_of_hits_set_hit_charge): a fifth of the hits overflow (pass), a tenth
    carry code 0 (never pass), the rest codes 2..79 (1628..2244 per cm, near the cut for the path-length factor
    at the hit): the candidates match the oracle run with the same codes, seed by seed; clearing the codes
    gives the run without them back bit for bit."""
    import torch
    ev, descs = _event(35, 3000, 2.0)
    rng = np.random.default_rng(35)
    total = int(ev["hit_offset"][-1])
    code = rng.integers(2, 80, total).astype(np.uint8)
    u = rng.random(total)
    code[u < 0.2] = 0xff
    code[(u >= 0.2) & (u < 0.3)] = 0
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))

    def run():
        b = mk.CEBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=30.0, max_cands=6)
        (ctx.find_tracks_standard if standard else ctx.find_tracks_clone_engine)(eoh, b)
        torch.cuda.synchronize()
        return b.numpy()
    without = run()
    eoh.set_hit_charge(_t(code, gpu_device))
    g = run()
    assert not np.array_equal(g["hot_idx"], without["hot_idx"])
    cpu, spread = clone_engine_reference(dict(ev, hit_charge=code), descs, 3, chi2_cut=30.0, max_cands=6,
                                         n_threads=8, standard=standard)
    assert_ce_parity("strip charge %s" % ("standard" if standard else "clone engine"), g, cpu, spread, 3)
    eoh.set_hit_charge(None)
    again = run()
    for k in ("n_cands", "hot_idx", "score"):
        assert np.array_equal(again[k], without[k]), k
    ctx.close()
