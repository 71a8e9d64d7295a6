"""GPU parity of best-hit finding: the HIP LayerOfHits build, SelectHitIndices and the fused
FindTracksBestHit loop (through the C ABI) against the CPU oracle on the same toy events.

Integer results (hit ranks, bin ranges, candidate lists, chosen hit indices) must agree exactly
except where a float32 rounding difference decides a comparison: a hit on a bin or window edge, or
two candidates whose chi2 agree to rounding.  Those cases are counted and bounded explicitly.
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from _common import RawDeviceBuffer, par_rel
from _find_common import (F_PHI_FINE, _brute_select, _keys, _random_hits, assert_bins, assert_find_parity,
                          best_hit_reference, edge_explained, pad_seed_hits)

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _event(seed, n_tracks, noise, pt_min=0.5):
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=seed, simulate_material=0, pt_min=pt_min, pt_max=10.0)
    ev = mk.generate_find_event(geo, p, n_tracks, noise_per_cm2=noise)
    return geo, ev, mk.barrel_layer_descs(geo)


@pytest.fixture(scope="module")
def setup(gpu_device):
    geo, ev, descs = _event(123, 5000, 3.0)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    hits = _t(ev["hits_flat"], gpu_device)
    eoh.load(ev["hit_offset"], hits)
    return ev, descs, ctx, eoh


def _fine_boundary(phi):
    """phi within a few ulps of a fine phi-bin edge (one ulp of phi moves the scaled value by ~4e-5)."""
    x = F_PHI_FINE * (phi + np.float32(np.pi))
    return np.abs(x - np.round(x)) < 3e-4


def test_loh_build_matches_oracle(setup):
    ev, descs, ctx, eoh = setup
    off = ev["hit_offset"]
    flips = 0
    for l in range(len(descs)):
        hl = ev["hits_flat"][off[l]:off[l + 1]]
        r_c, b_c, p_c, q_c, nq_c = oracle.loh_build(descs[l], hl)
        r_g, b_g, p_g, q_g, nq_g = eoh.layer_views(l)
        assert nq_g == nq_c and len(r_g) == len(hl)
        assert sorted(r_g.tolist()) == list(range(len(hl)))
        # per original hit: phi to an ulp (device vs host atan2f), q exact
        phi_c = np.empty_like(p_c); phi_c[r_c] = p_c
        phi_g = np.empty_like(p_g); phi_g[r_g] = p_g
        q_o = np.empty_like(q_g); q_o[r_g] = q_g
        assert np.array_equal(q_o, hl[:, 2])
        # device and host atan2f agree to a few ulp (both within their documented error bounds)
        dphi = np.abs(phi_g - phi_c)
        dphi = np.minimum(dphi, np.abs(dphi - np.float32(2 * np.pi)))
        assert np.all(dphi <= 4 * np.spacing(np.maximum(np.abs(phi_c), np.float32(1e-6)))), dphi.max()
        k_c, _ = _keys(descs[l], phi_c, hl[:, 2])
        k_g, _ = _keys(descs[l], phi_g, hl[:, 2])
        moved = k_c != k_g
        assert np.all(_fine_boundary(phi_c[moved])), "a key differs away from a bin edge"
        flips += int(moved.sum())
        if not moved.any():
            assert np.array_equal(r_g, r_c), l
            assert_bins(b_g, b_c, l)
        else:
            assert np.array_equal(np.argsort(k_g, kind="stable"), r_g)
    # a hit lands within a few ulp of one of the 1024 fine-bin edges with probability ~1e-4, and the two
    # atan2f implementations then disagree on the bin about a third of the time
    assert flips <= max(5, 1e-4 * off[-1]), (flips, off[-1])


def _crowded_event(descs, seed=5):
    """Hits of four barrel layers piled into few (q, fine phi) bins, in shuffled order: 10,000 copies of one
    hit (beyond the LDS sort), 3,000 of another (sorted in LDS), 150 positions x 40 copies (one thread's
    share exceeded many times), and hits at phi = +pi exactly (fine bin 1024, sorting last in its q row but
    filling coarse bin 0) and -pi, each layer with random hits among them."""
    rng = np.random.default_rng(seed)

    def hit(r, phi, z):
        return [r * np.cos(phi), r * np.sin(phi), z, 1e-4, 0.0, 1e-4, 0.0, 0.0, 1e-2]

    def rand(d, n):
        return [hit(d["propagate_to"], rng.uniform(-np.pi, np.pi), rng.uniform(d["zmin"], d["zmax"]))
                for _ in range(n)]
    layers = []
    d = descs[0]
    layers.append([hit(d["propagate_to"], 0.3, 1.0)] * 10000 + rand(d, 500))
    d = descs[1]
    layers.append([hit(d["propagate_to"], -2.0, -3.0)] * 3000 + rand(d, 200))
    d = descs[2]
    pos = [(rng.uniform(-np.pi, np.pi), rng.uniform(d["zmin"], d["zmax"])) for _ in range(150)]
    layers.append([hit(d["propagate_to"], p, z) for p, z in pos for _ in range(40)] + rand(d, 100))
    d = descs[3]
    r = d["propagate_to"]
    plus = [[-r, 0.0, 2.0, 1e-4, 0.0, 1e-4, 0.0, 0.0, 1e-2]] * 50          # atan2(+0, -r) = +pi
    minus = [[-r, -0.0, 2.0, 1e-4, 0.0, 1e-4, 0.0, 0.0, 1e-2]] * 50        # atan2(-0, -r) = -pi
    near = [hit(r, -np.pi + 1e-3, 2.0)] * 30                                # coarse bin 0 of the same row
    layers.append(plus + minus + near + rand(d, 100))
    out = []
    for h in layers:
        h = np.array(h, np.float32)
        out.append(h[rng.permutation(len(h))])
    return out


def test_loh_crowded_bins(gpu_device):
    """The LayerOfHits build keeps the reference's stable order and bin ranges however the hits crowd."""
    geo = mk.geometry(1)
    descs = mk.barrel_layer_descs(geo)[:4]
    layers = _crowded_event(descs)
    off = np.concatenate([[0], np.cumsum([len(h) for h in layers])]).astype(np.int32)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(off, _t(np.concatenate(layers), gpu_device))
    for l, hl in enumerate(layers):
        r_c, b_c, p_c, q_c, nq_c = oracle.loh_build(descs[l], hl)
        r_g, b_g, p_g, q_g, nq_g = eoh.layer_views(l)
        phi_c = np.empty_like(p_c); phi_c[r_c] = p_c
        phi_g = np.empty_like(p_g); phi_g[r_g] = p_g
        k_c, _ = _keys(descs[l], phi_c, hl[:, 2])
        k_g, _ = _keys(descs[l], phi_g, hl[:, 2])
        assert np.array_equal(np.argsort(k_g, kind="stable"), r_g), l
        if np.array_equal(k_c, k_g):
            assert np.array_equal(r_g, r_c), l
            assert_bins(b_g, b_c, l)
        else:
            assert np.all(_fine_boundary(phi_c[k_c != k_g])), l
    # fine bin 1024: the +pi hits own coarse bin 0 of their q row, after the row's other hits
    r_g, b_g, p_g, _, _ = eoh.layer_views(3)
    assert np.sum(p_g == np.float32(np.pi)) == 50
    ctx.close()


def test_loh_reload_smaller_event(setup, gpu_device):
    ev, descs, ctx, eoh = setup
    geo, ev2, _ = _event(7, 300, 0.5)
    e2 = mk.EventOfHits(ctx, descs)
    e2.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))     # large first, then the small one
    e2.load(ev2["hit_offset"], _t(ev2["hits_flat"], gpu_device))
    off = ev2["hit_offset"]
    for l in (0, 7, 16):
        hl = ev2["hits_flat"][off[l]:off[l + 1]]
        r_c, b_c, p_c, _, _ = oracle.loh_build(descs[l], hl)
        r_g, b_g, p_g, _, _ = e2.layer_views(l)
        # the layer holds exactly the small event's hits, ordered by the device's own keys
        assert sorted(r_g.tolist()) == list(range(len(hl)))
        phi_g = np.empty_like(p_g); phi_g[r_g] = p_g
        k_g, _ = _keys(descs[l], phi_g, hl[:, 2])
        assert np.array_equal(np.argsort(k_g, kind="stable"), r_g)
        phi_c = np.empty_like(p_c); phi_c[r_c] = p_c
        k_c, _ = _keys(descs[l], phi_c, hl[:, 2])
        if np.array_equal(k_g, k_c):            # no fine-bin flip: identical to the oracle
            assert_bins(b_g, b_c)
            assert np.array_equal(r_g, r_c)


def _random_states(rng, n, r, zspan):
    par = np.zeros((6, n), np.float32); err = np.zeros((21, n), np.float32)
    phi = rng.uniform(-np.pi, np.pi, n).astype(np.float32)
    par[0] = r * np.cos(phi); par[1] = r * np.sin(phi); par[2] = rng.uniform(-zspan, zspan, n)
    par[3] = rng.uniform(0.1, 2, n); par[4] = phi; par[5] = rng.uniform(0.3, 2.8, n)
    s = rng.uniform(1e-4, 0.3, n).astype(np.float32) ** 2
    err[0] = s; err[2] = s; err[1] = rng.uniform(-0.5, 0.5, n) * s; err[5] = rng.uniform(1e-4, 0.8, n) ** 2
    return par, err


@pytest.mark.parametrize("layer", [3, 9, 16])
def test_select_hit_indices_matches_oracle(setup, gpu_device, layer):
    import torch
    ev, descs, ctx, eoh = setup
    d = descs[layer]
    off = ev["hit_offset"]
    hl = ev["hits_flat"][off[layer]:off[layer + 1]]
    rng = np.random.default_rng(layer)
    n = 4000
    par, err = _random_states(rng, n, d["propagate_to"], d["zmax"] + 1.5)
    xs = torch.zeros(n, dtype=torch.int32, device=gpu_device)
    xa = torch.zeros((16, n), dtype=torch.int32, device=gpu_device)
    w = torch.zeros(n, dtype=torch.int32, device=gpu_device)
    eoh.select_hit_indices(layer, _t(par, gpu_device), _t(err, gpu_device), xs, xa, w)
    xs, xa, w = xs.cpu().numpy(), xa.cpu().numpy(), w.cpu().numpy()
    xs_c, xa_c, w_c = oracle.select_hit_indices(d, hl, par, err)
    wsr_diff = np.nonzero(w != w_c)[0]
    assert len(wsr_diff) <= 2
    ranks, _, phis, qs, _ = oracle.loh_build(d, hl)
    same = (xs == xs_c) & np.all(xa == xa_c, axis=0)
    bad = np.nonzero(~same & (w == w_c))[0]
    if len(bad):
        _, wins = _brute_select(d, hl, par[:, bad], err[:, bad], ranks, phis, qs)
        for k, t in enumerate(bad):
            diff = np.setxor1d(xa[:max(xs[t], 0), t], xa_c[:max(xs_c[t], 0), t])
            assert edge_explained(wins[k], par[:, t], hl, diff), (t, xa[:, t], xa_c[:, t])
    assert len(bad) <= n // 500
    assert (xs == 16).sum() > 0 and (xs == -1).sum() > 0 and (xs == 0).sum() > 0


def _run_find(ctx, eoh, ev, n_layers, first_layer, dev, chi2_cut=15.0):
    import torch
    b = mk.FindBatch(ev, n_layers, first_layer, device=dev, chi2_cut=chi2_cut)
    ctx.find_tracks_best_hit(eoh, b)
    torch.cuda.synchronize()
    return b, b.numpy()


def _compare_find(gpu, ev, descs, first_layer=3, **kw):
    """The GPU's best-hit result against the oracle's, seed by seed (assert_find_parity); returns the oracle's."""
    cpu, spread = best_hit_reference(ev, descs, first_layer, n_threads=8, bfield=3.8, **kw)
    assert_find_parity("best hit", gpu, cpu, spread, first_layer)
    return cpu


def _true_fraction(res, ev, first_layer):
    hot = res["hot_idx"][first_layer:]
    off = ev["hit_offset"][first_layer:-1]
    ok = hot >= 0
    truth = np.where(ok, ev["hit_mc"][off[:, None] + np.maximum(hot, 0)], -5)
    return (truth == ev["seed_mc"][None, :]).sum() / hot.size


def test_find_best_hit_matches_oracle(setup, gpu_device):
    ev, descs, ctx, eoh = setup
    _, gpu = _run_find(ctx, eoh, ev, len(descs), 3, gpu_device)
    cpu = _compare_find(gpu, ev, descs, chi2_cut=15.0)
    assert abs(_true_fraction(gpu, ev, 3) - _true_fraction(cpu, ev, 3)) < 2e-3
    assert _true_fraction(gpu, ev, 3) > 0.8


def test_find_best_hit_propagate_to_hit(setup, gpu_device):
    import torch
    ev, descs, ctx, eoh = setup
    b = mk.FindBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=15.0, prop_to_hit=True)
    ctx.find_tracks_best_hit(eoh, b)
    torch.cuda.synchronize()
    _compare_find(b.numpy(), ev, descs, chi2_cut=15.0, prop_to_hit=True)


def test_find_noise_free_exact_path(gpu_device):
    geo, ev, descs = _event(31, 3000, 0.0, pt_min=1.0)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
    _, gpu = _run_find(ctx, eoh, ev, len(descs), 3, gpu_device)
    _compare_find(gpu, ev, descs)
    assert _true_fraction(gpu, ev, 3) > 0.97


def test_find_repeatable(setup, gpu_device):
    ev, descs, ctx, eoh = setup
    b, r1 = _run_find(ctx, eoh, ev, len(descs), 3, gpu_device)
    b.reset()
    ctx.find_tracks_best_hit(eoh, b)
    r2 = b.numpy()
    for k in r1:
        assert np.array_equal(r1[k], r2[k]), k


def test_find_edge_cases(setup, gpu_device):
    import torch
    ev, descs, ctx, eoh = setup
    L = len(descs)
    # nothing to search: the seeds come back unchanged
    b, g = _run_find(ctx, eoh, pad_seed_hits(ev, L), L, L, gpu_device)
    assert np.array_equal(g["par"], ev["seed_par"]) and np.array_equal(g["hot_idx"][:3], ev["seed_hits"][:3])
    assert np.all(g["hot_idx"][3:] == -1)
    assert np.array_equal(g["n_found"], np.sum(ev["seed_hits"] >= 0, axis=0))
    # no seeds
    e0 = dict(ev, n_seeds=0)
    b0 = mk.FindBatch(e0, L, 3, device=gpu_device)
    ctx.find_tracks_best_hit(eoh, b0)
    # empty layers 5 and 6, then an event without hits at all
    off = ev["hit_offset"]
    keep = np.ones(off[-1], bool); keep[off[5]:off[7]] = False
    cnt = np.diff(off).copy(); cnt[5] = cnt[6] = 0
    ev2 = dict(ev, hits_flat=ev["hits_flat"][keep], hit_offset=np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32),
               hit_mc=ev["hit_mc"][keep])
    for evx in (ev2, dict(ev, hits_flat=np.zeros((0, 9), np.float32), hit_offset=np.zeros(L + 1, np.int32),
                          hit_mc=np.zeros(0, np.int32))):
        ex = mk.EventOfHits(ctx, descs)
        ex.load(evx["hit_offset"], _t(evx["hits_flat"], gpu_device) if len(evx["hits_flat"]) else
                torch.zeros((1, 9), device=gpu_device))
        _, g = _run_find(ctx, ex, evx, L, 3, gpu_device)
        _compare_find(g, evx, descs)
        assert np.all(g["hot_idx"][5:7] < 0)


def test_find_rejects_bad_arguments(setup, gpu_device):
    import torch
    ev, descs, ctx, eoh = setup
    b = mk.FindBatch(ev, len(descs), 3, device=gpu_device)
    b.first_layer = len(descs) + 1
    with pytest.raises(mk.MkfError, match="first_layer"):
        ctx.find_tracks_best_hit(eoh, b)
    b = mk.FindBatch(ev, len(descs), 3, device=gpu_device)
    with RawDeviceBuffer(2 * b.n * 4) as short:                  # too short for 17 layers
        b.hot_idx = short
        with pytest.raises(mk.MkfError, match="hot_idx"):
            ctx.find_tracks_best_hit(eoh, b)
    fresh = mk.EventOfHits(ctx, descs)
    with pytest.raises(mk.MkfError, match="not loaded"):
        ctx.find_tracks_best_hit(fresh, mk.FindBatch(ev, len(descs), 3, device=gpu_device))
    with pytest.raises(mk.MkfError, match="hit_offset"):
        fresh.load(np.array([0, 5, 3] + [3] * (len(descs) - 2), np.int32), torch.zeros((5, 9), device=gpu_device))


def test_find_on_endcap_disks(gpu_device):
    geo = mk.geometry(4)
    p = mk.default_toy_params(seed=14, simulate_material=0, eta_max=2.5, pt_min=1.0)
    ev = mk.generate_find_event(geo, p, 4000, noise_per_cm2=1.0)
    descs = mk.disk_layer_descs(geo)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
    for l in (0, 4, 8):
        r_c, b_c, _, _, _ = oracle.loh_build(descs[l], ev["hits_flat"][ev["hit_offset"][l]:ev["hit_offset"][l + 1]])
        r_g, b_g, _, _, _ = eoh.layer_views(l)
        assert np.mean(r_g == r_c) > 0.999
    _, gpu = _run_find(ctx, eoh, ev, len(descs), 3, gpu_device)
    _compare_find(gpu, ev, descs, chi2_cut=15.0)



def test_hit_mask_matches_oracle(gpu_device):
    """The iteration's hit mask (mkf_event_of_hits_set_hit_mask): best-hit and clone-engine finding with 40 %
    of the finding layers' hits masked match the oracle run with the same mask, seed by seed; clearing the
    mask and a new load give the unmasked result back bit for bit."""
    import torch
    from _find_common import assert_ce_parity, clone_engine_reference
    geo, ev, descs = _event(9, 3000, 2.0)
    rng = np.random.default_rng(9)
    total = int(ev["hit_offset"][-1])
    mask = (rng.random(total) < 0.4).astype(np.uint8)
    mask[:ev["hit_offset"][3]] = 0
    evm = dict(ev, hit_mask=mask)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs)
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))

    def run_bh():
        b = mk.FindBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=15.0)
        ctx.find_tracks_best_hit(eoh, b)
        torch.cuda.synchronize()
        return b.numpy()

    def run_ce():
        c = mk.CEBatch(ev, len(descs), 3, device=gpu_device, chi2_cut=30.0, max_cands=6)
        ctx.find_tracks_clone_engine(eoh, c)
        torch.cuda.synchronize()
        return c.numpy()
    plain_bh, plain_ce = run_bh(), run_ce()
    eoh.set_hit_mask(_t(mask, gpu_device))
    bh, ce = run_bh(), run_ce()
    off = ev["hit_offset"]
    for l in range(3, len(descs)):
        for h in (bh["hot_idx"][l], ce["hot_idx"][:, l]):
            assert not mask[h[h >= 0] + off[l]].any(), l
    _compare_find(bh, evm, descs, chi2_cut=15.0)
    cpu, spread = clone_engine_reference(evm, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=8)
    assert_ce_parity("masked clone engine", ce, cpu, spread, 3)
    eoh.set_hit_mask(None)
    for a, b in ((run_bh(), plain_bh), (run_ce(), plain_ce)):
        for k in a:
            assert np.array_equal(a[k], b[k]), k
    eoh.set_hit_mask(_t(mask, gpu_device))
    eoh.load(ev["hit_offset"], _t(ev["hits_flat"], gpu_device))
    a = run_bh()
    for k in a:
        assert np.array_equal(a[k], plain_bh[k]), k
    ctx.close()
