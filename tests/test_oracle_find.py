"""Pins of the oracle's best-hit finding (LayerOfHits, SelectHitIndices, AddBestHit, FindTracksBestHit).

The reference holds no golden vectors for finding, so each piece is checked against an independent
formulation:
  * LayerOfHits (HitStructures.cc:77-183): numpy stable argsort of the bin keys, bin ranges by scan;
  * SelectHitIndices (MkFinder.cc:227-684): brute-force window test over all hits of the layer, in the
    (q-bin, phi-bin, sorted position) scan order, with the 16-hit cap;
  * the FindTracksBestHit loop (MkBuilder.cc:1395-1478, MkFinder.cc:691-832): a numpy composition of the
    oracle's propagation / chi2 / update primitives with the stop rules and fake-hit codes written out
    per track, plus physics: on noise-free events nearly every picked hit is the generating track's.
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle

from _find_common import F_PHI_FINE, TWO_PI, _brute_select, _cdist, _keys, _layer, _random_hits, edge_explained


def test_loh_matches_stable_sort():
    rng = np.random.default_rng(11)
    desc = _layer()
    h = _random_hits(rng, 5000, ties=300)
    ranks, bins, phis, qs, nq = oracle.loh_build(desc, h)
    assert sorted(ranks.tolist()) == list(range(len(h)))
    # phi / q of the sorted hits are those of the ranked input hits (Hit::phi, Hit::z)
    np.testing.assert_allclose(phis, np.arctan2(h[ranks, 1], h[ranks, 0]), atol=3e-7)
    assert np.array_equal(qs, h[ranks, 2])
    key, nqb = _keys(desc, phis, qs)
    assert nqb == nq
    # stable sort: the oracle's order equals numpy's stable argsort of the same keys
    key_in = np.empty_like(key)
    key_in[ranks] = key
    assert np.array_equal(np.argsort(key_in, kind="stable"), ranks)
    # bins: each (q, phi) bin's [first, end) holds exactly the sorted hits with that coarse key
    coarse = (key >> 16) * 256 + ((key & 0x3ff) >> 2)
    first, end = bins[:, 0], bins[:, 1]
    for b in np.unique(coarse):
        idx = np.nonzero(coarse == b)[0]
        assert first[b] == idx[0] and end[b] == idx[-1] + 1
    empty = np.setdiff1d(np.arange(nq * 256), coarse)
    assert np.all(first[empty] == end[empty])


def test_loh_q_bins_clamped_and_counted():
    desc = _layer(q_bin=-7.0)   # q_bin < 0: a fixed number of bins over [zmin, zmax]
    _, _, _, _, nq = oracle.loh_build(desc, np.zeros((0, 9), np.float32))
    assert nq == 7
    desc = _layer(zmin=-59.0, zmax=60.0, q_bin=2.0)   # extent 119 -> 60 bins
    _, _, _, _, nq = oracle.loh_build(desc, np.zeros((0, 9), np.float32))
    assert nq == 60
    rng = np.random.default_rng(3)
    h = _random_hits(rng, 500, zmax=90.0)   # hits beyond the layer's z range land in the edge bins
    ranks, bins, phis, qs, nq = oracle.loh_build(desc, h)
    assert bins[:, 1].max() == len(h)


def test_select_hit_indices_matches_brute_force():
    rng = np.random.default_rng(5)
    desc = _layer()
    h = _random_hits(rng, 80000)
    n = 3000
    par = np.zeros((6, n), np.float32); err = np.zeros((21, n), np.float32)
    phi = rng.uniform(-np.pi, np.pi, n).astype(np.float32)
    par[0] = 20 * np.cos(phi); par[1] = 20 * np.sin(phi); par[2] = rng.uniform(-61, 61, n)
    par[3] = rng.uniform(0.1, 2, n); par[4] = phi; par[5] = rng.uniform(0.3, 2.8, n)
    s = rng.uniform(1e-4, 0.3, n).astype(np.float32) ** 2     # position variances: windows at both clamps
    err[0] = s; err[2] = s; err[1] = rng.uniform(-0.5, 0.5, n) * s; err[5] = rng.uniform(1e-4, 0.8, n) ** 2
    xs, xa, wsr = oracle.select_hit_indices(desc, h, par, err)
    ranks, bins, phis, qs, nq = oracle.loh_build(desc, h)
    ref, wins = _brute_select(desc, h, par, err, ranks, phis, qs)
    capped = edge = 0
    for t in range(n):
        if wsr[t] == 2:
            assert xs[t] == -1
            assert abs(par[2, t]) > 60.0
            continue
        got = xa[:xs[t], t]
        exp = ref[t][:16]
        if not np.array_equal(got, exp):
            # only hits on the window edge (float32 rounding of the limits) may differ
            assert edge_explained(wins[t], par[:, t], h, np.setxor1d(got, exp), eps=1e-5), (t, got, exp)
            edge += 1
        capped += len(ref[t]) > 16
    assert capped > 50 and edge <= 3
    assert set(np.unique(wsr)) <= {0, 1, 2}


def test_select_hit_indices_skips_masked_hits():
    """The iteration's hit mask (MkFinder.cc:476-482): a masked hit is passed over before the window test and
    does not count toward the 16-hit cap, so each list is the brute-force list without the masked hits, cut
    at 16; an all-zero mask changes nothing."""
    rng = np.random.default_rng(7)
    desc = _layer()
    h = _random_hits(rng, 40000)
    n = 1500
    par = np.zeros((6, n), np.float32); err = np.zeros((21, n), np.float32)
    phi = rng.uniform(-np.pi, np.pi, n).astype(np.float32)
    par[0] = 20 * np.cos(phi); par[1] = 20 * np.sin(phi); par[2] = rng.uniform(-59, 59, n)
    par[3] = rng.uniform(0.1, 2, n); par[4] = phi; par[5] = rng.uniform(0.3, 2.8, n)
    s = rng.uniform(1e-2, 0.3, n).astype(np.float32) ** 2
    err[0] = s; err[2] = s; err[5] = rng.uniform(1e-2, 0.8, n) ** 2
    mask = (rng.random(len(h)) < 0.4).astype(np.uint8)
    xs, xa, wsr = oracle.select_hit_indices(desc, h, par, err, mask=mask)
    xs0, xa0, _ = oracle.select_hit_indices(desc, h, par, err)
    xz, xaz, _ = oracle.select_hit_indices(desc, h, par, err, mask=np.zeros(len(h), np.uint8))
    assert np.array_equal(xz, xs0) and np.array_equal(xaz, xa0)
    ranks, bins, phis, qs, nq = oracle.loh_build(desc, h)
    ref, wins = _brute_select(desc, h, par, err, ranks, phis, qs)
    capped = edge = 0
    for t in range(n):
        got = xa[:xs[t], t]
        assert not mask[got].any()
        exp = np.array([i for i in ref[t] if not mask[i]][:16], np.int64)
        if not np.array_equal(got, exp):
            assert edge_explained(wins[t], par[:, t], h, np.setxor1d(got, exp), eps=1e-5), (t, got, exp)
            edge += 1
        capped += len(ref[t]) > 16 and xs0[t] == 16
    assert capped > 50 and edge <= 3


def test_finders_never_take_masked_hits():
    """Best-hit and clone-engine finding with a mask over half of every finding layer's hits: no masked hit
    in any candidate, an all-zero mask gives the unmasked result bit for bit, and masking every hit of the
    seeds' own tracks leaves those tracks without hits."""
    geo, ev = make_find_event()
    descs = mk.barrel_layer_descs(geo)
    rng = np.random.default_rng(8)
    total = ev["hit_offset"][-1]
    mask = (rng.random(total) < 0.5).astype(np.uint8)
    mask[:ev["hit_offset"][3]] = 0                   # seed layers
    for run in (lambda e: oracle.find_best_hit(e, descs, 3, chi2_cut=15.0, n_threads=4),
                lambda e: oracle.find_clone_engine(e, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=4)):
        plain = run(ev)
        zero = run(dict(ev, hit_mask=np.zeros(total, np.uint8)))
        for k in plain:
            assert np.array_equal(plain[k], zero[k]), k
        got = run(dict(ev, hit_mask=mask))
        hot = got["hot_idx"]
        off = ev["hit_offset"]
        for l in range(3, len(descs)):
            h = hot[..., l, :] if hot.ndim == 3 else hot[l]
            taken = h[h >= 0] + off[l]
            assert not mask[taken].any(), l
    own = np.isin(ev["hit_mc"], ev["seed_mc"]).astype(np.uint8)
    own[:ev["hit_offset"][3]] = 0
    got = oracle.find_best_hit(dict(ev, hit_mask=own), descs, 3, chi2_cut=15.0, n_threads=4)
    hot = got["hot_idx"][3:]
    off = ev["hit_offset"][3:-1]
    ok = hot >= 0
    truth = np.where(ok, ev["hit_mc"][off[:, None] + np.maximum(hot, 0)], -5)
    assert not (truth == ev["seed_mc"][None, :]).any()


def test_window_edge_counts_the_layer_thickness():
    """SelectHitIndices' z test (MkFinder.cc:310, 336) widens the window's z reach by
    the layer's half thickness seen along the track, 0.5 (rout - rin) / tan(theta), added in quadrature: on a
    2 cm thick layer a track at 45 degrees whose 3-sigma z window ends 0.3 cm inside the layer end is on the
    edge (that reach is 1.0 cm), while the same track crossing at 90 degrees (no reach) is inside, and one
    1.5 cm beyond the end is outside only when the reach does not cover it."""
    desc = _layer(rin=19.0, rout=21.0, zmin=-60.0, zmax=60.0)
    rng = np.random.default_rng(6)
    h = _random_hits(rng, 2000)
    cases = [  # (z, theta, wsr)
        (60.0 - 0.6, np.pi / 4, 1),     # dq 0.3: z + dq inside; with the 1.0 cm reach on the edge
        (60.0 - 0.6, np.pi / 2, 0),     # no reach: inside
        (-60.0 + 0.6, 3 * np.pi / 4, 1),
        (60.0 + 0.9, np.pi / 4, 1),     # beyond the end by more than dq, within the reach: edge
        (60.0 + 0.9, np.pi / 2, 2),     # without the reach: outside
        (60.0 + 1.5, np.pi / 4, 2),     # beyond the reach sqrt(0.3^2 + 1^2): outside
    ]
    n = len(cases)
    par = np.zeros((6, n), np.float32); err = np.zeros((21, n), np.float32)
    for i, (z, th, _) in enumerate(cases):
        par[:, i] = [20.0, 0.0, z, 0.5, 0.0, th]
    err[0] = err[2] = 1e-4
    err[5] = 0.1 ** 2                   # sigma_z 0.1 cm: dq = 0.3 cm
    _, _, wsr = oracle.select_hit_indices(desc, h, par, err)
    assert wsr.tolist() == [c[2] for c in cases], wsr.tolist()


def _find_composed(O, ev, descs, first_layer, chi2_cut):
    """FindTracksBestHit over barrel layers from the oracle's primitives (MkBuilder.cc:1395-1478)."""
    n = ev["n_seeds"]
    par = ev["seed_par"].copy(); err = ev["seed_err"].copy(); chg = ev["seed_chg"].copy()
    chi2 = np.zeros(n, np.float32)
    L = len(descs)
    hot = np.full((L, n), -9, np.int32)
    hot[:first_layer] = ev["seed_hits"][:first_layer]
    nfound = np.sum(hot[:first_layer] >= 0, axis=0).astype(np.int32)
    ninside = np.zeros(n, np.int32); ntail = np.zeros(n, np.int32)
    stopped = np.zeros(n, bool)
    snap = {}
    off = ev["hit_offset"]; hits = ev["hits_flat"]
    for l in range(first_layer, L):
        d = descs[l]
        hl = hits[off[l]:off[l + 1]]
        errP, parP, _ = O.propagate_to_r(err, par, chg, np.full(n, d["propagate_to"], np.float32), pflags=0)
        xs, xa, wsr = oracle.select_hit_indices(d, hl, parP, errP)
        r2 = parP[0] * parP[0] + parP[1] * parP[1]
        for t in range(n):
            if not stopped[t] and r2[t] < np.float32(d["rin"]) * np.float32(d["rin"]):
                stopped[t] = True
                snap[t] = (par[:, t].copy(), err[:, t].copy(), chg[t], chi2[t], hot[:, t].copy(), nfound[t])
            if stopped[t]:
                wsr[t] = 2; xs[t] = 0
        best = np.full(n, -1); minc = np.full(n, np.float32(chi2_cut), np.float32)
        for hc in range(16):
            m = xs > hc
            if not m.any():
                break
            ms_par = np.zeros((3, n), np.float32); ms_err = np.ones((6, n), np.float32)
            idx = xa[hc, m]
            ms_par[:, m] = hl[idx, :3].T; ms_err[:, m] = hl[idx, 3:].T
            _, _, c2 = O.kalman_operation(1, errP, parP, ms_err, ms_par)
            c2 = np.abs(c2)
            better = m & (c2 < minc)
            minc[better] = c2[better]; best[better] = xa[hc, better]
        ms_par = parP[:3].copy(); ms_err = np.tile(np.array([666, 0, 666, 0, 0, 666], np.float32)[:, None], (1, n))
        for t in range(n):
            if wsr[t] != 2 and best[t] >= 0:
                ms_par[:, t] = hl[best[t], :3]; ms_err[:, t] = hl[best[t], 3:]
        err_new, par_new, _ = O.kalman_operation(2, errP, parP, ms_err, ms_par)
        for t in range(n):
            if wsr[t] == 2:
                continue                         # AddBestHit adds no hit for lanes outside the layer
            if best[t] >= 0:
                chi2[t] += minc[t]; hot[l, t] = best[t]; nfound[t] += 1
                ninside[t] += ntail[t]; ntail[t] = 0
            else:
                code = -3 if wsr[t] == 1 else (-2 if ninside[t] + ntail[t] else -1)
                hot[l, t] = code
                ntail[t] += code == -1
        flip = par_new[3] < 0
        chg = np.where(flip, -chg, chg); par_new[3] = np.abs(par_new[3])
        par, err = par_new, err_new
        for t in range(n):
            if not stopped[t] and hot[l, t] == -2:
                stopped[t] = True
                snap[t] = (par[:, t].copy(), err[:, t].copy(), chg[t], chi2[t], hot[:, t].copy(), nfound[t])
    for t, (p, e, c, c2, hh, nf) in snap.items():
        par[:, t] = p; err[:, t] = e; chg[t] = c; chi2[t] = c2; hot[:, t] = hh; nfound[t] = nf
    return dict(par=par, err=err, chg=chg, chi2=chi2, hot_idx=hot, n_found=nfound)


def make_find_event():
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=77, simulate_material=0, pt_min=0.5, pt_max=10.0)
    ev = mk.generate_find_event(geo, p, 400, noise_per_cm2=3.0)
    return geo, ev


@pytest.fixture(scope="module")
def find_event():
    return make_find_event()


def test_find_loop_matches_composition(find_event):
    geo, ev = find_event
    descs = mk.barrel_layer_descs(geo)
    O = oracle.Oracle(bfield=3.8)
    got = oracle.find_best_hit(ev, descs, first_layer=3, chi2_cut=15.0, bfield=3.8, n_threads=2)
    ref = _find_composed(O, ev, descs, 3, 15.0)
    same = np.all(got["hot_idx"] == ref["hot_idx"], axis=0)
    assert same.all(), np.nonzero(~same)[0][:10]
    assert np.array_equal(got["n_found"][same], ref["n_found"][same])
    assert np.array_equal(got["chg"][same], ref["chg"][same])
    # the same pinned primitives in the same order: bit for bit, the fake-hit updates' covariance included
    assert np.array_equal(got["chi2"][same], ref["chi2"][same])
    assert np.array_equal(got["par"][:, same], ref["par"][:, same])
    assert np.array_equal(got["err"][:, same], ref["err"][:, same])
    assert (got["hot_idx"] == -1).sum() >= 10


def test_find_loop_matches_composition_short_barrel(find_event):
    """The same with the finding layers cut to |z| <= r / 2 + 2 cm: tracks near cot(theta) = 1/2 run along
    the layer ends, meeting window edges (-3) and leaving layers through their ends."""
    geo, ev = find_event
    descs = [dict(d, zmin=-(0.5 * d["propagate_to"] + 2), zmax=0.5 * d["propagate_to"] + 2) if i >= 3 else d
             for i, d in enumerate(mk.barrel_layer_descs(geo))]
    got = oracle.find_best_hit(ev, descs, first_layer=3, chi2_cut=15.0, bfield=3.8, n_threads=2)
    ref = _find_composed(oracle.Oracle(bfield=3.8), ev, descs, 3, 15.0)
    assert np.array_equal(got["hot_idx"], ref["hot_idx"])
    assert (got["hot_idx"] == -3).sum() >= 1 and (got["hot_idx"][3:] == -9).sum() >= 500
    assert np.array_equal(got["n_found"], ref["n_found"])
    assert np.array_equal(got["par"], ref["par"]) and np.array_equal(got["err"], ref["err"])


def test_find_invariants_and_threads(find_event):
    geo, ev = find_event
    descs = mk.barrel_layer_descs(geo)
    a = oracle.find_best_hit(ev, descs, first_layer=3, bfield=3.8, n_threads=1)
    b = oracle.find_best_hit(ev, descs, first_layer=3, bfield=3.8, n_threads=5)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    hot = a["hot_idx"]; nlh = np.diff(ev["hit_offset"])
    assert np.all((hot >= -3) | (hot == -9))
    assert np.all(hot < nlh[:, None])
    assert np.array_equal(a["n_found"], np.sum(hot >= 0, axis=0))
    assert np.all(a["chi2"] >= 0)
    # after a -2 (too many holes) or a -9 (stopped) nothing more is recorded
    for t in range(hot.shape[1]):
        col = hot[3:, t]
        stop = np.nonzero((col == -2) | (col == -9))[0]
        if len(stop):
            assert np.all(col[stop[0] + 1:] == -9)


def test_find_noise_free_picks_true_hits():
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=9, simulate_material=0, pt_min=1.0, pt_max=10.0)
    ev = mk.generate_find_event(geo, p, 2000, noise_per_cm2=0.0)
    descs = mk.barrel_layer_descs(geo)
    r = oracle.find_best_hit(ev, descs, first_layer=3, bfield=3.8, n_threads=4)
    hot = r["hot_idx"][3:]; off = ev["hit_offset"][3:-1]
    ok = hot >= 0
    truth = np.where(ok, ev["hit_mc"][off[:, None] + np.maximum(hot, 0)], -5)
    correct = truth == ev["seed_mc"][None, :]
    assert correct.sum() / ok.sum() > 0.995
    assert ok.sum() / hot.size > 0.97


def test_find_with_noise_efficiency():
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=10, simulate_material=0, pt_min=1.0, pt_max=10.0)
    ev = mk.generate_find_event(geo, p, 2000, noise_per_cm2=3.0)
    descs = mk.barrel_layer_descs(geo)
    r = oracle.find_best_hit(ev, descs, first_layer=3, bfield=3.8, n_threads=4)
    hot = r["hot_idx"][3:]; off = ev["hit_offset"][3:-1]
    ok = hot >= 0
    truth = np.where(ok, ev["hit_mc"][off[:, None] + np.maximum(hot, 0)], -5)
    assert (truth == ev["seed_mc"][None, :]).sum() / hot.size > 0.8   # 0.856 at 3 noise hits / cm2


def test_find_propagate_to_hit_option():
    """finding_requires_propagation_to_hit_pos: chi2 and update after propagating to each hit.  The toy
    hits lie on the layer radius where the state already is, so the choice changes little; with the
    propagation the chosen hits stay the generating track's."""
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=12, simulate_material=0, pt_min=1.0, pt_max=10.0)
    ev = mk.generate_find_event(geo, p, 1500, noise_per_cm2=2.0)
    descs = mk.barrel_layer_descs(geo)
    a = oracle.find_best_hit(ev, descs, first_layer=3, bfield=3.8, n_threads=4)
    b = oracle.find_best_hit(ev, descs, first_layer=3, bfield=3.8, n_threads=4, prop_to_hit=True)
    same = np.all(a["hot_idx"] == b["hot_idx"], axis=0).mean()
    assert 0.9 < same <= 1.0
    assert not np.array_equal(a["par"], b["par"])             # the extra propagation does act
    for r in (a, b):
        hot = r["hot_idx"][3:]; off = ev["hit_offset"][3:-1]
        ok = hot >= 0
        truth = np.where(ok, ev["hit_mc"][off[:, None] + np.maximum(hot, 0)], -5)
        assert (truth == ev["seed_mc"][None, :]).sum() / hot.size > 0.9


def test_find_on_endcap_disks():
    """Disk-only geometry: endcap layers (q = r bins, propagation to z, the endcap update)."""
    geo = mk.geometry(4)
    p = mk.default_toy_params(seed=13, simulate_material=0, eta_max=2.5, pt_min=1.0)
    ev = mk.generate_find_event(geo, p, 3000, noise_per_cm2=1.0)
    descs = mk.disk_layer_descs(geo)
    assert ev["n_seeds"] > 500
    for r in (oracle.find_best_hit(ev, descs, 3, chi2_cut=15.0, n_threads=4),
              {"hot_idx": oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, n_threads=4)["hot_idx"][0]}):
        hot = r["hot_idx"][3:]; off = ev["hit_offset"][3:-1]
        ok = hot >= 0
        truth = np.where(ok, ev["hit_mc"][off[:, None] + np.maximum(hot, 0)], -5)
        assert (truth == ev["seed_mc"][None, :]).sum() / ok.sum() > 0.9
        assert ok.sum() / hot.size > 0.5
