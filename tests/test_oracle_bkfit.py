"""Pins of the oracle's backward fit of found candidates (MkBuilder::BackwardFit, MkFinder::BkFit*).

The restatement (oracle/oracle_bkfit.cc) walks each candidate's HitOnTrack nodes from the last, over the
layer plan in reverse, in blocks of `group` candidates.  It is checked against
  * compositions through the separately pinned fit (orc_fit, tests/pins.py): each candidate becomes a fit
    track whose seed is its state with the covariance scaled by 100 and whose slots are, in reverse layer
    order, every layer on which some candidate of its block has a hit -- its own hit there, else a missing
    slot (the fake hit {666, 0, 666, 0, 0, 666} at the track position, without material);
  * a composition through the oracle's single steps (propagate, Kalman update) for the variant whose fake
    hits take material (BkFitFitTracksBH), which the fit has no slot kind for;
  * known answers of the candidate score after the fit (getScoreCand over the HitOnTrack counters);
  * properties: fake hits after a candidate's first hit change nothing, and noise-free candidates fit with
    chi2 per degree of freedom near one.
Events: toy barrel and disk events, candidates from the oracle's clone engine (slot 0 of each seed).
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle

F = np.float32
HIT_ENDCAP, HIT_MISSING = 1, 2
PF_PARAM_B, PF_MATERIAL = 1, 2


def _ce_cands(ev, descs, K=4, chi2_cut=30.0):
    ce = oracle.find_clone_engine(ev, descs, 3, chi2_cut=chi2_cut, max_cands=K, n_threads=4)
    keep = ce["n_cands"] > 0
    return {k: np.ascontiguousarray(ce[k][0][..., keep]) for k in ("par", "err", "chg", "hot_idx", "score")}


def _barrel(n=400, noise=2.0, seed=61):
    geo = mk.geometry(1)
    ev = mk.generate_find_event(geo, mk.default_toy_params(seed=seed, simulate_material=0, pt_min=0.5, pt_max=10.0),
                                n, noise_per_cm2=noise)
    descs = mk.barrel_layer_descs(geo)
    return geo, ev, descs, _ce_cands(ev, descs)


def _disk(n=1500, noise=1.0, seed=62):
    geo = mk.geometry(4)
    ev = mk.generate_find_event(geo, mk.default_toy_params(seed=seed, simulate_material=0, eta_max=2.5, pt_min=1.0),
                                n, noise_per_cm2=noise)
    descs = mk.disk_layer_descs(geo)
    return geo, ev, descs, _ce_cands(ev, descs)


@pytest.fixture(scope="module")
def barrel():
    return _barrel()


def _mixed():
    """A barrel and a disk event in one event of hits, their candidates interleaved (layer_base per
    candidate), so that blocks hold candidates of both layer kinds."""
    gb, eb, db, cb = _barrel(300)
    gd, ed, dd, cd = _disk(1200)
    L = min(len(db), len(dd))
    evs = []
    for ev in (eb, ed):
        evs.append(dict(ev, hit_offset=ev["hit_offset"][:L + 1], hits_flat=ev["hits_flat"][:ev["hit_offset"][L]],
                        hit_mc=ev["hit_mc"][:ev["hit_offset"][L]]))
    cat = mk.concat_find_events(evs, L)
    cands = [{k: (v[:L] if k == "hot_idx" else v) for k, v in c.items()} for c in (cb, cd)]
    n0, n1 = cands[0]["chg"].size, cands[1]["chg"].size
    cand = {k: np.ascontiguousarray(np.concatenate([c[k] for c in cands], axis=-1)) for k in cands[0]}
    base = np.concatenate([np.zeros(n0, np.int32), np.full(n1, L, np.int32)])
    perm = np.random.default_rng(5).permutation(n0 + n1)
    cand = {k: np.ascontiguousarray(v[..., perm]) for k, v in cand.items()}
    return cat, db[:L] + dd[:L], cand, base[perm], L


@pytest.fixture(scope="module")
def mixed():
    return _mixed()


def _real(ev, descs, hot, base):
    """[L][n] bool: the candidate has a hit of the layer there (an index within the layer's hits)."""
    nl, n = hot.shape
    off = np.asarray(ev["hit_offset"], np.int64)
    lay = base[None, :] + np.arange(nl)[:, None]
    nh = off[lay + 1] - off[lay]
    return (hot >= 0) & (hot < nh)


def _fit_composition(ev, descs, cand, group, pflags=PF_PARAM_B, pca_pflags=PF_PARAM_B, include_pca=True,
                     last_layer=0, layer_base=None, rl=None, xi=None):
    """The backward fit as a forward fit of the reversed hit lists through orc_fit."""
    hot = np.asarray(cand["hot_idx"]); nl, n = hot.shape
    base = np.zeros(n, np.int64) if layer_base is None else np.asarray(layer_base, np.int64)
    real = _real(ev, descs, hot, base)
    real[:last_layer] = False
    off = np.asarray(ev["hit_offset"], np.int64); hits = np.asarray(ev["hits_flat"], F)
    slots = []
    for g0 in range(0, n, group):
        ls = [l for l in range(nl - 1, last_layer - 1, -1) if real[l, g0:g0 + group].any()]
        slots.append(ls)
    S = max(1, max(len(s) for s in slots))
    hp = np.zeros((S, 3, n), F); he = np.zeros((S, 6, n), F); fl = np.zeros((S, n), np.uint8)
    nh = np.zeros(n, np.int32)
    for gi, ls in enumerate(slots):
        for t in range(gi * group, min(n, gi * group + group)):
            nh[t] = len(ls)
            for s, l in enumerate(ls):
                lay = base[t] + l
                f = 0 if descs[lay]["is_barrel"] else HIT_ENDCAP
                if real[l, t]:
                    h = hits[off[lay] + hot[l, t]]
                    hp[s, :, t] = h[:3]; he[s, :, t] = h[3:]
                else:
                    f |= HIT_MISSING
                fl[s, t] = f
    fev = dict(seed_par=np.array(cand["par"], F), seed_err=np.array(cand["err"], F) * F(100), chg=cand["chg"],
               hit_pos=hp, hit_err=he, n_hits=nh, hit_flags=fl)
    O = oracle.Oracle(3.8, rl, xi)
    ff = oracle.FIT_CHI2 | oracle.FIT_CHARGE_FLIP | (oracle.FIT_PROP_TO_PCA if include_pca else 0)
    par, err, chg, chi2 = O.fit(fev, pflags=pflags, fit_flags=ff, pca_pflags=pca_pflags, n_threads=4)
    return dict(par=par, err=err, chg=chg, chi2=chi2)


def _step_composition(ev, descs, cand, group, pflags, fake_material, rl=None, xi=None):
    """The backward fit through the oracle's single steps, layer by layer over all candidates (barrel only)."""
    hot = np.asarray(cand["hot_idx"]); nl, n = hot.shape
    real = _real(ev, descs, hot, np.zeros(n, np.int64))
    off = np.asarray(ev["hit_offset"], np.int64); hits = np.asarray(ev["hits_flat"], F)
    O = oracle.Oracle(3.8, rl, xi)
    par = np.array(cand["par"], F); err = np.array(cand["err"], F) * F(100); chg = np.array(cand["chg"], np.int32)
    chi2 = np.zeros(n, F)
    gid = np.arange(n) // group
    fake = np.array([666, 0, 666, 0, 0, 666], F)
    for l in range(nl - 1, -1, -1):
        assert descs[l]["is_barrel"]
        act_g = np.zeros(gid.max() + 1, bool)
        np.logical_or.at(act_g, gid, real[l])
        act = act_g[gid]
        if not act.any():
            continue
        mp = np.array(par[:3]); me = np.repeat(fake[:, None], n, axis=1)
        idx = off[l] + np.maximum(hot[l], 0)
        mp[:, real[l]] = hits[idx[real[l]], :3].T
        me[:, real[l]] = hits[idx[real[l]], 3:].T
        nomat = (~real[l] & (not fake_material)).astype(np.int32)
        r = np.hypot(mp[0], mp[1]).astype(F)
        e1, p1, _ = O.propagate_to_r(err, par, chg, r, pflags=pflags, no_mat=nomat)
        e2, p2, c2 = O.kalman_operation(3, e1, p1, me, mp)
        flip = p2[3] < 0
        chg2 = np.where(flip, -chg, chg); p2[3] = np.where(flip, -p2[3], p2[3])
        par[:, act] = p2[:, act]; err[:, act] = e2[:, act]; chg[act] = chg2[act]
        chi2[act] = (chi2 + c2)[act]
    return dict(par=par, err=err, chg=chg, chi2=chi2)


def _bk(ev, descs, cand, **kw):
    """oracle.backward_fit with the toy geometry's flags (Geoms/CylCowWLids.cc:228-231: backward fit and PCA
    propagation with the parametrised field) unless given."""
    kw.setdefault("pflags", PF_PARAM_B)
    kw.setdefault("pca_pflags", PF_PARAM_B)
    return oracle.backward_fit(ev, descs, cand, **kw)


def _same(got, ref, keys=("par", "err", "chg", "chi2")):
    for k in keys:
        a, b = np.asarray(got[k]), np.asarray(ref[k])
        same = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
        assert same.all(), (k, np.nonzero(~same.reshape(-1, a.shape[-1]).all(axis=0))[0][:10])


@pytest.mark.parametrize("group", [1, 2, 8, 16, 64])
def test_blocks_equal_fit_composition(barrel, group):
    geo, ev, descs, cand = barrel
    assert cand["chg"].size > 200
    got = _bk(ev, descs, cand, group=group, pflags=PF_PARAM_B, pca_pflags=PF_PARAM_B, n_threads=3)
    _same(got, _fit_composition(ev, descs, cand, group))


def test_without_pca_and_other_pca_flags(barrel):
    geo, ev, descs, cand = barrel
    got = _bk(ev, descs, cand, group=16, include_pca=False)
    _same(got, _fit_composition(ev, descs, cand, 16, include_pca=False))
    got = _bk(ev, descs, cand, group=16, pflags=PF_PARAM_B, pca_pflags=0)
    _same(got, _fit_composition(ev, descs, cand, 16, pflags=PF_PARAM_B, pca_pflags=0))


def test_last_layer_stops_the_plan(barrel):
    geo, ev, descs, cand = barrel
    got = _bk(ev, descs, cand, group=8, last_layer=3)
    _same(got, _fit_composition(ev, descs, cand, 8, last_layer=3))


def test_mixed_layer_kinds_with_layer_base(mixed):
    cat, descs, cand, base, L = mixed
    assert (base == 0).sum() > 100 and (base == L).sum() > 100
    for group in (1, 16):
        got = _bk(cat, descs, cand, group=group, layer_base=base, n_threads=3)
        _same(got, _fit_composition(cat, descs, cand, group, layer_base=base))


def test_material_and_fake_hit_material(barrel):
    """With material: fake hits of BkFitFitTracks take none (the fit's missing slot), those of
    BkFitFitTracksBH do (the step composition); with one candidate per block there are no fake hits and the
    two agree."""
    geo, ev, descs, cand = barrel
    rl, xi = mk.material_grid(geo)
    kw = dict(pflags=PF_PARAM_B | PF_MATERIAL, rl_grid=rl, xi_grid=xi)
    ce = _bk(ev, descs, cand, group=16, include_pca=False, **kw)
    _same(ce, _fit_composition(ev, descs, cand, 16, pflags=PF_PARAM_B | PF_MATERIAL, include_pca=False, rl=rl,
                               xi=xi))
    _same(ce, _step_composition(ev, descs, cand, 16, PF_PARAM_B | PF_MATERIAL, False, rl, xi))
    bh = _bk(ev, descs, cand, group=16, include_pca=False, fake_material=True, **kw)
    _same(bh, _step_composition(ev, descs, cand, 16, PF_PARAM_B | PF_MATERIAL, True, rl, xi))
    assert not np.array_equal(bh["err"], ce["err"])
    one_ce = _bk(ev, descs, cand, group=1, include_pca=False, **kw)
    one_bh = _bk(ev, descs, cand, group=1, include_pca=False, fake_material=True, **kw)
    _same(one_ce, one_bh)


def test_fake_hits_after_the_first_hit_change_nothing(barrel):
    """A fake hit sits at the track's own position with a 666 cm^2 variance.  Once a candidate's state has been
    propagated to a hit's layer, the fake hits of its block's other layers leave it unchanged to the bit, so a
    candidate with a hit on the plan's last layer fits the same in any block.  A candidate whose finding
    stopped earlier meets its block's fake hits first: the propagation to the cylinder through its own position
    projects the end-of-finding covariance onto that cylinder, a real change (MkFinder.cc:1635-1647)."""
    geo, ev, descs, cand = barrel
    one = _bk(ev, descs, cand, group=1)
    blk = _bk(ev, descs, cand, group=64)
    top = cand["hot_idx"][-1] >= 0
    assert top.sum() > 300 and (~top).sum() > 0
    _same({k: v[..., top] for k, v in blk.items()}, {k: v[..., top] for k, v in one.items()})
    assert not np.array_equal(blk["err"][:, ~top], one["err"][:, ~top])


def test_noise_free_chi2_per_degree_of_freedom():
    geo, ev, descs, cand = _barrel(600, noise=0.0, seed=17)
    got = _bk(ev, descs, cand, group=16)
    nh = (cand["hot_idx"] >= 0).sum(axis=0)
    ok = nh >= 10
    assert ok.sum() > 300
    # two measured coordinates per barrel hit, five track parameters
    ndof = 2 * nh[ok] - 5
    assert 0.7 < np.mean(got["chi2"][ok] / ndof) < 1.4


# (hit codes per layer, found hits, -1 holes before the last found hit)
SCORE_CASES = [
    ([5, 7, 9, 11, 13, 17, 19], 7, 0),
    ([5, 7, 9, -1, 13, -1, -1, 4, -2, -9], 5, 3),
    ([5, 7, 9, 2, -1, -1], 4, 0),
    ([5, 7, 9, -3, 2, -2, 8, -9, -9], 5, 0),
]


@pytest.mark.parametrize("case", range(len(SCORE_CASES)))
def test_score_known_answers(barrel, case):
    """getScoreCand after the fit: bonus 0.2 n + 4 per found hit (capped at 8), 8 per -1 hole before the last
    found hit (12 below pT 0.9 GeV), minus the fitted chi2; tail holes and -2/-3 codes cost nothing.  Checked on
    a candidate above and one below 0.9 GeV."""
    geo, ev, descs, cand = barrel
    codes, nfound, ninside = SCORE_CASES[case]
    L = len(descs)
    full = cand["hot_idx"][:L].min(axis=0) >= 0
    pt_in = np.abs(1.0 / cand["par"][3])
    picks = [int(np.nonzero(full & (pt_in > 1.2))[0][0]), int(np.nonzero(full & (pt_in < 0.8))[0][0])]
    expect = {  # (found, inside holes): score + chi2, for pT >= 0.9 and pT < 0.9
        (7, 0): (7 * (0.2 * 7 + 4), 7 * (0.2 * 7 + 4)),
        (5, 3): (5 * 5.0 - 24, 5 * 5.0 - 36),
        (4, 0): (4 * 4.8, 4 * 4.8),
        (5, 0): (25.0, 25.0),
    }[(nfound, ninside)]
    for t, high in zip(picks, (True, False)):
        one = {k: np.ascontiguousarray(v[..., t:t + 1]) for k, v in cand.items()}
        hot = np.full((L, 1), -9, np.int32)
        for l, c in enumerate(codes):
            hot[l, 0] = cand["hot_idx"][l, t] if c >= 0 else c
        one["hot_idx"] = hot
        got = _bk(ev, descs, one, group=1)
        chi2 = float(got["chi2"][0])
        pt = abs(1.0 / float(got["par"][3, 0]))
        assert (pt >= 0.9) == high, pt
        want = (expect[0] if high else expect[1]) - chi2
        assert abs(float(got["score"][0]) - want) <= 1e-5 * max(1.0, abs(want)), (got["score"][0], want, pt)


def test_candidates_left_untouched_outside_the_event(barrel):
    geo, ev, descs, cand = barrel
    base = np.zeros(cand["chg"].size, np.int32)
    base[::7] = len(descs)                          # the plan would run past the event's layers
    got = _bk(ev, descs, cand, group=16, layer_base=base)
    bad = base != 0
    assert np.array_equal(got["par"][:, bad], cand["par"][:, bad])
    assert np.array_equal(got["err"][:, bad], cand["err"][:, bad])
    ref = _fit_composition(ev, descs, {k: v[..., ~bad] for k, v in cand.items()}, 1)
    one = _bk(ev, descs, cand, group=1, layer_base=base)
    _same({k: v[..., ~bad] for k, v in one.items() if k in ref}, ref)
