"""Pins of the oracle's per-layer MkFinder steps (AddBestHit, FindCandidatesCloneEngine,
UpdateWithLastHit on one layer's lanes):
  * the best-hit layer loop of MkBuilder (MkBuilder.cc:1395-1478) written in Python over the per-layer
    steps reproduces the oracle's whole-event FindTracksBestHit exactly;
  * the clone-engine records equal a numpy composition of the propagation / chi2 primitives with the
    acceptance cuts, isStripQCompatible and getScoreStruct written out (MkFinder.cc:1114-1284);
  * UpdateWithLastHit equals propagate-to-hit + Kalman update + charge flip from the primitives.
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from test_oracle_ce import _score

F = np.float32


@pytest.fixture(scope="module")
def event():
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=91, simulate_material=0, pt_min=0.5, pt_max=10.0)
    ev = mk.generate_find_event(geo, p, 600, noise_per_cm2=3.0)
    return ev, mk.barrel_layer_descs(geo)


def _layer_hits(ev, layer):
    off = ev["hit_offset"]
    return ev["hits_flat"][off[layer]:off[layer + 1]]


def _best_hit_by_layers(O, ev, descs, first_layer, chi2_cut):
    n = ev["n_seeds"]
    L = len(descs)
    par = ev["seed_par"][:, :n].astype(F).copy(); err = ev["seed_err"][:, :n].astype(F).copy()
    chg = ev["seed_chg"][:n].astype(np.int32).copy(); chi2 = np.zeros(n, F)
    hot = np.full((L, n), -9, np.int32)
    hot[:first_layer] = ev["seed_hits"][:first_layer, :n]
    nfound = (hot[:first_layer] >= 0).sum(0).astype(np.int32)
    ninside = np.zeros(n, np.int32); ntail = np.zeros(n, np.int32)
    active = np.ones(n, bool)
    for layer in range(first_layer, L):
        d = descs[layer]
        a = np.nonzero(active)[0]
        if not len(a):
            break
        errP, parP, _ = O.propagate_to_r(err[:, a], par[:, a], chg[a], np.full(len(a), d["propagate_to"], F))
        reach = parP[0] * parP[0] + parP[1] * parP[1] >= F(d["rin"]) * F(d["rin"])
        active[a[~reach]] = False                           # MkBuilder.cc:1431-1455
        a, errP, parP = a[reach], errP[:, reach], parP[:, reach]
        if not len(a):
            break
        xs, xa, w = oracle.select_hit_indices(d, _layer_hits(ev, layer), parP, errP)
        out = oracle.add_best_hit(ev, descs, layer, dict(prop_par=parP, prop_err=errP, chg=chg[a], chi2=chi2[a],
                                                          n_found=nfound[a], n_inside_minus_one=ninside[a],
                                                          n_tail_minus_one=ntail[a]),
                                  xs, xa, w, chi2_cut=chi2_cut, bfield=3.8)
        par[:, a], err[:, a], chg[a], chi2[a] = out["par"], out["err"], out["chg"], out["chi2"]
        nfound[a], ninside[a], ntail[a] = out["n_found"], out["n_inside_minus_one"], out["n_tail_minus_one"]
        hot[layer, a] = out["hit_code"]
        active[a[out["hit_code"] == -2]] = False            # MkBuilder.cc:1465-1473
    return dict(par=par, err=err, chg=chg, chi2=chi2, hot_idx=hot, n_found=nfound)


def test_best_hit_layer_steps_compose_to_find_loop(event):
    ev, descs = event
    O = oracle.Oracle(bfield=3.8)
    got = _best_hit_by_layers(O, ev, descs, 3, 15.0)
    ref = oracle.find_best_hit(ev, descs, first_layer=3, chi2_cut=15.0, bfield=3.8, n_threads=2)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
    assert (ref["hot_idx"][3:] == -2).any() and (ref["hot_idx"][3:] >= 0).mean() > 0.5


def _lanes_at(O, ev, descs, layer, seed=5):
    """Seeds propagated to `layer` with random chi2 and hit counters."""
    rng = np.random.default_rng(seed)
    n = ev["n_seeds"]
    par = ev["seed_par"][:, :n].astype(F); err = ev["seed_err"][:, :n].astype(F)
    chg = ev["seed_chg"][:n].astype(np.int32)
    errP, parP, _ = O.propagate_to_r(err, par, chg, np.full(n, descs[layer]["propagate_to"], F))
    return dict(prop_par=parP, prop_err=errP, chg=chg, chi2=rng.uniform(0, 20, n).astype(F),
                n_found=rng.integers(3, 6, n).astype(np.int32),
                n_inside_minus_one=rng.integers(0, 3, n).astype(np.int32),
                n_tail_minus_one=rng.integers(0, 2, n).astype(np.int32))


def test_add_best_hit_skips_invalid_list_entries(event):
    ev, descs = event
    O = oracle.Oracle(bfield=3.8)
    layer = 4
    ln = _lanes_at(O, ev, descs, layer)
    hl = _layer_hits(ev, layer)
    xs, xa, w = oracle.select_hit_indices(descs[layer], hl, ln["prop_par"], ln["prop_err"])
    ref = oracle.add_best_hit(ev, descs, layer, ln, xs, xa, w, bfield=3.8)
    # interleave out-of-range entries; the valid ones keep their order
    xa2 = np.full_like(xa, -1)
    xs2 = xs.copy()
    rng = np.random.default_rng(1)
    for t in range(len(xs)):
        lst = [int(v) for v in xa[:max(xs[t], 0), t]]
        for _ in range(int(rng.integers(0, 3))):
            if len(lst) < 16:
                lst.insert(int(rng.integers(0, len(lst) + 1)), int(rng.choice([-5, len(hl), len(hl) + 7])))
        xa2[:len(lst), t] = lst
        if xs[t] >= 0:
            xs2[t] = len(lst)
    got = oracle.add_best_hit(ev, descs, layer, ln, xs2, xa2, w, bfield=3.8)
    for k in ("par", "err", "chg", "chi2", "n_found", "n_inside_minus_one", "n_tail_minus_one", "hit_code"):
        assert np.array_equal(got[k], ref[k]), k


def _dyn_cut(d, cut, invpt, theta):
    """getHitSelDynamicChi2Cut (MkFinder.cc:202-220)."""
    th = np.abs(theta - F(np.pi / 2))
    c = lambda k: F(d.get(k, 0.0))
    this = c("c_c2_0") * np.minimum(invpt, F(10)) + c("c_c2_1") * th + c("c_c2_2")
    return np.where(c("c_c2_sf") * this > F(cut), c("c_c2_sf") * this, F(cut))


@pytest.mark.parametrize("layer", [3, 9])
def test_find_candidates_records_match_numpy(event, layer):
    ev, descs = event
    O = oracle.Oracle(bfield=3.8)
    d = descs[layer]
    ln = _lanes_at(O, ev, descs, layer, seed=layer)
    hl = _layer_hits(ev, layer)
    parP, errP, chg = ln["prop_par"], ln["prop_err"], ln["chg"]
    xs, xa, w = oracle.select_hit_indices(d, hl, parP, errP)
    ne, ent = oracle.find_candidates_ce(ev, descs, layer, ln, xs, xa, w, chi2_cut=30.0, max_holes=3,
                                        max_consec_holes=1, bfield=3.8)
    n = parP.shape[1]
    exp = [[] for _ in range(n)]
    cut = _dyn_cut(d, 30.0, parP[3], parP[5])
    pt = np.abs(F(1) / parP[3])
    for hc in range(16):
        sel = xs > hc
        if not sel.any():
            break
        ms_par = np.zeros((3, n), F); ms_err = np.ones((6, n), F)
        ms_par[:, sel] = hl[xa[hc, sel], :3].T; ms_err[:, sel] = hl[xa[hc, sel], 3:].T
        rad = np.hypot(ms_par[0], ms_par[1]).astype(F)
        pe, pp, _ = O.propagate_to_r(errP, parP, chg, rad)
        _, _, c2 = O.kalman_operation(1, pe, pp, ms_err, ms_par)
        for t in np.nonzero(sel)[0]:
            c2t = abs(F(c2[t]))
            if not c2t < cut[t]:
                continue
            res = abs(ms_par[2, t] - pp[2, t])                      # isStripQCompatible, barrel
            if not np.sqrt(ms_err[5, t] * F(3)) + max(F(3) * np.sqrt(errP[5, t]), F(0.5)) > res:
                continue
            nh = int(ln["n_inside_minus_one"][t] + ln["n_tail_minus_one"][t])
            tot = F(ln["chi2"][t] + c2t)
            exp[t].append((int(xa[hc, t]), int(ln["n_found"][t]) + 1, 0, nh, tot, c2t,
                           _score(int(ln["n_found"][t]) + 1, 0, 0, nh, tot, pt[t], True), hc))
    for t in range(n):
        if w[t] == 2:
            continue
        ni, nt = int(ln["n_inside_minus_one"][t]), int(ln["n_tail_minus_one"][t])
        fake = -1 if (ni + nt < 3 and nt < 1) else -2
        if w[t] == 1:
            fake = -3
        ntl = nt + 1 if fake == -1 else nt
        exp[t].append((fake, int(ln["n_found"][t]), ntl, ni, F(ln["chi2"][t]), F(0),
                       _score(int(ln["n_found"][t]), ntl, 0, ni, F(ln["chi2"][t]), pt[t], True), -1))
    n_hit_records = 0
    for t in range(n):
        assert ne[t] == len(exp[t]), t
        for k, e in enumerate(exp[t]):
            r = ent[t, k]
            assert (r["trk_idx"], r["hit_idx"], r["nhits"], r["ntailholes"], r["nholes"], r["xhit_pos"]) == \
                (t, e[0], e[1], e[2], e[3], e[7]), (t, k)
            assert r["module"] == (0 if e[0] >= 0 else 0xffffffff)
            assert r["chi2"] == e[4] and r["chi2_hit"] == e[5] and r["pt"] == pt[t]
            np.testing.assert_allclose(r["score"], e[6], rtol=1e-6, atol=1e-5)
            n_hit_records += e[0] >= 0
    assert n_hit_records > n // 2


def test_update_with_last_hit_matches_primitives(event):
    ev, descs = event
    O = oracle.Oracle(bfield=3.8)
    n = ev["n_seeds"]
    layer = 3
    hl = _layer_hits(ev, layer)
    par = ev["seed_par"][:, :n].astype(F); err = ev["seed_err"][:, :n].astype(F)
    chg = ev["seed_chg"][:n].astype(np.int32)
    mc = ev["hit_mc"][ev["hit_offset"][layer]:ev["hit_offset"][layer + 1]]
    first = {}
    for i, m in enumerate(mc):
        if m >= 0 and m not in first:
            first[m] = i
    hit = np.array([first.get(int(m), -1) for m in ev["seed_mc"][:n]], np.int32)
    hit[::7] = -1
    hit[3::11] = len(hl) + 3
    p2, e2, c2 = oracle.update_with_last_hit(ev, descs, layer, dict(prop_par=par, prop_err=err, chg=chg), hit,
                                             pflags=0, bfield=3.8)
    ok = (hit >= 0) & (hit < len(hl))
    ms_par = np.ascontiguousarray(hl[hit[ok], :3].T); ms_err = np.ascontiguousarray(hl[hit[ok], 3:].T)
    pe, pp, _ = O.propagate_to_r(err[:, ok], par[:, ok], chg[ok], np.hypot(ms_par[0], ms_par[1]).astype(F))
    eu, pu, _ = O.kalman_operation(2, pe, pp, ms_err, ms_par)
    cu = chg[ok].copy()
    flip = pu[3] < 0
    cu[flip] = -cu[flip]; pu[3, flip] = -pu[3, flip]
    assert np.array_equal(p2[:, ok], pu) and np.array_equal(e2[:, ok], eu) and np.array_equal(c2[ok], cu)
    assert np.array_equal(p2[:, ~ok], par[:, ~ok]) and np.array_equal(e2[:, ~ok], err[:, ~ok])
    assert np.array_equal(c2[~ok], chg[~ok])
