"""Brute-force and known-answer pins of the oracle's finding bookkeeping.

The reference ships no vectors for track finding.  These pins check the oracle's clone-engine and
best-hit bookkeeping against formulations that share none of its control flow:

  * known answers: IdxChi2List scores (getScoreCalc, Track.h:648-670, with the constants of
    Config.h:355-359) worked out by hand for chosen hit counts, chi2 and pT, and the invalid-hit codes
    (-1 / -2 / -3, MkFinder.cc:1248-1262) for chosen hole counts and limits;
  * exhaustive enumeration: on tiny events (one seed, three layers after the seed) run with a candidate
    limit K above the number of records at every layer, nothing is ever pruned, so the clone engine's
    final candidates must be exactly the set of hit/hole paths through the layers, each path's state the
    chain of per-hit updates, and the list ordered by getScoreCand with the best short candidate merged
    as in CombCandidate::MergeCandsAndBestShortOne (HitStructures.cc:511-552);
  * greedy path: with K = 1 the clone engine must follow the single path that takes, layer by layer,
    the record of highest score (the first one on equal scores);
  * ties: duplicated hits must never be chosen over their first copy (strict '<' in AddBestHit,
    MkFinder.cc:753; stable score order in the clone engine).

The enumerator below walks a tree of paths with the oracle's separately pinned primitives
(propagation, Kalman chi2 / update, SelectHitIndices); it has no top-K selection, no extras
interleaving and no HitOnTrack node store.
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle

F = np.float32
FAKE_ERR = np.array([666, 0, 666, 0, 0, 666], F)

# ---------------------------------------------------------------------------
# Scores: the formula, pinned by the known-answer table below
# ---------------------------------------------------------------------------


def score(nfound, ntail, nmiss, chi2, pt, in_find):
    """getScoreCalc (Track.h:648-670) with Config.h:355-359: bonus 4 + 0.2 per found hit, hole penalty 8,
    tail-hole penalty 3, overlap bonus 0; below pT 0.9 the penalty grows x1.7 (finding) / x1.5 (final)
    and the bonus is x0.9 (finding), capped at 8.  Negative chi2 counts as 0."""
    bonus = F(0.2) * F(nfound) + F(4)
    penalty = F(8)
    if pt < 0.9:
        penalty = penalty * (F(1.7) if in_find else F(1.5))
        bonus = min(bonus * (F(0.9) if in_find else F(1.0)), F(8))
    return F(bonus * F(nfound) - penalty * F(nmiss) - F(3) * F(ntail) - F(max(chi2, 0)))


# (nfound, ninside, ntail, chi2, 1/pT, wsr, max_holes, max_consec) -> (code, nhits, ntailholes, nholes, score)
# Scores worked out by hand from the formula above.
INVALID_HIT_CASES = [
    ((5, 1, 0, 10.0, 0.5, 0, 4, 1), (-1, 5, 1, 1, 4.0)),       # 5*5 - 8*1 - 3*1 - 10
    ((5, 1, 0, 10.0, 2.0, 0, 4, 1), (-1, 5, 1, 1, -4.1)),      # pT 0.5: 4.5*5 - 13.6*1 - 3 - 10
    ((25, 0, 0, 0.0, 2.0, 0, 4, 1), (-1, 25, 1, 0, 197.0)),    # bonus 9 * 0.9 = 8.1, capped at 8: 8*25 - 3
    ((25, 0, 0, 0.0, 0.5, 0, 4, 1), (-1, 25, 1, 0, 222.0)),    # no cap above pT 0.9: 9*25 - 3
    ((6, 3, 1, 2.0, 0.5, 0, 4, 1), (-2, 6, 1, 3, 2.2)),        # 3 + 1 holes reach max_holes 4: 5.2*6 - 24 - 3 - 2
    ((4, 0, 1, 1.0, 0.25, 0, 4, 1), (-2, 4, 1, 0, 15.2)),      # a tail hole reaches max_consec 1: 4.8*4 - 3 - 1
    ((4, 0, 1, 1.0, 0.25, 0, 4, 2), (-1, 4, 2, 0, 12.2)),      # max_consec 2 allows it: 4.8*4 - 3*2 - 1
    ((4, 1, 0, 1.0, 0.25, 0, 1, 1), (-2, 4, 0, 1, 10.2)),      # max_holes 1: 4.8*4 - 8 - 1
    ((4, 1, 1, 1.0, 0.25, 0, 2, 2), (-2, 4, 1, 1, 7.2)),       # inside + tail holes reach max_holes 2: 19.2 - 8 - 3 - 1
    ((3, 0, 0, 5.0, 0.5, 1, 4, 1), (-3, 3, 0, 0, 8.8)),        # window edge: 4.6*3 - 5, no tail hole added
    ((3, 2, 1, 5.0, 0.5, 1, 1, 1), (-3, 3, 1, 2, -10.2)),      # edge wins over the hole limits: 13.8 - 16 - 3 - 5
    ((3, 0, 0, -1.0, 0.5, 0, 4, 1), (-1, 3, 1, 0, 10.8)),      # negative chi2 scores as 0: 13.8 - 3
    ((7, 0, 0, 3.0, 1.25, 0, 4, 1), (-1, 7, 1, 0, 28.02)),     # pT 0.8: bonus min(5.4*0.9, 8) = 4.86: 34.02 - 3 - 3
]


def test_score_formula_known_answers():
    for (nf, ni, nt, c2, ipt, _, _, _), (code, nhits, ntail, nholes, want) in INVALID_HIT_CASES:
        got = score(nhits, ntail, nholes, c2, abs(F(1) / F(ipt)), True)
        assert abs(got - want) < 1e-4, (nf, ni, nt, c2, ipt, got, want)
    # final scores (getScoreCand: no tail term, x1.5 penalty and no x0.9 below pT 0.9)
    assert abs(score(5, 0, 1, 10.0, 0.5, False) - (5 * 5 - 12 - 10)) < 1e-4
    assert abs(score(25, 0, 0, 0.0, 0.5, False) - 200.0) < 1e-4      # 9 capped at 8 also in the final score


def _toy_layers():
    geo = mk.geometry(1)
    descs = mk.barrel_layer_descs(geo)
    ev = mk.generate_find_event(geo, mk.default_toy_params(seed=3, simulate_material=0), 50, noise_per_cm2=0.0)
    return ev, descs


@pytest.mark.parametrize("case", range(len(INVALID_HIT_CASES)))
def test_invalid_hit_record_known_answers(case):
    """FindCandidatesCloneEngine's invalid-hit record (MkFinder.cc:1237-1270) for one lane with no hit in its
    window: code, counts and score against the hand-worked table."""
    (nf, ni, nt, c2, ipt, wsr, mh, mc), (code, nhits, ntail, nholes, want) = INVALID_HIT_CASES[case]
    ev, descs = _toy_layers()
    par = np.zeros((6, 1), F); par[:, 0] = [5.0, 0.0, 0.0, ipt, 0.0, 1.5]
    err = np.zeros((21, 1), F); err[[0, 2, 5, 9, 14, 20], 0] = 1e-3
    lanes = dict(prop_par=par, prop_err=err, chg=np.array([1]), chi2=np.array([c2], F), n_found=np.array([nf]),
                 n_inside_minus_one=np.array([ni]), n_tail_minus_one=np.array([nt]))
    ne, ent = oracle.find_candidates_ce(ev, descs, 5, lanes, np.zeros(1), np.full((16, 1), -1), np.array([wsr]),
                                        max_holes=mh, max_consec_holes=mc)
    assert ne[0] == 1
    e = ent[0, 0]
    assert (e["hit_idx"], e["nhits"], e["ntailholes"], e["nholes"]) == (code, nhits, ntail, nholes)
    assert e["xhit_pos"] == -1 and e["chi2_hit"] == 0 and e["chi2"] == F(c2)
    assert abs(e["score"] - want) < 1e-4, (e["score"], want)


def test_outside_lane_gets_no_record():
    ev, descs = _toy_layers()
    par = np.zeros((6, 2), F); par[:, :] = np.array([5.0, 0.0, 0.0, 0.5, 0.0, 1.5], F)[:, None]
    err = np.zeros((21, 2), F); err[[0, 2, 5, 9, 14, 20]] = 1e-3
    lanes = dict(prop_par=par, prop_err=err, chg=np.array([1, 1]), chi2=np.zeros(2, F), n_found=np.array([3, 3]))
    ne, ent = oracle.find_candidates_ce(ev, descs, 5, lanes, np.zeros(2), np.full((16, 2), -1), np.array([2, 0]))
    assert ne.tolist() == [0, 1]


def test_hit_record_counts_and_score():
    """Records of accepted hits: nhits + 1, the holes so far (inside + tail) as inside holes, no tail
    hole, chi2 = lane chi2 + hit chi2, score from those counts (MkFinder.cc:1210-1225)."""
    ev, descs = _toy_layers()
    L = 5
    d = descs[L]
    hl = ev["hits_flat"][ev["hit_offset"][L]:ev["hit_offset"][L + 1]]
    n = hl.shape[0]
    O = oracle.Oracle(bfield=3.8)
    # lanes sitting just inside the layer, aimed at each signal hit, with assorted counters
    par = np.zeros((6, n), F)
    par[0], par[1], par[2] = hl[:, 0] * F(0.999), hl[:, 1] * F(0.999), hl[:, 2]
    par[3] = np.where(np.arange(n) % 2 == 0, F(0.5), F(1.6))
    par[4] = np.arctan2(hl[:, 1], hl[:, 0]); par[5] = np.arctan2(np.hypot(hl[:, 0], hl[:, 1]), hl[:, 2])
    err = np.zeros((21, n), F); err[[0, 2, 5]] = F(1e-3); err[[9, 14, 20]] = F(1e-4)
    rng = np.random.default_rng(4)
    nf = rng.integers(3, 20, n); ni = rng.integers(0, 3, n); nt = rng.integers(0, 2, n)
    c2l = rng.uniform(0, 20, n).astype(F)
    lanes = dict(prop_par=par, prop_err=err, chg=np.ones(n), chi2=c2l, n_found=nf, n_inside_minus_one=ni,
                 n_tail_minus_one=nt)
    xs, xa, wsr = oracle.select_hit_indices(d, hl, par, err)
    ne, ent = oracle.find_candidates_ce(ev, descs, L, lanes, xs, xa, wsr, chi2_cut=1e9)
    n_hit_records = 0
    for t in range(n):
        for e in ent[t, :ne[t]]:
            if e["hit_idx"] < 0:
                continue
            n_hit_records += 1
            assert (e["nhits"], e["ntailholes"], e["nholes"]) == (nf[t] + 1, 0, ni[t] + nt[t])
            assert e["chi2"] == F(c2l[t] + e["chi2_hit"])
            assert abs(e["score"] - score(nf[t] + 1, 0, ni[t] + nt[t], e["chi2"], abs(F(1) / par[3, t]), True)) \
                <= 1e-5 * max(1.0, abs(float(e["score"])))
    assert n_hit_records >= n


# ---------------------------------------------------------------------------
# Exhaustive paths of the clone engine on tiny events
# ---------------------------------------------------------------------------


class Path:
    """One hit/hole path: state, counters (TrackCand::addHitIdx, HitStructures.h:873-889), layer codes."""

    def __init__(self, par, err, chg):
        self.par, self.err, self.chg = np.array(par, F), np.array(err, F), int(chg)
        self.chi2 = F(0)
        self.nfound = self.ninside = self.ntail = 0
        self.codes = {}            # layer -> hit index or fake code
        self.last = None
        self.score = F(-1e16)

    def child(self, layer, code, chi2_hit):
        c = Path(self.par, self.err, self.chg)
        c.chi2, c.nfound, c.ninside, c.ntail = self.chi2, self.nfound, self.ninside, self.ntail
        c.codes = dict(self.codes)
        c.codes[layer] = code
        c.last = code
        if code >= 0:
            c.nfound += 1
            c.chi2 = F(c.chi2 + F(chi2_hit))
            c.ninside += c.ntail
            c.ntail = 0
        elif code == -1:
            c.ntail += 1
        return c

    def pt(self):
        return abs(F(1) / self.par[3])


def _prop_r(O, paths, r, pflags=0):
    par = np.stack([p.par for p in paths], 1); err = np.stack([p.err for p in paths], 1)
    chg = np.array([p.chg for p in paths], np.int32)
    rad = np.broadcast_to(np.asarray(r, F), (len(paths),)).copy()
    e, p, _ = O.propagate_to_r(err, par, chg, rad, pflags=pflags)
    return e, p, chg


def _children(O, path, layer, d, hl, chi2_cut, max_holes, max_consec):
    """What one live path becomes at a layer: (kind, path, in-find score, position in the window's hit list)
    in generation order, kind one of 'hit', 'hole', 'short' (a -2 record), 'carry' (outside the layer: kept
    unchanged, no record), 'stopped' (barrel layer missed from inside: kept with a -2, no record); the
    position is 16 for all but hits."""
    errP, parP, chg = _prop_r(O, [path], d["propagate_to"])
    xs, xa, wsr = oracle.select_hit_indices(d, hl, parP, errP)
    if np.hypot(parP[0, 0], parP[1, 0]) < F(d["rin"]):
        c = path.child(layer, -2, 0.0)
        c.score = path.score
        return [("stopped", c, None, 16)]
    if wsr[0] == 2:
        return [("carry", path, None, 16)]
    out = []
    pt = abs(F(1) / parP[3, 0])
    for hc, h in enumerate(xa[:xs[0], 0]):
        ms_par = hl[h, :3].reshape(3, 1).copy(); ms_err = hl[h, 3:].reshape(6, 1).copy()
        pe, pp, _ = O.propagate_to_r(errP, parP, chg, np.hypot(ms_par[0], ms_par[1]).astype(F))
        _, _, c2 = O.kalman_operation(1, pe, pp, ms_err, ms_par)
        c2 = abs(F(c2[0]))
        if not c2 < F(chi2_cut):
            continue
        # isStripQCompatible (MkFinder.cc:837-862), barrel: |z_hit - z_track| within the hit's half length
        # plus max(3 sigma_z of the layer state, 0.5 cm)
        if not np.sqrt(ms_err[5, 0] * F(3)) + max(F(3) * np.sqrt(errP[5, 0]), F(0.5)) > abs(ms_par[2, 0] - pp[2, 0]):
            continue
        c = path.child(layer, int(h), c2)
        # UpdateWithLastHit: propagate the layer-entry state to the hit, update, flip a negative 1/pT
        ue, up, _ = O.propagate_to_r(np.stack([path.err], 1), np.stack([path.par], 1), np.array([path.chg]),
                                     np.hypot(ms_par[0], ms_par[1]).astype(F))
        e2, p2, _ = O.kalman_operation(2, ue, up, ms_err, ms_par)
        c.par, c.err = p2[:, 0].copy(), e2[:, 0].copy()
        if c.par[3] < 0:
            c.chg, c.par[3] = -c.chg, -c.par[3]
        out.append(("hit", c, score(path.nfound + 1, 0, path.ninside + path.ntail, c.chi2, pt, True), hc))
    holes = path.ninside + path.ntail
    code = -1 if holes < max_holes and path.ntail < max_consec else -2
    if wsr[0] == 1:
        code = -3
    c = path.child(layer, code, 0.0)
    sc = score(path.nfound, path.ntail + 1 if code == -1 else path.ntail, path.ninside, path.chi2, pt, True)
    out.append(("short" if code == -2 else "hole", c, sc, 16))
    return out


def _apex_stop(p):
    """find_tracks_unroll_candidates (MkBuilder.cc:1521-1533): a low-pT candidate near its helix apex stops."""
    if p.pt() < 1.2:
        dphi = abs(np.arctan2(p.par[1], p.par[0]) - p.par[4])
        return p.par[0] * p.par[0] + p.par[1] * p.par[1] > F(625) and F(1.371) < dphi < F(4.512)
    return False


def enumerate_paths(O, ev, descs, s, first_layer, chi2_cut, max_holes, max_consec, greedy=False):
    """Every path of seed s (or, greedy, the one of best record score per layer).  Returns (final candidates
    in final order, largest number of records plus held-back candidates at one layer)."""
    off, hits = ev["hit_offset"], ev["hits_flat"]
    seed = Path(ev["seed_par"][:, s], ev["seed_err"][:, s], ev["seed_chg"][s])
    for layer in range(first_layer):
        seed = seed.child(layer, int(ev["seed_hits"][layer, s]), 0.0)
    cands, best_short, finished, max_records = [seed], None, False, 0
    for layer in range(first_layer, len(descs)):
        if finished:
            break
        d = descs[layer]
        hl = hits[off[layer]:off[layer + 1]]
        live, kept = [], []
        for p in cands:
            if p.last == -2:
                kept.append(p)
            elif _apex_stop(p):
                kept.append(p.child(layer, -2, 0.0))
            else:
                live.append(p)
        if not live:
            finished = True            # CombCandidate::Finished: the candidates stay as they are
            cands = kept
            continue
        records, extras = [], []
        for p in live:
            for kind, c, sc, _ in _children(O, p, layer, d, hl, chi2_cut, max_holes, max_consec):
                if kind in ("carry", "stopped"):
                    extras.append(c)
                else:
                    records.append((kind, c, sc))
        max_records = max(max_records, len(records) + len(extras))
        if greedy and records:
            records = [records[int(np.argmax([sc for _, _, sc in records]))]]     # first of equal maxima
        for kind, c, sc in records:
            c.score = sc
            if kind == "short" and (best_short is None or sc > best_short.score):
                best_short = c
        if records:
            cands = [c for kind, c, _ in records if kind != "short"] + extras
        else:
            cands = extras
    # MergeCandsAndBestShortOne with final scores
    for c in cands:
        c.score = score(c.nfound, 0, c.ninside, c.chi2, c.pt(), False)
    cands = sorted(cands, key=lambda c: -c.score)
    if best_short is not None:
        if not cands:
            cands = [best_short]      # taken as it is: its score stays the one of its -2 record
        else:
            best_short.score = score(best_short.nfound, 0, best_short.ninside, best_short.chi2, best_short.pt(),
                                     False)
            if best_short.score > cands[-1].score:
                at = next(i for i, c in enumerate(cands) if not c.score > best_short.score)
                if greedy:
                    cands = cands[:-1]
                cands.insert(at, best_short)
    return cands, max_records


def _tiny_events(seed, n, noise, n_layers=7, pt_min=0.5, zmax=None):
    """Toy barrel events cut to their first n_layers layers; zmax shortens the finding layers (from the fourth
    on) to |z| <= zmax, so that tracks leave through the layer ends (window edges, misses)."""
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=seed, simulate_material=0, pt_min=pt_min, pt_max=10.0)
    ev = mk.generate_find_event(geo, p, n, noise_per_cm2=noise)
    descs = mk.barrel_layer_descs(geo)[:n_layers]
    if zmax is not None:
        descs = [dict(d, zmin=-zmax, zmax=zmax) if i >= 3 else d for i, d in enumerate(descs)]
    off = ev["hit_offset"][:n_layers + 1]
    ev = dict(ev, hit_offset=off, hits_flat=ev["hits_flat"][:off[-1]], hit_mc=ev["hit_mc"][:off[-1]])
    return ev, descs


def _one_seed(ev, s):
    return dict(ev, n_seeds=1, seed_par=ev["seed_par"][:, s:s + 1], seed_err=ev["seed_err"][:, s:s + 1],
                seed_chg=ev["seed_chg"][s:s + 1], seed_hits=ev["seed_hits"][:, s:s + 1])


def _rows(cands, L):
    rows = []
    for c in cands:
        r = np.full(L, -9, np.int32)
        for layer, code in c.codes.items():
            r[layer] = code
        rows.append(r)
    return rows


def _compare(got, cands, L, K):
    nc = int(got["n_cands"][0])
    assert nc == min(len(cands), K), (nc, len(cands))
    rows = _rows(cands, L)
    for i in range(nc):
        c = cands[i]
        assert np.array_equal(got["hot_idx"][i, :, 0], rows[i]), (i, got["hot_idx"][i, :, 0], rows[i])
        assert got["n_found"][i, 0] == c.nfound and got["chg"][i, 0] == c.chg
        assert got["chi2"][i, 0] == c.chi2
        assert abs(got["score"][i, 0] - c.score) <= 1e-5 * max(1.0, abs(float(c.score)))
        np.testing.assert_allclose(got["par"][i, :, 0], c.par, rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(got["err"][i, :, 0], c.err, rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("holes", [(4, 1, None), (1, 1, None), (3, 2, None), (4, 1, 20.0), (2, 2, 20.0)])
def test_clone_engine_equals_all_paths(holes):
    """K = 16 above every layer's record count: the oracle's candidates are all paths, best short merged.
    Three layers after the seed, noise 1 hit / cm^2: 50-60 of the 100 seeds stay below K.  With the finding
    layers cut to |z| <= 20 cm, tracks meet window edges (-3) and leave the layers (held back unchanged)."""
    max_holes, max_consec, zmax = holes
    ev, descs = _tiny_events(61, 100, 1.0, n_layers=6, zmax=zmax)
    O = oracle.Oracle(bfield=3.8)
    L, K = len(descs), 16
    checked = branched = shorts = edges = 0
    for s in range(ev["n_seeds"]):
        cands, rec = enumerate_paths(O, ev, descs, s, 3, 30.0, max_holes, max_consec)
        if rec >= K:
            continue
        one = _one_seed(ev, s)
        got = oracle.find_clone_engine(one, descs, 3, chi2_cut=30.0, max_cands=K, max_holes=max_holes,
                                       max_consec_holes=max_consec)
        _compare(got, cands, L, K)
        checked += 1
        # two candidates with different hits on one layer: a window held two accepted hits
        rows = np.array(_rows(cands, L))
        branched += any(len(set(rows[:, l][rows[:, l] >= 0])) > 1 for l in range(3, L))
        shorts += any(c.last == -2 for c in cands)
        edges += any(-3 in c.codes.values() for c in cands)
    assert checked >= 50 and branched >= 5, (checked, branched)
    if max_holes == 1:
        assert shorts >= 5, shorts
    if zmax is not None:
        assert edges >= 5, edges


@pytest.mark.parametrize("zmax", [None, 30.0])
def test_clone_engine_k1_is_greedy(zmax):
    ev, descs = _tiny_events(62, 80, 1.0, n_layers=8, zmax=zmax)
    O = oracle.Oracle(bfield=3.8)
    L = len(descs)
    got = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=1)
    for s in range(ev["n_seeds"]):
        cands, _ = enumerate_paths(O, ev, descs, s, 3, 30.0, 4, 1, greedy=True)
        one = {k: (v[..., s:s + 1] if k != "n_cands" else v[s:s + 1]) for k, v in got.items()}
        _compare(one, cands, L, 1)


def enumerate_ce_k(O, ev, descs, s, first_layer, chi2_cut, max_holes, max_consec, K):
    """Seed s through the layers with CandCloner's K-limited selection (CandCloner.cc:37-191), built on the
    same per-path steps as enumerate_paths: per layer the records of all live candidates in generation order
    (accepted hits by window position, then candidate; then the invalid-hit records by candidate) ranked by
    score (earlier generated first among equals); the best K of them are taken in rank order, a -2 one
    competing for the best short candidate and every other one placed after the held-back candidates (in
    candidate order) that score higher than it, the held-back ones filling what is left of the K places.
    Returns (final candidates, number of held-back candidates placed ahead of a new one by their score)."""
    off, hits = ev["hit_offset"], ev["hits_flat"]
    seed = Path(ev["seed_par"][:, s], ev["seed_err"][:, s], ev["seed_chg"][s])
    for layer in range(first_layer):
        seed = seed.child(layer, int(ev["seed_hits"][layer, s]), 0.0)
    seed.score = score(seed.nfound, 0, seed.ninside, seed.chi2, seed.pt(), False)
    cands, best_short, ahead = [seed], None, 0
    for layer in range(first_layer, len(descs)):
        d = descs[layer]
        hl = hits[off[layer]:off[layer + 1]]
        live = []
        for i, p in enumerate(cands):
            if p.last == -2:
                continue
            if _apex_stop(p):
                cands[i] = p.child(layer, -2, 0.0)
                cands[i].score = p.score
                continue
            live.append(p)
        if not live:
            break                                   # CombCandidate::Finished: the candidates stay
        made, fakes, extras = [], [], []
        for lane, p in enumerate(live):
            for kind, c, sc, hc in _children(O, p, layer, d, hl, chi2_cut, max_holes, max_consec):
                if kind in ("carry", "stopped"):
                    extras.append(c)
                elif kind == "hit":
                    made.append((hc, lane, kind, c, sc))
                else:
                    fakes.append((kind, c, sc))
        records = [(k, c, sc) for _, _, k, c, sc in sorted(made, key=lambda x: (x[0], x[1]))] + fakes
        if not records:
            cands = extras
            continue
        ranked = sorted(records, key=lambda r: -r[2])[:K]            # stable
        placed, xi = [], 0
        for kind, c, sc in ranked:
            c.score = sc
            if kind == "short":
                if best_short is None or sc > best_short.score:
                    best_short = c
                continue
            while xi < len(extras) and extras[xi].score > sc and len(placed) < K:
                placed.append(extras[xi])
                xi += 1
                ahead += 1
            if len(placed) >= K:
                break
            placed.append(c)
        placed += extras[xi:xi + max(0, K - len(placed))]
        cands = placed
    # MergeCandsAndBestShortOne with final scores, at most K candidates
    for c in cands:
        c.score = score(c.nfound, 0, c.ninside, c.chi2, c.pt(), False)
    cands = sorted(cands, key=lambda c: -c.score)
    if best_short is not None:
        if not cands:
            cands = [best_short]
        else:
            best_short.score = score(best_short.nfound, 0, best_short.ninside, best_short.chi2, best_short.pt(),
                                     False)
            if best_short.score > cands[-1].score:
                at = next(i for i, c in enumerate(cands) if not c.score > best_short.score)
                cands.insert(at, best_short)
    return cands[:K], ahead


@pytest.mark.parametrize("zmax", [30.0, None])
def test_clone_engine_k_limited(zmax):
    """Fewer places than candidates (K = 2): the K-limited selection with held-back candidates (tracks
    leaving the shortened layers keep their place by score among the new candidates) and the best short
    candidate."""
    ev, descs = _tiny_events(63, 80, 1.5, n_layers=8, zmax=zmax)
    _check_k_limited(ev, descs, 2, range(ev["n_seeds"]))


def test_clone_engine_held_back_goes_ahead():
    """Held-back candidates that outscore a new one take their place ahead of it (CandCloner.cc:103-109): on
    layers cut to |z| <= 20 cm with noise 3 / cm^2 this happens to about one seed in fifteen of those that
    pass within 0.5 cm of a layer end (the seeds checked here)."""
    ev, descs = _tiny_events(63, 1000, 3.0, n_layers=7, zmax=20.0)
    sp = ev["seed_par"]
    r0, cot = np.hypot(sp[0], sp[1]), 1 / np.tan(sp[5])
    near = np.zeros(ev["n_seeds"], bool)
    for d in descs[3:]:
        near |= np.abs(np.abs(sp[2] + (d["propagate_to"] - r0) * cot) - 20.0) < 0.5
    ahead = _check_k_limited(ev, descs, 3, np.nonzero(near)[0])
    assert ahead >= 3, ahead


def _check_k_limited(ev, descs, K, seeds):
    O = oracle.Oracle(bfield=3.8)
    L = len(descs)
    got = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=K, n_threads=4)
    ahead = 0
    for s in seeds:
        cands, a = enumerate_ce_k(O, ev, descs, s, 3, 30.0, 4, 1, K)
        ahead += a
        one = {k: (v[..., s:s + 1] if k != "n_cands" else v[s:s + 1]) for k, v in got.items()}
        _compare(one, cands, L, K)
    return ahead


def _short_layer_event(noise, n):
    """Seven toy barrel layers, the fifth cut to |z| <= 8 cm and 6 cm thick (a wide window edge)."""
    ev, descs = _tiny_events(66, n, noise, n_layers=7)
    return ev, [dict(d, zmin=-8.0, zmax=8.0, rin=d["propagate_to"] - 3.0, rout=d["propagate_to"] + 3.0) if i == 4
                else d for i, d in enumerate(descs)]


def test_clone_engine_across_a_short_layer():
    """One finding layer cut to |z| <= 8 cm (and 6 cm thick, so that its window edge is wide) between full
    ones: paths take a window-edge record (-3) there and hits after it, which count it neither as a tail
    hole nor as an inside hole (TrackCand::addHitIdx)."""
    ev, descs = _short_layer_event(0.2, 300)
    O = oracle.Oracle(bfield=3.8)
    L, K = len(descs), 16
    checked = edge_then_hit = 0
    for s in range(ev["n_seeds"]):
        cands, rec = enumerate_paths(O, ev, descs, s, 3, 30.0, 4, 1)
        if rec >= K:
            continue
        got = oracle.find_clone_engine(_one_seed(ev, s), descs, 3, chi2_cut=30.0, max_cands=K)
        _compare(got, cands, L, K)
        checked += 1
        edge_then_hit += any(c.codes.get(4) == -3 and c.codes.get(5, -1) >= 0 for c in cands)
    assert checked >= 50 and edge_then_hit >= 5, (checked, edge_then_hit)


def test_apex_stop_before_a_disk():
    """The apex stop (MkBuilder.cc:1521-1533) applies whatever the next layer is: before an endcap disk,
    where the propagation to z does not fail as the one to the next barrel radius would, a candidate of
    pT 1 GeV at r = 30 cm whose momentum turns 1.45 rad from its position's azimuth still stops there."""
    ev, descs = _tiny_events(64, 4, 0.0, n_layers=5)
    disks = mk.disk_layer_descs(mk.geometry(4))
    descs = descs + [dict(dd) for dd in disks[:3]]
    L = len(descs)
    one = _one_seed(ev, 0)
    off = np.concatenate([ev["hit_offset"], np.full(L - 5, ev["hit_offset"][-1])]).astype(np.int32)
    one = dict(one, hit_offset=off)
    for z, theta in ((0.0, np.pi / 4), (0.0, 3 * np.pi / 4)):
        sp = one["seed_par"].copy()
        sp[:, 0] = [30.0, 0.0, z, 1.0, 1.45, theta]
        o = dict(one, seed_par=sp.astype(F), seed_hits=np.zeros((5, 1), np.int32))
        got = oracle.find_clone_engine(o, descs, 5, chi2_cut=30.0, max_cands=6)
        assert got["n_cands"][0] == 1 and got["hot_idx"][0, 5, 0] == -2, got["hot_idx"][0, :, 0]
        assert np.all(got["hot_idx"][0, 6:, 0] == -9)
        np.testing.assert_array_equal(got["par"][0, :, 0], sp[:, 0].astype(F))


def test_wrong_charge_seeds_flip_to_the_true_charge():
    """Seeds given the wrong charge, 1/pT = 0.01 and a 1/pT uncertainty of 1 / GeV: the updates pull 1/pT
    through zero, and UpdateWithLastHit's update moves the sign to the charge (KalmanUtilsMPlex.cc:497-504),
    so the best candidates end with the true charge, 1/pT >= 0 everywhere, and as many hits as with the
    right charge."""
    geo = mk.geometry(1)
    descs = mk.barrel_layer_descs(geo)
    ev = mk.generate_find_event(geo, mk.default_toy_params(seed=8, simulate_material=0, pt_min=2.0, pt_max=10.0),
                                300, noise_per_cm2=0.5)
    sp = ev["seed_par"].copy(); se = ev["seed_err"].copy()
    sp[3] = F(0.01)
    se[9] = F(1.0)
    wrong = dict(ev, seed_par=sp, seed_err=se, seed_chg=(-ev["seed_chg"]).astype(np.int32))
    r = oracle.find_clone_engine(wrong, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=8)
    K = r["par"].shape[0]
    valid = np.arange(K)[:, None] < r["n_cands"][None, :]
    assert np.all(r["par"][:, 3][valid] >= 0)
    assert (r["chg"][0] == ev["seed_chg"]).mean() > 0.95
    right = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=8)
    assert np.median(r["n_found"][0]) >= np.median(right["n_found"][0]) - 1


def _duplicate_hits(ev, n_layers):
    """Every hit of every layer appended once more after the layer's own hits (same values)."""
    off = ev["hit_offset"]
    chunks, mcs = [], []
    for l in range(n_layers):
        h = ev["hits_flat"][off[l]:off[l + 1]]
        chunks.append(np.concatenate([h, h]))
        m = ev["hit_mc"][off[l]:off[l + 1]]
        mcs.append(np.concatenate([m, m]))
    new_off = np.concatenate([[0], np.cumsum([len(c) for c in chunks])]).astype(np.int32)
    return dict(ev, hits_flat=np.concatenate(chunks), hit_offset=new_off, hit_mc=np.concatenate(mcs))


def test_duplicated_hits_resolve_to_first_copy():
    """Equal chi2 / equal score: best-hit keeps the first hit in scan order (chi2 < minChi2, MkFinder.cc:753),
    the clone engine with K = 1 the first record of equal score; the copies (indices past the layer's own
    hits) are never chosen and the results equal those on the event without copies."""
    ev, descs = _tiny_events(63, 300, 0.0, n_layers=17)
    L = len(descs)
    dup = _duplicate_hits(ev, L)
    a = oracle.find_best_hit(ev, descs, 3, chi2_cut=15.0)
    b = oracle.find_best_hit(dup, descs, 3, chi2_cut=15.0)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    a = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=1)
    b = oracle.find_clone_engine(dup, descs, 3, chi2_cut=30.0, max_cands=1)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # with K = 2 every first-copy path has its twin through the copy, scored equal and ranked after it
    c = oracle.find_clone_engine(dup, descs, 3, chi2_cut=30.0, max_cands=2)
    nl = np.diff(ev["hit_offset"])
    two = c["n_cands"] == 2
    assert two.mean() > 0.9
    assert np.all(c["score"][0, two] == c["score"][1, two])
    first, second = c["hot_idx"][0][:, two], c["hot_idx"][1][:, two]
    assert np.all(first < nl[:, None])
    moved = second != first
    assert np.all(second[moved] == first[moved] + nl[:, None].repeat(two.sum(), 1)[moved])


@pytest.mark.parametrize("case", [
    # (pT, momentum phi, stopped) at r = 30 cm, position phi 0: find_tracks_unroll_candidates stops a
    # candidate of pT < 1.2 beyond r = 25 cm whose momentum turns 1.371..4.512 rad away from its position's
    # azimuth (MkBuilder.cc:1521-1533).  Such a state also cannot be propagated to the next layer
    # (1/cos(1.371) > 5 fails the helix step, PropagationMPlex.icc), so for a lone candidate the stop and
    # the barrel-miss replica (MkBuilder.cc:1595-1609) give the same output: the candidate unchanged with
    # -2 on the layer and nothing after it.
    (1.0, 1.45, True),
    (1.0, -1.45, True),
    (1.0, 1.30, False),
    (0.8, 1.00, False),
])
def test_apex_stop_rule(case):
    pt, phimom, stopped = case
    ev, descs = _tiny_events(64, 4, 0.0, n_layers=8)
    one = _one_seed(ev, 0)
    sp = one["seed_par"].copy()
    sp[:, 0] = [30.0, 0.0, 0.0, 1.0 / pt, phimom, np.pi / 2]
    one["seed_par"] = sp.astype(F)
    one["seed_hits"] = np.zeros((5, 1), np.int32)
    got = oracle.find_clone_engine(one, descs, 5, chi2_cut=30.0, max_cands=6)
    first = got["hot_idx"][:got["n_cands"][0], 5, 0]
    if stopped:
        assert got["n_cands"][0] == 1 and first[0] == -2 and np.all(got["hot_idx"][0, 6:, 0] == -9)
        np.testing.assert_array_equal(got["par"][0, :, 0], one["seed_par"][:, 0])
    else:
        assert np.all(first != -2), first


def _gap_event(standard, masked=True):
    """One noise-free seed through 7 barrel layers, with (masked) its own hits on layers 4 and 5 masked (an
    earlier iteration took them), and the run of the combinatorial finder with optional dead regions."""
    ev, descs = _tiny_events(67, 20, 0.0, n_layers=7)
    s = 0
    off = ev["hit_offset"]
    own = [int(np.nonzero(ev["hit_mc"][off[l]:off[l + 1]] == ev["seed_mc"][s])[0][0]) for l in range(7)]
    mask = np.zeros(off[-1], np.uint8)
    if masked:
        mask[off[4] + own[4]] = mask[off[5] + own[5]] = 1
    one = dict(_one_seed(ev, s), hit_mask=mask)
    h4 = ev["hits_flat"][off[4] + own[4]]
    phi4, z4 = float(np.arctan2(h4[1], h4[0])), float(h4[2])
    find = oracle.find_standard if standard else oracle.find_clone_engine

    def run(dead):
        return find(dict(one, dead_regions=np.array(dead, np.float32).reshape(-1, 5)), descs, 3, chi2_cut=30.0,
                    max_cands=6, max_holes=4, max_consec_holes=1)
    return run, phi4, z4, own, ev, descs


@pytest.mark.parametrize("standard", [False, True])
def test_dead_module_gap(standard):
    """Dead modules (LayerOfHits::SuckInDeads, MkFinder.cc:451-462, 1092-1096 / 1263-1267): a candidate whose
    window's central q row crosses a dead bin and that takes no hit there gets -7 instead of a hole, which
    does not count toward the consecutive-hole limit.  The seed's own hits on layers 4 and 5 are masked,
    max_consec_holes 1: without a dead region it takes -1 on layer 4 and -2 on layer 5 (stop); with one over
    its path on layer 4 it takes -7 there and -1 on layer 5, and goes on to layer 6.  A dead region away
    from the window changes no candidate; without the mask, where the seed's own hit on layer 4 is found in
    the dead region, no candidate takes -7 (nHitsAdded > 0)."""
    run, phi4, z4, own, ev, descs = _gap_event(standard)
    best = lambda r: r["hot_idx"][0, :, 0]
    plain = best(run([]))
    assert plain[4] == -1 and plain[5] == -2 and plain[6] == -9, plain
    gap = best(run([[4, phi4 - 0.05, phi4 + 0.05, z4 - 5.0, z4 + 5.0]]))
    assert gap[4] == -7 and gap[5] == -1 and gap[6] != -9, gap
    all_plain = run([])["hot_idx"]
    away = run([[4, phi4 - 0.05, phi4 + 0.05, z4 + 20.0, z4 + 30.0]])["hot_idx"]
    assert np.array_equal(away, all_plain)
    # unmasked: the hit on layer 4 is found inside the dead region, so no candidate takes -7 there (the one
    # continuing from the invalid hit keeps -1 and takes the hits after it)
    run_u = _gap_event(standard, masked=False)[0]
    unmasked = run_u([])["hot_idx"]
    assert (unmasked[:, 4, 0] == -1).any(), unmasked[:, :, 0]
    found = run_u([[4, phi4 - 0.05, phi4 + 0.05, z4 - 5.0, z4 + 5.0]])["hot_idx"]
    assert np.array_equal(found, unmasked)
