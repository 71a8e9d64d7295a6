"""Pins of the oracle's clone-engine finding (FindTracksCloneEngine, CandCloner, UpdateWithLastHit).

The reference holds no vectors for finding; the C++ restatement is checked against
  * a per-seed numpy composition of the same layer loop built from the oracle's separately pinned
    primitives (propagation, Kalman chi2/update, SelectHitIndices), with the candidate bookkeeping
    (IdxChi2List scores, extras, best short candidate, final merge) written out independently;
  * invariants of the output (sorted by score, hit counts, distinct candidates);
  * physics: on noise-free events the best candidate carries the generating track's hits, and with
    noise it finds more true hits than best-hit finding.
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle

F = np.float32


def _score(nfound, ntail, noverlap, nmiss, chi2, pt, in_find):
    """getScoreCalc (Track.h:648-672) with the Config.h:355-359 constants."""
    bonus = F(0.2) * F(nfound) + F(4)
    penalty = F(8)
    if pt < 0.9:
        penalty = penalty * (F(1.7) if in_find else F(1.5))
        bonus = min(bonus * (F(0.9) if in_find else F(1.0)), F(8))
    return F(bonus * F(nfound) + F(0) * F(noverlap) - penalty * F(nmiss) - F(3) * F(ntail) - F(max(chi2, 0)))


class _Cand:
    def __init__(self, par, err, chg):
        self.par, self.err, self.chg = par.astype(F), err.astype(F), int(chg)
        self.chi2 = F(0); self.score = F(-1e16)
        self.nfound = self.nmissing = self.ninside = self.ntail = 0
        self.hots = []                     # (layer, index), oldest first

    def copy(self):
        c = _Cand(self.par.copy(), self.err.copy(), self.chg)
        c.chi2, c.score, c.nfound, c.nmissing, c.ninside, c.ntail = (self.chi2, self.score, self.nfound,
                                                                     self.nmissing, self.ninside, self.ntail)
        c.hots = list(self.hots)
        return c

    def add_hit(self, idx, layer, chi2):   # TrackCand::addHitIdx (HitStructures.h:873-889)
        self.hots.append((layer, idx))
        if idx >= 0 or idx == -9:
            self.nfound += 1; self.chi2 = F(self.chi2 + F(chi2))
            self.ninside += self.ntail; self.ntail = 0
        else:
            self.nmissing += 1
            if idx == -1:
                self.ntail += 1

    def last_idx(self):
        return self.hots[-1][1] if self.hots else -1

    def pt(self):
        return abs(F(1) / self.par[3])


def _ce_composed(O, ev, descs, first_layer, chi2_cut, K, max_holes=4, max_consec=1, bfield=3.8):
    n = ev["n_seeds"]
    off, hits = ev["hit_offset"], ev["hits_flat"]
    seeds = []
    for s in range(n):
        c = _Cand(ev["seed_par"][:, s], ev["seed_err"][:, s], ev["seed_chg"][s])
        for l in range(first_layer):
            c.add_hit(int(ev["seed_hits"][l, s]), l, 0.0)
        c.score = _score(c.nfound, 0, 0, c.ninside, c.chi2, c.pt(), False)
        seeds.append(dict(cands=[c], best_short=None, state=0, pickup=first_layer - 1))
    prev = first_layer - 1
    fake_err = np.array([666, 0, 666, 0, 0, 666], F)
    for layer in range(first_layer, len(descs)):
        d = descs[layer]
        hl = hits[off[layer]:off[layer + 1]]
        work = []                          # (seed, cand index) in seed order
        for s, sd in enumerate(seeds):
            if sd["state"] == 0 and sd["pickup"] == prev:
                sd["state"] = 1
            if sd["state"] != 1:
                continue
            active = False
            for ic, c in enumerate(sd["cands"]):
                if c.last_idx() == -2:
                    continue
                if c.pt() < 1.2:
                    dphi = abs(np.arctan2(c.par[1], c.par[0]) - c.par[4])
                    if c.par[0] * c.par[0] + c.par[1] * c.par[1] > F(625) and F(1.371) < dphi < F(4.512):
                        c.add_hit(-2, layer, 0.0)
                        continue
                active = True
                work.append((s, ic))
            if not active:
                sd["state"] = 2
        prev = layer
        if not work:
            continue
        m = len(work)
        cands = [seeds[s]["cands"][ic] for s, ic in work]
        par = np.stack([c.par for c in cands], 1); err = np.stack([c.err for c in cands], 1)
        chg = np.array([c.chg for c in cands], np.int32)
        errP, parP, _ = O.propagate_to_r(err, par, chg, np.full(m, d["propagate_to"], F), pflags=0)
        xs, xa, wsr = oracle.select_hit_indices(d, hl, parP, errP)
        extras = {}
        for k, (s, ic) in enumerate(work):
            if np.hypot(parP[0, k], parP[1, k]) < F(d["rin"]):
                xs[k] = 0; wsr[k] = 2
                e = cands[k].copy(); e.add_hit(-2, layer, 0.0)
                extras.setdefault(s, []).append(e)
            elif wsr[k] == 2:
                extras.setdefault(s, []).append(cands[k].copy())
        entries = {}
        for hc in range(16):
            sel = xs > hc
            if not sel.any():
                break
            ms_par = np.zeros((3, m), F); ms_err = np.ones((6, m), F)
            ms_par[:, sel] = hl[xa[hc, sel], :3].T; ms_err[:, sel] = hl[xa[hc, sel], 3:].T
            rad = np.hypot(ms_par[0], ms_par[1]).astype(F)
            pe, pp, _ = O.propagate_to_r(errP, parP, chg, rad, pflags=0)
            _, _, c2 = O.kalman_operation(1, pe, pp, ms_err, ms_par)
            for k in np.nonzero(sel)[0]:
                s, ic = work[k]
                c2k = abs(F(c2[k]))
                if not c2k < F(chi2_cut):
                    continue
                res = abs(ms_par[2, k] - pp[2, k])               # isStripQCompatible, barrel
                if not np.sqrt(ms_err[5, k] * F(3)) + max(F(3) * np.sqrt(errP[5, k]), F(0.5)) > res:
                    continue
                c = cands[k]
                sc = _score(c.nfound + 1, 0, 0, c.ninside + c.ntail, F(c.chi2 + c2k), abs(F(1) / parP[3, k]), True)
                entries.setdefault(s, []).append((sc, ic, int(xa[hc, k]), c2k))
        for k, (s, ic) in enumerate(work):
            if wsr[k] == 2:
                continue
            c = cands[k]
            fake = -1 if (c.ninside + c.ntail < max_holes and c.ntail < max_consec) else -2
            if wsr[k] == 1:
                fake = -3
            nt = c.ntail + 1 if fake == -1 else c.ntail
            sc = _score(c.nfound, nt, 0, c.ninside, c.chi2, abs(F(1) / parP[3, k]), True)
            entries.setdefault(s, []).append((sc, ic, fake, F(0)))
        updates = []
        for s in sorted(set(s for s, _ in work)):
            sd = seeds[s]
            ex = extras.get(s, [])
            es = sorted(entries.get(s, []), key=lambda e: -e[0])        # stable
            if not es:
                sd["cands"] = ex
                continue
            cv, xi = [], 0
            for sc, ic, hit, c2 in es[:min(len(es), K)]:
                tc = sd["cands"][ic].copy(); tc.add_hit(hit, layer, c2); tc.score = sc
                if hit == -2:
                    if sd["best_short"] is None or sc > sd["best_short"].score:
                        sd["best_short"] = tc
                    continue
                while xi < len(ex) and ex[xi].score > tc.score and len(cv) < K:
                    cv.append(ex[xi]); xi += 1
                if len(cv) >= K:
                    break
                cv.append(tc)
                if hit >= 0:
                    updates.append(tc)
            while xi < len(ex) and len(cv) < K:
                cv.append(ex[xi]); xi += 1
            sd["cands"] = cv
        if updates:
            u = len(updates)
            par = np.stack([c.par for c in updates], 1); err = np.stack([c.err for c in updates], 1)
            chg = np.array([c.chg for c in updates], np.int32)
            hidx = np.array([c.last_idx() for c in updates])
            ms_par = np.ascontiguousarray(hl[hidx, :3].T); ms_err = np.ascontiguousarray(hl[hidx, 3:].T)
            pe, pp, _ = O.propagate_to_r(err, par, chg, np.hypot(ms_par[0], ms_par[1]).astype(F), pflags=0)
            e2, p2, _ = O.kalman_operation(2, pe, pp, ms_err, ms_par)
            for k, c in enumerate(updates):
                c.par, c.err = p2[:, k].copy(), e2[:, k].copy()
                if c.par[3] < 0:
                    c.chg = -c.chg; c.par[3] = -c.par[3]
    out = []
    for sd in seeds:
        cs, bs = sd["cands"], sd["best_short"]
        if cs:
            for c in cs:
                c.score = _score(c.nfound, 0, 0, c.ninside, c.chi2, c.pt(), False)
            if bs is not None:
                bs.score = _score(bs.nfound, 0, 0, bs.ninside, bs.chi2, bs.pt(), False)
            cs = sorted(cs, key=lambda c: -c.score)
            if bs is not None and bs.score > cs[-1].score:
                ci = 0
                while cs[ci].score > bs.score:
                    ci += 1
                if len(cs) >= K:
                    cs.pop()
                cs.insert(ci, bs)
        elif bs is not None:
            cs = [bs]
        out.append(cs)
    return out


def _event(seed, n, noise, pt_min=0.5):
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=seed, simulate_material=0, pt_min=pt_min, pt_max=10.0)
    return mk.generate_find_event(geo, p, n, noise_per_cm2=noise), mk.barrel_layer_descs(geo)


def _hot_rows(cands, L):
    rows = []
    for c in cands:
        r = np.full(L, -9, np.int32)
        for layer, idx in c.hots:
            if r[layer] == -9:
                r[layer] = idx
        rows.append(r)
    return rows


def _true_fraction(hot, ev, first_layer):
    h = hot[first_layer:]; off = ev["hit_offset"][first_layer:-1]
    ok = h >= 0
    truth = np.where(ok, ev["hit_mc"][off[:, None] + np.maximum(h, 0)], -5)
    return (truth == ev["seed_mc"][None, :]).sum() / h.size


@pytest.mark.parametrize("K", [1, 6])
def test_ce_matches_composition(K):
    ev, descs = _event(41, 250, 2.0)
    got = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=K, n_threads=4)
    ref = _ce_composed(oracle.Oracle(bfield=3.8), ev, descs, 3, 30.0, K)
    L = len(descs)
    same = 0
    for s in range(ev["n_seeds"]):
        rows = _hot_rows(ref[s], L)
        ok = got["n_cands"][s] == len(rows) and all(
            np.array_equal(got["hot_idx"][i, :, s], rows[i]) for i in range(len(rows)))
        if ok:
            same += 1
            for i, c in enumerate(ref[s]):
                assert got["n_found"][i, s] == c.nfound and got["chg"][i, s] == c.chg
                np.testing.assert_allclose(got["par"][i, :, s], c.par, rtol=2e-4, atol=2e-5)
                np.testing.assert_allclose(got["score"][i, s], c.score, rtol=1e-5, atol=1e-3)
    assert same == ev["n_seeds"], same


def test_ce_invariants():
    ev, descs = _event(42, 1500, 2.0)
    K = 6
    r = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=K, n_threads=8)
    r1 = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=K, n_threads=1)
    for k in r:
        assert np.array_equal(r[k], r1[k]), k
    nlh = np.diff(ev["hit_offset"])
    for s in range(ev["n_seeds"]):
        nc = r["n_cands"][s]
        assert 1 <= nc <= K
        sc = r["score"][:nc, s]
        assert np.all(sc[:-1] >= sc[1:])
        hots = r["hot_idx"][:nc, :, s]
        assert np.all(hots < nlh[None, :])
        assert np.array_equal(r["n_found"][:nc, s], np.sum(hots >= 0, axis=1))
        assert len({tuple(h) for h in hots}) == nc            # candidates are distinct hit lists
        assert np.all(r["chi2"][:nc, s] >= 0)


def test_ce_noise_free_true_hits():
    ev, descs = _event(43, 1500, 0.0, pt_min=1.0)
    r = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=8)
    assert _true_fraction(r["hot_idx"][0], ev, 3) > 0.985


def test_ce_beats_best_hit_with_noise():
    ev, descs = _event(44, 1500, 2.0)
    ce = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=8)
    bh = oracle.find_best_hit(ev, descs, 3, chi2_cut=30.0, n_threads=8)
    f_ce, f_bh = _true_fraction(ce["hot_idx"][0], ev, 3), _true_fraction(bh["hot_idx"], ev, 3)
    assert f_ce > 0.95 and f_ce > f_bh + 0.02, (f_ce, f_bh)
