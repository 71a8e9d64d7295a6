"""Shared helpers of the finding tests: layer descriptions, random hits, brute-force selection."""
import numpy as np

import oracle

TWO_PI = np.float32(2 * np.pi)
F_PHI_FINE = np.float32(1024 / np.float32(2 * np.pi))


def _layer(**kw):
    d = dict(is_barrel=1, rin=19.99, rout=20.01, zmin=-60.0, zmax=60.0, propagate_to=20.0, q_bin=2.0,
             min_dphi=0.002, max_dphi=0.1, min_dq=0.3, max_dq=2.0, c_dp_sf=1.1, c_dq_sf=1.1, c_c2_sf=1.1)
    d.update(kw)
    return d


def _random_hits(rng, n, r=20.0, zmax=62.0, ties=0):
    phi = rng.uniform(-np.pi, np.pi, n)
    z = rng.uniform(-zmax, zmax, n)
    if ties:   # duplicate positions so that equal keys must keep their input order
        k = rng.integers(0, n, ties)
        phi[-ties:] = phi[k]
        z[-ties:] = z[k]
    h = np.zeros((n, 9), np.float32)
    h[:, 0] = r * np.cos(phi); h[:, 1] = r * np.sin(phi); h[:, 2] = z
    h[:, 3] = 1e-6; h[:, 5] = 1e-6; h[:, 8] = 1e-4
    return h


def _keys(desc, phis, qs):
    nq = np.int32(0)
    _, _, _, _, nq = oracle.loh_build(desc, np.zeros((0, 9), np.float32))
    # setup_bins (HitStructures.cc:11-32), bin width q_bin > 0
    qmin, qmax, dq = np.float32(desc["zmin"]), np.float32(desc["zmax"]), np.float32(desc["q_bin"])
    extent = qmax - qmin
    nqb = int(np.ceil(extent / dq))
    extra = np.float32(0.5) * (np.float32(nqb) * dq - extent)
    fq = np.float32(nqb) / (qmax - qmin)
    qb = np.clip(((qs - (qmin - extra)) * fq).astype(np.int64), 0, nqb - 1)
    pf = np.floor(F_PHI_FINE * (phis + np.float32(np.pi))).astype(np.int64)
    return pf + (qb << 16), nqb


def _cdist(a):
    return np.where(a > np.float32(np.pi), TWO_PI - a, a)


def _brute_select(desc, h, par, err, ranks, phis, qs):
    """Window test over every hit; returns per-track lists in the reference's scan order."""
    n = par.shape[1]
    key, nqb = _keys(desc, phis, qs)
    pos_of = np.empty(len(h), np.int64)
    pos_of[ranks] = np.arange(len(h))
    out, wins = [], []
    for t in range(n):
        x, y, z = par[0, t], par[1, t], par[2, t]
        r2 = x * x + y * y
        dphidx, dphidy = -y / r2, x / r2
        dphi2 = dphidx * dphidx * err[0, t] + dphidy * dphidy * err[2, t] + 2 * dphidx * dphidy * err[1, t]
        phi = np.arctan2(y, x)
        dphi = np.clip(max(3 * np.sqrt(abs(dphi2)), desc["min_dphi"]), desc["min_dphi"], desc["max_dphi"])
        dq = np.clip(abs(3 * np.sqrt(err[5, t])), desc["min_dq"], desc["max_dq"])
        ddq = np.abs(z - h[:, 2])
        ddphi = _cdist(np.abs(phi - np.arctan2(h[:, 1], h[:, 0])))
        inside = np.nonzero((ddq < dq) & (ddphi < dphi))[0]
        # scan order: q bin, then phi bin counted from the window's first phi bin, then sorted position
        pb1 = int(np.floor(F_PHI_FINE * (phi - dphi + np.float32(np.pi)))) >> 2
        kk = key[pos_of[inside]]
        order = np.lexsort((pos_of[inside], (((kk & 0x3ff) >> 2) - pb1) % 256, kk >> 16))
        out.append(inside[order])
        wins.append((dq, dphi))
    return out, wins



def assert_bins(b_g, b_c, tag=None):
    """The device bin table against the oracle's (the reference's m_phi_bin_infos): every bin with hits has
    the same [first, end); a bin without hits is empty on the device too, and there bins 1..255 of a row
    hold their insertion point, so that those bins tile the row (bin b ends where bin b + 1 starts) and a
    run of them is one contiguous range."""
    g = b_g.reshape(-1, 256, 2)
    c = b_c.reshape(-1, 256, 2)
    full = c[..., 1] > c[..., 0]
    assert np.array_equal(g[full], c[full]), tag
    assert np.all(g[~full][:, 0] == g[~full][:, 1]), tag
    assert np.all(g[:, 1:255, 1] == g[:, 2:256, 0]), tag


def edge_explained(desc_win, par_t, h, diff, eps=2e-5):
    """True if every hit in `diff` lies within eps of the track's dq / dphi window edge."""
    dq, dphi = desc_win
    ph = np.arctan2(h[diff, 1], h[diff, 0])
    near_q = np.abs(np.abs(par_t[2] - h[diff, 2]) - dq) < eps
    near_p = np.abs(_cdist(np.abs(np.arctan2(par_t[1], par_t[0]) - ph)) - dphi) < eps
    return bool(np.all(near_q | near_p))


def pad_seed_hits(ev, first_layer):
    """The event with seed_hits extended to first_layer rows; layers past the seed layers hold -1 (no hit),
    for searches that start beyond the seed layers."""
    sh = ev["seed_hits"]
    if sh.shape[0] >= first_layer:
        return ev
    pad = np.full((first_layer - sh.shape[0], sh.shape[1]), -1, np.int32)
    return dict(ev, seed_hits=np.ascontiguousarray(np.vstack([sh, pad]), np.int32))


# Decision-margin thresholds for explaining a GPU / oracle difference in finding.  The GPU's propagated
# states differ from the oracle's by FP32 rounding (positions ~1e-6 relative, chi2 ~1e-5): a decision whose
# margin lies below these can go either way (chi2: the states differ after several updates, so chi2 values
# of one hit move by up to ~2e-3 of themselves between FP32 implementations).
WIN_EPS = 1e-4      # cm: a hit this close to a window edge
BIN_EPS = 1e-4      # bin fraction: a window limit or a hit's phi this close to a bin edge
CHI2_EPS = 5e-3     # relative: two chi2 values, or a chi2 and the cut, this close
SCORE_EPS = 5e-4    # two candidate scores this close, relative to the larger of their chi2 sums


def _classes(m, l, t, ce):
    out = []
    if m[0, l, t] < WIN_EPS:
        out.append("window")
    if m[1, l, t] < BIN_EPS:
        out.append("bin")
    if m[2, l, t] < CHI2_EPS:
        out.append("chi2")
    if ce and m[3, l, t] < SCORE_EPS:
        out.append("score")
    return out


def explain_best_hit(gpu_hot, cpu_hot, margins, first_layer):
    """Per seed whose hit list differs ([n_layers][n] each), the first differing layer and the decision there
    that the oracle took within a margin of rounding (window edge, bin edge, chi2 tie or cut).  Returns
    (n_differing, {class: count}, unexplained [(seed, layer)])."""
    diff = np.nonzero(np.any(gpu_hot != cpu_hot, axis=0))[0]
    counts, bad = {}, []
    for t in diff:
        l = int(np.nonzero(gpu_hot[:, t] != cpu_hot[:, t])[0][0])
        cls = _classes(margins, l, t, False) if l >= first_layer else []
        if not cls:
            bad.append((int(t), l))
        for c in cls[:1]:
            counts[c] = counts.get(c, 0) + 1
    return len(diff), counts, bad


def explain_clone_engine(gpu, cpu, first_layer, alt_lists=()):
    """As explain_best_hit for the clone engine: gpu / cpu dicts with n_cands and hot_idx [K][L][n], cpu with the
    oracle's margins [4][L][n] and merge_gap [n].  A seed differs when its candidate lists (in order) differ.
    Explained when the lists equal those of the oracle with its helix-step threshold ties taken the other way
    (alt_lists: (n_cands, hot_idx) of such runs).  Same set in another order: explained by a tie of the final
    merge.  Otherwise the first layer where the candidate sets differ (a hit-list prefix of one side missing
    on the other) must hold a decision within rounding (window, bin, chi2, score)."""
    n = cpu["n_cands"].shape[0]
    m = cpu["margins"]
    counts, bad, n_diff = {}, [], 0
    for t in range(n):
        ng, nc = int(gpu["n_cands"][t]), int(cpu["n_cands"][t])
        rg = [tuple(gpu["hot_idx"][i, :, t]) for i in range(ng)]
        rc = [tuple(cpu["hot_idx"][i, :, t]) for i in range(nc)]
        if rg == rc:
            continue
        n_diff += 1
        if any(int(an[t]) == ng and all(tuple(ah[i, :, t]) == rg[i] for i in range(ng)) for an, ah in alt_lists):
            counts["tie"] = counts.get("tie", 0) + 1
            continue
        if sorted(rg) == sorted(rc):
            # the order comes from the final merge's sort, whose equal scores keep the order of the last
            # layer's selection: a tie in either
            if cpu["merge_gap"][t] < SCORE_EPS or np.any(m[3, :, t] < SCORE_EPS):
                counts["merge"] = counts.get("merge", 0) + 1
            else:
                bad.append((t, -1))
            continue
        L = cpu["hot_idx"].shape[1]
        first = L
        for l in range(first_layer, L + 1):
            if {r[:l] for r in rg} != {r[:l] for r in rc}:
                first = l - 1
                break
        # the candidate lists are kept K at a time, so a path that split off at `first` can be dropped by
        # the selection of any later layer: the decision within rounding may lie at `first` or after it
        cls = []
        for l in range(max(first, first_layer), L):
            cls = _classes(m, l, t, True)
            if cls:
                break
        if not cls:
            bad.append((t, first))
        for c in cls[:1]:
            counts[c] = counts.get(c, 0) + 1
    return n_diff, counts, bad


def _state_spread(cpu, variants):
    """[3, n] per seed and group (parameters, covariance, chi2): the largest distance, in tolerance units, of a
    rounding variant's final state from the oracle's, over the variants that found the same hit list."""
    from oracle.parity import group_error
    s = np.zeros((3, cpu["hot_idx"].shape[-1]))
    for v in variants:
        same = np.all(v["hot_idx"] == cpu["hot_idx"], axis=0)
        g = group_error((v["par"], v["err"], v["chi2"]), (cpu["par"], cpu["err"], cpu["chi2"]))
        s = np.maximum(s, np.where(same[None, :], g, 0.0))
    return s


def _variants(run, ev, n_variants, seed0):
    """The finder under legal changes of rounding: the no-FMA build, then one-ulp seed perturbations
    alternating the builds."""
    from oracle.parity import perturbed_seeds
    out = []
    for i in range(n_variants):
        e = perturbed_seeds(ev, seed0 + i) if i else ev
        if i % 2 == 0:
            with oracle.use_lib(oracle.default_lib_path(contract=False)):
                out.append(run(e))
        else:
            out.append(run(e))
    return out


def _alternatives(run, ev):
    """The finder with the helix step's threshold ties taken the other way (oracle/parity.py): the threshold
    moved by 1, 2, 4 and 16 float spacings either way (a finder's states pass more helix steps than a fit's,
    and no per-track margins are recorded for them)."""
    from oracle.parity import TIE_SHIFTS
    out = []
    for s in TIE_SHIFTS + (16.0, -16.0):
        with oracle.tie_shift(s):
            out.append(run(ev))
    return out


def best_hit_reference(ev, descs, first_layer, n_variants=16, **kw):
    """The oracle's best-hit result with its decision margins, and the checker data of its final states: the
    per-seed rounding spread and the results with threshold ties taken the other way (alt)."""
    cpu = oracle.find_best_hit(ev, descs, first_layer, margins=True, **kw)
    run = lambda e: oracle.find_best_hit(e, descs, first_layer, **kw)
    spread = _state_spread(cpu, _variants(run, ev, n_variants, 500))
    alt = []
    for a in _alternatives(run, ev):
        same = np.all(a["hot_idx"] == cpu["hot_idx"], axis=0)
        alt.append(tuple(np.where(same, a[k], cpu[k]) for k in ("par", "err", "chi2")))
    return cpu, {"spread": spread, "alt": alt}


def assert_find_parity(name, gpu, cpu, spread, first_layer):
    """Best-hit finding: every seed's hit list equals the oracle's, or its first difference is a decision the
    oracle took within rounding (explain_best_hit); seeds with equal lists have final states within tolerance
    of the oracle's, or within four times their rounding spread where it reaches a third of it (at most
    0.02 % of the seeds beyond, by at most 2x: oracle/parity.py).  Returns (same mask, summary line)."""
    from oracle.parity import describe, flagged_parity
    n_diff, counts, bad = explain_best_hit(gpu["hot_idx"], cpu["hot_idx"], cpu["margins"], first_layer)
    same = np.all(gpu["hot_idx"] == cpu["hot_idx"], axis=0)
    sel = lambda x: x[..., same]
    chk = {"spread": spread["spread"][:, same], "tie": None, "alt": [tuple(sel(x) for x in a) for a in spread["alt"]],
           "tie_all": True}
    r = flagged_parity((gpu["par"][:, same], gpu["err"][:, same], gpu["chi2"][same]),
                       (cpu["par"][:, same], cpu["err"][:, same], cpu["chi2"][same]),
                       (cpu["par"][:, same], cpu["err"][:, same], cpu["chi2"][same]), chk)
    line = "%s: %d seeds, %d differ (%s), unexplained %s; equal lists: %s" % (
        name, len(same), n_diff, counts, bad[:10], describe(r))
    print(line)
    assert not bad, line
    assert r["ok"], (line, r["beyond"])
    assert np.array_equal(gpu["n_found"][same], cpu["n_found"][same])
    assert np.array_equal(gpu["chg"][same], cpu["chg"][same])
    return same, line


def clone_engine_reference(ev, descs, first_layer, n_variants=16, **kw):
    """The oracle's clone-engine result with its decision margins, and the checker data of its final candidate
    states: per candidate slot and seed the rounding spread [K][3][n] (variants as in best_hit_reference; a
    variant counts for a seed when its candidate lists equal the oracle's) and the results with threshold
    ties taken the other way (alt)."""
    from oracle.parity import group_error
    cpu = oracle.find_clone_engine(ev, descs, first_layer, margins=True, **kw)
    K, n = cpu["par"].shape[0], cpu["n_cands"].shape[0]
    run = lambda e: oracle.find_clone_engine(e, descs, first_layer, **kw)
    same_as = lambda v: (v["n_cands"] == cpu["n_cands"]) & np.all(v["hot_idx"] == cpu["hot_idx"], axis=(0, 1))
    spread = np.zeros((K, 3, n))
    for v in _variants(run, ev, n_variants, 700):
        same = same_as(v)
        for k in range(K):
            g = group_error((v["par"][k], v["err"][k], v["chi2"][k]), (cpu["par"][k], cpu["err"][k], cpu["chi2"][k]))
            spread[k] = np.maximum(spread[k], np.where((same & (cpu["n_cands"] > k))[None, :], g, 0.0))
    alt, lists = [], []
    for a in _alternatives(run, ev):
        # per candidate: the alternative's candidate in the same slot with the same hit list
        row = np.all(a["hot_idx"] == cpu["hot_idx"], axis=1) & (a["n_cands"] > np.arange(K)[:, None])
        alt.append({"par": np.where(row[:, None, :], a["par"], cpu["par"]),
                    "err": np.where(row[:, None, :], a["err"], cpu["err"]),
                    "chi2": np.where(row, a["chi2"], cpu["chi2"])})
        lists.append((a["n_cands"], a["hot_idx"]))
    return cpu, {"spread": spread, "alt": alt, "alt_lists": lists}


def assert_ce_parity(name, gpu, cpu, spread, first_layer):
    """Clone-engine finding: every seed's candidate lists equal the oracle's, or their difference is explained
    by a decision within rounding (explain_clone_engine); the candidates of seeds with equal lists have final
    states within tolerance, or within four times their rounding spread where it reaches a third of it (at
    most 0.02 % beyond, by at most 2x: oracle/parity.py).  Returns (same mask, summary line)."""
    from oracle.parity import describe, flagged_parity
    n_diff, counts, bad = explain_clone_engine(gpu, cpu, first_layer, spread.get("alt_lists", ()))
    same = (gpu["n_cands"] == cpu["n_cands"]) & np.all(gpu["hot_idx"] == cpu["hot_idx"], axis=(0, 1))
    K = cpu["par"].shape[0]
    cols = [(k, np.nonzero(same & (cpu["n_cands"] > k))[0]) for k in range(K)]
    cat = lambda d, key: np.concatenate([d[key][k][..., idx] for k, idx in cols], axis=-1)
    g = (cat(gpu, "par"), cat(gpu, "err"), cat(gpu, "chi2"))
    c = (cat(cpu, "par"), cat(cpu, "err"), cat(cpu, "chi2"))
    sp = np.concatenate([spread["spread"][k][:, idx] for k, idx in cols], axis=-1)
    alt = [(cat(a, "par"), cat(a, "err"), cat(a, "chi2")) for a in spread["alt"]]
    r = flagged_parity(g, c, c, {"spread": sp, "tie": None, "alt": alt, "tie_all": True})
    line = "%s: %d seeds, %d differ (%s), unexplained %s; candidates of equal seeds: %s" % (
        name, len(same), n_diff, counts, bad[:10], describe(r))
    print(line)
    assert not bad, line
    assert r["ok"], (line, r["beyond"])
    for k, idx in cols:
        assert np.array_equal(gpu["n_found"][k, idx], cpu["n_found"][k, idx])
        assert np.array_equal(gpu["chg"][k, idx], cpu["chg"][k, idx])
    return same, line
