"""Per-track parity of a fit result against the FP32 and FP64 oracle -- TEST INFRASTRUCTURE.

The contract's tolerances are relative to the reference's FP32 CPU result:
parameters 1e-5, covariance 1e-4, chi2 1e-4.  FP32 itself cannot hold them
on every track: the reference's own FP32 result sits further than that from
the exact (FP64) fit on ill-conditioned tracks (a Jacobian element 1 + (~-1)
near phi = +-90 deg, a drifting low-pT fit), and any other FP32 rounding
order lands elsewhere on those.  Such tracks are *flagged*, one by one, by
the oracle alone: a track is flagged when its FP32 oracle result is beyond a
tolerance of its FP64 oracle result in any parameter, covariance element or
chi2.  Every element of every unflagged track must then be within its
tolerance of the FP32 oracle or, where it is not, at least as near the FP64
value as the FP32 oracle is; flagged tracks are counted and reported with
their errors, never silently accepted.

Metrics (as in the parity tests):
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


def flagged_tracks(cpu32, cpu64, tol=(TOL_PAR, TOL_COV, TOL_CHI2)):
    """Boolean [n]: tracks whose FP32 oracle result is beyond a tolerance of the FP64 oracle in some element."""
    c64 = _elem_errors(cpu32, cpu64)
    flagged = np.zeros(c64[0].shape[1], bool)
    for k in range(3):
        flagged |= (c64[k] > tol[k]).any(axis=0)
    return flagged


def flagged_parity(gpu, cpu32, cpu64, tol=(TOL_PAR, TOL_COV, TOL_CHI2)):
    """gpu, cpu32, cpu64: (par [6,n], err [21,n], chi2 [n] or None).

    A track is *flagged* (FP32 conditioning) when the FP32 oracle itself is
    beyond a tolerance of the FP64 oracle in any element.  An element of an
    unflagged track passes when the GPU value is within the tolerance of the
    FP32 oracle, or -- where it is not -- when the GPU value is at least as
    near the FP64 value as the FP32 oracle's is (the two FP32 roundings sit
    on opposite sides of the exact value; the GPU is then no less accurate
    than the reference).  `n_unexplained` counts unflagged tracks with an
    element that passes neither way: it must be 0.

    Returns counts, the largest errors over unflagged and over all tracks,
    and the flagged tracks' distances to FP64 (GPU beside FP32 oracle)."""
    g32 = _elem_errors(gpu, cpu32)
    c64 = _elem_errors(cpu32, cpu64)
    g64 = _elem_errors(gpu, cpu64)
    n = g32[0].shape[1]
    flagged = np.zeros(n, bool)
    within = np.ones(n, bool)
    passed = np.ones(n, bool)
    for k in range(3):
        flagged |= (c64[k] > tol[k]).any(axis=0)
        w = g32[k] <= tol[k]
        within &= w.all(axis=0)
        passed &= (w | (g64[k] <= c64[k])).all(axis=0)
    unf = ~flagged
    names = ("par", "cov", "chi2")
    mx = lambda e, m: float(e.max(axis=0)[m].max()) if m.any() else 0.0
    res = {
        "n": int(n),
        "n_flagged": int(flagged.sum()),
        "n_unflagged_within_tol": int(np.sum(unf & within)),
        "n_unflagged_nearer_fp64": int(np.sum(unf & ~within & passed)),
        "n_unexplained": int(np.sum(unf & ~passed)),
        "unflagged_max": {nm: mx(g32[k], unf) for k, nm in enumerate(names)},
        "all_max": {nm: mx(g32[k], np.ones(n, bool)) for k, nm in enumerate(names)},
        "frac_tracks_within_tol": float(within.mean()),
        "flagged_vs_fp64_max": {"gpu": {nm: mx(g64[k], flagged) for k, nm in enumerate(names)},
                                "cpu32": {nm: mx(c64[k], flagged) for k, nm in enumerate(names)}},
        "flag_rule": "flagged: the FP32 oracle beyond a tolerance of the FP64 oracle in some element",
        "unexplained_idx": np.where(unf & ~passed)[0][:20].tolist(),
    }
    return res


def describe(res):
    return ("%d tracks: %d flagged (FP32 conditioning); unflagged: %d within tolerance of the FP32 oracle, %d nearer "
            "the FP64 fit than it, %d unexplained; unflagged max par %.3g cov %.3g chi2 %.3g; all-track max par %.3g" % (
                res["n"], res["n_flagged"], res["n_unflagged_within_tol"], res["n_unflagged_nearer_fp64"],
                res["n_unexplained"], res["unflagged_max"]["par"], res["unflagged_max"]["cov"],
                res["unflagged_max"]["chi2"], res["all_max"]["par"]))
