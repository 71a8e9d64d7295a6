"""The FP64 covariance fallback of the fit (ORC_FIT_FP64_FALLBACK / MKF_FIT_FP64_FALLBACK).

After each FP32 update the covariance is tested (negative or NaN variance, or C_ij^2 > (1+1e-5) C_ii C_jj);
where the test fails, the update's covariance is recomputed in the Joseph form in FP64 from the FP32
pre-update state.  Pinned here against an independent numpy evaluation of the Joseph form.
"""
import numpy as np

import mkfit_amd as mk
import oracle

SYM = [[(i * (i + 1) // 2 + j) if i >= j else (j * (j + 1) // 2 + i) for j in range(6)] for i in range(6)]


def _full(e):
    return np.array([[e[SYM[i][j]] for j in range(6)] for i in range(6)], np.float64)


def _ill(c):
    for i in range(6):
        if not c[SYM[i][i]] >= 0:
            return True
    for i in range(1, 6):
        for j in range(i):
            if float(c[SYM[i][j]]) ** 2 > float(c[SYM[i][i]]) * float(c[SYM[j][j]]) * 1.00001:
                return True
    return False


def _joseph(C, p, m, v):
    """Barrel update covariance (I-KH) C (I-KH)^T + K V K^T in float64 (tangent-plane projection)."""
    r = np.hypot(float(m[0]), float(m[1]))
    a00, a01 = -(float(m[1]) + float(p[1])) / (2 * r), (float(m[0]) + float(p[0])) / (2 * r)
    H = np.zeros((2, 6)); H[0, 0], H[0, 1], H[1, 2] = a00, a01, 1.0
    V3 = np.array([[v[0], v[1], v[3]], [v[1], v[2], v[4]], [v[3], v[4], v[5]]], np.float64)
    H3 = H[:, :3]
    V = H3 @ V3 @ H3.T
    S = H @ C @ H.T + V
    K = C @ H.T @ np.linalg.inv(S)
    A = np.eye(6) - K @ H
    return A @ C @ A.T + K @ V @ K.T


def _ill_event(n=3000):
    geo = mk.geometry(1)
    ev = mk.generate(geo, mk.default_toy_params(seed=3), n, 3)
    ev = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in ev.items()}
    ev["seed_err"][5] = 1e4                         # sigma_z 100 cm before the first hit
    ev["seed_err"][0] *= 1e4; ev["seed_err"][2] *= 1e4
    ev["hit_err"][:, 5] = 1e-10                     # the hit fixes z to 1e-5 cm
    ev["hit_err"][:, :3] *= 1e-6
    ev["n_hits"][:] = 1
    return ev


def test_fallback_inactive_on_healthy_tracks():
    geo = mk.geometry(3)
    rl, xi = mk.material_grid(geo)
    O = oracle.Oracle(bfield=3.8, rl_grid=rl, xi_grid=xi)
    ev = mk.generate(geo, mk.default_toy_params(seed=5, eta_max=2.5, hits_min=8, hits_max=30, missing_frac=0.05),
                     20000, 30)
    a = O.fit(ev, pflags=2, fit_flags=1, n_threads=8)
    b = O.fit(ev, pflags=2, fit_flags=1 | 4, n_threads=8)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_fallback_matches_numpy_joseph():
    ev = _ill_event()
    O = oracle.Oracle(bfield=3.8)
    std = O.fit(ev, 0, 1, n_threads=4)
    fb = O.fit(ev, 0, 1 | 4, n_threads=4)
    n = ev["n_hits"].size
    hp, he = ev["hit_pos"][0], ev["hit_err"][0]
    rad = np.hypot(hp[0], hp[1]).astype(np.float32)
    errP, parP, _ = O.propagate_to_r(ev["seed_err"], ev["seed_par"], ev["chg"], rad, pflags=0)
    n_ill = 0
    for t in range(n):
        if _ill(std[1][:, t]):
            n_ill += 1
            C = _full(errP[:, t])
            J = _joseph(C, parP[:, t], hp[:, t], he[:, t])
            got = _full(fb[1][:, t])
            # float rounding of J, plus double rounding of the O(C) terms that cancel in the Joseph form
            tol = 1e-6 * np.abs(J) + 1e-12 * np.sqrt(np.outer(np.diag(C), np.diag(C)))
            assert np.all(np.abs(got - J) <= tol), t
        else:
            assert np.array_equal(fb[1][:, t], std[1][:, t])
    assert n_ill > n // 2
    # parameters, chi2 and charge are untouched by the fallback
    for k in (0, 2, 3):
        assert np.array_equal(fb[k], std[k])


def test_fallback_restores_measured_variance():
    ev = _ill_event()
    O = oracle.Oracle(bfield=3.8)
    std = O.fit(ev, 0, 1, n_threads=4)
    fb = O.fit(ev, 0, 1 | 4, n_threads=4)
    ref = O.fit(ev, 0, 1, dtype=np.float64, n_threads=4)
    zz = 5
    e_std = np.median(np.abs(std[1][zz] - ref[1][zz]) / ref[1][zz])
    e_fb = np.median(np.abs(fb[1][zz] - ref[1][zz]) / ref[1][zz])
    assert e_std > 0.5 and e_fb < 0.05, (e_std, e_fb)
    assert (fb[1][zz] > 0).mean() > 0.95 and (std[1][zz] > 0).mean() < (fb[1][zz] > 0).mean()
