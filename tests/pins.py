"""Pins of the CPU oracle against what the mathematics fixes.

Each `pin_*` function takes a factory make(bfield, rl_grid, xi_grid) -> oracle
Oracle and raises AssertionError when
the restatement disagrees with an independent statement of the same quantity:

  pin_helix_r_analytic   helix-to-cylinder step vs the exact helix of the toy
                         MC (double precision circle intersection)
  pin_helix_z_analytic   helix-to-plane step vs the exact helix
  pin_jacobian_fd        propagation Jacobian vs central finite differences of
                         the propagation itself (the reference's Jacobian is an
                         approximation, so the bound is loose but element-wise)
  pin_similarity_dense   propagated covariance == J C J^T with a dense product
  pin_kalman_textbook    barrel and endcap update/chi2 vs the textbook gain
                         K = C H^T (H C H^T + V)^-1 in FP64 with an explicit H
  pin_material_pdg       covariance/momentum changes vs the PDG Highland and
                         Bethe (density-corrected) formulas written from the PDG
  pin_fail_semantics     unreachable or backward targets set the fail flag and
                         return the input unchanged
  pin_fit_precision      FP32 fit vs FP64 fit of the same restatement

tests/test_oracle_pins.py runs them on the oracle; tests/test_oracle_mutants.py
plants mistakes in a copy of the oracle source and requires a pin to fail.
"""
import numpy as np

import mkfit_amd as mk

SYM_IJ = [(i, j) for i in range(6) for j in range(i + 1)]


def full(c):
    n = c.shape[1]
    m = np.zeros((n, 6, 6))
    for k, (i, j) in enumerate(SYM_IJ):
        m[:, i, j] = c[k]
        m[:, j, i] = c[k]
    return m


def pack(m):
    return np.stack([m[:, i, j] for (i, j) in SYM_IJ])


def _ev(preset=2, n=3000, seed=3, **kw):
    geo = mk.geometry(preset)
    p = mk.default_toy_params(seed=seed, simulate_material=0, **kw)
    S = 30 if preset == 3 else geo.n_barrel
    return geo, p, mk.generate(geo, p, n, S, truth=True)


def _steps(ev, k, endcap=False):
    """Truth states at crossing k and k+1 (both barrel, or k+1 an endcap)."""
    fl = ev["hit_flags"]
    ok = ev["n_hits"] > k + 1
    ok &= ((fl[k + 1] & mk.HIT_ENDCAP) != 0) if endcap else ((fl[k + 1] & mk.HIT_ENDCAP) == 0)
    a = ev["truth_par"][k][:, ok].astype(np.float64)
    b = ev["truth_par"][k + 1][:, ok].astype(np.float64)
    return a, b, ev["chg"][ok]


def _dphi(a, b):
    d = np.abs(a - b)
    return np.minimum(d, 2 * np.pi - d)


def pin_helix_r_analytic(make):
    orc = make(3.8, None, None)
    geo, p, ev = _ev(2)
    for k in (0, 5, 12):
        a, b, chg = _steps(ev, k)
        r = np.hypot(b[0], b[1])
        err = np.zeros((21, a.shape[1]))
        _, op, fl = orc.propagate_to_r(err, a, chg, r, dtype=np.float64)
        assert not fl.any()
        dpos = np.abs(op[:3] - b[:3]).max()
        dang = max(_dphi(op[4], b[4]).max(), np.abs(op[5] - b[5]).max(), np.abs(op[3] - b[3]).max())
        # sincos4 truncation over one 4.5 cm step leaves O(1e-6) cm
        assert dpos < 2e-5, dpos
        assert dang < 2e-6, dang
        _, op32, _ = orc.propagate_to_r(err.astype(np.float32), a.astype(np.float32), chg, r.astype(np.float32))
        assert np.abs(op32[:3] - b[:3]).max() < 2e-3


def pin_helix_z_analytic(make):
    orc = make(3.8, None, None)
    geo, p, ev = _ev(3, n=20000, seed=4, eta_max=2.5)
    a, b, chg = _steps(ev, 2, endcap=True)
    assert a.shape[1] > 200
    err = np.zeros((21, a.shape[1]))
    _, op = orc.propagate_to_z(err, a, chg, b[2], dtype=np.float64)
    # the step length along z can be 30 cm or more: sincos4 of the half angle
    # leaves an error growing like alpha^5
    assert np.median(np.abs(op[:2] - b[:2])) < 1e-5
    assert np.abs(op[:2] - b[:2]).max() < 5e-3
    assert _dphi(op[4], b[4]).max() < 1e-6
    assert np.abs(op[2] - b[2]).max() == 0


def pin_jacobian_fd(make):
    orc = make(3.8, None, None)
    geo, p, ev = _ev(2, n=2000, seed=5)
    a, b, chg = _steps(ev, 4)
    r = np.hypot(b[0], b[1])
    _, J, _ = orc.helix_at_r(a, chg, r, dtype=np.float64)
    h = np.array([1e-5, 1e-5, 1e-5, 1e-7, 1e-7, 1e-7])
    Jfd = np.zeros_like(J)
    for j in range(6):
        ap, am = a.copy(), a.copy()
        ap[j] += h[j]
        am[j] -= h[j]
        pp, _, _ = orc.helix_at_r(ap, chg, r, dtype=np.float64)
        pm, _, _ = orc.helix_at_r(am, chg, r, dtype=np.float64)
        Jfd[:, j] = (pp - pm) / (2 * h[j])
    rel = np.abs(J - Jfd) / np.maximum(np.abs(Jfd), 1e-3)
    med = np.median(rel, axis=2)
    # Elements of the reference's approximate Jacobian that track the true
    # derivative (the others, J(2,3) and J(2,4), are dropped by the approximation).
    for (i, j) in [(0, 0), (0, 1), (0, 3), (0, 4), (1, 0), (1, 1), (1, 3), (1, 4), (2, 0), (2, 1), (2, 5),
                   (4, 0), (4, 1), (4, 3), (4, 4)]:
        assert med[i, j] < 0.06, ((i, j), med[i, j])
    for (i, j) in [(2, 2), (3, 3), (5, 5)]:
        assert np.all(J[i, j] == 1)
    # Sub-micron steps (|r - r0| < 1e-4 from the first iteration on): here the
    # accumulated derivative terms (PropagationMPlex.icc:115-144) carry the
    # whole derivative, so J - I must follow the finite differences closely,
    # including the tiny phi-phi term.
    r0 = np.hypot(a[0], a[1])
    rs = r0 + 5e-5
    _, J, _ = orc.helix_at_r(a, chg, rs, dtype=np.float64)
    for j in range(6):
        ap, am = a.copy(), a.copy()
        ap[j] += h[j]
        am[j] -= h[j]
        pp, _, _ = orc.helix_at_r(ap, chg, rs, dtype=np.float64)
        pm, _, _ = orc.helix_at_r(am, chg, rs, dtype=np.float64)
        Jfd[:, j] = (pp - pm) / (2 * h[j])
    eye = np.eye(6)[:, :, None]
    rel = np.abs((J - eye) - (Jfd - eye)) / np.maximum(np.abs(Jfd - eye), 1e-12)
    med = np.median(rel, axis=2)
    bounds = {(0, 0): 5e-3, (0, 1): 5e-3, (1, 0): 5e-3, (1, 1): 5e-3, (2, 0): 5e-3, (2, 1): 5e-3,
              (4, 0): 5e-3, (4, 1): 5e-3, (0, 4): 5e-3, (1, 4): 5e-3, (4, 3): 5e-3, (2, 4): 0.05,
              (2, 5): 1e-3, (4, 4): 0.5}
    for (i, j), bnd in bounds.items():
        assert med[i, j] < bnd, ("sub-micron", (i, j), med[i, j])
    # endcap: helixAtZ's Jacobian is analytic -> tight agreement
    geo, p, ev = _ev(3, n=20000, seed=6, eta_max=2.5)
    a, b, chg = _steps(ev, 2, endcap=True)
    z = b[2]
    _, J = orc.helix_at_z(a, chg, z, dtype=np.float64)
    Jfd = np.zeros_like(J)
    for j in range(6):
        ap, am = a.copy(), a.copy()
        ap[j] += h[j]
        am[j] -= h[j]
        pp, _ = orc.helix_at_z(ap, chg, z, dtype=np.float64)
        pm, _ = orc.helix_at_z(am, chg, z, dtype=np.float64)
        Jfd[:, j] = (pp - pm) / (2 * h[j])
    # sincos4 of alpha/2 makes the closed form slightly inexact for long steps
    rel = np.abs(J - Jfd) / np.maximum(np.abs(Jfd), 1e-2)
    med = np.median(rel, axis=2)
    assert med.max() < 2e-3, med


def pin_similarity_dense(make):
    orc = make(3.8, None, None)
    geo, p, ev = _ev(2, n=1000, seed=7)
    a, b, chg = _steps(ev, 3)
    n = a.shape[1]
    rng = np.random.default_rng(1)
    L = rng.normal(size=(n, 6, 6)) * 0.1
    C = L @ L.transpose(0, 2, 1) + np.eye(6) * 1e-3
    r = np.hypot(b[0], b[1])
    oe, op, fl = orc.propagate_to_r(pack(C), a, chg, r, dtype=np.float64)
    _, J, _ = orc.helix_at_r(a, chg, r, dtype=np.float64)
    Jn = J.transpose(2, 0, 1)
    want = Jn @ C @ Jn.transpose(0, 2, 1)
    np.testing.assert_allclose(full(oe), want, rtol=1e-11, atol=1e-14)
    # endcap
    geo, p, ev = _ev(3, n=20000, seed=8, eta_max=2.5)
    a, b, chg = _steps(ev, 2, endcap=True)
    n = a.shape[1]
    L = rng.normal(size=(n, 6, 6)) * 0.1
    C = L @ L.transpose(0, 2, 1) + np.eye(6) * 1e-3
    oe, op = orc.propagate_to_z(pack(C), a, chg, b[2], dtype=np.float64)
    _, J = orc.helix_at_z(a, chg, b[2], dtype=np.float64)
    Jn = J.transpose(2, 0, 1)
    want = Jn @ C @ Jn.transpose(0, 2, 1)
    np.testing.assert_allclose(full(oe), want, rtol=1e-11, atol=1e-14)


def _spd(n, rng, scale):
    L = rng.normal(size=(n, 6, 6)) * scale
    return L @ L.transpose(0, 2, 1) + np.eye(6) * 1e-4


def pin_kalman_textbook(make):
    orc = make(3.8, None, None)
    rng = np.random.default_rng(2)
    n = 500
    phi = rng.uniform(-np.pi, np.pi, n)
    R = rng.uniform(5, 100, n)
    ps = np.stack([R * np.cos(phi), R * np.sin(phi), rng.normal(0, 20, n), rng.uniform(0.1, 2, n),
                   phi + rng.normal(0, 0.1, n), rng.uniform(0.5, 2.6, n)])
    C = _spd(n, rng, 0.05)
    hp = ps[:3] + rng.normal(0, 0.02, (3, n))
    V = np.zeros((n, 3, 3))
    s = rng.uniform(1e-5, 1e-3, (n, 3))
    for i in range(3):
        V[:, i, i] = s[:, i]
    V[:, 0, 1] = V[:, 1, 0] = 0.3 * np.sqrt(s[:, 0] * s[:, 1])
    vpack = np.stack([V[:, 0, 0], V[:, 1, 0], V[:, 1, 1], V[:, 2, 0], V[:, 2, 1], V[:, 2, 2]])
    for endcap in (False, True):
        oe, op, c2 = orc.kalman_operation(3, pack(C), ps, vpack, hp, endcap=endcap, dtype=np.float64)
        for t in range(n):
            if endcap:
                H = np.zeros((2, 6)); H[0, 0] = 1; H[1, 1] = 1
                res = hp[:2, t] - ps[:2, t]
                Vl = V[t, :2, :2]
            else:
                rh = np.hypot(hp[0, t], hp[1, t])
                r00 = -(hp[1, t] + ps[1, t]) / (2 * rh)
                r01 = (hp[0, t] + ps[0, t]) / (2 * rh)
                Rm = np.array([[r00, r01, 0], [0, 0, 1]])
                H = np.zeros((2, 6)); H[:, :3] = Rm
                res = Rm @ (hp[:, t] - ps[:3, t])
                Vl = Rm @ V[t] @ Rm.T
            Sm = H @ C[t] @ H.T + Vl
            Si = np.linalg.inv(Sm)
            K = C[t] @ H.T @ Si
            pn = ps[:, t] + K @ res
            # squashPhiMPlex uses the float constants Config::PI / TwoPI (Config.h:77-78)
            if pn[4] >= float(np.float32(np.pi)):
                pn[4] -= float(np.float32(2 * np.pi))
            if pn[4] < -float(np.float32(np.pi)):
                pn[4] += float(np.float32(2 * np.pi))
            Cn = C[t] - K @ H @ C[t]
            chi = res @ Si @ res
            np.testing.assert_allclose(op[:, t], pn, rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(full(oe[:, t:t + 1])[0], Cn, rtol=1e-8, atol=1e-13)
            np.testing.assert_allclose(c2[t], chi, rtol=1e-9)


def _pdg_material(theta, ipt, x_over_x0, xi, barrel, sign):
    """PDG Highland (eq. 34.16) and mean Bethe loss with the high-energy density correction."""
    m_pi, m_e, I = 0.140, 0.0005, 16.0e-9 * 10.75
    pt = 1 / ipt
    p = pt / np.sin(theta)
    E = np.sqrt(p * p + m_pi ** 2)
    beta, gamma = p / E, E / m_pi
    path = 1 / np.sin(theta) if barrel else 1 / np.abs(np.cos(theta))
    t = x_over_x0 * path
    theta0_sq = (0.0136 / (beta * p) * (1 + 0.038 * np.log(t))) ** 2 * t
    wmax = 2 * m_e * beta ** 2 * gamma ** 2 / (1 + 2 * gamma * m_e / m_pi + (m_e / m_pi) ** 2)
    hw_plasma = 28.816e-9 * np.sqrt(2.33 * 0.498)  # silicon plasma energy, GeV
    half_delta = np.log(hw_plasma / I) + np.log(beta * gamma) - 0.5
    dE = 2 * xi * path / beta ** 2 * (0.5 * np.log(2 * m_e * beta ** 2 * gamma ** 2 * wmax / I ** 2) - beta ** 2
                                      - half_delta)
    dP = sign * dE / beta
    return theta0_sq, dP, p, pt


def pin_material_pdg(make):
    rl = np.full((300, 120), 0.03, np.float32)
    xi = np.full((300, 120), 6e-5, np.float32)
    orc = make(3.8, rl, xi)
    orc0 = make(3.8, None, None)
    geo, p, ev = _ev(2, n=2000, seed=9)
    for endcap in (False, True):
        if endcap:
            geo, p, ev = _ev(3, n=20000, seed=10, eta_max=2.5)
        a, b, chg = _steps(ev, 2 if endcap else 3, endcap=endcap)
        n = a.shape[1]
        C = np.zeros((n, 6, 6))
        for i in range(6):
            C[:, i, i] = 1e-4
        if endcap:
            oe, op = orc.propagate_to_z(pack(C), a, chg, b[2], pflags=2, dtype=np.float64)
            oe0, op0 = orc0.propagate_to_z(pack(C), a, chg, b[2], pflags=0, dtype=np.float64)
            _, J = orc0.helix_at_z(a, chg, b[2], dtype=np.float64)
            sign = np.where(np.abs(b[2]) > np.abs(a[2]), 1.0, -1.0)
        else:
            r = np.hypot(b[0], b[1])
            oe, op, _ = orc.propagate_to_r(pack(C), a, chg, r, pflags=2, dtype=np.float64)
            oe0, op0, _ = orc0.propagate_to_r(pack(C), a, chg, r, pflags=0, dtype=np.float64)
            _, J, _ = orc0.helix_at_r(a, chg, r, dtype=np.float64)
            sign = np.where(r > np.hypot(a[0], a[1]), 1.0, -1.0)
        th0, dP, pp, pt = _pdg_material(a[5], a[3], 0.03, 6e-5, not endcap, sign)
        Q = np.zeros((n, 6, 6))
        Q[:, 4, 4] = th0
        Q[:, 5, 5] = th0
        Q[:, 3, 3] = dP ** 2 / (pp ** 2 * pt ** 2)
        Jn = J.transpose(2, 0, 1)
        want = Jn @ (C + Q) @ Jn.transpose(0, 2, 1)
        # the restatement keeps the reference's float-valued constants (m_pi = 0.140f,
        # 0.0136f, ...) in its FP64 path: agreement to ~1e-8, bounded at 1e-6
        np.testing.assert_allclose(full(oe), want, rtol=1e-6, atol=1e-16)
        np.testing.assert_allclose(op[3], a[3] * pp / (pp + dP), rtol=1e-6)
        assert np.abs(op[3] - a[3]).max() > 1e-6  # energy loss is visible at this thickness
        for i in (0, 1, 2, 4, 5):
            np.testing.assert_array_equal(op[i], op0[i])


def pin_fail_semantics(make):
    orc = make(3.8, None, None)
    rng = np.random.default_rng(3)
    n = 200
    phi = rng.uniform(-np.pi, np.pi, n)
    R = rng.uniform(10, 50, n)
    chg = np.where(rng.uniform(size=n) < 0.5, 1, -1).astype(np.int32)
    C = pack(_spd(n, rng, 0.05))
    # (1) target beyond the helix diameter: pT = 0.5 GeV -> 2R ~ 88 cm
    par = np.stack([R * np.cos(phi), R * np.sin(phi), np.zeros(n), np.full(n, 2.0), phi, np.full(n, 1.2)])
    for dt in (np.float32, np.float64):
        oe, op, fl = orc.propagate_to_r(C, par, chg, np.full(n, 500.0), dtype=dt)
        assert fl.all()
        np.testing.assert_array_equal(op, par.astype(dt))
        np.testing.assert_array_equal(oe, C.astype(dt))
    # (2) momentum pointing back towards the beam line: 1/cos < 0
    par2 = par.copy()
    par2[3] = 0.2
    par2[4] = phi + np.pi
    par2[4] = np.where(par2[4] >= np.pi, par2[4] - 2 * np.pi, par2[4])
    oe, op, fl = orc.propagate_to_r(C, par2, chg, R + 5, dtype=np.float64)
    assert fl.all()
    np.testing.assert_array_equal(op, par2)
    # (3) a normal outward step does not fail
    par3 = par.copy()
    par3[3] = 0.2
    oe, op, fl = orc.propagate_to_r(C, par3, chg, R + 5, dtype=np.float64)
    assert not fl.any()


def pin_fit_precision(make):
    geo = mk.geometry(2)
    rl, xi = mk.material_grid(geo)
    p = mk.default_toy_params(seed=11)
    ev = mk.generate(geo, p, 5000, 25)
    orc = make(p.bfield, rl, xi)
    a = orc.fit(ev, pflags=2, n_threads=4)
    b = orc.fit(ev, pflags=2, dtype=np.float64, n_threads=4)
    fl = np.array([1.0, 1.0, 1.0, 0.1, 1.0, 1.0])[:, None]
    d = np.abs(a[0] - b[0])
    d[4] = np.minimum(d[4], 2 * np.pi - d[4])
    rel = (d / np.maximum(np.abs(b[0]), fl)).max(axis=0)
    # FP32 rounding over ~25 layers: ~1e-6 typical, ~1e-5 at the 99th percentile;
    # low-pT tracks with large chi2 are ill-conditioned and spread further.
    assert np.median(rel) < 3e-6, np.median(rel)
    assert np.percentile(rel, 99) < 3e-5, np.percentile(rel, 99)
    c = np.abs(a[3] - b[3]) / np.maximum(b[3], 1)
    assert np.percentile(c, 99) < 1e-3


ALL_PINS = [pin_helix_r_analytic, pin_helix_z_analytic, pin_jacobian_fd, pin_similarity_dense, pin_kalman_textbook,
            pin_material_pdg, pin_fail_semantics, pin_fit_precision]
