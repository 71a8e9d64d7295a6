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
