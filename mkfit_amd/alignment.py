"""Synthetic events for the material and alignment gradient (the config-2 barrel with per-layer
x/X0 and injected alignment offsets), and the weak modes a track fit cannot see."""
import numpy as np

from . import default_toy_params, generate, geometry


def make_event(n_tracks, seed=5, x0_range=(0.01, 0.05), sigma_align=50e-4, pt_min=1.0, pt_max=10.0,
               xi_per_x0=2.2e-3):
    """Tracks through the 25-layer barrel with per-layer x/X0 drawn from x0_range and per-layer
    r-phi / z offsets drawn from N(0, sigma_align) cm.  Hits are reported in the nominal
    frame: a hit truly at t is stored at t - d_rphi e_phi - d_z e_z, so moving it back by
    the true offsets (the model's correction) restores t.

    Returns (event, geometry, true_theta [L, 3])."""
    rng = np.random.default_rng(seed)
    geo = geometry(2)
    L = geo.n_barrel
    x0 = rng.uniform(*x0_range, L)
    for l in range(L):
        geo.x_over_x0[l] = x0[l]
        geo.xi[l] = xi_per_x0 * x0[l]
    p = default_toy_params(seed=seed, pt_min=pt_min, pt_max=pt_max, simulate_material=1)
    ev = generate(geo, p, n_tracks, L)
    d_rphi = rng.normal(0.0, sigma_align, L)
    d_z = rng.normal(0.0, sigma_align, L)
    hp = ev["hit_pos"]
    for l in range(L):
        x, y = hp[l, 0].astype(np.float64), hp[l, 1].astype(np.float64)
        r = np.hypot(x, y)
        ir = np.divide(1.0, r, out=np.zeros_like(r), where=r > 0)    # unused slots hold zeros
        hp[l, 0] = (x + d_rphi[l] * y * ir).astype(np.float32)
        hp[l, 1] = (y - d_rphi[l] * x * ir).astype(np.float32)
        hp[l, 2] = (hp[l, 2].astype(np.float64) - d_z[l]).astype(np.float32)
    theta = np.stack([x0, d_rphi, d_z], 1)
    return ev, geo, theta


def weak_mode_residual(d_est, d_true, radii, kind):
    """Offsets differ by what the helix absorbs: r-phi by {1, r, r^2} (impact parameter, phi,
    curvature), z by {1, r} (z0, cot theta).  Returns the difference with those modes fitted out."""
    r = np.asarray(radii, np.float64)
    basis = np.stack([np.ones_like(r), r, r * r] if kind == "rphi" else [np.ones_like(r), r], 1)
    diff = np.asarray(d_est, np.float64) - np.asarray(d_true, np.float64)
    coef, *_ = np.linalg.lstsq(basis, diff, rcond=None)
    return diff - basis @ coef
