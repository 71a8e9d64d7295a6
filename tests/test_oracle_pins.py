"""The CPU oracle against closed forms, textbook algebra and invariants (see tests/pins.py)."""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
import pins


def _make(bfield, rl, xi):
    return oracle.Oracle(bfield, rl, xi)


@pytest.mark.parametrize("pin", pins.ALL_PINS, ids=lambda f: f.__name__)
def test_pin(pin):
    pin(_make)


def test_physics_pulls_and_chi2():
    """Fitted state at the last hit vs the toy-MC truth: pulls ~ N(0,1), chi2/ndf ~ 1.
    Energy loss is off in the truth and in the fit (the reference adds energy
    moving outward, PropagationMPlex.cc:762-763, which would bias 1/pT), and
    pT >= 1 GeV: the reference scales x/X0 by 1/sin(theta) only
    (PropagationMPlex.cc:740, "fixme works only for barrel geom"), which
    underestimates the scattering of low-pT tracks crossing layers at large
    transverse angles."""
    geo = mk.geometry(1)
    for lyr in range(geo.n_barrel):
        geo.xi[lyr] = 0.0
    rl, xi = mk.material_grid(geo)
    p = mk.default_toy_params(seed=17, pt_min=1.0)
    ev = mk.generate(geo, p, 20000, geo.n_barrel, truth=True)
    par, err, chg, chi2 = oracle.Oracle(p.bfield, rl, xi).fit(ev, pflags=2, dtype=np.float64, n_threads=8)
    nh = ev["n_hits"]
    ok = nh >= 5
    idx = np.arange(len(nh))[ok]
    tr = ev["truth_par"][nh[ok] - 1, :, idx].T
    d = par[:, ok] - tr
    d[4] = (d[4] + np.pi) % (2 * np.pi) - np.pi
    for e, dg in ((3, 9), (4, 14), (5, 20)):
        pull = d[e] / np.sqrt(err[dg, ok])
        assert abs(pull.mean()) < 0.1, (e, pull.mean())
        assert 0.85 < pull.std() < 1.25, (e, pull.std())
    # z: the hit smearing dominates; rphi position pull via the local frame
    pz = d[2] / np.sqrt(err[5, ok])
    assert abs(pz.mean()) < 0.1 and 0.8 < pz.std() < 1.3
    c = chi2[ok] / (2 * nh[ok])
    assert 0.9 < c.mean() < 1.15, c.mean()


def test_missing_hit_is_nearly_a_skipped_layer():
    """A missing layer (fake hit, variance 666, at the track position, no material)
    leaves the fit almost as if the layer were absent (MkFinder.cc:1635-1647)."""
    geo = mk.geometry(2)
    rl, xi = mk.material_grid(geo)
    p = mk.default_toy_params(seed=18, simulate_material=0)
    ev = mk.generate(geo, p, 4000, 25)
    orc = oracle.Oracle(p.bfield, rl, xi)
    k = 10
    ok = ev["n_hits"] > k + 2
    ev = {key: np.ascontiguousarray(v[..., ok]) for key, v in ev.items()}
    miss = {key: v.copy() for key, v in ev.items()}
    miss["hit_flags"][k] |= mk.HIT_MISSING
    skip = {key: v.copy() for key, v in ev.items()}
    for key in ("hit_pos", "hit_err", "hit_flags"):
        skip[key] = np.ascontiguousarray(np.delete(skip[key], k, axis=0))
    skip["n_hits"] = skip["n_hits"] - 1
    a = orc.fit(miss, pflags=0, dtype=np.float64)
    b = orc.fit(skip, pflags=0, dtype=np.float64)
    d = np.abs(a[0] - b[0])
    sig = np.sqrt(b[1][[0, 2, 5, 9, 14, 20]])
    assert (d / sig).max() < 0.05  # far below one standard deviation
    full = orc.fit(ev, pflags=0, dtype=np.float64)
    assert np.abs(full[0] - a[0]).max() > 10 * np.abs(a[0] - b[0]).max()  # the flag does change the fit


def test_lanes_are_independent():
    """Fitting a subset of tracks gives bit-identical results (block packing, padding)."""
    geo = mk.geometry(3)
    rl, xi = mk.material_grid(geo)
    p = mk.default_toy_params(seed=19, eta_max=2.5, hits_min=8, hits_max=30, missing_frac=0.05)
    ev = mk.generate(geo, p, 3001, 30)
    orc = oracle.Oracle(p.bfield, rl, xi)
    a = orc.fit(ev, pflags=2, fit_flags=3, n_threads=3)
    sel = np.arange(5, 3001, 7)
    sub = {k: np.ascontiguousarray(v[..., sel]) for k, v in ev.items()}
    b = orc.fit(sub, pflags=2, fit_flags=3)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x[..., sel], y)


def test_both_isa_builds_agree():
    """The AVX-512 (NN=16) and AVX2 (NN=8) builds of the oracle agree to FP32 rounding."""
    import os
    if not oracle._cpu_has_avx512():
        pytest.skip("no AVX-512 on this host")
    geo = mk.geometry(2)
    rl, xi = mk.material_grid(geo)
    p = mk.default_toy_params(seed=20)
    ev = mk.generate(geo, p, 2000, 25)
    a = oracle.Oracle(p.bfield, rl, xi, lib_path=os.path.join(os.path.dirname(oracle.__file__),
                                                               "liborc_avx512.so")).fit(ev, pflags=2)
    b = oracle.Oracle(p.bfield, rl, xi, lib_path=os.path.join(os.path.dirname(oracle.__file__),
                                                               "liborc_avx2.so")).fit(ev, pflags=2)
    rel = np.abs(a[0] - b[0]) / np.maximum(np.abs(b[0]), np.array([1, 1, 1, 0.1, 1, 1.0])[:, None])
    assert np.mean(rel.max(axis=0) < 1e-5) > 0.999
