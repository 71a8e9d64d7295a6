"""Toy Monte Carlo: determinism, geometry and hit placement."""
import numpy as np
import pytest

import mkfit_amd as mk


def test_deterministic_and_thread_independent():
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=99)
    a = mk.generate(geo, p, 3000, 17, n_threads=1)
    b = mk.generate(geo, p, 3000, 17, n_threads=7)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    c = mk.generate(geo, mk.default_toy_params(seed=100), 3000, 17)
    assert not np.array_equal(a["hit_pos"], c["hit_pos"])


@pytest.mark.parametrize("preset", [1, 2])
def test_barrel_hits_on_cylinders(preset):
    geo = mk.geometry(preset)
    p = mk.default_toy_params(seed=5)
    ev = mk.generate(geo, p, 5000, geo.n_barrel)
    radii = np.array(geo.barrel_r[: geo.n_barrel])
    for s in range(geo.n_barrel):
        on = ev["n_hits"] > s
        lyr = ev["hit_layer"][s][on]
        r = np.hypot(ev["hit_pos"][s][0][on], ev["hit_pos"][s][1][on])
        np.testing.assert_allclose(r, radii[lyr], rtol=2e-6)
        assert (np.abs(ev["hit_pos"][s][2][on]) < 121).all()
    # layers are crossed in order from the innermost
    nh = ev["n_hits"]
    for t in range(200):
        np.testing.assert_array_equal(ev["hit_layer"][: nh[t], t], np.arange(nh[t]))


def test_hit_counts_and_missing_fraction():
    geo = mk.geometry(3)
    p = mk.default_toy_params(seed=6, eta_max=2.5, hits_min=8, hits_max=30, missing_frac=0.05)
    ev = mk.generate(geo, p, 20000, 30)
    nh = ev["n_hits"]
    assert nh.max() <= 30
    assert (nh >= 1).all()
    fl = ev["hit_flags"]
    used = np.arange(30)[:, None] < nh[None, :]
    miss = (fl & mk.HIT_MISSING) != 0
    assert not (miss & ~used).any()
    assert not miss[0].any()
    frac = miss[1:][used[1:]].mean()
    assert 0.04 < frac < 0.06
    endcap = (fl & mk.HIT_ENDCAP) != 0
    assert endcap[used].mean() > 0.1  # |eta| up to 2.5 reaches the disks


def test_material_grid_covers_every_hit():
    geo = mk.geometry(3)
    rl, xi = mk.material_grid(geo)
    p = mk.default_toy_params(seed=8, eta_max=2.5)
    ev = mk.generate(geo, p, 5000, 30)
    for s in range(30):
        on = (ev["n_hits"] > s) & ((ev["hit_flags"][s] & mk.HIT_MISSING) == 0)
        x, y, z = ev["hit_pos"][s][:, on]
        zb = (np.abs(z) * 300 / 300.0).astype(int)
        rb = (np.hypot(x, y) * 120 / 120.0).astype(int)
        assert (rl[zb, rb] == np.float32(0.02)).all()


def test_seed_is_smeared_truth_at_first_layer():
    geo = mk.geometry(2)
    p = mk.default_toy_params(seed=9)
    ev = mk.generate(geo, p, 20000, 25, truth=True)
    d = ev["seed_par"] - ev["truth_par"][0]
    d[4] = (d[4] + np.pi) % (2 * np.pi) - np.pi
    sig = np.sqrt(ev["seed_err"][[0, 2, 5, 9, 14, 20]])
    pull = d / sig
    assert np.all(np.abs(pull.mean(axis=1)) < 0.05)
    assert np.all(np.abs(pull.std(axis=1) - 1) < 0.05)
