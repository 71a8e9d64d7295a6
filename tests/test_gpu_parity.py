"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Every test generates seeded toy-MC inputs, runs the same operation in the
oracle's FP32 restatement of the reference Matriplex code and on the GPU, and
bounds the difference with the metrics of tests/_common.py.  The FP64 oracle
error of both is reported alongside.
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from oracle.parity import fit_reference
from _common import (TOL_CHI2, TOL_COV, TOL_PAR, assert_parity, assert_track_parity, chi2_rel, cov_rel, event,
                     par_rel, summary)

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _np(t):
    return t.detach().cpu().numpy()


def _states(ev, slot, ok):
    """Fitted-like states: seeds propagated part way, as inputs for single ops."""
    return ev["seed_par"][:, ok], ev["seed_err"][:, ok], ev["chg"][ok]


def test_abi_rejects_host_pointer(gpu_device):
    import torch
    ctx = mk.Context(bfield=3.8)
    a = torch.zeros(6, 8)  # host memory
    with pytest.raises(mk.MkfError, match="not a HIP allocation|not device memory"):
        ctx.propagate_helix_to_r(torch.zeros(21, 8), a, torch.zeros(8, dtype=torch.int32), torch.ones(8),
                                 torch.zeros(21, 8), torch.zeros(6, 8))


def _hip_malloc(nbytes):
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
    return hip, p


def test_abi_rejects_short_buffer(gpu_device):
    import ctypes
    import torch
    ctx = mk.Context(bfield=3.8)
    d = gpu_device
    n = 1024
    hip, short = _hip_malloc(5 * n * 4)  # in_par needs 6*n floats
    try:
        z = lambda *s: torch.zeros(*s, device=d)
        rc = mk.lib().mkf_propagate_helix_to_r(ctx.h, n, mk._ptr(z(21, n)), short,
                                               mk._ptr(torch.zeros(n, dtype=torch.int32, device=d)),
                                               mk._ptr(z(n)), mk._ptr(z(21, n)), mk._ptr(z(6, n)), 0, None, None,
                                               None)
        assert rc == 4, mk.lib().mkf_last_error()
        assert b"needed past the pointer" in mk.lib().mkf_last_error()
    finally:
        hip.hipFree(short)


@pytest.mark.parametrize("pflags", [0, mk.PF_APPLY_MATERIAL, mk.PF_USE_PARAM_B_FIELD | mk.PF_APPLY_MATERIAL])
def test_propagate_to_r(gpu_device, pflags):
    geo, p, ev, rl, xi = event(2, 20000, 11)
    ok = ev["n_hits"] > 3
    par, err, chg = ev["seed_par"][:, ok], ev["seed_err"][:, ok], ev["chg"][ok]
    hp = ev["hit_pos"][3][:, ok]
    r = np.hypot(hp[0], hp[1]).astype(np.float32)
    n = par.shape[1]
    orc = oracle.Oracle(p.bfield, rl, xi)
    oe, op, of = orc.propagate_to_r(err, par, chg, r, pflags=pflags)
    de, dp, _ = orc.propagate_to_r(err, par, chg, r, pflags=pflags, dtype=np.float64)
    ctx = mk.Context(p.bfield, rl, xi)
    d = gpu_device
    ge, gp = _t(np.zeros((21, n), np.float32), d), _t(np.zeros((6, n), np.float32), d)
    gf = _t(np.zeros(n, np.int32), d)
    ctx.propagate_helix_to_r(_t(err, d), _t(par, d), _t(chg, d), _t(r, d), ge, gp, pflags=pflags, out_fail=gf)
    ge, gp, gf = _np(ge), _np(gp), _np(gf)
    assert (gf == of).all()
    assert_parity("propagate_to_r", gp, op, dp, "par", TOL_PAR, frac=1.0)
    assert_parity("propagate_to_r", ge, oe, de, "cov", TOL_COV)


@pytest.mark.parametrize("pflags", [0, mk.PF_APPLY_MATERIAL])
def test_propagate_to_z(gpu_device, pflags):
    geo, p, ev, rl, xi = event(3, 20000, 12)
    flags = ev["hit_flags"]
    # tracks whose slot 2 is an endcap hit
    ok = (ev["n_hits"] > 2) & ((flags[2] & mk.HIT_ENDCAP) != 0) & ((flags[2] & mk.HIT_MISSING) == 0)
    assert ok.sum() > 500
    par, err, chg = ev["seed_par"][:, ok], ev["seed_err"][:, ok], ev["chg"][ok]
    z = ev["hit_pos"][2][2, ok].astype(np.float32)
    n = par.shape[1]
    orc = oracle.Oracle(p.bfield, rl, xi)
    oe, op = orc.propagate_to_z(err, par, chg, z, pflags=pflags)
    de, dp = orc.propagate_to_z(err, par, chg, z, pflags=pflags, dtype=np.float64)
    ctx = mk.Context(p.bfield, rl, xi)
    d = gpu_device
    ge, gp = _t(np.zeros((21, n), np.float32), d), _t(np.zeros((6, n), np.float32), d)
    ctx.propagate_helix_to_z(_t(err, d), _t(par, d), _t(chg, d), _t(z, d), ge, gp, pflags=pflags)
    assert_parity("propagate_to_z", _np(gp), op, dp, "par", TOL_PAR, frac=1.0)
    assert_parity("propagate_to_z", _np(ge), oe, de, "cov", TOL_COV)


@pytest.mark.parametrize("endcap", [False, True])
@pytest.mark.parametrize("op", [1, 2, 3])
def test_kalman_operation(gpu_device, endcap, op):
    geo, p, ev, rl, xi = event(2, 20000, 13)
    ok = ev["n_hits"] > 0
    par, err = ev["seed_par"][:, ok], ev["seed_err"][:, ok]
    hp, he = ev["hit_pos"][0][:, ok], ev["hit_err"][0][:, ok]
    n = par.shape[1]
    orc = oracle.Oracle(p.bfield)
    oe, opar, oc2 = orc.kalman_operation(op, err, par, he, hp, endcap=endcap)
    de, dpar, dc2 = orc.kalman_operation(op, err, par, he, hp, endcap=endcap, dtype=np.float64)
    ctx = mk.Context(p.bfield)
    d = gpu_device
    ge, gp, gc = (_t(np.zeros((21, n), np.float32), d), _t(np.zeros((6, n), np.float32), d),
                  _t(np.zeros(n, np.float32), d))
    ctx.kalman_operation(op, _t(err, d), _t(par, d), _t(he, d), _t(hp, d), ge, gp, gc, endcap=endcap)
    if op & 1:
        assert_parity("kalman", _np(gc), oc2, dc2, "chi2", TOL_CHI2)
    if op & 2:
        assert_parity("kalman", _np(gp), opar, dpar, "par", TOL_PAR)
        assert_parity("kalman", _np(ge), oe, de, "cov", TOL_COV)


def _fit_gpu(ctx, ev, d, pflags, fit_flags, batched, with_flags=True):
    b = mk.FitBatch.from_numpy(ev, d, pflags=pflags, fit_flags=fit_flags, with_flags=with_flags)
    (ctx.fit_batched if batched else ctx.fit_plain)(b)
    import torch
    torch.cuda.synchronize()
    return _np(b.par), _np(b.err), _np(b.chg), _np(b.chi2)


FIT_CASES = [
    # (preset, n, seed, pflags, fit_flags, kw)
    (1, 20000, 21, mk.PF_APPLY_MATERIAL, mk.FIT_CHI2, {}),
    (2, 20000, 22, mk.PF_APPLY_MATERIAL, mk.FIT_CHI2, {}),
    (2, 20000, 23, mk.PF_APPLY_MATERIAL | mk.PF_USE_PARAM_B_FIELD, mk.FIT_CHI2 | mk.FIT_CHARGE_FLIP, {}),
    (3, 20000, 24, mk.PF_APPLY_MATERIAL, mk.FIT_CHI2 | mk.FIT_CHARGE_FLIP,
     dict(hits_min=8, hits_max=30, missing_frac=0.05)),
]


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("case", range(len(FIT_CASES)))
def test_fit_parity(gpu_device, case, batched):
    preset, n, seed, pflags, fflags, kw = FIT_CASES[case]
    geo, p, ev, rl, xi = event(preset, n, seed, **kw)
    o, o64, chk = fit_reference(p.bfield, rl, xi, ev, pflags, fflags)
    ctx = mk.Context(p.bfield, rl, xi)
    g = _fit_gpu(ctx, ev, gpu_device, pflags, fflags, batched)
    print()
    # Every track: unflagged ones within tolerance of the FP32 oracle, flagged
    # ones (rounding-sensitive, or with a threshold tie) within their bound
    # (assert_track_parity).
    r = assert_track_parity("fit", (g[0], g[1], g[3]), (o[0], o[1], o[3]), (o64[0], o64[1], o64[3]), chk)
    # Charge flips of rounding-sensitive tracks can differ between FP32
    # roundings; every unflagged track must keep the oracle's charge.
    from oracle.parity import flag_mask
    unflagged = ~flag_mask(chk)[0]
    assert (g[2][unflagged] == o[2][unflagged]).all(), "charges differ on tracks with unflagged parameters"
    assert r["par"]["n_flagged"] < n


def test_plain_equals_batched(gpu_device):
    geo, p, ev, rl, xi = event(3, 20000, 31, hits_min=8, hits_max=30, missing_frac=0.05)
    ctx = mk.Context(p.bfield, rl, xi)
    fl = mk.PF_APPLY_MATERIAL
    a = _fit_gpu(ctx, ev, gpu_device, fl, mk.FIT_CHI2 | mk.FIT_CHARGE_FLIP, False)
    b = _fit_gpu(ctx, ev, gpu_device, fl, mk.FIT_CHI2 | mk.FIT_CHARGE_FLIP, True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("n", [1, 63, 65, 1000])
def test_fit_ragged_sizes(gpu_device, n):
    geo, p, ev, rl, xi = event(1, 1000, 41)
    sub = {k: np.ascontiguousarray(v[..., :n]) for k, v in ev.items()}
    o, o64, chk = fit_reference(p.bfield, rl, xi, sub, mk.PF_APPLY_MATERIAL, mk.FIT_CHI2)
    ctx = mk.Context(p.bfield, rl, xi)
    for batched in (False, True):
        g = _fit_gpu(ctx, sub, gpu_device, mk.PF_APPLY_MATERIAL, mk.FIT_CHI2, batched)
        assert_track_parity("ragged n=%d" % n, (g[0], g[1], g[3]), (o[0], o[1], o[3]), (o64[0], o64[1], o64[3]),
                            chk)


def test_fit_zero_hits_and_empty(gpu_device):
    import torch
    geo, p, ev, rl, xi = event(1, 256, 42)
    ev = {k: v.copy() for k, v in ev.items()}
    ev["n_hits"][::3] = 0  # these tracks must come back unchanged
    ctx = mk.Context(p.bfield, rl, xi)
    g = _fit_gpu(ctx, ev, gpu_device, mk.PF_APPLY_MATERIAL, mk.FIT_CHI2, True)
    np.testing.assert_array_equal(g[0][:, ::3], ev["seed_par"][:, ::3])
    np.testing.assert_array_equal(g[1][:, ::3], ev["seed_err"][:, ::3])
    e0 = {k: np.ascontiguousarray(v[..., :0]) for k, v in ev.items()}
    b = mk.FitBatch.from_numpy(e0, gpu_device)
    ctx.fit_plain(b)
    ctx.fit_batched(b)
    torch.cuda.synchronize()


def test_propagation_failure_resets(gpu_device):
    # A step that starts with the momentum pointing back to the beam line
    # (1/cos < 0) or beyond the helix apex (1/cos > 5) sets the fail flag and
    # returns the input unchanged (PropagationMPlex.icc:84-92, .cc:497-504).
    rng = np.random.default_rng(5)
    n = 512
    phi = rng.uniform(-np.pi, np.pi, n)
    R = rng.uniform(10, 50, n).astype(np.float32)
    chg = np.where(rng.uniform(size=n) < 0.5, 1, -1).astype(np.int32)
    back = rng.uniform(size=n) < 0.5
    mom = np.where(back, phi + np.pi, phi + np.pi / 2 - 0.05 * np.sign(rng.normal(size=n)))
    mom = (mom + np.pi) % (2 * np.pi) - np.pi
    par = np.stack([R * np.cos(phi), R * np.sin(phi), np.zeros(n), np.full(n, 0.5), mom,
                    np.full(n, 1.2)]).astype(np.float32)
    err = np.zeros((21, n), np.float32)
    for dg in (0, 2, 5, 9, 14, 20):
        err[dg] = 1e-3
    r = (R + 3).astype(np.float32)
    orc = oracle.Oracle(3.8)
    oe, op, of = orc.propagate_to_r(err, par, chg, r)
    assert of.all()
    ctx = mk.Context(3.8)
    d = gpu_device
    ge, gp, gf = (_t(np.zeros((21, n), np.float32), d), _t(np.zeros((6, n), np.float32), d),
                  _t(np.zeros(n, np.int32), d))
    ctx.propagate_helix_to_r(_t(err, d), _t(par, d), _t(chg, d), _t(r, d), ge, gp, out_fail=gf)
    assert _np(gf).all()
    np.testing.assert_array_equal(_np(gp), par)
    np.testing.assert_array_equal(_np(ge), err)


@pytest.mark.parametrize("fallback", [False, True])
def test_all_batched_variants_identical(gpu_device, fallback):
    """Every launch shape / register-budget variant the autotuner may pick gives the same bits, with and
    without the FP64 fallback."""
    import torch
    geo, p, ev, rl, xi = event(3, 5000, 45, hits_min=8, hits_max=30, missing_frac=0.05)
    ctx = mk.Context(p.bfield, rl, xi)
    ff = mk.FIT_CHI2 | mk.FIT_CHARGE_FLIP | (mk.FIT_FP64_FALLBACK if fallback else 0)
    ref = None
    for v in range(mk.lib().mkf_num_batched_variants()):
        for block, bpc, tile in ((256, 0, 0), (64, 4, 0), (512, 1, 0), (64, 0, 1024), (256, 0, 4096),
                                 (128, 2, 16384)):
            ctx.set_batched_shape(block, bpc, v, tile)
            g = _fit_gpu(ctx, ev, gpu_device, mk.PF_APPLY_MATERIAL, ff, True)
            if ref is None:
                ref = g
            for name, x, y in zip(("par", "err", "chg", "chi2"), g, ref):
                ndiff = int(np.sum(~((x == y) | (np.isnan(x) & np.isnan(y)))))
                if ndiff:
                    bad = np.where(~((x == y) | (np.isnan(x) & np.isnan(y))))
                    print("variant %d shape (%d,%d,%d): %s differs in %d entries, max abs %.3g, tracks %s" % (
                        v, block, bpc, tile, name, ndiff, np.nanmax(np.abs(x.astype(np.float64) - y)),
                        np.unique(bad[-1])[:10]))
                assert ndiff == 0


def test_batched_graph_replay(gpu_device):
    import torch
    geo, p, ev, rl, xi = event(2, 5000, 44)
    ctx = mk.Context(p.bfield, rl, xi)
    b = mk.FitBatch.from_numpy(ev, gpu_device, pflags=mk.PF_APPLY_MATERIAL)
    st = b.clone_state()
    outs = []
    for _ in range(3):
        b.restore_state(st)
        ctx.fit_batched(b)
        torch.cuda.synchronize()
        outs.append((_np(b.par).copy(), _np(b.err).copy(), _np(b.chi2).copy()))
        _ = torch.ones(1000, device=gpu_device).sum().item()  # eager work between replays
    for o in outs[1:]:
        for x, y in zip(o, outs[0]):
            np.testing.assert_array_equal(x, y)


def test_full_size_properties(gpu_device):
    """1M tracks, 25 layers (the plain-at-scale configuration): size-independent checks."""
    import torch
    geo, p, ev, rl, xi = event(2, 1000000, 51)
    ctx = mk.Context(p.bfield, rl, xi)
    a = _fit_gpu(ctx, ev, gpu_device, mk.PF_APPLY_MATERIAL, mk.FIT_CHI2, False)
    b = _fit_gpu(ctx, ev, gpu_device, mk.PF_APPLY_MATERIAL, mk.FIT_CHI2, True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    par, err, chg, chi2 = b
    nh = ev["n_hits"]
    ok = nh > 0
    assert np.isfinite(par).all() and np.isfinite(err).all() and np.isfinite(chi2).all()
    assert (chi2 >= 0).all()
    for dgi in (0, 2, 5, 9, 14, 20):
        assert (err[dgi][ok] > 0).all()
    ndf = 2 * nh[ok]
    m = float(np.mean(chi2[ok] / ndf))
    print("chi2/ndf mean", m)
    assert 0.7 < m < 1.6
    # a seeded sample through the oracle
    idx = np.random.default_rng(0).choice(len(nh), 20000, replace=False)
    sub = {k: np.ascontiguousarray(v[..., idx]) for k, v in ev.items()}
    o, o64, chk = fit_reference(p.bfield, rl, xi, sub, mk.PF_APPLY_MATERIAL, mk.FIT_CHI2)
    assert_track_parity("full-size sample", (par[:, idx], err[:, idx], chi2[idx]), (o[0], o[1], o[3]),
                        (o64[0], o64[1], o64[3]), chk)
