"""Pins of the oracle's strip charge cut (passStripChargePCMfromTrack, MkFinder.cc:863-896, with the charge
code of Hit::PackedData, Hit.h:251-282):
  * the 8-bit code's encode / decode known answers (1620 per cm and below -> 0, steps of 8, 0xff overflow);
  * the cut for one state and hit against the path-length factor written as a quadratic form,
    qSF = |sin theta| sqrt|1 + v^T P v| with v = (cos phi, -sin phi) and P the hit's xy covariance over its
    trace (barrel, when det P < 0.1; else 1), |cos theta| on a disk;
  * in FindCandidatesCloneEngine's records for one layer: with charge codes the records are those without
    codes whose hit passes the cut at the state propagated to the hit, in the same order;
  * in the whole-event finders: all-overflow codes change nothing, all-zero codes take no hit.
"""
import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from test_oracle_finder_layer import _lanes_at, _layer_hits

F = np.float32


def _pass64(barrel, code, par, herr):
    """The cut in float64 from the quadratic form; returns (passes, corrected charge per cm)."""
    pcm = float(oracle.charge_pcm(code))
    if pcm >= 1620 + (0xfe << 3):
        return True, pcm
    par = np.asarray(par, np.float64); e = np.asarray(herr, np.float64)
    if barrel:
        P = np.array([[e[0], e[1]], [e[1], e[2]]]) / (e[0] + e[2])
        v = np.array([np.cos(par[4]), -np.sin(par[4])])
        qsf = abs(np.sin(par[5])) * np.sqrt(abs(1.0 + v @ P @ v)) if abs(np.linalg.det(P)) < 0.1 else 1.0
    else:
        qsf = abs(np.cos(par[5]))
    return pcm * qsf > 1620, pcm * qsf


def test_charge_code_known_answers():
    pcm = np.array([0, 1000, 1619, 1620, 1627, 1628, 2000, 3651, 3652, 3660, 10000])
    code = oracle.charge_code(pcm)
    assert code.tolist() == [0, 0, 0, 1, 1, 2, 48, 254, 255, 255, 255]
    assert oracle.charge_pcm(code).tolist() == [0, 0, 0, 1620, 1620, 1628, 1996, 3644, 3652, 3652, 3652]
    assert oracle.charge_pcm(0xff) == 1620 + (0xfe << 3)               # Hit::maxChargePerCM


def _strip_err(rng, phi, barrel):
    """A strip hit's covariance: tight along one direction, long along the strip (rank-deficient in xy for a
    barrel strip along z), or round in xy (a pixel-like hit, det P = 1/4)."""
    if rng.random() < 0.25:
        s = rng.uniform(1e-4, 1e-2)
        return np.array([s, 0, s, 0, 0, rng.uniform(1e-4, 1e-2)], F)
    c, s = np.cos(phi), np.sin(phi)
    a, b = rng.uniform(1e-6, 1e-4), rng.uniform(0.5, 5.0)             # across, along
    if barrel:
        # along the z axis: xy error only across the strip, direction (-sin, cos)
        return np.array([a * s * s, -a * s * c, a * c * c, 0, 0, b], F)
    return np.array([a * c * c + b * s * s, (a - b) * s * c, a * s * s + b * c * c, 0, 0, 1e-4], F)


@pytest.mark.parametrize("barrel", [True, False])
def test_cut_matches_quadratic_form(barrel):
    rng = np.random.default_rng(17 + barrel)
    n_pass = n_fail = 0
    for _ in range(3000):
        phi = rng.uniform(-np.pi, np.pi)
        par = np.array([50 * np.cos(phi), 50 * np.sin(phi), rng.uniform(-100, 100), rng.uniform(-2, 2),
                        rng.uniform(-np.pi, np.pi), rng.uniform(0.1, np.pi - 0.1)], F)
        herr = _strip_err(rng, phi, barrel)
        code = int(rng.integers(0, 256))
        ok, q = oracle.pass_strip_charge(barrel, code, par, herr)
        ok64, q64 = _pass64(barrel, code, par, herr)
        assert q == pytest.approx(q64, rel=2e-6, abs=1e-3), (code, par, herr)
        if abs(q64 - 1620) > 1e-4 * 1620:
            assert ok == ok64, (code, par, herr, q, q64)
        n_pass += ok; n_fail += not ok
    assert n_pass > 300 and n_fail > 300
    # the overflow code passes whatever the direction; code 0 never does
    par = np.array([50, 0, 0, 1, 0, 0.05], F)
    assert oracle.pass_strip_charge(barrel, 0xff, par, _strip_err(rng, 0.0, barrel)) == (True, 3652.0)
    assert oracle.pass_strip_charge(barrel, 0, par, _strip_err(rng, 0.0, barrel))[0] is False


def make_event():
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=91, simulate_material=0, pt_min=0.5, pt_max=10.0)
    ev = mk.generate_find_event(geo, p, 600, noise_per_cm2=3.0)
    rng = np.random.default_rng(23)
    total = int(ev["hit_offset"][-1])
    # a fifth of the hits overflow (pass), a tenth carry code 0 (fail), the rest lie near the cut
    code = rng.integers(2, 80, total).astype(np.uint8)      # code 1 is 1620 per cm: qSF = 1 would sit on the cut
    u = rng.random(total)
    code[u < 0.2] = 0xff
    code[(u >= 0.2) & (u < 0.3)] = 0
    return ev, mk.barrel_layer_descs(geo), code


@pytest.fixture(scope="module")
def event():
    return make_event()


@pytest.mark.parametrize("layer", [3, 9])
def test_clone_engine_records_with_charge(event, layer):
    ev, descs, code = event
    O = oracle.Oracle(bfield=3.8)
    d = descs[layer]
    ln = _lanes_at(O, ev, descs, layer, seed=layer)
    hl = _layer_hits(ev, layer)
    lc = code[ev["hit_offset"][layer]:ev["hit_offset"][layer + 1]]
    parP, errP, chg = ln["prop_par"], ln["prop_err"], ln["chg"]
    xs, xa, w = oracle.select_hit_indices(d, hl, parP, errP)
    kw = dict(chi2_cut=30.0, max_holes=3, max_consec_holes=1, bfield=3.8)
    ne0, ent0 = oracle.find_candidates_ce(ev, descs, layer, ln, xs, xa, w, **kw)
    ne1, ent1 = oracle.find_candidates_ce(dict(ev, hit_charge=code), descs, layer, ln, xs, xa, w, **kw)
    n = parP.shape[1]
    kept = dropped = 0
    for t in range(n):
        exp = []
        for r in ent0[t, :ne0[t]]:
            if r["hit_idx"] >= 0:
                h = hl[r["hit_idx"]]
                pe, pp, _ = O.propagate_to_r(errP[:, t:t + 1], parP[:, t:t + 1], chg[t:t + 1],
                                             np.array([np.hypot(h[0], h[1])], F))
                ok, q = _pass64(True, lc[r["hit_idx"]], pp[:, 0], h[3:])
                if abs(q - 1620) < 2e-5 * 1620:
                    pytest.skip("a corrected charge within rounding of the cut")
                if not ok:
                    dropped += 1
                    continue
                kept += 1
            exp.append(r)
        assert ne1[t] == len(exp), t
        assert all(ent1[t, k].tobytes() == e.tobytes() for k, e in enumerate(exp)), t
    assert kept > 20 and dropped > 20, (kept, dropped)


@pytest.mark.parametrize("standard", [False, True])
def test_whole_event_overflow_and_zero_codes(event, standard):
    ev, descs, code = event
    kw = dict(chi2_cut=30.0, max_cands=4, n_threads=8, standard=standard)
    ref = oracle.find_clone_engine(ev, descs, 3, **kw)
    ovf = oracle.find_clone_engine(dict(ev, hit_charge=np.full_like(code, 0xff)), descs, 3, **kw)
    for k in ref:
        assert np.array_equal(ref[k], ovf[k]), k
    assert (ref["hot_idx"][:, 3:] >= 0).any()
    none = oracle.find_clone_engine(dict(ev, hit_charge=np.zeros_like(code)), descs, 3, **kw)
    nc = none["n_cands"]
    for s in range(ev["n_seeds"]):
        assert (none["hot_idx"][:nc[s], 3:, s] < 0).all(), s
    # with the mixed codes some found hits go, and the finders still find hits
    mixed = oracle.find_clone_engine(dict(ev, hit_charge=code), descs, 3, **kw)
    assert (mixed["hot_idx"][:, 3:] >= 0).any() and not np.array_equal(mixed["hot_idx"], ref["hot_idx"])
