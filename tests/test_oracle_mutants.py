"""The oracle pins have teeth: each planted mistake must make some pin fail.

Every mutant is a one-place textual change to a copy of the oracle source (a
dropped term, a flipped sign, a transposed operand, a wrong index, a changed
constant or comparison), compiled into a temporary library and run through the
pins.  Fit mutants (oracle_core.h) run through every pin of tests/pins.py;
finding mutants (oracle_find_core.h, oracle_find.cc, oracle_ce.cc) through the
finding pins of test_oracle_find_pins.py, test_oracle_find.py and
test_oracle_ce.py, cheapest first, until one fails.  A mutant that passes all
pins would mean that part of the restatement is not pinned; the test fails and
names it.
"""
import concurrent.futures
import os
import shutil
import subprocess
import tempfile

import pytest

import oracle
import pins

ORACLE_DIR = os.path.dirname(oracle.__file__)

MUTANTS = {
    # propagation: helix step and its Jacobian (PropagationMPlex.icc)
    "jac03_sign": ("J(0, 3) = k * (cosPorT", "J(0, 3) = -k * (cosPorT"),
    "jac14_drop_term": ("+ sinPorT * cosa + cosPorT * sina - sinPorT) * pt;", "+ sinPorT * cosa - sinPorT) * pt;"),
    "jac44_sign": ("J(4, 4) = T(1) + dadphi;", "J(4, 4) = T(1) - dadphi;"),
    "jac25_power": ("J(2, 5) = -k * alpha * pt * sinPorT * sinPorT;", "J(2, 5) = -k * alpha * pt * sinPorT;"),
    "dDdphi_sign": ("dDdphi += k * (x * (pysa", "dDdphi -= k * (x * (pysa"),
    "sincos4_drop_x4": ("c = T(1) - T(0.5f) * x2 + T(0.04166667f) * x2 * x2;", "c = T(1) - T(0.5f) * x2;"),
    "fail_no_negative": ("if (oodotp > T(5.0f) || oodotp < T(0)) {", "if (oodotp > T(5.0f)) {"),
    "helixz_j05": ("J(0, 5) = deltaZ * ipt * pxcaMpysa * icos2T;", "J(0, 5) = deltaZ * ipt * pxcaMpysa;"),
    "helixz_j43_sign": ("J(4, 3) = tanT * deltaZ * kinv;", "J(4, 3) = -tanT * deltaZ * kinv;"),
    # similarity
    "similarity_transposed_blocked": ("const T* a = J + (j * 6 + k) * NN;", "const T* a = J + (k * 6 + j) * NN;"),
    # material (PropagationMPlex.cc:716-767)
    "highland_log_term": ("(T(1) + T(0.038f) * std::log(radL))", "(T(1))"),
    "eloss_sign": ("const T dP = propSign[n] * dEdx / beta;", "const T dP = -propSign[n] * dEdx / beta;"),
    "ms_theta_only": ("outErr[sym(4, 4) * NN + n] += thetaMSC2;", ""),
    # Kalman update (KalmanUtilsMPlex.cc)
    "rotation_sign": ("const T a00 = -(my + py) / (2 * r);", "const T a00 = (my + py) / (2 * r);"),
    "gain_index": ("K[i][1] = C(i, 0) * h[1] + C(i, 1) * h[4] + C(i, 2) * h[7];",
                   "K[i][1] = C(i, 0) * h[1] + C(i, 1) * h[4] + C(i, 2) * h[6];"),
    "chi2_cross_term": ("2 * (s1 * l1 * l0)", "(s1 * l1 * l0)"),
    "khc_drop_z": ("KH[i][0] * C(0, j) + KH[i][1] * C(1, j) + KH[i][2] * C(2, j);",
                   "KH[i][0] * C(0, j) + KH[i][1] * C(1, j);"),
    "endcap_gain_swap": ("K[i][1] = C(i, 0) * s1 + C(i, 1) * s2;", "K[i][1] = C(i, 1) * s1 + C(i, 0) * s2;"),
    "inverse_offdiag_sign": ("a1 *= -s;", "a1 *= s;"),
}


# Finding restatement: (file, original text, mutated text).  Scores and their constants (Config.h:355-359,
# Track.h:648-670), invalid-hit codes and counts (MkFinder.cc:1237-1270), hit records (:1210-1225), the
# candidate bookkeeping (CandCloner.cc:37-191, HitStructures.h:873-889), the best short candidate and the
# final merge (HitStructures.cc:511-552), the stop rules (MkBuilder.cc:1521-1533, 1568-1630), best-hit
# selection and fake codes (MkFinder.cc:691-832) and the window / bin arithmetic (MkFinder.cc:227-684,
# HitStructures.cc:77-183).  Left out as equivalent, no input telling them apart: dropping the pop_back
# before the best short candidate's insertion (HitStructures.cc:540-545), since the output keeps the
# first max_cands either way.
FIND_MUTANTS = {
    # dead-module gaps (MkFinder.cc:451-462, 1263-1267)
    "gap_ignores_hits": ("oracle_ce.cc",
                         "else if (in_gap && in_gap[i] && added[i] == 0) fake = -7;   // dead module (MkFinder.cc:1263-1267)",
                         "else if (in_gap && in_gap[i]) fake = -7;"),
    "gap_counts_as_hole": ("oracle_ce.cc", "if (idx == -1) ++c.ntail;", "if (idx == -1 || idx == -7) ++c.ntail;"),
    "score_valid_bonus": ("oracle_ce.cc", "kValidHitBonus = 4,", "kValidHitBonus = 5,"),
    "score_valid_slope": ("oracle_ce.cc", "kValidHitSlope = 0.2f", "kValidHitSlope = 0.25f"),
    "score_hole_penalty": ("oracle_ce.cc", "kMissingHitPenalty = 8,", "kMissingHitPenalty = 7,"),
    "score_tail_penalty": ("oracle_ce.cc", "kTailMissingHitPenalty = 3;", "kTailMissingHitPenalty = 2;"),
    "score_lowpt_find_penalty": ("oracle_ce.cc", "penalty *= in_find ? 1.7f : 1.5f;",
                                 "penalty *= in_find ? 1.5f : 1.5f;"),
    "score_lowpt_final_penalty": ("oracle_ce.cc", "penalty *= in_find ? 1.7f : 1.5f;",
                                  "penalty *= in_find ? 1.7f : 1.7f;"),
    "score_lowpt_bonus": ("oracle_ce.cc", "bonus * (in_find ? 0.9f : 1.0f)", "bonus * (in_find ? 1.0f : 1.0f)"),
    "score_max_bonus": ("oracle_ce.cc", "const float maxBonus = 8.0;", "const float maxBonus = 9.0;"),
    "score_lowpt_threshold": ("oracle_ce.cc", "if (pt < 0.9) {", "if (pt < 0.8) {"),
    "score_negative_chi2": ("oracle_ce.cc", "  if (c < 0) c = 0.f;\n", ""),
    "count_inside_holes": ("oracle_ce.cc", "      c.ninside += c.ntail;\n", ""),
    "count_tail_any_fake": ("oracle_ce.cc", "if (idx == -1) ++c.ntail;", "if (idx < 0) ++c.ntail;"),
    "fake_consec_le": ("oracle_ce.cc", "lc[i].ntail < P.max_consec_holes) ? -1 : -2;",
                       "lc[i].ntail <= P.max_consec_holes) ? -1 : -2;"),
    "fake_holes_inside_only": ("oracle_ce.cc", "(lc[i].ninside + lc[i].ntail < P.max_holes",
                               "(lc[i].ninside < P.max_holes"),
    "fake_edge_code": ("oracle_ce.cc", "if (wsr[i] == WSR_Edge) fake = -3;\n    else if (in_gap && in_gap[i] && added[i] == 0) fake = -7;   // dead module (MkFinder.cc:1263-1267)",
                       "if (wsr[i] == WSR_Edge) fake = -1;\n    else if (in_gap && in_gap[i] && added[i] == 0) fake = -7;"),
    "fake_tailholes": ("oracle_ce.cc", "e.ntailholes = fake == -1 ? lc[i].ntail + 1 : lc[i].ntail;",
                       "e.ntailholes = lc[i].ntail + 1;"),
    "hit_record_holes": ("oracle_ce.cc", "e.nholes = lc[i].ninside + lc[i].ntail;", "e.nholes = lc[i].ninside;"),
    "ce_chi2_cut": ("oracle_ce.cc", "if (!(c2 < max_c2)) continue;", "if (!(c2 < 0.5f * max_c2)) continue;"),
    "charge_cross_sign": ("oracle_ce.cc", "sinP * sinP * proj[2] -", "sinP * sinP * proj[2] +"),
    "charge_decode": ("oracle_ce.cc", "((code - 1u) << 3) + 1620u", "(code << 3) + 1620u"),
    "charge_endcap_sin": ("oracle_ce.cc", "qSF = std::abs(std::cos(P(5)));", "qSF = std::abs(std::sin(P(5)));"),
    "charge_wrong_hit": ("oracle_ce.cc", "if (charge && !pass_strip_charge(barrel, charge_pcm_of(charge[xa[hc * NN + i]])",
                         "if (charge && !pass_strip_charge(barrel, charge_pcm_of(charge[hc * NN + i])"),
    "best_short_compare": ("oracle_ce.cc", "if (h.score > cc.best_short.score) cc.best_short = tc;",
                           "if (h.score < cc.best_short.score) cc.best_short = tc;"),
    "short_kept_live": ("oracle_ce.cc", "if (h.score > cc.best_short.score) cc.best_short = tc;\n          continue;",
                        "if (h.score > cc.best_short.score) cc.best_short = tc;"),
    "extras_at_end": ("oracle_ce.cc", "while (xi < extras.size() && extras[xi].score > tc.score",
                      "while (false && xi < extras.size() && extras[xi].score > tc.score"),
    "top_k_one_short": ("oracle_ce.cc", "const int num_hits = std::min((int)entries.size(), max_cands);",
                        "const int num_hits = std::min((int)entries.size(), std::max(1, max_cands - 1));"),
    "merge_no_rescore": ("oracle_ce.cc", "    for (auto& c : cc.cands) c.score = score_cand(c);\n", ""),
    "apex_pt": ("oracle_ce.cc", "if (c.pT() < 1.2) {", "if (c.pT() < 0.6) {"),
    "apex_no_dphi_test": ("oracle_ce.cc", "625.f && dphi > 1.371f && dphi < 4.512f", "625.f"),
    "barrel_miss_not_stopped": ("oracle_ce.cc", "cc.add_hit(held.back(), -2, layer, 0);", ""),
    "update_no_charge_flip": ("oracle_ce.cc", "      chg[n] = -chg[n];\n", ""),
    "bh_chi2_le": ("oracle_find.cc", "if (c2 < minChi2[it]) {", "if (c2 <= minChi2[it]) {"),
    "bh_fake_inside_only": ("oracle_find.cc", "else if (B.ninside[it] + B.ntail[it])", "else if (B.ninside[it])"),
    "bh_edge_code": ("oracle_find.cc", "fake_idx = -3;", "fake_idx = -1;"),
    "bh_chi2_not_summed": ("oracle_find.cc", "      B.chi2[it] += minChi2[it];\n", ""),
    "bh_count_inside": ("oracle_find.cc", "      B.ninside[it] += B.ntail[it];\n      B.ntail[it] = 0;",
                        "      B.ntail[it] = 0;"),
    "bh_fake_hit_error": ("oracle_find.cc", "const float f[6] = {666, 0, 666, 0, 0, 666};",
                          "const float f[6] = {66, 0, 66, 0, 0, 66};"),
    "bh_stop_code": ("oracle_find.cc", "B.hot_idx[std::min(B.nhits[i], kMaxTrkHits) - 1][i] == -2",
                     "B.hot_idx[std::min(B.nhits[i], kMaxTrkHits) - 1][i] == -3"),
    "bh_barrel_miss": ("oracle_find.cc", "if (x * x + y * y < r_min_sqr) {", "if (x * x + y * y < 0.8f * r_min_sqr) {"),
    "cdist_no_wrap": ("oracle_find_core.h", "return a > kPI ? kTwoPI - a : a;", "return a;"),
    "nsigma_phi": ("oracle_find_core.h", "const float nSigmaPhi = 3, nSigmaZ = 3", "const float nSigmaPhi = 2.5, nSigmaZ = 3"),
    "phi_bin_end": ("oracle_find_core.h", "pb2v[it] = phi_bin(phi + dphi) + 1;", "pb2v[it] = phi_bin(phi + dphi);"),
    "q_bin_end": ("oracle_find_core.h", "qb2v[it] = q_bin_checked(L.b, q + dq) + 1;",
                  "qb2v[it] = q_bin_checked(L.b, q + dq);"),
    "window_phi_cut": ("oracle_find_core.h", "if (ddphi >= dphi) continue;", "if (ddphi > 1.1f * dphi) continue;"),
    "loh_key_shift": ("oracle_find_core.h", "(q_bin_checked(L.b, q[i]) << 16)", "(q_bin_checked(L.b, q[i]) << 15)"),
    "edge_correction": ("oracle_find_core.h",
                        "const float edgeCorr = std::abs(0.5f * (d.rout - d.rin) / std::tan(par[5 * NN + it]));",
                        "const float edgeCorr = 0.f;"),
    "wsr_inside_margin": ("oracle_find_core.h", "if (z < d.zmax - dz && z > d.zmin + dz) return WSR_Inside;",
                          "if (z < d.zmax && z > d.zmin) return WSR_Inside;"),
}

# Standard combinatorial finding (oracle_ce.cc, standard mode): new candidates' states (MkFinder.cc:1015-1060,
# 1082-1106), the propagation flags of their update, their scores, and the per-seed selection
# (MkBuilder.cc:1760-1828).
STD_MUTANTS = {
    "std_fake_parent_state": ("nc.par[k] = B.parP[k * NN + i];", "(void)0;"),
    "std_hit_not_updated": ("nc.par[k] = oPar[k * NN + i];", "nc.par[k] = B.parP[k * NN + i];"),
    "std_update_inter_flags": ("n_proc, oErr, oPar, P.intra_pflags);", "n_proc, oErr, oPar, ev.pflags);"),
    "std_no_tail_penalty": ("return score_calc(c.nfound, c.ntail, c.noverlap,", "return score_calc(c.nfound, 0, c.noverlap,"),
    "std_shorts_placed": ("if (cc.last_idx(tc) != -2) {", "if (true) {"),
    "std_best_short_always": ("if (tc.score > cc.best_short.score) cc.best_short = tc;", "cc.best_short = tc;"),
    "std_held_dropped": ("std::vector<Cand>& held = standard ? tmp : extras;", "std::vector<Cand>& held = extras;"),
    "std_fake_consec_le": ("c.ntail < P.max_consec_holes) ? -1 : -2;", "c.ntail <= P.max_consec_holes) ? -1 : -2;"),
    "std_fake_edge_code": ("if (wsr[i] == WSR_Edge) fake = -3;\n    else if (in_gap && in_gap[i] && added[i] == 0) fake = -7;   // dead module (MkFinder.cc:1092-1096)",
                           "if (wsr[i] == WSR_Edge) fake = -1;\n    else if (in_gap && in_gap[i] && added[i] == 0) fake = -7;"),
    "std_charge_ignored": ("(!charge || pass_strip_charge(", "(true || pass_strip_charge("),
    "std_sort_ascending": ("std::stable_sort(tmp.begin(), tmp.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });",
                           "std::stable_sort(tmp.begin(), tmp.end(), [](const Cand& a, const Cand& b) { return a.score < b.score; });"),
}

# Backward fit of found candidates (oracle_bkfit.cc): the input scaling (MkFinder.cc:1368-1399), the node walk,
# fake hits and block counts (MkFinder.cc:1573-1742), the plan limits (SteeringParams.h:138-151), the charge
# flip, the PCA step and the output score (MkFinder.cc:1425-1446, HitStructures.h:574-585,873-889).
BKFIT_MUTANTS = {
    "bk_scale": ("* T(100.0f);   // Scale(100.0f)", "* T(10.0f);   // Scale(100.0f)"),
    "bk_fake_err": ("const T fake[6] = {666, 0, 666, 0, 0, 666};", "const T fake[6] = {666, 0, 666, 0, 0, 66.6};"),
    "bk_fake_material": ("k.nomat[ln] = p->fake_material ? 0 : 1;", "k.nomat[ln] = 0;"),
    "bk_skip_empty_layers": ("    if (here_count == 0) continue;\n", ""),
    "bk_skip_only_minus_one": ("while (c >= 0 && H[c].index < 0) c = H[c].prev;",
                               "while (c >= 0 && H[c].index == -1) c = H[c].prev;"),
    "bk_forward_order": ("for (int l = L - 1; l >= p->last_layer; --l) {",
                         "for (int l = p->last_layer; l < L; ++l) {"),
    "bk_last_layer_excluded": ("for (int l = L - 1; l >= p->last_layer; --l) {",
                               "for (int l = L - 1; l > p->last_layer; --l) {"),
    "bk_pca_flags": ("k.errP, k.parP, cnt, p->pca_pflags, nullptr, ctx);", "k.errP, k.parP, cnt, p->pflags, nullptr, ctx);"),
    "bk_endcap_as_barrel": ("k.ec[ln] = valid[i] ? !ev->layers[base[i] + l].is_barrel : 0;", "k.ec[ln] = 0;"),
    "bk_score_inside": ("            ninside += ntail;\n", ""),
    "bk_score_tail_any": ("          } else if (h == -1) {", "          } else {"),
    "bk_score_lowpt": ("    penalty *= 1.5f;", "    penalty *= 1.7f;"),
    "bk_layer_offset": ("ev->hits + (size_t)(ev->hit_offset[base[i] + l] + H[c].index) * 9;",
                        "ev->hits + (size_t)(ev->hit_offset[l] + H[c].index) * 9;"),
}

ORACLE_FILES = ("oracle.cc", "oracle_find.cc", "oracle_ce.cc", "oracle_align.cc", "oracle_bkfit.cc", "oracle.h",
                "oracle_core.h", "oracle_find_core.h")


def _build(name, old, new, workdir, target="oracle_core.h"):
    d = os.path.join(workdir, name)
    os.makedirs(d)
    for f in ORACLE_FILES:
        shutil.copy(os.path.join(ORACLE_DIR, f), d)
    p = os.path.join(d, target)
    src = open(p).read()
    assert src.count(old) == 1, "%s: pattern matches %d times" % (name, src.count(old))
    open(p, "w").write(src.replace(old, new))
    so = os.path.join(d, "liborc_mut.so")
    srcs = [os.path.join(d, "oracle.cc")] if target == "oracle_core.h" else \
        [os.path.join(d, f) for f in ORACLE_FILES if f.endswith(".cc")]
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-pthread", "-DORC_NN=8", "-o", so] + srcs,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return so


def _first_caught(pin_list):
    """Name of the first pin that fails (an assertion, or a numpy error on the mutant's output), else None."""
    for pin_name, pin in pin_list:
        try:
            pin()
        except Exception as e:
            return "%s: %s" % (pin_name, type(e).__name__)
    return None


def _fit_pins(so):
    def make(bfield, rl, xi):
        return oracle.Oracle(bfield, rl, xi, lib_path=so)
    return [(pin.__name__, (lambda pin=pin: pin(make))) for pin in pins.ALL_PINS]


def _param_values(test):
    """The values of a test's pytest.mark.parametrize (one argument: values; several: tuples)."""
    return [m for m in test.pytestmark if m.name == "parametrize"][0].args[1]


def _find_pins():
    import test_oracle_ce as ce
    import test_oracle_charge as ch
    import test_oracle_find as bh
    import test_oracle_find_pins as fp
    return [
        ("invalid_hit_records", lambda: [fp.test_invalid_hit_record_known_answers(c)
                                         for c in range(len(fp.INVALID_HIT_CASES))]),
        ("outside_lane", fp.test_outside_lane_gets_no_record),
        ("hit_records", fp.test_hit_record_counts_and_score),
        ("loh_stable_sort", bh.test_loh_matches_stable_sort),
        ("select_brute_force", bh.test_select_hit_indices_matches_brute_force),
        ("duplicated_hits", fp.test_duplicated_hits_resolve_to_first_copy),
        ("score_formula", fp.test_score_formula_known_answers),
        ("apex_stop_rule", lambda: [fp.test_apex_stop_rule(c) for c in _param_values(fp.test_apex_stop_rule)]),
        ("ce_greedy", lambda: [fp.test_clone_engine_k1_is_greedy(z) for z in (None, 30.0)]),
        ("ce_all_paths_4_1", lambda: fp.test_clone_engine_equals_all_paths((4, 1, None))),
        ("ce_all_paths_1_1", lambda: fp.test_clone_engine_equals_all_paths((1, 1, None))),
        ("ce_all_paths_3_2", lambda: fp.test_clone_engine_equals_all_paths((3, 2, None))),
        ("ce_all_paths_edges", lambda: fp.test_clone_engine_equals_all_paths((2, 2, 20.0))),
        ("bh_loop_composition", lambda: bh.test_find_loop_matches_composition(bh.make_find_event())),
        ("bh_short_barrel", lambda: bh.test_find_loop_matches_composition_short_barrel(bh.make_find_event())),
        ("window_edge", bh.test_window_edge_counts_the_layer_thickness),
        ("ce_apex_before_disk", fp.test_apex_stop_before_a_disk),
        ("ce_short_layer", fp.test_clone_engine_across_a_short_layer),
        ("ce_k_limited", lambda: [fp.test_clone_engine_k_limited(z) for z in _param_values(fp.test_clone_engine_k_limited)]),
        ("ce_held_back_ahead", fp.test_clone_engine_held_back_goes_ahead),
        ("ce_charge_flip", fp.test_wrong_charge_seeds_flip_to_the_true_charge),
        ("dead_module_gap", lambda: [fp.test_dead_module_gap(v) for v in (False, True)]),
        ("ce_composition_k6", lambda: ce.test_ce_matches_composition(6)),
        ("strip_charge_cut", lambda: [ch.test_cut_matches_quadratic_form(b) for b in (True, False)]),
        ("strip_charge_records", lambda: [ch.test_clone_engine_records_with_charge(ev, la)
                                          for ev in [ch.make_event()] for la in (3, 9)]),
    ]


def _bkfit_pins():
    import test_oracle_bkfit as bk
    events = {}

    def ev(name):                           # built on first use, once per process
        if name not in events:
            events[name] = bk._barrel() if name == "barrel" else bk._mixed()
        return events[name]
    return [
        ("score_known_answers", lambda: [bk.test_score_known_answers(ev("barrel"), c)
                                         for c in range(len(bk.SCORE_CASES))]),
        ("blocks_1", lambda: bk.test_blocks_equal_fit_composition(ev("barrel"), 1)),
        ("blocks_16", lambda: bk.test_blocks_equal_fit_composition(ev("barrel"), 16)),
        ("pca_flags", lambda: bk.test_without_pca_and_other_pca_flags(ev("barrel"))),
        ("last_layer", lambda: bk.test_last_layer_stops_the_plan(ev("barrel"))),
        ("material", lambda: bk.test_material_and_fake_hit_material(ev("barrel"))),
        ("mixed", lambda: bk.test_mixed_layer_kinds_with_layer_base(ev("mixed"))),
    ]


def _std_pins():
    import test_oracle_charge as ch
    import test_oracle_std_pins as sp
    return [
        ("strip_charge_std", lambda: ch.test_whole_event_overflow_and_zero_codes(ch.make_event(), True)),
        ("selection_k1", lambda: sp.test_standard_selection(1, None, 0)),
        ("selection_k3_intra", lambda: sp.test_standard_selection(3, None, 1)),
        ("all_paths_4_1", lambda: sp.test_standard_equals_all_paths((4, 1, None))),
        ("all_paths_1_1", lambda: sp.test_standard_equals_all_paths((1, 1, None))),
        ("all_paths_2_2_short", lambda: sp.test_standard_equals_all_paths((2, 2, 20.0))),
    ]


def _pins_of(kind, so):
    return _fit_pins(so) if kind == "fit" else {"find": _find_pins, "bkfit": _bkfit_pins, "std": _std_pins}[kind]()


def _verdict(kind, so, only=None):
    """Runs in a worker process: the first pin of the mutant's kind that catches it (only: just that pin),
    or None."""
    with oracle.use_lib(so):
        pl = _pins_of(kind, so)
        return _first_caught(pl if only is None else [pl[only]])




# Every mutant: (kind, file, original text, mutated text).
ALL_MUTANTS = {}
ALL_MUTANTS.update({("fit", k): ("oracle_core.h", a, b) for k, (a, b) in MUTANTS.items()})
ALL_MUTANTS.update({("find", k): (f, a, b) for k, (f, a, b) in FIND_MUTANTS.items()})
ALL_MUTANTS.update({("bkfit", k): ("oracle_bkfit.cc", a, b) for k, (a, b) in BKFIT_MUTANTS.items()})
ALL_MUTANTS.update({("std", k): ("oracle_ce.cc", a, b) for k, (a, b) in STD_MUTANTS.items()})
# Controls: the unchanged oracle, built and run the same way, passes every pin of each kind, so a
# caught mutant is caught by its change and not by the runner.
for _kind, _file in (("fit", "oracle_core.h"), ("find", "oracle_find.cc"), ("bkfit", "oracle_bkfit.cc"),
                     ("std", "oracle_ce.cc")):
    _head = open(os.path.join(ORACLE_DIR, _file)).read()[:80]
    ALL_MUTANTS[(_kind, "control")] = (_file, _head, _head)


@pytest.fixture(scope="module")
def verdicts():
    """All mutants built on 8 threads, then run through their pins in 8 forked worker processes (the
    pins are deterministic and share nothing between mutants)."""
    import multiprocessing
    work = tempfile.mkdtemp(prefix="orc_mut_")
    n = min(8, os.cpu_count() or 1)
    try:
        with concurrent.futures.ThreadPoolExecutor(max_workers=n) as ex:
            futs = {k: ex.submit(_build, "%s_%s" % k, a, b, work, f) for k, (f, a, b) in ALL_MUTANTS.items()}
            libs = {k: f.result() for k, f in futs.items()}
        with concurrent.futures.ProcessPoolExecutor(max_workers=n,
                                                    mp_context=multiprocessing.get_context("fork")) as ex:
            # controls run every pin, so each pin is a job of its own; mutants stop at their first catch
            futs = {}
            for k, so in libs.items():
                if k[1] == "control":
                    futs[k] = [ex.submit(_verdict, k[0], so, i) for i in range(len(_pins_of(k[0], so)))]
                else:
                    futs[k] = [ex.submit(_verdict, k[0], so)]
            out = {}
            for k, fl in futs.items():
                v = [f.result() for f in fl]
                out[k] = next((x for x in v if x), None)
            yield out
    finally:
        shutil.rmtree(work, ignore_errors=True)


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_is_caught(verdicts, name):
    assert verdicts[("fit", name)], "mutant %s passes every pin" % name


@pytest.mark.parametrize("name", sorted(FIND_MUTANTS))
def test_find_mutant_is_caught(verdicts, name):
    assert verdicts[("find", name)], "mutant %s passes every finding pin" % name


@pytest.mark.parametrize("name", sorted(BKFIT_MUTANTS))
def test_bkfit_mutant_is_caught(verdicts, name):
    assert verdicts[("bkfit", name)], "mutant %s passes every backward-fit pin" % name


@pytest.mark.parametrize("name", sorted(STD_MUTANTS))
def test_std_mutant_is_caught(verdicts, name):
    assert verdicts[("std", name)], "mutant %s passes every standard-finding pin" % name


@pytest.mark.parametrize("kind", ["fit", "find", "bkfit", "std"])
def test_unchanged_oracle_passes_every_pin(verdicts, kind):
    assert verdicts[(kind, "control")] is None, verdicts[(kind, "control")]
