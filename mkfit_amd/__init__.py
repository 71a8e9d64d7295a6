"""mkfit_amd -- MI355X (gfx950) Kalman track fit behind mkFit's Matriplex interface.

Thin Python harness over the C ABI in include/mkfit_amd.h (libmkfit_amd.so) and
the toy Monte Carlo in include/mkfit_toymc.h (libmkfit_toymc.so).  Torch is
used only to own device memory and streams; every computation runs in the HIP
library.  Importing this package never falls back to CPU code: if the native
library is missing the import of `lib()` raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_TOY = None

PF_USE_PARAM_B_FIELD = 0x1
PF_APPLY_MATERIAL = 0x2
KFO_CALCULATE_CHI2 = 1
KFO_UPDATE_PARAMS = 2
FIT_CHI2 = 1
FIT_CHARGE_FLIP = 2
FIT_FP64_FALLBACK = 4
FIT_PROP_TO_PCA = 8
HIT_ENDCAP = 1
HIT_MISSING = 2
NBINS_Z, NBINS_R = 300, 120
MAX_LAYERS = 64

_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_void_p = ctypes.c_void_p


class MkfConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("bfield", ctypes.c_float),
                ("rl_grid", _c_void_p), ("xi_grid", _c_void_p)]


class MkfFitBatch(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int), ("n_slots", ctypes.c_int),
                ("par", _c_void_p), ("err", _c_void_p), ("chg", _c_void_p), ("chi2", _c_void_p),
                ("hit_pos", _c_void_p), ("hit_err", _c_void_p), ("n_hits", _c_void_p),
                ("hit_flags", _c_void_p), ("pflags", ctypes.c_int), ("fit_flags", ctypes.c_int),
                ("seed_par", _c_void_p), ("seed_err", _c_void_p), ("seed_chg", _c_void_p),
                ("pca_pflags", ctypes.c_int)]


class MkfFindShape(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("waves_per_simd", ctypes.c_int)]


class MkfLaunchShape(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("blocks_per_cu", ctypes.c_int), ("variant", ctypes.c_int),
                ("sort_tile", ctypes.c_int)]


class MkfGeometry(ctypes.Structure):
    _fields_ = [("n_barrel", ctypes.c_int),
                ("barrel_r", ctypes.c_float * MAX_LAYERS), ("barrel_zmax", ctypes.c_float * MAX_LAYERS),
                ("n_disk", ctypes.c_int),
                ("disk_z", ctypes.c_float * MAX_LAYERS), ("disk_rmin", ctypes.c_float * MAX_LAYERS),
                ("disk_rmax", ctypes.c_float * MAX_LAYERS),
                ("x_over_x0", ctypes.c_float * MAX_LAYERS), ("xi", ctypes.c_float * MAX_LAYERS)]


class MkfToyParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("bfield", ctypes.c_float),
                ("pt_min", ctypes.c_float), ("pt_max", ctypes.c_float), ("eta_max", ctypes.c_float),
                ("vtx_sxy", ctypes.c_float), ("vtx_sz", ctypes.c_float),
                ("sig_rphi", ctypes.c_float), ("sig_z", ctypes.c_float), ("sig_r_endcap", ctypes.c_float),
                ("simulate_material", ctypes.c_int), ("hits_min", ctypes.c_int), ("hits_max", ctypes.c_int),
                ("missing_frac", ctypes.c_float),
                ("seed_sig_pos", ctypes.c_float), ("seed_sig_z", ctypes.c_float),
                ("seed_rel_ipt", ctypes.c_float), ("seed_sig_ang", ctypes.c_float)]


class MkfToyFindOutput(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int), ("layer_capacity", ctypes.c_int), ("hits", _c_void_p),
                ("n_layer_hits", _c_void_p), ("hit_mc", _c_void_p), ("n_seeds", ctypes.c_int),
                ("seed_par", _c_void_p), ("seed_err", _c_void_p), ("seed_chg", _c_void_p),
                ("seed_hits", _c_void_p), ("seed_mc", _c_void_p)]


class MkfToyOutput(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int), ("n_slots", ctypes.c_int),
                ("seed_par", _c_void_p), ("seed_err", _c_void_p), ("chg", _c_void_p),
                ("hit_pos", _c_void_p), ("hit_err", _c_void_p), ("n_hits", _c_void_p),
                ("hit_flags", _c_void_p), ("hit_layer", _c_void_p), ("truth_par", _c_void_p)]


ABI_SYMBOLS = [
    "mkf_context_create", "mkf_context_destroy", "mkf_last_error",
    "mkf_propagate_helix_to_r", "mkf_propagate_helix_to_z",
    "mkf_kalman_operation", "mkf_kalman_operation_endcap",
    "mkf_fit_plain", "mkf_fit_batched",
    "mkf_set_batched_shape", "mkf_get_batched_shape", "mkf_autotune_batched", "mkf_num_batched_variants",
    "mkf_fit_order",
    "mkf_event_of_hits_create", "mkf_event_of_hits_destroy", "mkf_event_of_hits_load", "mkf_event_of_hits_layer",
    "mkf_select_hit_indices", "mkf_find_tracks_best_hit", "mkf_find_tracks_clone_engine",
    "mkf_add_best_hit", "mkf_find_candidates_clone_engine", "mkf_update_with_last_hit",
    "mkf_fit_adjoint", "mkf_fit_adjoint_noalloc", "mkf_backward_fit", "mkf_find_tracks_standard",
    "mkf_set_find_shape", "mkf_get_find_shape", "mkf_autotune_find", "mkf_event_of_hits_set_hit_mask",
    "mkf_event_of_hits_set_dead_regions", "mkf_event_of_hits_set_hit_charge",
]
CE_MAX_CANDS = 8

LAYER_DESC_FIELDS = ["is_barrel", "rin", "rout", "zmin", "zmax", "propagate_to", "q_bin", "min_dphi", "max_dphi",
                     "min_dq", "max_dq", "c_dp_sf", "c_dp_0", "c_dp_1", "c_dp_2", "c_dq_sf", "c_dq_0", "c_dq_1",
                     "c_dq_2", "c_c2_sf", "c_c2_0", "c_c2_1", "c_c2_2"]
MAX_FIND_LAYERS = 64
HIT_IDX_MAX = 16
WSR_INSIDE, WSR_EDGE, WSR_OUTSIDE = 0, 1, 2


class MkfLayerDesc(ctypes.Structure):
    _fields_ = [("is_barrel", ctypes.c_int)] + [(f, ctypes.c_float) for f in LAYER_DESC_FIELDS[1:]]


class MkfCEBatch(ctypes.Structure):
    _fields_ = [("n_seeds", ctypes.c_int), ("seed_par", _c_void_p), ("seed_err", _c_void_p), ("seed_chg", _c_void_p),
                ("seed_hits", _c_void_p), ("first_layer", ctypes.c_int), ("chi2_cut", ctypes.c_float),
                ("pflags", ctypes.c_int), ("intra_pflags", ctypes.c_int), ("prop_to_hit", ctypes.c_int),
                ("max_cands", ctypes.c_int), ("max_holes", ctypes.c_int), ("max_consec_holes", ctypes.c_int),
                ("n_cands", _c_void_p), ("par", _c_void_p), ("err", _c_void_p), ("chg", _c_void_p),
                ("chi2", _c_void_p), ("score", _c_void_p), ("n_found", _c_void_p), ("hot_idx", _c_void_p),
                ("n_layers", ctypes.c_int), ("layer_base", _c_void_p)]


BKFIT_PCA, BKFIT_FAKE_MATERIAL = 1, 2


class MkfBkFitBatch(ctypes.Structure):
    _fields_ = [("n_cands", ctypes.c_int), ("in_par", _c_void_p), ("in_err", _c_void_p), ("in_chg", _c_void_p),
                ("hot_idx", _c_void_p), ("n_layers", ctypes.c_int), ("last_layer", ctypes.c_int),
                ("layer_base", _c_void_p), ("group", ctypes.c_int), ("pflags", ctypes.c_int),
                ("pca_pflags", ctypes.c_int), ("flags", ctypes.c_int), ("par", _c_void_p), ("err", _c_void_p),
                ("chg", _c_void_p), ("chi2", _c_void_p), ("score", _c_void_p)]


class MkfAdjointBatch(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int), ("n_layers", ctypes.c_int)] + [
        (k, _c_void_p) for k in ("seed_par", "seed_err", "seed_chg", "hit_pos", "hit_err", "n_hits", "theta")] + [
        ("pflags", ctypes.c_int), ("xi_per_x0", ctypes.c_float), ("checkpoint_every", ctypes.c_int),
        ("chi2", _c_void_p), ("grad", _c_void_p), ("fp64", ctypes.c_int), ("tangent", ctypes.c_int)]


class MkfFinderLanes(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int)] + [(k, ctypes.c_void_p) for k in (
        "prop_par", "prop_err", "par", "err", "chg", "chi2", "n_found", "n_inside_minus_one", "n_tail_minus_one")]


# mkf_idx_chi2: IdxChi2List (Track.h:28-42) plus the record's position in its lane's hit list.
IDX_CHI2_FIELDS = [("trk_idx", "i4"), ("hit_idx", "i4"), ("module", "u4"), ("nhits", "i4"), ("ntailholes", "i4"),
                   ("noverlaps", "i4"), ("nholes", "i4"), ("pt", "f4"), ("chi2", "f4"), ("chi2_hit", "f4"),
                   ("score", "f4"), ("xhit_pos", "i4")]
IDX_CHI2_DTYPE = np.dtype([(k, "<" + t) for k, t in IDX_CHI2_FIELDS])
CE_ENTRIES_PER_LANE = 17


class MkfFindBatch(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int), ("par", _c_void_p), ("err", _c_void_p), ("chg", _c_void_p),
                ("chi2", _c_void_p), ("hot_idx", _c_void_p), ("n_found", _c_void_p), ("first_layer", ctypes.c_int),
                ("chi2_cut", ctypes.c_float), ("pflags", ctypes.c_int), ("prop_to_hit", ctypes.c_int),
                ("intra_pflags", ctypes.c_int), ("n_layers", ctypes.c_int), ("layer_base", _c_void_p)]


def lib_path():
    # MKFIT_AMD_LIB selects another build of the same library (A/B timing of kernel variants).
    return os.environ.get("MKFIT_AMD_LIB") or os.path.join(_HERE, "libmkfit_amd.so")


def lib():
    """The HIP library; raises if it has not been built (no CPU fallback exists)."""
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            raise RuntimeError("libmkfit_amd.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(p)
        L.mkf_last_error.restype = ctypes.c_char_p
        L.mkf_context_create.argtypes = [ctypes.POINTER(MkfConfig), ctypes.POINTER(_c_void_p)]
        L.mkf_context_destroy.argtypes = [_c_void_p]
        L.mkf_propagate_helix_to_r.argtypes = [_c_void_p, ctypes.c_int] + [_c_void_p] * 6 + [ctypes.c_int] + [_c_void_p] * 3
        L.mkf_propagate_helix_to_z.argtypes = [_c_void_p, ctypes.c_int] + [_c_void_p] * 6 + [ctypes.c_int] + [_c_void_p] * 2
        L.mkf_kalman_operation.argtypes = [_c_void_p, ctypes.c_int, ctypes.c_int] + [_c_void_p] * 8
        L.mkf_kalman_operation_endcap.argtypes = [_c_void_p, ctypes.c_int, ctypes.c_int] + [_c_void_p] * 8
        L.mkf_fit_plain.argtypes = [_c_void_p, ctypes.POINTER(MkfFitBatch), _c_void_p]
        L.mkf_fit_batched.argtypes = [_c_void_p, ctypes.POINTER(MkfFitBatch), _c_void_p]
        L.mkf_set_batched_shape.argtypes = [_c_void_p, ctypes.POINTER(MkfLaunchShape)]
        L.mkf_fit_order.argtypes = [_c_void_p, ctypes.POINTER(MkfFitBatch), ctypes.c_int, _c_void_p, _c_void_p]
        L.mkf_get_batched_shape.argtypes = [_c_void_p, ctypes.POINTER(MkfLaunchShape)]
        L.mkf_event_of_hits_create.argtypes = [_c_void_p, ctypes.c_int, ctypes.POINTER(MkfLayerDesc),
                                               ctypes.POINTER(_c_void_p)]
        L.mkf_event_of_hits_destroy.argtypes = [_c_void_p]
        L.mkf_event_of_hits_load.argtypes = [_c_void_p, _c_void_p, _c_void_p, _c_void_p]
        L.mkf_event_of_hits_set_hit_mask.argtypes = [_c_void_p, _c_void_p, _c_void_p]
        L.mkf_event_of_hits_set_hit_charge.argtypes = [_c_void_p, _c_void_p, _c_void_p]
        L.mkf_event_of_hits_set_dead_regions.argtypes = [_c_void_p, ctypes.c_int, _c_void_p, _c_void_p]
        L.mkf_event_of_hits_layer.argtypes = [_c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_int)] + [ctypes.POINTER(_c_void_p)] * 4
        L.mkf_select_hit_indices.argtypes = [_c_void_p, ctypes.c_int, ctypes.c_int] + [_c_void_p] * 6
        L.mkf_find_tracks_best_hit.argtypes = [_c_void_p, _c_void_p, ctypes.POINTER(MkfFindBatch), _c_void_p]
        L.mkf_find_tracks_clone_engine.argtypes = [_c_void_p, _c_void_p, ctypes.POINTER(MkfCEBatch), _c_void_p]
        L.mkf_fit_adjoint.argtypes = [_c_void_p, ctypes.POINTER(MkfAdjointBatch), _c_void_p]
        L.mkf_add_best_hit.argtypes = [_c_void_p, _c_void_p, ctypes.c_int, ctypes.POINTER(MkfFinderLanes)] + \
            [_c_void_p] * 3 + [ctypes.c_float, ctypes.c_int, ctypes.c_int, _c_void_p, _c_void_p]
        L.mkf_find_candidates_clone_engine.argtypes = [_c_void_p, _c_void_p, ctypes.c_int,
                                                       ctypes.POINTER(MkfFinderLanes)] + [_c_void_p] * 3 + \
            [ctypes.c_float] + [ctypes.c_int] * 4 + [_c_void_p] * 3
        L.mkf_update_with_last_hit.argtypes = [_c_void_p, _c_void_p, ctypes.c_int, ctypes.POINTER(MkfFinderLanes),
                                               _c_void_p, ctypes.c_int, ctypes.c_int, _c_void_p]
        L.mkf_autotune_batched.argtypes = [_c_void_p, ctypes.POINTER(MkfFitBatch), _c_void_p, ctypes.c_int,
                                           ctypes.POINTER(MkfLaunchShape), _c_float_p]
        L.mkf_set_find_shape.argtypes = [_c_void_p, ctypes.c_int, ctypes.POINTER(MkfFindShape)]
        L.mkf_get_find_shape.argtypes = [_c_void_p, ctypes.c_int, ctypes.POINTER(MkfFindShape)]
        L.mkf_autotune_find.argtypes = [_c_void_p, _c_void_p, ctypes.POINTER(MkfCEBatch), ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(MkfFindShape), _c_float_p, _c_void_p]
        _LIB = L
    return _LIB


def toy():
    global _TOY
    if _TOY is None:
        p = os.path.join(_HERE, "libmkfit_toymc.so")
        if not os.path.exists(p):
            raise RuntimeError("libmkfit_toymc.so is not built (run __graft_entry__.build())")
        T = ctypes.CDLL(p)
        T.mkf_toymc_geometry_preset.argtypes = [ctypes.c_int, ctypes.POINTER(MkfGeometry)]
        T.mkf_toymc_default_params.argtypes = [ctypes.POINTER(MkfToyParams)]
        T.mkf_toymc_generate.argtypes = [ctypes.POINTER(MkfGeometry), ctypes.POINTER(MkfToyParams),
                                         ctypes.POINTER(MkfToyOutput), ctypes.c_int]
        T.mkf_toymc_material_grid.argtypes = [ctypes.POINTER(MkfGeometry), _c_void_p, _c_void_p]
        T.mkf_toymc_find_event.argtypes = [ctypes.POINTER(MkfGeometry), ctypes.POINTER(MkfToyParams), ctypes.c_float,
                                           ctypes.c_int, ctypes.POINTER(MkfToyFindOutput), ctypes.c_int]
        T.mkf_toymc_noise_count.argtypes = [ctypes.POINTER(MkfGeometry), ctypes.c_int, ctypes.c_float]
        _TOY = T
    return _TOY


class MkfError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise MkfError("mkfit_amd error %d: %s" % (rc, lib().mkf_last_error().decode()))


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


# ----------------------------------------------------------------------------
# Toy Monte Carlo
# ----------------------------------------------------------------------------

def geometry(preset):
    g = MkfGeometry()
    if toy().mkf_toymc_geometry_preset(int(preset), ctypes.byref(g)) != 0:
        raise ValueError("unknown geometry preset %r" % (preset,))
    return g


def default_toy_params(**kw):
    p = MkfToyParams()
    toy().mkf_toymc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def material_grid(geo):
    rl = np.zeros((NBINS_Z, NBINS_R), np.float32)
    xi = np.zeros((NBINS_Z, NBINS_R), np.float32)
    toy().mkf_toymc_material_grid(ctypes.byref(geo), rl.ctypes.data_as(_c_void_p), xi.ctypes.data_as(_c_void_p))
    return rl, xi


def generate(geo, params, n_tracks, n_slots, n_threads=None, truth=False):
    """Runs the toy MC; returns a dict of numpy arrays in the fit's layout."""
    n, S = int(n_tracks), int(n_slots)
    out = {
        "seed_par": np.zeros((6, n), np.float32), "seed_err": np.zeros((21, n), np.float32),
        "chg": np.zeros(n, np.int32), "hit_pos": np.zeros((S, 3, n), np.float32),
        "hit_err": np.zeros((S, 6, n), np.float32), "n_hits": np.zeros(n, np.int32),
        "hit_flags": np.zeros((S, n), np.uint8), "hit_layer": np.zeros((S, n), np.int32),
    }
    if truth:
        out["truth_par"] = np.zeros((S, 6, n), np.float32)
    o = MkfToyOutput()
    o.n_tracks, o.n_slots = n, S
    for k in ("seed_par", "seed_err", "chg", "hit_pos", "hit_err", "n_hits", "hit_flags", "hit_layer"):
        setattr(o, k, out[k].ctypes.data_as(_c_void_p))
    o.truth_par = out["truth_par"].ctypes.data_as(_c_void_p) if truth else None
    if n_threads is None:
        n_threads = min(16, os.cpu_count() or 1)
    rc = toy().mkf_toymc_generate(ctypes.byref(geo), ctypes.byref(params), ctypes.byref(o), int(n_threads))
    if rc != 0:
        raise RuntimeError("toy MC generation failed (%d)" % rc)
    return out


def fit_order(ev):
    """The order in which a fit batch packs its tracks: by hit count, then by the endcap/barrel pattern of the
    slots (bit s for an endcap hit in slot s), stable.  Tracks of one length and one pattern then share
    wavefronts, so a wavefront's slot loop lasts as long as its tracks and takes one branch per slot (the
    same grouping the reference's fit driver makes when it splits blocks on a change of layer signature,
    MkBuilder.cc:594-625, and FitTracksSteered's one barrel/endcap pattern per block, MkFitter.cc:473-503).
    Returns the permutation: position i of the ordered batch holds track perm[i]."""
    nh = np.asarray(ev["n_hits"])
    S = ev["hit_flags"].shape[0] if "hit_flags" in ev else 0
    mask = np.zeros(nh.shape[0], np.uint64)
    for s in range(min(S, 40)):
        ec = ((ev["hit_flags"][s] & HIT_ENDCAP) != 0) & (s < nh)
        mask |= ec.astype(np.uint64) << np.uint64(s)
    key = (np.minimum(nh, 1023).astype(np.uint64) << np.uint64(40)) | mask
    return np.argsort(key, kind="stable")


def ordered_for_fit(ev):
    """The event's tracks in fit_order: (ordered event, perm); results of the ordered batch at position i
    belong to track perm[i] of the original event."""
    perm = fit_order(ev)
    return {k: np.ascontiguousarray(v[..., perm]) for k, v in ev.items()}, perm


def generate_find_event(geo, params, n_tracks, noise_per_cm2=2.0, seed_layers=3, n_threads=None):
    """A barrel finding event: hits of every layer (signal + noise) and 3-hit seeds.

    Returns a dict: hits_flat [total][9], hit_offset [L+1], hit_mc [total], n_seeds,
    seed_par [6][n], seed_err [21][n], seed_chg [n], seed_hits [seed_layers][n], seed_mc [n].
    """
    L = geo.n_barrel if geo.n_barrel else geo.n_disk     # barrel layers, or the +z disks of a disk-only geometry
    cap = n_tracks + max(toy().mkf_toymc_noise_count(ctypes.byref(geo), l, float(noise_per_cm2)) for l in range(L))
    hits = np.zeros((L, cap, 9), np.float32)
    nlh = np.zeros(L, np.int32)
    mc = np.zeros((L, cap), np.int32)
    n = int(n_tracks)
    sp = np.zeros((6, n), np.float32); se = np.zeros((21, n), np.float32); sc = np.zeros(n, np.int32)
    sh = np.zeros((seed_layers, n), np.int32); smc = np.zeros(n, np.int32)
    o = MkfToyFindOutput()
    o.n_tracks, o.layer_capacity = n, cap
    o.hits, o.n_layer_hits, o.hit_mc = (a.ctypes.data_as(_c_void_p) for a in (hits, nlh, mc))
    o.seed_par, o.seed_err, o.seed_chg, o.seed_hits, o.seed_mc = (a.ctypes.data_as(_c_void_p)
                                                                  for a in (sp, se, sc, sh, smc))
    if n_threads is None:
        n_threads = min(16, os.cpu_count() or 1)
    rc = toy().mkf_toymc_find_event(ctypes.byref(geo), ctypes.byref(params), float(noise_per_cm2), int(seed_layers),
                                    ctypes.byref(o), int(n_threads))
    if rc:
        raise RuntimeError("find-event generation failed (%d)" % rc)
    off = np.concatenate([[0], np.cumsum(nlh)]).astype(np.int32)
    flat = np.concatenate([hits[l, :nlh[l]] for l in range(L)]).astype(np.float32)
    mcf = np.concatenate([mc[l, :nlh[l]] for l in range(L)]).astype(np.int32)
    ns = o.n_seeds
    return dict(hits_flat=flat, hit_offset=off, hit_mc=mcf, n_seeds=ns, seed_par=np.ascontiguousarray(sp[:, :ns]),
                seed_err=np.ascontiguousarray(se[:, :ns]), seed_chg=sc[:ns].copy(),
                seed_hits=np.ascontiguousarray(sh[:, :ns]), seed_mc=smc[:ns].copy())


def concat_find_events(events, n_layers):
    """Several finding events as one: the layers of event e become layers e*n_layers .. of a shared event
    of hits, the seeds are concatenated, and layer_base [n] gives each seed its event's first layer.
    Hit indices stay layer-local, so the finders' outputs read as for the separate events."""
    counts, hits, mc = [], [], []
    for ev in events:
        counts.append(np.diff(ev["hit_offset"]).astype(np.int64))
        hits.append(ev["hits_flat"])
        mc.append(ev["hit_mc"])
    cnt = np.concatenate(counts)
    out = dict(hit_offset=np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32),
               hits_flat=np.ascontiguousarray(np.concatenate(hits), np.float32),
               hit_mc=np.concatenate(mc).astype(np.int32),
               n_seeds=int(sum(ev["n_seeds"] for ev in events)))
    for k in ("seed_par", "seed_err", "seed_hits"):
        out[k] = np.ascontiguousarray(np.concatenate([ev[k][..., :ev["n_seeds"]] for ev in events], axis=-1))
    for k in ("seed_chg", "seed_mc"):
        out[k] = np.concatenate([ev[k][:ev["n_seeds"]] for ev in events])
    out["layer_base"] = np.concatenate([np.full(ev["n_seeds"], e * n_layers, np.int32)
                                        for e, ev in enumerate(events)])
    out["event_seed_offset"] = np.concatenate([[0], np.cumsum([ev["n_seeds"] for ev in events])]).astype(np.int64)
    return out


def disk_layer_descs(geo, q_bin=1.5, min_dphi=0.01, max_dphi=0.05, min_dq=1.0, max_dq=2.0, half_thickness=0.01):
    """LayerInfo + selection limits for the +z disks of a disk-only geometry (endcap layers: q = r).

    Defaults are the toy cylinder geometry's endcap settings (Geoms/CylCowWLids.cc:97-123: q_bin 1.5 cm,
    selection limits 0.01/0.05 rad and 1/2 cm); the track is propagated to the disk's z, where the toy
    hits lie."""
    out = []
    for l in range(geo.n_disk):
        z = geo.disk_z[l]
        out.append(dict(is_barrel=0, rin=geo.disk_rmin[l], rout=geo.disk_rmax[l], zmin=z - half_thickness,
                        zmax=z + half_thickness, propagate_to=z, q_bin=q_bin, min_dphi=min_dphi, max_dphi=max_dphi,
                        min_dq=min_dq, max_dq=max_dq, c_dp_sf=1.1, c_dq_sf=1.1, c_c2_sf=1.1))
    return out


def barrel_layer_descs(geo, q_bin=2.0, min_dphi=0.01, max_dphi=0.05, min_dq=1.0, max_dq=2.0, half_thickness=0.01):
    """Per-layer LayerInfo + selection limits for a barrel geometry of thin layers at r.

    Defaults are the toy cylinder geometry's (Geoms/CylCowWLids.cc:64-79: q_bin 2 cm, selection
    limits 0.01/0.05 rad and 1/2 cm, dynamic-window coefficients zero), except that the track is
    propagated to the layer's centre radius, where the toy hits lie, rather than to r_in.
    """
    out = []
    for l in range(geo.n_barrel):
        r, zm = geo.barrel_r[l], geo.barrel_zmax[l]
        out.append(dict(is_barrel=1, rin=r - half_thickness, rout=r + half_thickness, zmin=-zm, zmax=zm,
                        propagate_to=r, q_bin=q_bin,
                        min_dphi=min_dphi, max_dphi=max_dphi, min_dq=min_dq, max_dq=max_dq,
                        c_dp_sf=1.1, c_dq_sf=1.1, c_c2_sf=1.1))
    return out


# ----------------------------------------------------------------------------
# Device-side harness
# ----------------------------------------------------------------------------

def layer_desc_array(descs):
    arr = (MkfLayerDesc * len(descs))()
    for i, d in enumerate(descs):
        for k, v in d.items():
            setattr(arr[i], k, v)
    return arr


class EventOfHits:
    """mkf_event_of_hits: the device LayerOfHits of every layer of one event."""

    def __init__(self, ctx, descs):
        self.ctx = ctx
        self.n_layers = len(descs)
        self._descs = layer_desc_array(descs)
        h = _c_void_p()
        _check(lib().mkf_event_of_hits_create(ctx.h, self.n_layers, self._descs, ctypes.byref(h)))
        self.h = h
        import weakref
        if not hasattr(ctx, "_events"):
            ctx._events = []
        ctx._events.append(weakref.ref(self))

    def close(self):
        if getattr(self, "h", None):
            lib().mkf_event_of_hits_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, hit_offset, hits, stream=None):
        """hit_offset: host ints [n_layers+1]; hits: device float tensor [total, 9]."""
        off = np.ascontiguousarray(hit_offset, np.int32)
        if off.shape != (self.n_layers + 1,):
            raise ValueError("hit_offset must have n_layers + 1 entries")
        self._hits = hits
        _check(lib().mkf_event_of_hits_load(self.h, off.ctypes.data_as(_c_void_p), _ptr(hits),
                                            self.ctx._stream(stream)))

    def set_hit_mask(self, mask, stream=None):
        """mkf_event_of_hits_set_hit_mask: mask, a device uint8 tensor with one byte per loaded hit in load
        order (nonzero: taken by an earlier iteration, never selected), or None to clear it."""
        if mask is not None:
            import torch
            if mask.dtype != torch.uint8 or not mask.is_contiguous():
                raise ValueError("mask must be a contiguous uint8 device tensor")
        self._mask = mask
        _check(lib().mkf_event_of_hits_set_hit_mask(self.h, _ptr(mask), self.ctx._stream(stream)))

    def set_hit_charge(self, code, stream=None):
        """mkf_event_of_hits_set_hit_charge: code, a device uint8 tensor with one byte per loaded hit in load
        order, Hit::PackedData's charge code (0xff: pixel-like, passes), or None to clear it."""
        if code is not None:
            import torch
            if code.dtype != torch.uint8 or not code.is_contiguous():
                raise ValueError("charge codes must be a contiguous uint8 device tensor")
        self._charge = code
        _check(lib().mkf_event_of_hits_set_hit_charge(self.h, _ptr(code), self.ctx._stream(stream)))

    def set_dead_regions(self, regions, stream=None):
        """mkf_event_of_hits_set_dead_regions: regions [n][5] (layer, phi1, phi2, q1, q2), or None to clear."""
        r = np.zeros((0, 5), np.float32) if regions is None else np.ascontiguousarray(regions, np.float32)
        r = r.reshape(-1, 5)
        _check(lib().mkf_event_of_hits_set_dead_regions(self.h, len(r), r.ctypes.data_as(_c_void_p) if len(r)
                                                        else None, self.ctx._stream(stream)))

    def layer_views(self, layer):
        """Copies of (ranks, bins [nq*256, 2], phis, qs) of one loaded layer, as numpy arrays."""
        import torch
        n = ctypes.c_int(); nq = ctypes.c_int()
        pr, pb, pp, pq = _c_void_p(), _c_void_p(), _c_void_p(), _c_void_p()
        _check(lib().mkf_event_of_hits_layer(self.h, int(layer), ctypes.byref(n), ctypes.byref(nq), ctypes.byref(pr),
                                             ctypes.byref(pb), ctypes.byref(pp), ctypes.byref(pq)))
        torch.cuda.synchronize()
        dev = torch.device("cuda", self.ctx.device)

        def grab(ptr, count, dtype):
            out = torch.empty(count, dtype=dtype, device=dev)
            if count:
                _check_hip(out.data_ptr(), ptr.value, count * out.element_size())
            return out.cpu().numpy()
        ranks = grab(pr, n.value, torch.int32)
        bins = grab(pb, nq.value * 256 * 2, torch.int32).reshape(-1, 2)
        phis = grab(pp, n.value, torch.float32)
        qs = grab(pq, n.value, torch.float32)
        return ranks, bins, phis, qs, nq.value

    def select_hit_indices(self, layer, par, err, xsize, xarr, wsr, stream=None):
        n = par.shape[-1]
        _check(lib().mkf_select_hit_indices(self.h, int(layer), n, _ptr(par), _ptr(err), _ptr(xsize), _ptr(xarr),
                                            _ptr(wsr), self.ctx._stream(stream)))


class FinderLanes:
    """Per-lane MkFinder state on the device for the per-layer steps (mkf_finder_lanes): the state at the
    layer (prop_par, prop_err), the updated state (par, err), charge, chi2 and the hit counters."""

    KEYS = ("prop_par", "prop_err", "par", "err", "chg", "chi2", "n_found", "n_inside_minus_one",
            "n_tail_minus_one")

    def __init__(self, prop_par, prop_err, chg, chi2=None, n_found=None, n_inside_minus_one=None,
                 n_tail_minus_one=None):
        import torch
        dev = prop_par.device
        n = prop_par.shape[-1]
        self.n = n
        self.prop_par = prop_par.contiguous()
        self.prop_err = prop_err.contiguous()
        self.par = torch.zeros((6, n), dtype=torch.float32, device=dev)
        self.err = torch.zeros((21, n), dtype=torch.float32, device=dev)
        self.chg = chg.contiguous().clone()
        zi = lambda v: v.contiguous().clone() if v is not None else torch.zeros(n, dtype=torch.int32, device=dev)
        self.chi2 = chi2.contiguous().clone() if chi2 is not None else torch.zeros(n, dtype=torch.float32, device=dev)
        self.n_found = zi(n_found)
        self.n_inside_minus_one = zi(n_inside_minus_one)
        self.n_tail_minus_one = zi(n_tail_minus_one)

    def c_struct(self):
        return MkfFinderLanes(self.n, *[_ptr(getattr(self, k)) for k in self.KEYS])

    def numpy(self):
        return {k: getattr(self, k).cpu().numpy() for k in self.KEYS}


def _check_hip(dst, src, nbytes):
    """Device-to-device copy through the HIP runtime torch already loaded."""
    global _HIPRT
    if _HIPRT is None:
        _HIPRT = ctypes.CDLL("libamdhip64.so")
        _HIPRT.hipMemcpy.argtypes = [_c_void_p, _c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIPRT.hipDeviceSynchronize.argtypes = []
    rc = _HIPRT.hipMemcpy(_c_void_p(dst), _c_void_p(src), ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    if rc != 0:
        raise MkfError("hipMemcpy failed (%d)" % rc)


_HIPRT = None


class FindBatch:
    """Device arrays of one best-hit finding batch (seeds in, found tracks out)."""

    def __init__(self, ev, n_layers, first_layer, device="cuda:0", chi2_cut=15.0, pflags=0, prop_to_hit=False,
                 intra_pflags=0, layer_base=None):
        """n_layers: layers per event.  layer_base [n]: each seed's first layer in a shared event of hits
        (several events in one; see concat_find_events), or None."""
        import torch
        n = int(ev["n_seeds"])
        dev = torch.device(device)
        self.n = n
        self.n_layers = int(n_layers)
        self.layer_base = None if layer_base is None else \
            torch.from_numpy(np.ascontiguousarray(layer_base, np.int32)).to(dev)
        self.par = torch.from_numpy(np.ascontiguousarray(ev["seed_par"][:, :n], np.float32)).to(dev)
        self.err = torch.from_numpy(np.ascontiguousarray(ev["seed_err"][:, :n], np.float32)).to(dev)
        self.chg = torch.from_numpy(np.ascontiguousarray(ev["seed_chg"][:n], np.int32)).to(dev)
        self.chi2 = torch.zeros(n, dtype=torch.float32, device=dev)
        if ev["seed_hits"].shape[0] < first_layer:
            raise ValueError("seed_hits has %d rows; first_layer %d needs one per layer before it"
                             % (ev["seed_hits"].shape[0], first_layer))
        hot = np.full((n_layers, n), -9, np.int32)
        hot[:first_layer] = ev["seed_hits"][:first_layer, :n]
        self.hot_idx = torch.from_numpy(hot).to(dev)
        self.n_found = torch.zeros(n, dtype=torch.int32, device=dev)
        self.first_layer = int(first_layer)
        self.chi2_cut = float(chi2_cut)
        self.pflags = int(pflags)
        self.prop_to_hit, self.intra_pflags = int(bool(prop_to_hit)), int(intra_pflags)
        self._seed = [t.clone() for t in (self.par, self.err, self.chg, self.chi2, self.hot_idx)]

    def reset(self):
        for dst, src in zip((self.par, self.err, self.chg, self.chi2, self.hot_idx), self._seed):
            dst.copy_(src)

    def c_struct(self):
        b = MkfFindBatch()
        b.n_tracks = self.n
        b.par, b.err, b.chg, b.chi2 = _ptr(self.par), _ptr(self.err), _ptr(self.chg), _ptr(self.chi2)
        b.hot_idx, b.n_found = _ptr(self.hot_idx), _ptr(self.n_found)
        b.first_layer, b.chi2_cut, b.pflags = self.first_layer, self.chi2_cut, self.pflags
        b.prop_to_hit, b.intra_pflags = self.prop_to_hit, self.intra_pflags
        b.n_layers, b.layer_base = self.n_layers, _ptr(self.layer_base)
        return b

    def numpy(self):
        return dict(par=self.par.cpu().numpy(), err=self.err.cpu().numpy(), chg=self.chg.cpu().numpy(),
                    chi2=self.chi2.cpu().numpy(), hot_idx=self.hot_idx.cpu().numpy(),
                    n_found=self.n_found.cpu().numpy())


class CEBatch:
    """Device arrays of one clone-engine finding batch: seeds in, up to max_cands candidates per seed out."""

    def __init__(self, ev, n_layers, first_layer, device="cuda:0", chi2_cut=30.0, pflags=0, intra_pflags=0,
                 prop_to_hit=True, max_cands=6, max_holes=4, max_consec_holes=1, layer_base=None):
        import torch
        n = int(ev["n_seeds"])
        dev = torch.device(device)
        K, L = int(max_cands), int(n_layers)
        self.n, self.K, self.L = n, K, L
        self.layer_base = None if layer_base is None else \
            torch.from_numpy(np.ascontiguousarray(layer_base, np.int32)).to(dev)
        f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).to(dev)
        self.seed_par = f(ev["seed_par"][:, :n], np.float32)
        self.seed_err = f(ev["seed_err"][:, :n], np.float32)
        self.seed_chg = f(ev["seed_chg"][:n], np.int32)
        if ev["seed_hits"].shape[0] < first_layer:
            raise ValueError("seed_hits has %d rows; first_layer %d needs one per layer before it"
                             % (ev["seed_hits"].shape[0], first_layer))
        self.seed_hits = f(ev["seed_hits"][:first_layer, :n], np.int32)
        z = lambda shape, dt: torch.zeros(shape, dtype=dt, device=dev)
        self.n_cands = z(n, torch.int32)
        self.par = z((K, 6, n), torch.float32)
        self.err = z((K, 21, n), torch.float32)
        self.chg = z((K, n), torch.int32)
        self.chi2 = z((K, n), torch.float32)
        self.score = z((K, n), torch.float32)
        self.n_found = z((K, n), torch.int32)
        self.hot_idx = z((K, L, n), torch.int32)
        self.first_layer, self.chi2_cut, self.pflags, self.intra_pflags = int(first_layer), float(chi2_cut), \
            int(pflags), int(intra_pflags)
        self.prop_to_hit, self.max_holes, self.max_consec_holes = int(bool(prop_to_hit)), int(max_holes), \
            int(max_consec_holes)

    def c_struct(self):
        b = MkfCEBatch()
        b.n_seeds = self.n
        b.seed_par, b.seed_err, b.seed_chg, b.seed_hits = (_ptr(self.seed_par), _ptr(self.seed_err),
                                                           _ptr(self.seed_chg), _ptr(self.seed_hits))
        b.first_layer, b.chi2_cut, b.pflags, b.intra_pflags = self.first_layer, self.chi2_cut, self.pflags, \
            self.intra_pflags
        b.prop_to_hit, b.max_cands, b.max_holes, b.max_consec_holes = self.prop_to_hit, self.K, self.max_holes, \
            self.max_consec_holes
        for k in ("n_cands", "par", "err", "chg", "chi2", "score", "n_found", "hot_idx"):
            setattr(b, k, _ptr(getattr(self, k)))
        b.n_layers, b.layer_base = self.L, _ptr(self.layer_base)
        return b

    def numpy(self):
        return {k: getattr(self, k).cpu().numpy()
                for k in ("n_cands", "par", "err", "chg", "chi2", "score", "n_found", "hot_idx")}


class BkFitBatch:
    """Device arrays of a backward fit of found candidates (mkf_backward_fit): the candidates' state and
    HitOnTrack lists in, the fitted state, chi2 and score out.  `cand` holds device tensors or numpy arrays par
    [6][n], err [21][n], chg [n], hot_idx [n_layers][n] and optionally score [n] -- e.g. slot 0 of a CEBatch."""

    def __init__(self, cand, device="cuda:0", group=1, pflags=0, pca_pflags=0, include_pca=True,
                 fake_material=False, last_layer=0, layer_base=None):
        import torch
        dev = torch.device(device)

        def f(a, dt):
            if isinstance(a, torch.Tensor):
                return a.to(device=dev, dtype=dt).contiguous()
            return torch.from_numpy(np.ascontiguousarray(a, {torch.float32: np.float32,
                                                            torch.int32: np.int32}[dt])).to(dev)
        self.in_par = f(cand["par"], torch.float32)
        self.in_err = f(cand["err"], torch.float32)
        self.in_chg = f(cand["chg"], torch.int32)
        self.hot_idx = f(cand["hot_idx"], torch.int32)
        self.L, self.n = int(self.hot_idx.shape[0]), int(self.hot_idx.shape[1])
        self.layer_base = None if layer_base is None else f(layer_base, torch.int32)
        # outputs start as the inputs: a candidate the fit does not write (outside its event) reads as given
        self.par = self.in_par.clone()
        self.err = self.in_err.clone()
        self.chg = self.in_chg.clone()
        self.chi2 = torch.zeros(self.n, dtype=torch.float32, device=dev)
        self.score = f(cand["score"], torch.float32).clone() if "score" in cand else \
            torch.zeros(self.n, dtype=torch.float32, device=dev)
        self.group, self.pflags, self.pca_pflags, self.last_layer = int(group), int(pflags), int(pca_pflags), \
            int(last_layer)
        self.flags = (BKFIT_PCA if include_pca else 0) | (BKFIT_FAKE_MATERIAL if fake_material else 0)

    def c_struct(self):
        b = MkfBkFitBatch()
        b.n_cands, b.n_layers, b.last_layer, b.group = self.n, self.L, self.last_layer, self.group
        b.pflags, b.pca_pflags, b.flags = self.pflags, self.pca_pflags, self.flags
        for k in ("in_par", "in_err", "in_chg", "hot_idx", "layer_base", "par", "err", "chg", "chi2", "score"):
            setattr(b, k, _ptr(getattr(self, k)))
        return b

    def numpy(self):
        return {k: getattr(self, k).cpu().numpy() for k in ("par", "err", "chg", "chi2", "score")}


class AlignmentData:
    """Device copy of a fit batch for the gradient with respect to per-layer material and alignment.

    Slot l of every track must be its hit on layer l (the barrel generator's layout).
    """

    def __init__(self, ctx, ev, device="cuda:0", pflags=PF_APPLY_MATERIAL, xi_per_x0=2.2e-3, checkpoint_every=5,
                 fp64=False, tangent=False):
        import torch
        dev = torch.device(device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.ctx = ctx
        self.n = int(ev["seed_par"].shape[1])
        self.n_layers = int(ev["hit_pos"].shape[0])
        self.seed_par, self.seed_err, self.seed_chg = t(ev["seed_par"]), t(ev["seed_err"]), t(ev["chg"])
        self.hit_pos, self.hit_err, self.n_hits = t(ev["hit_pos"]), t(ev["hit_err"]), t(ev["n_hits"])
        self.chi2 = torch.zeros(self.n, dtype=torch.float32, device=dev)
        self.grad = torch.zeros((self.n_layers, 3), dtype=torch.float32, device=dev)
        self.pflags, self.xi_per_x0, self.checkpoint_every = int(pflags), float(xi_per_x0), int(checkpoint_every)
        self.fp64 = int(bool(fp64))
        self.tangent = int(bool(tangent))   # step derivatives from forward tangents (cross-check) instead of reverse

    def run(self, theta, stream=None, noalloc=False):
        """Per-track chi2 [n] and d(sum chi2)/d theta [n_layers, 3] at theta (device float32 [n_layers, 3]).
        noalloc: through mkf_fit_adjoint_noalloc (fixed device scratch, no allocation; at most 64 MiB of it)."""
        th = theta.detach().to(self.grad.device, dtype=self.grad.dtype).contiguous()
        b = MkfAdjointBatch()
        b.n_tracks, b.n_layers = self.n, self.n_layers
        b.seed_par, b.seed_err, b.seed_chg = _ptr(self.seed_par), _ptr(self.seed_err), _ptr(self.seed_chg)
        b.hit_pos, b.hit_err, b.n_hits, b.theta = _ptr(self.hit_pos), _ptr(self.hit_err), _ptr(self.n_hits), _ptr(th)
        b.pflags, b.xi_per_x0, b.checkpoint_every = self.pflags, self.xi_per_x0, self.checkpoint_every
        b.chi2, b.grad = _ptr(self.chi2), _ptr(self.grad)
        b.fp64, b.tangent = self.fp64, self.tangent
        f = lib().mkf_fit_adjoint_noalloc if noalloc else lib().mkf_fit_adjoint
        _check(f(self.ctx.h, ctypes.byref(b), self.ctx._stream(stream)))
        return self.chi2.clone(), self.grad.clone()


class Context:
    """Owns an mkf_context (B field + material grid on one device)."""

    def __init__(self, bfield=3.8112, rl_grid=None, xi_grid=None, device=0):
        cfg = MkfConfig()
        cfg.device = int(device)
        cfg.bfield = float(bfield)
        self._rl = None if rl_grid is None else np.ascontiguousarray(rl_grid, np.float32)
        self._xi = None if xi_grid is None else np.ascontiguousarray(xi_grid, np.float32)
        cfg.rl_grid = None if self._rl is None else self._rl.ctypes.data_as(_c_void_p)
        cfg.xi_grid = None if self._xi is None else self._xi.ctypes.data_as(_c_void_p)
        h = _c_void_p()
        _check(lib().mkf_context_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.device = int(device)

    def close(self):
        """Destroys the context, after every event of hits made on it that is still open."""
        for ref in getattr(self, "_events", []):
            e = ref()
            if e is not None:
                e.close()
        if getattr(self, "h", None):
            lib().mkf_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, stream):
        if stream is None:
            import torch
            stream = torch.cuda.current_stream(self.device)
        return _c_void_p(stream.cuda_stream)

    def propagate_helix_to_r(self, in_err, in_par, in_chg, ms_rad, out_err, out_par, pflags=0, no_mat=None,
                             out_fail=None, stream=None):
        n = in_par.shape[-1]
        _check(lib().mkf_propagate_helix_to_r(self.h, n, _ptr(in_err), _ptr(in_par), _ptr(in_chg), _ptr(ms_rad),
                                              _ptr(out_err), _ptr(out_par), int(pflags), _ptr(no_mat),
                                              _ptr(out_fail), self._stream(stream)))

    def propagate_helix_to_z(self, in_err, in_par, in_chg, ms_z, out_err, out_par, pflags=0, no_mat=None,
                             stream=None):
        n = in_par.shape[-1]
        _check(lib().mkf_propagate_helix_to_z(self.h, n, _ptr(in_err), _ptr(in_par), _ptr(in_chg), _ptr(ms_z),
                                              _ptr(out_err), _ptr(out_par), int(pflags), _ptr(no_mat),
                                              self._stream(stream)))

    def kalman_operation(self, kf_op, ps_err, ps_par, ms_err, ms_par, out_err=None, out_par=None, out_chi2=None,
                         endcap=False, stream=None):
        n = ps_par.shape[-1]
        f = lib().mkf_kalman_operation_endcap if endcap else lib().mkf_kalman_operation
        _check(f(self.h, int(kf_op), n, _ptr(ps_err), _ptr(ps_par), _ptr(ms_err), _ptr(ms_par), _ptr(out_err),
                 _ptr(out_par), _ptr(out_chi2), self._stream(stream)))

    def fit_plain(self, batch, stream=None):
        _check(lib().mkf_fit_plain(self.h, ctypes.byref(batch.c_struct()), self._stream(stream)))

    def fit_batched(self, batch, stream=None):
        _check(lib().mkf_fit_batched(self.h, ctypes.byref(batch.c_struct()), self._stream(stream)))

    def find_tracks_best_hit(self, eoh, batch, stream=None):
        _check(lib().mkf_find_tracks_best_hit(self.h, eoh.h, ctypes.byref(batch.c_struct()), self._stream(stream)))

    def find_tracks_clone_engine(self, eoh, batch, stream=None):
        _check(lib().mkf_find_tracks_clone_engine(self.h, eoh.h, ctypes.byref(batch.c_struct()),
                                                  self._stream(stream)))

    def find_tracks_standard(self, eoh, batch, stream=None):
        """mkf_find_tracks_standard: standard combinatorial finding of a CEBatch (same inputs and outputs)."""
        _check(lib().mkf_find_tracks_standard(self.h, eoh.h, ctypes.byref(batch.c_struct()), self._stream(stream)))

    def backward_fit(self, eoh, batch, stream=None):
        """mkf_backward_fit of a BkFitBatch over the layers of `eoh`."""
        _check(lib().mkf_backward_fit(self.h, eoh.h, ctypes.byref(batch.c_struct()), self._stream(stream)))

    def add_best_hit(self, eoh, layer, lanes, xsize, xarr, wsr, hit_code, chi2_cut=15.0, prop_to_hit=False,
                     intra_pflags=0, stream=None):
        st = lanes.c_struct()
        _check(lib().mkf_add_best_hit(self.h, eoh.h, int(layer), ctypes.byref(st), _ptr(xsize), _ptr(xarr), _ptr(wsr),
                                      float(chi2_cut), int(bool(prop_to_hit)), int(intra_pflags), _ptr(hit_code),
                                      self._stream(stream)))

    def find_candidates_clone_engine(self, eoh, layer, lanes, xsize, xarr, wsr, n_entries, entries, chi2_cut=30.0,
                                     prop_to_hit=True, intra_pflags=0, max_holes=4, max_consec_holes=1, stream=None):
        """entries: a device byte tensor of n * 17 records of IDX_CHI2_DTYPE (48 bytes each)."""
        st = lanes.c_struct()
        _check(lib().mkf_find_candidates_clone_engine(
            self.h, eoh.h, int(layer), ctypes.byref(st), _ptr(xsize), _ptr(xarr), _ptr(wsr), float(chi2_cut),
            int(bool(prop_to_hit)), int(intra_pflags), int(max_holes), int(max_consec_holes), _ptr(n_entries),
            _ptr(entries), self._stream(stream)))

    def update_with_last_hit(self, eoh, layer, lanes, hit_idx, prop_to_hit=True, pflags=0, stream=None):
        st = lanes.c_struct()
        _check(lib().mkf_update_with_last_hit(self.h, eoh.h, int(layer), ctypes.byref(st), _ptr(hit_idx),
                                              int(bool(prop_to_hit)), int(pflags), self._stream(stream)))

    def fit_order(self, batch, sort_tile, order, stream=None):
        """Writes the batched fit's track order for `sort_tile` into `order` (int32 device tensor, n_tracks)."""
        _check(lib().mkf_fit_order(self.h, ctypes.byref(batch.c_struct()), int(sort_tile), _ptr(order),
                                   self._stream(stream)))

    def set_batched_shape(self, block, blocks_per_cu, variant, sort_tile=0):
        s = MkfLaunchShape(int(block), int(blocks_per_cu), int(variant), int(sort_tile))
        _check(lib().mkf_set_batched_shape(self.h, ctypes.byref(s)))

    def get_batched_shape(self):
        """(block, blocks_per_cu, variant, sort_tile)"""
        s = MkfLaunchShape()
        _check(lib().mkf_get_batched_shape(self.h, ctypes.byref(s)))
        return (s.block, s.blocks_per_cu, s.variant, s.sort_tile)

    def set_find_shape(self, standard, slots=0, waves_per_simd=2):
        """Launch shape of the clone-engine (standard=False) or standard finder (mkf_set_find_shape)."""
        _check(lib().mkf_set_find_shape(self.h, int(bool(standard)), ctypes.byref(MkfFindShape(slots, waves_per_simd))))

    def get_find_shape(self, standard):
        s = MkfFindShape()
        _check(lib().mkf_get_find_shape(self.h, int(bool(standard)), ctypes.byref(s)))
        return (s.slots, s.waves_per_simd)

    def autotune_find(self, eoh, batch, standard=False, reps=2, stream=None):
        """mkf_autotune_find: times every finder shape on `batch` (a CEBatch), keeps the fastest; returns
        ((slots, waves_per_simd), ms per launch).  The batch holds a finding result afterwards."""
        s = MkfFindShape()
        ms = ctypes.c_float()
        _check(lib().mkf_autotune_find(self.h, eoh.h, ctypes.byref(batch.c_struct()), int(bool(standard)), int(reps),
                                       ctypes.byref(s), ctypes.byref(ms), self._stream(stream)))
        return (s.slots, s.waves_per_simd), ms.value

    def autotune_batched(self, batch, reps=3):
        import torch
        scratch = torch.empty(29 * batch.n_tracks, dtype=torch.float32, device=batch.par.device)
        s = MkfLaunchShape()
        ms = ctypes.c_float()
        torch.cuda.synchronize(self.device)
        _check(lib().mkf_autotune_batched(self.h, ctypes.byref(batch.c_struct()), _ptr(scratch), int(reps),
                                          ctypes.byref(s), ctypes.byref(ms)))
        return (s.block, s.blocks_per_cu, s.variant, s.sort_tile), ms.value


class FitBatch:
    """Device-resident batch (torch tensors) in the layout of mkf_fit_batch."""

    def __init__(self, par, err, chg, chi2, hit_pos, hit_err, n_hits, hit_flags=None, pflags=0,
                 fit_flags=FIT_CHI2, seed=None, pca_pflags=0):
        self.pca_pflags = int(pca_pflags)
        self.par, self.err, self.chg, self.chi2 = par, err, chg, chi2
        self.seed = seed  # optional (seed_par, seed_err, seed_chg) for an out-of-place fit
        self.hit_pos, self.hit_err, self.n_hits, self.hit_flags = hit_pos, hit_err, n_hits, hit_flags
        self.pflags, self.fit_flags = int(pflags), int(fit_flags)
        self.n_tracks = int(par.shape[-1])
        self.n_slots = int(hit_pos.shape[0])

    @classmethod
    def from_numpy(cls, ev, device, pflags=0, fit_flags=FIT_CHI2, with_flags=True, out_of_place=False, pca_pflags=0):
        import torch
        d = torch.device(device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(d)
        n = ev["seed_par"].shape[1]
        sp, se, sc = t(ev["seed_par"]), t(ev["seed_err"]), t(ev["chg"])
        if out_of_place:
            par, err, chg = torch.zeros_like(sp), torch.zeros_like(se), torch.zeros_like(sc)
            seed = (sp, se, sc)
        else:
            par, err, chg, seed = sp, se, sc, None
        return cls(par, err, chg, torch.zeros(n, dtype=torch.float32, device=d),
                   t(ev["hit_pos"]), t(ev["hit_err"]), t(ev["n_hits"]),
                   t(ev["hit_flags"]) if with_flags else None, pflags=pflags, fit_flags=fit_flags, seed=seed,
                   pca_pflags=pca_pflags)

    def clone_state(self):
        return [x.clone() for x in (self.par, self.err, self.chg, self.chi2)]

    def restore_state(self, st):
        for dst, src in zip((self.par, self.err, self.chg, self.chi2), st):
            dst.copy_(src)

    def c_struct(self):
        b = MkfFitBatch()
        b.n_tracks, b.n_slots = self.n_tracks, self.n_slots
        b.par, b.err, b.chg, b.chi2 = _ptr(self.par), _ptr(self.err), _ptr(self.chg), _ptr(self.chi2)
        b.hit_pos, b.hit_err, b.n_hits = _ptr(self.hit_pos), _ptr(self.hit_err), _ptr(self.n_hits)
        b.hit_flags = _ptr(self.hit_flags)
        b.pflags, b.fit_flags, b.pca_pflags = self.pflags, self.fit_flags, self.pca_pflags
        if self.seed is not None:
            b.seed_par, b.seed_err, b.seed_chg = (_ptr(x) for x in self.seed)
        return b
