"""CPU oracle for mkFit's Matriplex Kalman hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, as the checker and as the timed CPU baseline.  The product
package (mkfit_amd) never imports or links it.

The C++ restatement lives in oracle_core.h / oracle.cc (reference citations
there); this module is a ctypes wrapper over liborc_avx512.so / liborc_avx2.so.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

FIT_CHI2 = 1
FIT_PROP_TO_PCA = 8
FIT_CHARGE_FLIP = 2


class OrcCtx(ctypes.Structure):
    _fields_ = [("bfield", ctypes.c_float), ("rl_grid", ctypes.c_void_p), ("xi_grid", ctypes.c_void_p)]


class _Batch32(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int), ("n_slots", ctypes.c_int),
                ("par", ctypes.c_void_p), ("err", ctypes.c_void_p), ("chg", ctypes.c_void_p),
                ("chi2", ctypes.c_void_p), ("hit_pos", ctypes.c_void_p), ("hit_err", ctypes.c_void_p),
                ("n_hits", ctypes.c_void_p), ("hit_flags", ctypes.c_void_p),
                ("pflags", ctypes.c_int), ("fit_flags", ctypes.c_int), ("pca_pflags", ctypes.c_int)]


_Batch64 = _Batch32  # same field layout (pointers differ only in pointee type)


def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            return "avx512f" in f.read()
    except OSError:
        return False


_OVERRIDE = []


def default_lib_path(contract=True):
    """The ISA-matched build; contract=False selects the build without FMA contraction (-ffp-contract=off)."""
    if _OVERRIDE and contract:
        return _OVERRIDE[-1]
    name = "liborc_avx512" if _cpu_has_avx512() else "liborc_avx2"
    return os.path.join(_HERE, name + ("" if contract else "_nofma") + ".so")


class tie_shift:
    """Context manager: every oracle call inside moves the helix step's 1e-4 cm threshold by `shift` float
    spacings of r (the result with the threshold ties taken the other way; checker only)."""

    def __init__(self, shift):
        self.shift = float(shift)

    def __enter__(self):
        lib().orc_set_tie_shift(ctypes.c_float(self.shift))
        return self

    def __exit__(self, *exc):
        lib().orc_set_tie_shift(ctypes.c_float(0.0))


class use_lib:
    """Context manager: every oracle call inside uses the library at `path` (the mutation tests' builds)."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        _OVERRIDE.append(self.path)
        return self

    def __exit__(self, *exc):
        _OVERRIDE.pop()


def lib(path=None):
    """The oracle shared library (the ISA-matched build unless `path` is given)."""
    p = path or default_lib_path()
    if p not in _LIBS:
        if not os.path.exists(p):
            raise RuntimeError("%s is not built (run __graft_entry__.build())" % p)
        L = ctypes.CDLL(p)
        L.orc_simd_width.restype = ctypes.c_int
        _LIBS[p] = L
    return _LIBS[p]


def simd_width():
    return lib().orc_simd_width()


def isa_name():
    return "avx512" if simd_width() == 16 else "avx2"


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """Holds the B field and material grid (Config:: globals of the reference)."""

    def __init__(self, bfield=3.8112, rl_grid=None, xi_grid=None, lib_path=None, contract=True):
        self.L = lib(lib_path or default_lib_path(contract))
        self._rl = None if rl_grid is None else np.ascontiguousarray(rl_grid, np.float32)
        self._xi = None if xi_grid is None else np.ascontiguousarray(xi_grid, np.float32)
        self.ctx = OrcCtx(float(bfield), _p(self._rl), _p(self._xi))

    def _dt(self, dtype):
        return ("f64", np.float64) if np.dtype(dtype) == np.float64 else ("f32", np.float32)

    def propagate_to_r(self, in_err, in_par, in_chg, ms_rad, pflags=0, no_mat=None, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = in_par.shape[-1]
        ie = np.ascontiguousarray(in_err, T); ip = np.ascontiguousarray(in_par, T)
        ch = np.ascontiguousarray(in_chg, np.int32); r = np.ascontiguousarray(ms_rad, T)
        oe = np.zeros((21, n), T); op = np.zeros((6, n), T); fl = np.zeros(n, np.int32)
        nm = None if no_mat is None else np.ascontiguousarray(no_mat, np.int32)
        getattr(self.L, "orc_propagate_to_r_" + sfx)(ctypes.byref(self.ctx), n, _p(ie), _p(ip), _p(ch), _p(r),
                                                    _p(oe), _p(op), int(pflags), _p(nm), _p(fl))
        return oe, op, fl

    def propagate_to_z(self, in_err, in_par, in_chg, ms_z, pflags=0, no_mat=None, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = in_par.shape[-1]
        ie = np.ascontiguousarray(in_err, T); ip = np.ascontiguousarray(in_par, T)
        ch = np.ascontiguousarray(in_chg, np.int32); z = np.ascontiguousarray(ms_z, T)
        oe = np.zeros((21, n), T); op = np.zeros((6, n), T)
        nm = None if no_mat is None else np.ascontiguousarray(no_mat, np.int32)
        getattr(self.L, "orc_propagate_to_z_" + sfx)(ctypes.byref(self.ctx), n, _p(ie), _p(ip), _p(ch), _p(z),
                                                    _p(oe), _p(op), int(pflags), _p(nm))
        return oe, op

    def helix_at_r(self, in_par, in_chg, ms_rad, pflags=0, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = in_par.shape[-1]
        ip = np.ascontiguousarray(in_par, T); ch = np.ascontiguousarray(in_chg, np.int32)
        r = np.ascontiguousarray(ms_rad, T)
        op = np.zeros((6, n), T); J = np.zeros((36, n), T); fl = np.zeros(n, np.int32)
        getattr(self.L, "orc_helix_at_r_" + sfx)(ctypes.byref(self.ctx), n, _p(ip), _p(ch), _p(r), _p(op), _p(J),
                                                _p(fl), int(pflags))
        return op, J.reshape(6, 6, n), fl

    def helix_at_z(self, in_par, in_chg, ms_z, pflags=0, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = in_par.shape[-1]
        ip = np.ascontiguousarray(in_par, T); ch = np.ascontiguousarray(in_chg, np.int32)
        z = np.ascontiguousarray(ms_z, T)
        op = np.zeros((6, n), T); J = np.zeros((36, n), T)
        getattr(self.L, "orc_helix_at_z_" + sfx)(ctypes.byref(self.ctx), n, _p(ip), _p(ch), _p(z), _p(op), _p(J),
                                                int(pflags))
        return op, J.reshape(6, 6, n)

    def kalman_operation(self, op, ps_err, ps_par, ms_err, ms_par, endcap=False, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = ps_par.shape[-1]
        pe = np.ascontiguousarray(ps_err, T); pp = np.ascontiguousarray(ps_par, T)
        me = np.ascontiguousarray(ms_err, T); mp = np.ascontiguousarray(ms_par, T)
        oe = np.zeros((21, n), T); opar = np.zeros((6, n), T); c2 = np.zeros(n, T)
        getattr(self.L, "orc_kalman_op_" + sfx)(int(op), int(bool(endcap)), n, _p(pe), _p(pp), _p(me), _p(mp),
                                               _p(oe), _p(opar), _p(c2))
        return oe, opar, c2

    def fit(self, ev, pflags=0, fit_flags=FIT_CHI2, dtype=np.float32, n_threads=1, with_flags=True, pca_pflags=0,
            ties=False, tie_shift=0.0):
        """Whole fit of a toy-MC event dict; returns (par, err, chg, chi2), and with ties=True also [3][n] margins
        of each track's steps to their thresholds: the helix step's |r - r0| test and its |oodotp - 5| failure
        test, and material bin edges (the first and third in float spacings).  tie_shift moves the helix step's
        1e-4 cm threshold by that many float spacings of r (FP32 only)."""
        sfx, T = self._dt(dtype)
        n = ev["seed_par"].shape[1]
        S = ev["hit_pos"].shape[0]
        par = np.array(ev["seed_par"], T, order="C"); err = np.array(ev["seed_err"], T, order="C")
        chg = np.array(ev["chg"], np.int32, order="C"); chi2 = np.zeros(n, T)
        hp = np.ascontiguousarray(ev["hit_pos"], np.float32); he = np.ascontiguousarray(ev["hit_err"], np.float32)
        nh = np.ascontiguousarray(ev["n_hits"], np.int32)
        hf = np.ascontiguousarray(ev["hit_flags"], np.uint8) if with_flags else None
        b = _Batch32(n, S, _p(par), _p(err), _p(chg), _p(chi2), _p(hp), _p(he), _p(nh), _p(hf), int(pflags),
                     int(fit_flags), int(pca_pflags))
        if ties:
            if sfx != "f32":
                raise ValueError("threshold margins are recorded by the FP32 fit")
            tie = np.zeros((3, n), np.float32)
            rc = self.L.orc_fit_ties_f32(ctypes.byref(self.ctx), ctypes.byref(b), int(n_threads), _p(tie))
        elif tie_shift:
            if sfx != "f32":
                raise ValueError("the threshold shift applies to the FP32 fit")
            rc = self.L.orc_fit_tie_shift_f32(ctypes.byref(self.ctx), ctypes.byref(b), int(n_threads),
                                              ctypes.c_float(tie_shift))
        else:
            rc = getattr(self.L, "orc_fit_" + sfx)(ctypes.byref(self.ctx), ctypes.byref(b), int(n_threads))
        if rc != 0:
            raise RuntimeError("oracle fit failed (%d)" % rc)
        if ties:
            return par, err, chg, chi2, tie
        return par, err, chg, chi2


# ---------------------------------------------------------------------------
# Best-hit finding (oracle_find.cc)
# ---------------------------------------------------------------------------

class LayerDesc(ctypes.Structure):
    _fields_ = [("is_barrel", ctypes.c_int), ("rin", ctypes.c_float), ("rout", ctypes.c_float),
                ("zmin", ctypes.c_float), ("zmax", ctypes.c_float), ("propagate_to", ctypes.c_float),
                ("q_bin", ctypes.c_float), ("min_dphi", ctypes.c_float), ("max_dphi", ctypes.c_float),
                ("min_dq", ctypes.c_float), ("max_dq", ctypes.c_float)] + [
        (nm, ctypes.c_float) for nm in ("c_dp_sf", "c_dp_0", "c_dp_1", "c_dp_2", "c_dq_sf", "c_dq_0", "c_dq_1",
                                        "c_dq_2", "c_c2_sf", "c_c2_0", "c_c2_1", "c_c2_2")]


class _FindEvent(ctypes.Structure):
    _fields_ = [("n_layers", ctypes.c_int), ("layers", ctypes.c_void_p), ("hit_offset", ctypes.c_void_p),
                ("hits", ctypes.c_void_p), ("first_layer", ctypes.c_int), ("chi2_cut", ctypes.c_float),
                ("pflags", ctypes.c_int), ("prop_to_hit", ctypes.c_int), ("intra_pflags", ctypes.c_int),
                ("hit_mask", ctypes.c_void_p), ("n_dead", ctypes.c_int), ("dead_regions", ctypes.c_void_p),
                ("hit_charge", ctypes.c_void_p)]


def _dead_regions(ev):
    """The event's dead-module regions [n][5] (layer, phi1, phi2, q1, q2) as float32, or None."""
    d = ev.get("dead_regions")
    return None if d is None or len(d) == 0 else np.ascontiguousarray(d, np.float32).reshape(-1, 5)


def _hit_mask(ev):
    """The event's iteration hit mask as a uint8 array (kept alive by the caller), or None."""
    m = ev.get("hit_mask")
    return None if m is None else np.ascontiguousarray(m, np.uint8)


def _hit_charge(ev):
    """The event's per-hit strip charge codes (Hit::PackedData's 8-bit charge_pcm, Hit.h:251-274) as uint8,
    or None: then the combinatorial finders' charge cut passes every hit."""
    c = ev.get("hit_charge")
    return None if c is None else np.ascontiguousarray(c, np.uint8)


def charge_code(pcm):
    """Hit::PackedData::set_charge_pcm (Hit.h:264-268): charge per cm -> 8-bit code."""
    pcm = np.asarray(pcm, np.int64)
    return np.where(pcm < 1620, 0, np.minimum(0xff, ((pcm - 1620) >> 3) + 1)).astype(np.uint8)


def charge_pcm(code):
    """Hit::PackedData::get_charge_pcm (Hit.h:270-274): 8-bit code -> charge per cm."""
    code = np.asarray(code, np.int64)
    return np.where(code == 0, 0, ((code - 1) << 3) + 1620).astype(np.uint32)


class _FindTracks(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int), ("par", ctypes.c_void_p), ("err", ctypes.c_void_p),
                ("chg", ctypes.c_void_p), ("chi2", ctypes.c_void_p), ("hot_idx", ctypes.c_void_p),
                ("n_found", ctypes.c_void_p), ("margins", ctypes.c_void_p), ("work", ctypes.c_void_p)]


def layer_array(descs):
    arr = (LayerDesc * len(descs))()
    for i, d in enumerate(descs):
        for k, v in d.items():
            setattr(arr[i], k, v)
    return arr


def loh_build(desc, hits):
    """LayerOfHits of one layer: (ranks, bins [nq*256, 2], phis, qs, nq)."""
    L = lib()
    d = layer_array([desc])
    nq = ctypes.c_int()
    L.orc_loh_geometry(ctypes.byref(d[0]), ctypes.byref(nq), None, None, None)
    h = np.ascontiguousarray(hits, np.float32)
    n = h.shape[0]
    ranks = np.zeros(n, np.int32); bins = np.zeros((nq.value * 256, 2), np.int32)
    phis = np.zeros(n, np.float32); qs = np.zeros(n, np.float32)
    L.orc_loh_build(ctypes.byref(d[0]), n, _p(h), _p(ranks), _p(bins), _p(phis), _p(qs))
    return ranks, bins, phis, qs, nq.value


def select_hit_indices(desc, hits, par, err, mask=None):
    """SelectHitIndices on one layer; mask: optional per-hit flags of hits taken by an earlier iteration."""
    L = lib()
    d = layer_array([desc])
    h = np.ascontiguousarray(hits, np.float32)
    p = np.ascontiguousarray(par, np.float32); e = np.ascontiguousarray(err, np.float32)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    n = p.shape[1]
    xs = np.zeros(n, np.int32); xa = np.zeros((16, n), np.int32); w = np.zeros(n, np.int32)
    L.orc_select_hit_indices_masked(ctypes.byref(d[0]), h.shape[0], _p(h), _p(m), n, _p(p), _p(e), _p(xs), _p(xa),
                                    _p(w))
    return xs, xa, w


def find_best_hit(ev, descs, first_layer, chi2_cut=15.0, pflags=0, bfield=3.8, n_threads=1, prop_to_hit=False,
                  intra_pflags=0, margins=False, work=False):
    """Best-hit finding over a toy event dict (see mkfit_amd.generate_find_event).  margins=True adds the
    checker's decision margins [3][n_layers][n] (window edge cm, bin-edge fraction, relative chi2 gap);
    work=True the work units [3][n] (propagations to a layer, window hits evaluated, updates)."""
    L = lib()
    layers = layer_array(descs)
    n = ev["n_seeds"]
    par = np.array(ev["seed_par"][:, :n], np.float32, order="C")
    err = np.array(ev["seed_err"][:, :n], np.float32, order="C")
    chg = np.array(ev["seed_chg"][:n], np.int32, order="C"); chi2 = np.zeros(n, np.float32)
    nl = len(descs)
    if ev["seed_hits"].shape[0] < first_layer:
        raise ValueError("seed_hits has %d rows; first_layer %d needs one per layer before it"
                         % (ev["seed_hits"].shape[0], first_layer))
    hot = np.full((nl, n), -9, np.int32)
    hot[:first_layer] = ev["seed_hits"][:first_layer, :n]
    nfound = np.zeros(n, np.int32)
    off = np.ascontiguousarray(ev["hit_offset"], np.int32)
    hits = np.ascontiguousarray(ev["hits_flat"], np.float32)
    hm = _hit_mask(ev)
    fe = _FindEvent(nl, ctypes.cast(layers, ctypes.c_void_p), _p(off), _p(hits), first_layer, chi2_cut, pflags,
                    int(bool(prop_to_hit)), int(intra_pflags), _p(hm))
    mg = np.full((3, nl, n), 1e30, np.float32) if margins else None
    wk = np.zeros((3, n), np.int32) if work else None
    ft = _FindTracks(n, _p(par), _p(err), _p(chg), _p(chi2), _p(hot), _p(nfound), _p(mg), _p(wk))
    ctx = OrcCtx(float(bfield), None, None)
    rc = L.orc_find_best_hit(ctypes.byref(ctx), ctypes.byref(fe), ctypes.byref(ft), int(n_threads))
    if rc:
        raise RuntimeError("orc_find_best_hit failed (%d)" % rc)
    out = dict(par=par, err=err, chg=chg, chi2=chi2, hot_idx=hot, n_found=nfound)
    if margins:
        out["margins"] = mg
    if work:
        out["work"] = wk
    return out


class _CEParams(ctypes.Structure):
    _fields_ = [("max_cands", ctypes.c_int), ("max_holes", ctypes.c_int), ("max_consec_holes", ctypes.c_int),
                ("prop_to_hit", ctypes.c_int), ("intra_pflags", ctypes.c_int)]


class _CETracks(ctypes.Structure):
    _fields_ = [("n_seeds", ctypes.c_int)] + [(k, ctypes.c_void_p) for k in (
        "seed_par", "seed_err", "seed_chg", "seed_hits", "n_cands", "par", "err", "chg", "chi2", "score", "n_found",
        "hot_idx", "margins", "merge_gap", "work")]


def find_clone_engine(ev, descs, first_layer, chi2_cut=30.0, pflags=0, bfield=3.8, max_cands=6, max_holes=4,
                      max_consec_holes=1, prop_to_hit=True, intra_pflags=0, n_threads=1, margins=False,
                      standard=False, work=False):
    """Clone-engine finding over a toy event dict; per seed up to max_cands candidates, best first.  margins=True
    adds the checker's decision margins [4][n_layers][n] (window edge cm, bin-edge fraction, chi2 vs cut,
    score gap) and merge_gap [n].  standard=True runs the standard combinatorial finding
    (MkBuilder::FindTracksStandard) with the same inputs and outputs.  work=True adds the work units [3][n]
    (candidates propagated to a layer, window hits evaluated, candidates updated with their hit)."""
    L = lib()
    layers = layer_array(descs)
    n = ev["n_seeds"]
    K = int(max_cands)
    nl = len(descs)
    sp = np.ascontiguousarray(ev["seed_par"][:, :n], np.float32)
    se = np.ascontiguousarray(ev["seed_err"][:, :n], np.float32)
    sc = np.ascontiguousarray(ev["seed_chg"][:n], np.int32)
    if ev["seed_hits"].shape[0] < first_layer:
        raise ValueError("seed_hits has %d rows; first_layer %d needs one per layer before it"
                         % (ev["seed_hits"].shape[0], first_layer))
    sh = np.ascontiguousarray(ev["seed_hits"][:first_layer, :n], np.int32)
    out = dict(n_cands=np.zeros(n, np.int32), par=np.zeros((K, 6, n), np.float32),
               err=np.zeros((K, 21, n), np.float32), chg=np.zeros((K, n), np.int32), chi2=np.zeros((K, n), np.float32),
               score=np.zeros((K, n), np.float32), n_found=np.zeros((K, n), np.int32),
               hot_idx=np.zeros((K, nl, n), np.int32))
    off = np.ascontiguousarray(ev["hit_offset"], np.int32)
    hits = np.ascontiguousarray(ev["hits_flat"], np.float32)
    hm, dr, hq = _hit_mask(ev), _dead_regions(ev), _hit_charge(ev)
    fe = _FindEvent(nl, ctypes.cast(layers, ctypes.c_void_p), _p(off), _p(hits), first_layer, chi2_cut, pflags, 0, 0,
                    _p(hm), 0 if dr is None else len(dr), _p(dr), _p(hq))
    pr = _CEParams(K, int(max_holes), int(max_consec_holes), int(bool(prop_to_hit)), int(intra_pflags))
    if margins:
        out["margins"] = np.full((4, nl, n), 1e30, np.float32)
        out["merge_gap"] = np.full(n, 1e30, np.float32)
    if work:
        out["work"] = np.zeros((3, n), np.int32)
    tr = _CETracks(n, _p(sp), _p(se), _p(sc), _p(sh), *[_p(out.get(k)) for k in (
        "n_cands", "par", "err", "chg", "chi2", "score", "n_found", "hot_idx", "margins", "merge_gap", "work")])
    ctx = OrcCtx(float(bfield), None, None)
    f = L.orc_find_standard if standard else L.orc_find_clone_engine
    rc = f(ctypes.byref(ctx), ctypes.byref(fe), ctypes.byref(pr), ctypes.byref(tr), int(n_threads))
    if rc:
        raise RuntimeError("oracle %s finding failed (%d)" % ("standard" if standard else "clone-engine", rc))
    return out


def find_standard(ev, descs, first_layer, **kw):
    """Standard combinatorial finding (MkBuilder::FindTracksStandard, MkFinder::FindCandidates): arguments and
    outputs as find_clone_engine."""
    return find_clone_engine(ev, descs, first_layer, standard=True, **kw)


class _FinderLanes(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int)] + [(k, ctypes.c_void_p) for k in (
        "prop_par", "prop_err", "par", "err", "chg", "chi2", "n_found", "n_inside_minus_one", "n_tail_minus_one")]


IDX_CHI2_DTYPE = np.dtype([("trk_idx", "<i4"), ("hit_idx", "<i4"), ("module", "<u4"), ("nhits", "<i4"),
                           ("ntailholes", "<i4"), ("noverlaps", "<i4"), ("nholes", "<i4"), ("pt", "<f4"),
                           ("chi2", "<f4"), ("chi2_hit", "<f4"), ("score", "<f4"), ("xhit_pos", "<i4")])

LANE_KEYS = ("chg", "chi2", "n_found", "n_inside_minus_one", "n_tail_minus_one")


def _lanes(lanes):
    """Copies of a finder-lane dict (prop_par, prop_err, chg, chi2 and the hit counters) plus output arrays."""
    n = lanes["prop_par"].shape[1]
    out = dict(prop_par=np.ascontiguousarray(lanes["prop_par"], np.float32),
               prop_err=np.ascontiguousarray(lanes["prop_err"], np.float32),
               par=np.zeros((6, n), np.float32), err=np.zeros((21, n), np.float32))
    for k in LANE_KEYS:
        dt = np.float32 if k == "chi2" else np.int32
        out[k] = np.array(lanes.get(k, np.zeros(n, dt)), dt, order="C")
    st = _FinderLanes(n, *[_p(out[k]) for k in ("prop_par", "prop_err", "par", "err") + LANE_KEYS])
    return out, st


def _layer_event(ev, descs, chi2_cut, pflags, prop_to_hit, intra_pflags):
    layers = layer_array(descs)
    off = np.ascontiguousarray(ev["hit_offset"], np.int32)
    hits = np.ascontiguousarray(ev["hits_flat"], np.float32)
    hq = _hit_charge(ev)
    fe = _FindEvent(len(descs), ctypes.cast(layers, ctypes.c_void_p), _p(off), _p(hits), 0, chi2_cut, pflags,
                    int(bool(prop_to_hit)), int(intra_pflags), None, 0, None, _p(hq))
    return fe, (layers, off, hits, hq)


def add_best_hit(ev, descs, layer, lanes, xsize, xarr, wsr, chi2_cut=15.0, prop_to_hit=False, intra_pflags=0,
                 bfield=3.8):
    """MkFinder::AddBestHit on one layer; returns the lane dict after the step plus hit_code."""
    out, st = _lanes(lanes)
    fe, keep = _layer_event(ev, descs, chi2_cut, 0, prop_to_hit, intra_pflags)
    xs = np.ascontiguousarray(xsize, np.int32); xa = np.ascontiguousarray(xarr, np.int32)
    w = np.ascontiguousarray(wsr, np.int32)
    code = np.zeros(st.n, np.int32)
    ctx = OrcCtx(float(bfield), None, None)
    rc = lib().orc_add_best_hit(ctypes.byref(ctx), ctypes.byref(fe), int(layer), ctypes.byref(st), _p(xs), _p(xa),
                                _p(w), _p(code))
    if rc:
        raise RuntimeError("orc_add_best_hit failed (%d)" % rc)
    out["hit_code"] = code
    return out


def find_candidates_ce(ev, descs, layer, lanes, xsize, xarr, wsr, chi2_cut=30.0, max_holes=4, max_consec_holes=1,
                       prop_to_hit=True, intra_pflags=0, bfield=3.8):
    """MkFinder::FindCandidatesCloneEngine on one layer: (n_entries [n], entries [n][17] of IDX_CHI2_DTYPE)."""
    keep_lanes, st = _lanes(lanes)
    fe, keep = _layer_event(ev, descs, chi2_cut, 0, prop_to_hit, intra_pflags)
    pr = _CEParams(1, int(max_holes), int(max_consec_holes), int(bool(prop_to_hit)), int(intra_pflags))
    xs = np.ascontiguousarray(xsize, np.int32); xa = np.ascontiguousarray(xarr, np.int32)
    w = np.ascontiguousarray(wsr, np.int32)
    ne = np.zeros(st.n, np.int32)
    ent = np.zeros((st.n, 17), IDX_CHI2_DTYPE)
    ctx = OrcCtx(float(bfield), None, None)
    rc = lib().orc_find_candidates_ce(ctypes.byref(ctx), ctypes.byref(fe), ctypes.byref(pr), int(layer),
                                      ctypes.byref(st), _p(xs), _p(xa), _p(w), _p(ne), _p(ent))
    if rc:
        raise RuntimeError("orc_find_candidates_ce failed (%d)" % rc)
    return ne, ent


def pass_strip_charge(barrel, code, par, herr):
    """passStripChargePCMfromTrack (MkFinder.cc:867-896) for one state par[6] and hit covariance herr[6]
    given the hit's 8-bit charge code: (passes, corrected charge per cm)."""
    L = lib()
    p = np.ascontiguousarray(par, np.float32)
    e = np.ascontiguousarray(herr, np.float32)
    q = ctypes.c_float()
    ok = L.orc_pass_strip_charge(int(bool(barrel)), int(code), _p(p), _p(e), ctypes.byref(q))
    return bool(ok), q.value


def update_with_last_hit(ev, descs, layer, lanes, hit_idx, pflags=0, prop_to_hit=True, bfield=3.8):
    """MkFinder::UpdateWithLastHit on one layer: (par, err, chg)."""
    out, st = _lanes(lanes)
    fe, keep = _layer_event(ev, descs, 30.0, pflags, prop_to_hit, 0)
    pr = _CEParams(1, 4, 1, int(bool(prop_to_hit)), 0)
    hi = np.ascontiguousarray(hit_idx, np.int32)
    ctx = OrcCtx(float(bfield), None, None)
    rc = lib().orc_update_with_last_hit(ctypes.byref(ctx), ctypes.byref(fe), ctypes.byref(pr), int(layer),
                                        ctypes.byref(st), _p(hi))
    if rc:
        raise RuntimeError("orc_update_with_last_hit failed (%d)" % rc)
    return out["par"], out["err"], out["chg"]


class _AlignBatch(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int), ("n_layers", ctypes.c_int)] + [
        (k, ctypes.c_void_p) for k in ("seed_par", "seed_err", "seed_chg", "hit_pos", "hit_err", "n_hits")] + [
        ("pflags", ctypes.c_int), ("xi_per_x0", ctypes.c_float)]


def align_chi2(ev, theta, pflags=2, xi_per_x0=2.2e-3, dtype=np.float64, bfield=3.8, n_threads=1):
    """Per-track sum of update chi2 of the fit with per-layer theta[l] = (x/X0, d_rphi, d_z)."""
    L = lib()
    th = np.ascontiguousarray(theta, np.float64 if np.dtype(dtype) == np.float64 else np.float32)
    nl = th.shape[0]
    n = ev["seed_par"].shape[1]
    arrs = [np.ascontiguousarray(ev[k]) for k in ("seed_par", "seed_err", "chg", "hit_pos", "hit_err", "n_hits")]
    b = _AlignBatch(n, nl, *[_p(a) for a in arrs], int(pflags), float(xi_per_x0))
    out = np.zeros(n, np.float64)
    f = L.orc_align_chi2_f64 if th.dtype == np.float64 else L.orc_align_chi2_f32
    rc = f(ctypes.byref(OrcCtx(float(bfield), None, None)), ctypes.byref(b), _p(th), _p(out), int(n_threads))
    if rc:
        raise RuntimeError("orc_align_chi2 failed (%d)" % rc)
    return out


# ---------------------------------------------------------------------------
# Backward fit of found candidates (oracle_bkfit.cc)
# ---------------------------------------------------------------------------

class _BkParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int) for k in ("group", "pflags", "pca_pflags", "include_pca", "fake_material",
                                            "last_layer")]


class _BkTracks(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("n_layers", ctypes.c_int)] + [(k, ctypes.c_void_p) for k in (
        "hot_idx", "layer_base", "par", "err", "chg", "chi2", "score")]


def backward_fit(ev, descs, cand, group=1, pflags=0, pca_pflags=0, include_pca=True, fake_material=False,
                 last_layer=0, bfield=3.8, rl_grid=None, xi_grid=None, layer_base=None, dtype=np.float32,
                 n_threads=1, ties=False, tie_shift=0.0):
    """MkBuilder::BackwardFit over one candidate per seed: cand holds par [6][n], err [21][n], chg [n], hot_idx
    [n_layers][n] and optionally score [n] (a clone-engine result's slot 0, or a best-hit result).  Returns
    dict(par, err, chg, chi2, score) of the fitted candidates, plus tie [3][n] with ties=True (FP32 only)."""
    T = np.float64 if np.dtype(dtype) == np.float64 else np.float32
    hot = np.ascontiguousarray(cand["hot_idx"], np.int32)
    nl, n = hot.shape
    out = dict(par=np.array(cand["par"], T, order="C"), err=np.array(cand["err"], T, order="C"),
               chg=np.array(cand["chg"], np.int32, order="C"), chi2=np.zeros(n, T),
               score=np.array(cand["score"], T, order="C") if "score" in cand else np.zeros(n, T))
    lb = None if layer_base is None else np.ascontiguousarray(layer_base, np.int32)
    layers = layer_array(descs)
    off = np.ascontiguousarray(ev["hit_offset"], np.int32)
    hits = np.ascontiguousarray(ev["hits_flat"], np.float32)
    fe = _FindEvent(len(descs), ctypes.cast(layers, ctypes.c_void_p), _p(off), _p(hits), 0, 0.0, 0, 0, 0)
    pr = _BkParams(int(group), int(pflags), int(pca_pflags), int(bool(include_pca)), int(bool(fake_material)),
                   int(last_layer))
    tr = _BkTracks(n, nl, _p(hot), _p(lb), *[_p(out[k]) for k in ("par", "err", "chg", "chi2", "score")])
    rl = None if rl_grid is None else np.ascontiguousarray(rl_grid, np.float32)
    xi = None if xi_grid is None else np.ascontiguousarray(xi_grid, np.float32)
    ctx = OrcCtx(float(bfield), _p(rl), _p(xi))
    L = lib()
    if T == np.float64:
        if ties or tie_shift:
            raise ValueError("threshold margins and shifts apply to the FP32 fit")
        rc = L.orc_backward_fit_f64(ctypes.byref(ctx), ctypes.byref(fe), ctypes.byref(pr), ctypes.byref(tr),
                                    int(n_threads))
    else:
        tie = np.zeros((3, n), np.float32) if ties else None
        rc = L.orc_backward_fit_f32(ctypes.byref(ctx), ctypes.byref(fe), ctypes.byref(pr), ctypes.byref(tr),
                                    int(n_threads), _p(tie), ctypes.c_float(tie_shift))
        if ties:
            out["tie"] = tie
    if rc:
        raise RuntimeError("orc_backward_fit failed (%d)" % rc)
    return out
