"""mkFit binary event files (Event::write_out / Event::read_in, Event.cc:123-330) through the C ABI of
include/mkfit_event_file.h.

    with EventFileReader(path) as f:          # DataFile::OpenRead
        for ev in f:                          # Event::read_in, one event at a time
            eoh.load(ev["hit_offset"], torch.from_numpy(ev["hits_flat"]).to(dev))

An event is a dict in the layout the rest of the package uses: hits_flat [total][9] (x y z, then the
packed 3x3 covariance), hit_offset [n_layers + 1], hit_mc [total] (Hit::mcHitID), hit_packed [total];
per track set ("sim", "seed", "cmssw") par [6][n], err [21][n], chg, chi2, score, label, n_found,
status and the HitOnTrack lists as hot_offset [n + 1], hot_index, hot_layer; mc_hit_info [n][4];
optionally track_states (par, err, chg), hit_masks and beam_spot.
"""
import ctypes

import numpy as np

from . import MkfError, lib

SIM_TRACK_STATES, SEEDS, CMSSW_TRACKS, HIT_ITER_MASKS, BEAM_SPOT = 0x1, 0x2, 0x4, 0x8, 0x10
TRACK_SETS = ("sim", "seed", "cmssw")
_END = 100
_c_void_p = ctypes.c_void_p


class _Header(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int) for k in ("magic", "format_version", "sizeof_track", "sizeof_hit", "sizeof_hot",
                                            "n_layers", "n_events", "extra_sections")]


class _Counts(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int * 3), ("n_hots", ctypes.c_int * 3), ("n_track_states", ctypes.c_int),
                ("n_layers", ctypes.c_int), ("n_hits", ctypes.c_int), ("n_mc_hit_info", ctypes.c_int),
                ("has_masks", ctypes.c_int), ("has_beam_spot", ctypes.c_int)]


_TRACK_FIELDS = ("par", "err", "chg", "chi2", "score", "label", "n_found", "status", "hot_offset", "hot_index",
                 "hot_layer")


class _Tracks(ctypes.Structure):
    _fields_ = [(k, _c_void_p) for k in _TRACK_FIELDS]


class _EventData(ctypes.Structure):
    _fields_ = [("n_layers", ctypes.c_int), ("hit_offset", _c_void_p), ("hits", _c_void_p), ("hit_mc", _c_void_p),
                ("hit_packed", _c_void_p), ("hit_masks", _c_void_p), ("n_tracks", ctypes.c_int * 3)] + \
        [(k, _c_void_p * 3) for k in ("par", "err", "chg", "chi2", "score", "label", "status", "hot_offset",
                                      "hot_index", "hot_layer")] + \
        [("n_track_states", ctypes.c_int), ("ts_par", _c_void_p), ("ts_err", _c_void_p), ("ts_chg", _c_void_p),
         ("n_mc_hit_info", ctypes.c_int), ("mc_hit_info", _c_void_p), ("beam_spot", _c_void_p)]


def _L():
    L = lib()
    if not getattr(L, "_ef_ready", False):
        L.mkf_event_file_last_error.restype = ctypes.c_char_p
        L.mkf_event_file_open_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(_c_void_p), ctypes.POINTER(_Header)]
        L.mkf_event_file_open_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.POINTER(_c_void_p)]
        L.mkf_event_file_close.argtypes = [_c_void_p]
        L.mkf_event_file_read_event.argtypes = [_c_void_p, ctypes.POINTER(_Counts)]
        L.mkf_event_file_get_hits.argtypes = [_c_void_p] + [_c_void_p] * 5
        L.mkf_event_file_get_tracks.argtypes = [_c_void_p, ctypes.c_int, ctypes.POINTER(_Tracks)]
        L.mkf_event_file_get_track_states.argtypes = [_c_void_p] + [_c_void_p] * 3
        L.mkf_event_file_get_mc_hit_info.argtypes = [_c_void_p, _c_void_p]
        L.mkf_event_file_get_beam_spot.argtypes = [_c_void_p, _c_void_p]
        L.mkf_event_file_write_event.argtypes = [_c_void_p, ctypes.POINTER(_EventData)]
        L._ef_ready = True
    return L


def _check(rc):
    if rc != 0:
        raise MkfError("event file: " + _L().mkf_event_file_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(_c_void_p)


class EventFileReader:
    def __init__(self, path):
        L = _L()
        self.h = _c_void_p()
        hdr = _Header()
        _check(L.mkf_event_file_open_read(str(path).encode(), ctypes.byref(self.h), ctypes.byref(hdr)))
        self.header = {k: getattr(hdr, k) for k, _ in _Header._fields_}

    def close(self):
        if self.h:
            _L().mkf_event_file_close(self.h)
            self.h = _c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __iter__(self):
        while True:
            ev = self.read_event()
            if ev is None:
                return
            yield ev

    def read_event(self):
        """The next event as a dict, or None past the last one."""
        L = _L()
        c = _Counts()
        rc = L.mkf_event_file_read_event(self.h, ctypes.byref(c))
        if rc == _END:
            return None
        _check(rc)
        nl, nh = c.n_layers, c.n_hits
        ev = dict(n_layers=nl, hit_offset=np.zeros(nl + 1, np.int32), hits_flat=np.zeros((nh, 9), np.float32),
                  hit_mc=np.zeros(nh, np.int32), hit_packed=np.zeros(nh, np.uint32))
        masks = np.zeros(nh, np.uint64) if c.has_masks else None
        _check(L.mkf_event_file_get_hits(self.h, _p(ev["hit_offset"]), _p(ev["hits_flat"]), _p(ev["hit_mc"]),
                                         _p(ev["hit_packed"]), _p(masks)))
        if masks is not None:
            ev["hit_masks"] = masks
        for w, name in enumerate(TRACK_SETS):
            n, nh_ = c.n_tracks[w], c.n_hots[w]
            t = dict(par=np.zeros((6, n), np.float32), err=np.zeros((21, n), np.float32), chg=np.zeros(n, np.int32),
                     chi2=np.zeros(n, np.float32), score=np.zeros(n, np.float32), label=np.zeros(n, np.int32),
                     n_found=np.zeros(n, np.int32), status=np.zeros(n, np.uint32),
                     hot_offset=np.zeros(n + 1, np.int32), hot_index=np.zeros(nh_, np.int32),
                     hot_layer=np.zeros(nh_, np.int32))
            st = _Tracks(*[_p(t[k]) for k in _TRACK_FIELDS])
            _check(L.mkf_event_file_get_tracks(self.h, w, ctypes.byref(st)))
            ev[name] = t
        if self.header["extra_sections"] & SIM_TRACK_STATES:
            n = c.n_track_states
            ts = dict(par=np.zeros((6, n), np.float32), err=np.zeros((21, n), np.float32), chg=np.zeros(n, np.int32))
            _check(L.mkf_event_file_get_track_states(self.h, _p(ts["par"]), _p(ts["err"]), _p(ts["chg"])))
            ev["track_states"] = ts
        ev["mc_hit_info"] = np.zeros((c.n_mc_hit_info, 4), np.int32)
        _check(L.mkf_event_file_get_mc_hit_info(self.h, _p(ev["mc_hit_info"])))
        if c.has_beam_spot:
            ev["beam_spot"] = np.zeros(8, np.float32)
            _check(L.mkf_event_file_get_beam_spot(self.h, _p(ev["beam_spot"])))
        return ev


class EventFileWriter:
    def __init__(self, path, n_layers, n_events=0, extra_sections=0):
        self.h = _c_void_p()
        self.extra = int(extra_sections)
        _check(_L().mkf_event_file_open_write(str(path).encode(), int(n_layers), int(n_events), self.extra,
                                              ctypes.byref(self.h)))

    def close(self):
        if self.h:
            h, self.h = self.h, _c_void_p()
            _check(_L().mkf_event_file_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def write_event(self, ev):
        """Event::write_out of one event dict (missing track sets are written empty)."""
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a.ctypes.data_as(_c_void_p)

        d = _EventData()
        d.n_layers = int(ev["n_layers"]) if "n_layers" in ev else len(ev["hit_offset"]) - 1
        d.hit_offset = arr(ev["hit_offset"], np.int32)
        d.hits = arr(ev["hits_flat"], np.float32)
        d.hit_mc = arr(ev.get("hit_mc"), np.int32)
        d.hit_packed = arr(ev.get("hit_packed"), np.uint32)
        d.hit_masks = arr(ev.get("hit_masks"), np.uint64)
        for w, name in enumerate(TRACK_SETS):
            t = ev.get(name)
            if t is None:
                t = dict(par=np.zeros((6, 0), np.float32), err=np.zeros((21, 0), np.float32),
                         chg=np.zeros(0, np.int32), hot_offset=np.zeros(1, np.int32))
            d.n_tracks[w] = t["par"].shape[1]
            d.par[w] = arr(t["par"], np.float32)
            d.err[w] = arr(t["err"], np.float32)
            d.chg[w] = arr(t["chg"], np.int32)
            d.chi2[w] = arr(t.get("chi2"), np.float32)
            d.score[w] = arr(t.get("score"), np.float32)
            d.label[w] = arr(t.get("label"), np.int32)
            d.status[w] = arr(t.get("status"), np.uint32)
            d.hot_offset[w] = arr(t["hot_offset"], np.int32)
            d.hot_index[w] = arr(t.get("hot_index", np.zeros(0, np.int32)), np.int32)
            d.hot_layer[w] = arr(t.get("hot_layer", np.zeros(0, np.int32)), np.int32)
        ts = ev.get("track_states")
        if ts is not None:
            d.n_track_states = ts["par"].shape[1]
            d.ts_par, d.ts_err, d.ts_chg = arr(ts["par"], np.float32), arr(ts["err"], np.float32), \
                arr(ts["chg"], np.int32)
        mc = ev.get("mc_hit_info")
        if mc is not None:
            d.n_mc_hit_info = len(mc)
            d.mc_hit_info = arr(mc, np.int32)
        d.beam_spot = arr(ev.get("beam_spot"), np.float32)
        _check(_L().mkf_event_file_write_event(self.h, ctypes.byref(d)))


def find_event_to_file_event(ev, first_layer=3):
    """A toy finding event (mkfit_amd.generate_find_event) as an event-file dict: hits with their
    generating track as mcHitID, the seeds with their seed hits as HitOnTrack lists."""
    ns = ev["n_seeds"]
    idx = np.ascontiguousarray(ev["seed_hits"][:first_layer, :ns].T).reshape(-1)
    lay = np.tile(np.arange(first_layer, dtype=np.int32), ns)
    seed = dict(par=ev["seed_par"][:, :ns], err=ev["seed_err"][:, :ns], chg=ev["seed_chg"][:ns],
                label=ev["seed_mc"][:ns], hot_offset=np.arange(ns + 1, dtype=np.int32) * first_layer,
                hot_index=idx.astype(np.int32), hot_layer=lay)
    return dict(n_layers=len(ev["hit_offset"]) - 1, hit_offset=ev["hit_offset"], hits_flat=ev["hits_flat"],
                hit_mc=ev["hit_mc"], seed=seed)


def file_event_to_find_event(fev, first_layer=3):
    """The inverse of find_event_to_file_event: seeds in the element-major finder layout, the seed hit
    on each of the first `first_layer` layers taken from the seeds' HitOnTrack lists (-1 if none)."""
    s = fev["seed"]
    ns = s["par"].shape[1]
    seed_hits = np.full((first_layer, ns), -1, np.int32)
    for t in range(ns):
        for k in range(s["hot_offset"][t], s["hot_offset"][t + 1]):
            if 0 <= s["hot_layer"][k] < first_layer:
                seed_hits[s["hot_layer"][k], t] = s["hot_index"][k]
    return dict(hits_flat=fev["hits_flat"], hit_offset=fev["hit_offset"], hit_mc=fev["hit_mc"], n_seeds=ns,
                seed_par=s["par"], seed_err=s["err"], seed_chg=s["chg"], seed_hits=seed_hits, seed_mc=s["label"])
