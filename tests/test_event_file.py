"""The binary event-file reader and writer (include/mkfit_event_file.h) against an independent numpy
restatement of the format (DataFileHeader, Event.h:74-127; Event::write_out / read_in and
write_tracks / read_tracks, Event.cc:123-330, 389-440; record layouts Track.h:74-143, 360-367,
619-620 and Hit.h:147-160, 181-182, 301-303, 314-320, 340-346).

The reference ships no event file, so the format is pinned by the layouts of its classes: numpy
structured dtypes with the field offsets of the x86-64 memory images write and read files that the
library must read and write identically, byte for byte.
"""
import os

import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from mkfit_amd import event_file as ef

TRACK = np.dtype({"names": ["par", "err", "chg", "valid", "chi2", "score", "last_hit", "n_found", "status", "label"],
                  "formats": [("<f4", 6), ("<f4", 21), "<i2", "u1", "<f4", "<f4", "<i2", "<i2", "<u4", "<i4"],
                  "offsets": [0, 24, 108, 110, 112, 116, 120, 122, 124, 128], "itemsize": 160})
STATE = np.dtype({"names": ["par", "err", "chg", "valid"], "formats": [("<f4", 6), ("<f4", 21), "<i2", "u1"],
                  "offsets": [0, 24, 108, 110], "itemsize": 112})
HIT = np.dtype({"names": ["pos", "err", "mc", "packed"], "formats": [("<f4", 3), ("<f4", 6), "<i4", "<u4"],
                "offsets": [0, 12, 36, 40], "itemsize": 44})


def _hot_words(index, layer):
    return ((np.asarray(index, np.int64) & 0xFFFFFF) | ((np.asarray(layer, np.int64) & 0xFF) << 24)).astype("<u4")


def _np_tracks(t):
    n = t["par"].shape[1]
    rec = np.zeros(n, TRACK)
    rec["par"] = t["par"].T
    rec["err"] = t["err"].T
    rec["chg"] = t["chg"]
    rec["valid"] = 1
    rec["chi2"] = t.get("chi2", np.zeros(n))
    rec["score"] = t.get("score", np.zeros(n))
    nh = np.diff(t["hot_offset"])
    rec["last_hit"] = nh - 1
    found = np.array([(t["hot_index"][t["hot_offset"][i]:t["hot_offset"][i + 1]] >= 0).sum() for i in range(n)])
    rec["n_found"] = found
    rec["status"] = t.get("status", np.zeros(n))
    rec["label"] = t.get("label", np.full(n, -1))
    hots = _hot_words(t["hot_index"], t["hot_layer"])
    size = 8 + rec.nbytes + hots.nbytes
    return np.array([n, size], "<i4").tobytes() + rec.tobytes() + hots.tobytes()


def np_write(path, events, n_layers, extra):
    """Event::write_out restated with numpy records."""
    out = [np.array([0xBEEF, 6, 160, 44, 4, n_layers, len(events), extra], "<i4").tobytes()]
    for ev in events:
        body = [_np_tracks(ev["sim"])]
        if extra & ef.SIM_TRACK_STATES:
            ts = ev["track_states"]
            rec = np.zeros(ts["par"].shape[1], STATE)
            rec["par"], rec["err"], rec["chg"], rec["valid"] = ts["par"].T, ts["err"].T, ts["chg"], 1
            body.append(np.array([len(rec)], "<i4").tobytes() + rec.tobytes())
        off = ev["hit_offset"]
        body.append(np.array([len(off) - 1], "<i4").tobytes())
        for l in range(len(off) - 1):
            h = ev["hits_flat"][off[l]:off[l + 1]]
            rec = np.zeros(len(h), HIT)
            rec["pos"], rec["err"] = h[:, :3], h[:, 3:]
            rec["mc"] = ev["hit_mc"][off[l]:off[l + 1]]
            rec["packed"] = ev["hit_packed"][off[l]:off[l + 1]]
            body.append(np.array([len(h)], "<i4").tobytes() + rec.tobytes())
        if extra & ef.HIT_ITER_MASKS:
            body.append(np.ascontiguousarray(ev["hit_masks"], "<u8").tobytes())
        mc = np.ascontiguousarray(ev["mc_hit_info"], "<i4")
        body.append(np.array([len(mc)], "<i4").tobytes() + mc.tobytes())
        if extra & ef.SEEDS:
            body.append(_np_tracks(ev["seed"]))
        if extra & ef.CMSSW_TRACKS:
            body.append(_np_tracks(ev["cmssw"]))
        if extra & ef.BEAM_SPOT:
            body.append(np.ascontiguousarray(ev["beam_spot"], "<f4").tobytes())
        b = b"".join(body)
        out.append(np.array([len(b) + 4], "<i4").tobytes() + b)
    with open(path, "wb") as f:
        f.write(b"".join(out))


def _tracks(rng, n, n_layers):
    nh = rng.integers(0, 12, n)
    off = np.concatenate([[0], np.cumsum(nh)]).astype(np.int32)
    idx = rng.integers(-3, 5000, off[-1]).astype(np.int32)
    lay = rng.integers(0, n_layers, off[-1]).astype(np.int32)
    return dict(par=rng.normal(size=(6, n)).astype(np.float32), err=rng.normal(size=(21, n)).astype(np.float32),
                chg=rng.choice([-1, 1], n).astype(np.int32), chi2=rng.uniform(0, 50, n).astype(np.float32),
                score=rng.normal(size=n).astype(np.float32), label=rng.integers(-1, 1000, n).astype(np.int32),
                status=rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32), hot_offset=off, hot_index=idx,
                hot_layer=lay)


def _event(rng, n_layers=7):
    nh = rng.integers(0, 300, n_layers)
    nh[2] = 0
    off = np.concatenate([[0], np.cumsum(nh)]).astype(np.int32)
    tot = off[-1]
    return dict(n_layers=n_layers, hit_offset=off, hits_flat=rng.normal(size=(tot, 9)).astype(np.float32),
                hit_mc=rng.integers(-1, 10000, tot).astype(np.int32),
                hit_packed=rng.integers(0, 2**26, tot).astype(np.uint32),
                hit_masks=rng.integers(0, 2**62, tot, dtype=np.uint64),
                sim=_tracks(rng, 40, n_layers), seed=_tracks(rng, 25, n_layers), cmssw=_tracks(rng, 0, n_layers),
                track_states=dict(par=rng.normal(size=(6, 9)).astype(np.float32),
                                  err=rng.normal(size=(21, 9)).astype(np.float32),
                                  chg=rng.choice([-1, 1], 9).astype(np.int32)),
                mc_hit_info=rng.integers(0, 100, (33, 4)).astype(np.int32),
                beam_spot=rng.normal(size=8).astype(np.float32))


def _assert_tracks_equal(a, b):
    for k in ("par", "err", "chg", "chi2", "score", "label", "status", "hot_offset", "hot_index", "hot_layer"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k]).astype(np.asarray(a[k]).dtype)), k
    found = [(b["hot_index"][b["hot_offset"][i]:b["hot_offset"][i + 1]] >= 0).sum() for i in range(len(b["chg"]))]
    assert np.array_equal(a["n_found"], found)


def _assert_events_equal(got, ref, extra):
    for k in ("hit_offset", "hits_flat", "hit_mc", "hit_packed", "mc_hit_info"):
        assert np.array_equal(got[k], ref[k]), k
    _assert_tracks_equal(got["sim"], ref["sim"])
    if extra & ef.SEEDS:
        _assert_tracks_equal(got["seed"], ref["seed"])
    if extra & ef.CMSSW_TRACKS:
        _assert_tracks_equal(got["cmssw"], ref["cmssw"])
    if extra & ef.HIT_ITER_MASKS:
        assert np.array_equal(got["hit_masks"], ref["hit_masks"])
    if extra & ef.SIM_TRACK_STATES:
        for k in ("par", "err", "chg"):
            assert np.array_equal(got["track_states"][k], ref["track_states"][k])
    if extra & ef.BEAM_SPOT:
        assert np.array_equal(got["beam_spot"], ref["beam_spot"])


@pytest.mark.parametrize("extra", [0, ef.SEEDS, 0x1F])
def test_reader_reads_numpy_written_file(tmp_path, extra):
    rng = np.random.default_rng(extra)
    evs = [_event(rng) for _ in range(3)]
    path = tmp_path / "ev.bin"
    np_write(path, evs, 7, extra)
    with ef.EventFileReader(path) as f:
        assert f.header == dict(magic=0xBEEF, format_version=6, sizeof_track=160, sizeof_hit=44, sizeof_hot=4,
                                n_layers=7, n_events=3, extra_sections=extra)
        got = list(f)
    assert len(got) == 3
    for g, r in zip(got, evs):
        _assert_events_equal(g, r, extra)


@pytest.mark.parametrize("extra", [0, ef.SEEDS | ef.BEAM_SPOT, 0x1F])
def test_writer_matches_numpy_bytes(tmp_path, extra):
    rng = np.random.default_rng(10 + extra)
    evs = [_event(rng) for _ in range(2)]
    a, b = tmp_path / "a.bin", tmp_path / "b.bin"
    with ef.EventFileWriter(a, 7, n_events=0, extra_sections=extra) as w:
        for ev in evs:
            w.write_event(ev)
    np_write(b, evs, 7, extra)
    assert a.read_bytes() == b.read_bytes()


def test_reader_rejects_foreign_files(tmp_path):
    rng = np.random.default_rng(3)
    good = tmp_path / "g.bin"
    np_write(good, [_event(rng)], 7, ef.SEEDS)
    raw = bytearray(good.read_bytes())
    cases = {"magic": (0, 0xBEEE), "version": (4, 7), "Track": (8, 168), "Hit": (12, 40), "HitOnTrack": (16, 8)}
    for what, (off, val) in cases.items():
        bad = bytearray(raw)
        bad[off:off + 4] = np.array([val], "<i4").tobytes()
        p = tmp_path / ("bad_%s.bin" % what)
        p.write_bytes(bytes(bad))
        with pytest.raises(mk.MkfError):
            ef.EventFileReader(p)
    # a truncated event is an error, not a short read
    p = tmp_path / "trunc.bin"
    p.write_bytes(bytes(raw[:-10]))
    with ef.EventFileReader(p) as f:
        with pytest.raises(mk.MkfError, match="truncated"):
            f.read_event()
    # a section size that disagrees with its tracks is an error
    bad = bytearray(raw)
    sz = np.frombuffer(bytes(bad[40:44]), "<i4")[0]
    bad[40:44] = np.array([sz + 4], "<i4").tobytes()
    p = tmp_path / "size.bin"
    p.write_bytes(bytes(bad))
    with ef.EventFileReader(p) as f:
        with pytest.raises(mk.MkfError):
            f.read_event()
    # version 4 files are accepted (DataFile::OpenRead's range)
    v4 = bytearray(raw)
    v4[4:8] = np.array([4], "<i4").tobytes()
    p = tmp_path / "v4.bin"
    p.write_bytes(bytes(v4))
    with ef.EventFileReader(p) as f:
        assert f.header["format_version"] == 4 and f.read_event() is not None and f.read_event() is None


def test_hit_on_track_packing(tmp_path):
    """HitOnTrack {index:24, layer:8} keeps negative fake codes and layers up to 127."""
    rng = np.random.default_rng(4)
    ev = _event(rng)
    idx = np.array([-9, -3, -2, -1, 0, 1, 2**23 - 1], np.int32)
    lay = np.array([-1, 0, 5, 127, 64, 3, 1], np.int32)
    ev["sim"] = dict(par=np.zeros((6, 1), np.float32), err=np.zeros((21, 1), np.float32), chg=np.ones(1, np.int32),
                     hot_offset=np.array([0, len(idx)], np.int32), hot_index=idx, hot_layer=lay)
    p = tmp_path / "hot.bin"
    with ef.EventFileWriter(p, 7) as w:
        w.write_event(ev)
    with ef.EventFileReader(p) as f:
        got = f.read_event()
    assert np.array_equal(got["sim"]["hot_index"], idx) and np.array_equal(got["sim"]["hot_layer"], lay)
    assert got["sim"]["n_found"][0] == 3
    raw = p.read_bytes()
    words = np.frombuffer(raw[32 + 4 + 8 + 160:32 + 4 + 8 + 160 + 4 * len(idx)], "<u4")
    assert np.array_equal(words, _hot_words(idx, lay))


def test_toy_event_round_trip_finds_the_same_tracks(tmp_path):
    geo = mk.geometry(1)
    p = mk.default_toy_params(seed=17, simulate_material=0, pt_min=0.5, pt_max=10.0)
    ev = mk.generate_find_event(geo, p, 300, noise_per_cm2=2.0)
    path = tmp_path / "toy.bin"
    with ef.EventFileWriter(path, geo.n_barrel, extra_sections=ef.SEEDS) as w:
        w.write_event(ef.find_event_to_file_event(ev))
    with ef.EventFileReader(path) as f:
        back = ef.file_event_to_find_event(f.read_event())
    for k in ("hits_flat", "hit_offset", "hit_mc", "seed_par", "seed_err", "seed_chg", "seed_hits", "seed_mc"):
        assert np.array_equal(back[k], np.asarray(ev[k])[..., :back["n_seeds"]] if k.startswith("seed") else ev[k]), k
    descs = mk.barrel_layer_descs(geo)
    a = oracle.find_best_hit(ev, descs, 3, chi2_cut=15.0, bfield=3.8, n_threads=2)
    b = oracle.find_best_hit(back, descs, 3, chi2_cut=15.0, bfield=3.8, n_threads=2)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
