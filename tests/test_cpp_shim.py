"""The C++ interface (include/mkfit_amd.hpp): the reference's records (Track, Hit, HitOnTrack as byte images,
checked by static_assert when the callers compile) and MkFitter / MkFinder / MkBuilder with the reference's
signatures.  On the GPU two callers written against those records reproduce the oracle:
  tests/cpp/fit_shim_demo.cc   runFittingTestPlex's shape -- one call over the range, one launch per slot,
                               blocks of 16 tracks as the Matriplex loop, the steered fit -- all bit-identical
                               and within the fit tolerances of the oracle, track by track;
  tests/cpp/find_shim_demo.cc  MkBuilder's best-hit search in one call and as a layer loop over MkFinder's
                               steps, the clone engine, and the backward fits (MkBuilder::BackwardFit[BH] and
                               MkFinder::BkFit*), each against the oracle run on the same inputs."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import mkfit_amd as mk
from _common import assert_track_parity
from _find_common import assert_ce_parity, best_hit_reference, clone_engine_reference, explain_best_hit
from oracle.parity import bkfit_reference, describe, fit_reference, flag_mask, flagged_parity

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


@pytest.mark.parametrize("name", ["fit_shim_demo", "find_shim_demo"])
def test_shim_compiles_and_links(name):
    from mkfit_amd import _build
    if not os.path.exists(os.path.join(ROOT, "mkfit_amd", "libmkfit_amd.so")):
        pytest.skip("library not built")
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "demo")
        r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I",
                            os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", "-o", out,
                            os.path.join(CPP, name + ".cc"), "-L", _build.PKG,
                            "-lmkfit_amd", "-lmkfit_toymc", "-L", "/opt/rocm/lib", "-lamdhip64"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_record_byte_images_match_event_file_sizes():
    """The records' sizes are the ones the event-file header carries (sizeof Track, Hit, HitOnTrack)."""
    src = r'''
#include <cstdio>
#include "mkfit_amd.hpp"
int main() {
  mkfit_amd::HitOnTrack h(-5, 7);
  unsigned raw;
  static_assert(sizeof(h) == sizeof(raw), "");
  std::memcpy(&raw, &h, 4);
  mkfit_amd::Hit s;
  s.setPackedData(7u | (2u << 12) | (3u << 20));   // detid 7, charge code 2, span_rows 3
  mkfit_amd::Hit p;
  p.setPackedData(0xffu << 12);                      // setupAsPixel's charge code
  std::printf("%zu %zu %zu %zu %08x %u %u %u %u\n", sizeof(mkfit_amd::Track), sizeof(mkfit_amd::Hit),
              sizeof(mkfit_amd::HitOnTrack), sizeof(mkfit_amd::TrackState), raw, s.detIDinLayer(),
              s.chargePerCM(), p.chargePerCM(), mkfit_amd::Hit::maxChargePerCM());
}
'''
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "s.cc"), "w") as f:
            f.write(src)
        exe = os.path.join(d, "s")
        r = subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                            "-I", "/opt/rocm/include", "-o", exe, os.path.join(d, "s.cc")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert out[:4] == ["160", "44", "4", "112"]
    raw = int(out[4], 16)
    assert raw & 0xffffff == (-5) & 0xffffff and raw >> 24 == 7   # index low 24 bits, layer high 8
    # Hit::PackedData (Hit.h:251-282): code 2 is 1628 per cm; the pixel code 0xff is maxChargePerCM
    assert out[5:] == ["7", "1628", "3652", "3652"]


def _read_arrays(path):
    out, raw, i = {}, open(path, "rb").read(), 0
    while i < len(raw):
        (ln,) = np.frombuffer(raw, np.int32, 1, i); i += 4
        name = raw[i:i + ln].decode(); i += ln
        kind, rank = np.frombuffer(raw, np.int32, 2, i); i += 8
        dims = np.frombuffer(raw, np.int32, rank, i); i += 4 * rank
        cnt = int(np.prod(dims))
        out[name] = np.frombuffer(raw, np.float32 if kind else np.int32, cnt, i).reshape(dims).copy()
        i += 4 * cnt
    return out


@pytest.mark.gpu
def test_shim_fit_matches_oracle(gpu_device, tmp_path):
    n, seed = 20000, 99
    res = {}
    for form in ("batched", "plain", "blocks", "steered"):
        out = str(tmp_path / ("fit_%s.bin" % form))
        cmd = [os.path.join(CPP, "fit_shim_demo"), out, str(n), str(seed)] + ([] if form == "batched" else [form])
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (form, r.stderr)
        res[form] = np.fromfile(out, np.float32).reshape(28, n)
    for form in ("plain", "blocks", "steered"):
        assert np.array_equal(res[form], res["batched"]), form
    geo = mk.geometry(1)
    rl, xi = mk.material_grid(geo)
    ev = mk.generate(geo, mk.default_toy_params(seed=seed), n, geo.n_barrel, n_threads=4)
    o, o64, chk = fit_reference(3.8, rl, xi, ev, 2, 1)
    g = res["batched"]
    assert_track_parity("shim fit", (g[:6], g[6:27], g[27]), (o[0], o[1], o[3]), (o64[0], o64[1], o64[3]), chk)


@pytest.fixture(scope="module")
def find_run(tmp_path_factory):
    n, seed = 3000, 55
    out = str(tmp_path_factory.mktemp("find") / "find.bin")
    r = subprocess.run([os.path.join(CPP, "find_shim_demo"), out, str(n), str(seed)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a = _read_arrays(out)
    geo = mk.geometry(1)
    ev = mk.generate_find_event(geo, mk.default_toy_params(seed=seed, simulate_material=0, pt_min=0.5, pt_max=10.0),
                                n, noise_per_cm2=3.0)
    assert ev["n_seeds"] == a["n"][0]
    return a, ev, mk.barrel_layer_descs(geo)


@pytest.mark.gpu
def test_shim_best_hit_matches_oracle(gpu_device, find_run):
    """MkBuilder::FindTracksBestHit and the layer loop over MkFinder's steps, each against the oracle:
    hit lists explained seed by seed, parameters track by track."""
    a, ev, descs = find_run
    cpu, chk = best_hit_reference(ev, descs, 3, chi2_cut=15.0, bfield=3.8, n_threads=8)
    for name in ("bh_fused", "bh_loop"):
        hot, par = a[name + "_hot"], a[name + "_par"]
        n_diff, counts, bad = explain_best_hit(hot, cpu["hot_idx"], cpu["margins"], 3)
        for t, l in bad[:3]:
            print(name, "seed", t, "layer", l, "gpu", hot[:, t].tolist(), "oracle", cpu["hot_idx"][:, t].tolist(),
                  "margins", cpu["margins"][:, l, t].tolist())
        assert not bad, (name, n_diff, counts, bad[:10])
        same = np.all(hot == cpu["hot_idx"], axis=0)
        sel = lambda x: x[..., same]
        r = flagged_parity((sel(par), sel(a[name + "_err"]), sel(a[name + "_chi2"])),
                           (sel(cpu["par"]), sel(cpu["err"]), sel(cpu["chi2"])),
                           (sel(cpu["par"]), sel(cpu["err"]), sel(cpu["chi2"])),
                           {"spread": chk["spread"][:, same], "alt": [tuple(sel(x) for x in al) for al in chk["alt"]],
                            "tie_all": True})
        print(name, n_diff, counts, describe(r))
        assert r["ok"], (name, r["beyond"])
        assert np.array_equal(a[name + "_chg"][same], cpu["chg"][same]), name
    assert np.array_equal(a["bh_fused_hot"], a["bh_loop_hot"])


@pytest.mark.gpu
def test_shim_clone_engine_matches_oracle(gpu_device, find_run):
    """MkBuilder::FindTracksCloneEngine with the propagation to each hit, every candidate of every seed from
    export_comb_cands."""
    a, ev, descs = find_run
    g = {k[3:]: v for k, v in a.items() if k.startswith("ce_") and not k.startswith("ce_bkfit")}
    cpu, spread = clone_engine_reference(ev, descs, 3, chi2_cut=30.0, max_cands=5, prop_to_hit=True, n_threads=8)
    assert_ce_parity("shim clone engine", g, cpu, spread, 3)


@pytest.mark.gpu
def test_shim_strip_charges_match_oracle(gpu_device, find_run):
    """EventOfHits::UseStripCharges over charge codes in the hits' packed data: the clone engine's hit lists
    match the oracle run with the same codes, seed by seed (as far as the candidates' decision margins allow)."""
    a, ev, descs = find_run
    g = np.arange(int(ev["hit_offset"][-1]))
    r = g % 10
    code = np.where(r < 2, 0xff, np.where(r == 2, 0, 2 + (g * 7919) % 78)).astype(np.uint8)
    cpu, spread = clone_engine_reference(dict(ev, hit_charge=code), descs, 3, chi2_cut=30.0, max_cands=5,
                                         prop_to_hit=True, n_threads=8)
    gq = {"n_cands": a["ceq_n_cands"], "hot_idx": a["ceq_hot_idx"]}
    assert not np.array_equal(a["ceq_hot_idx"], a["ce_hot_idx"])
    same = np.array([gq["n_cands"][t] == cpu["n_cands"][t] and
                     np.array_equal(gq["hot_idx"][:cpu["n_cands"][t], :, t], cpu["hot_idx"][:cpu["n_cands"][t], :, t])
                     for t in range(ev["n_seeds"])])
    assert same.mean() > 0.999, (same.mean(), np.nonzero(~same)[0][:10])


def _bkfit_check(name, a, prefix, ev, descs, cand, **kw):
    c32, c64, chk = bkfit_reference(ev, descs, cand, bfield=3.8, n_threads=8, **kw)
    r = assert_track_parity(name, (a[prefix + "_par"], a[prefix + "_err"], a[prefix + "_chi2"]),
                            (c32[0], c32[1], c32[3]), (c64[0], c64[1], c64[3]), chk)
    # BkFitOutputTracks copies no charge out: the found charge stays
    assert np.array_equal(a[prefix + "_chg"], cand["chg"]), name
    # getScoreCand with the same counts and pT class: scores differ by the chi2 difference
    unflagged = ~flag_mask(chk)[0]
    pt_g, pt_c = np.abs(1 / a[prefix + "_par"][3]), np.abs(1 / c32[0][3])
    same_class = unflagged & ((pt_g < 0.9) == (pt_c < 0.9)) & np.isfinite(c32[3])
    ds = np.abs((a[prefix + "_score"] + a[prefix + "_chi2"]) - (c32[4] + c32[3]))[same_class]
    assert ds.max() <= 1e-4 * max(1.0, np.abs(c32[4]).max()), (name, ds.max())
    return r


@pytest.mark.gpu
def test_shim_backward_fits_match_oracle(gpu_device, find_run):
    """MkBuilder::BackwardFitBH over the best-hit tracks (fake hits with material) and BackwardFit over the
    best candidates (blocks of 16, PCA), against the oracle's backward fit of the same found tracks; the
    same fit through MkFinder::BkFitInputTracks / BkFitFitTracks / BkFitPropTracksToPCA / BkFitOutputTracks
    gives the same bits as MkBuilder's."""
    a, ev, descs = find_run
    kw = dict(group=16, pflags=1, pca_pflags=1, include_pca=True, last_layer=0)
    bh = {"par": a["bh_fused_par"], "err": a["bh_fused_err"], "chg": a["bh_fused_chg"],
          "hot_idx": a["bh_fused_hot"]}
    _bkfit_check("shim BackwardFitBH", a, "bh_bkfit", ev, descs, bh, fake_material=True, **kw)
    ce = {"par": a["ce_par"][0], "err": a["ce_err"][0], "chg": a["ce_chg"][0], "hot_idx": a["ce_hot_idx"][0]}
    _bkfit_check("shim BackwardFit", a, "ce_bkfit", ev, descs, ce, fake_material=False, **kw)
    for key in ("par", "err", "chi2", "score", "chg", "hot"):
        assert np.array_equal(a["mf_bkfit_" + key], a["ce_bkfit_" + key]), key
