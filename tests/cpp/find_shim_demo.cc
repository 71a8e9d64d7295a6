// find_shim_demo.cc -- a caller written against the reference's records and
// classes (include/mkfit_amd.hpp: Track, Hit, HitOnTrack, EventOfHits,
// MkBuilder, MkFinder, Config), driving one toy event the way the
// reference's MkBuilder does:
//   bh_fused   MkBuilder::FindTracksBestHit + export_tracks
//   bh_loop    the best-hit layer loop over MkFinder's per-layer steps
//              (PropagateTracksToR, SelectHitIndices, AddBestHit) with the
//              stop rules of MkBuilder.cc:1395-1478
//   bh_bkfit   MkBuilder::BackwardFitBH of the best-hit tracks
//   ce         MkBuilder::FindTracksCloneEngine (propagation to each hit), every
//              candidate of every seed
//   ce_bkfit   MkBuilder::BackwardFit of the best candidates
//   mf_bkfit   the same fit through MkFinder::BkFitInputTracks /
//              BkFitFitTracks / BkFitPropTracksToPCA / BkFitOutputTracks
// Writes named arrays: per array int32 name length, name, int32 kind (0 int32,
// 1 float32), int32 rank, int32 dims, data.
//   usage: find_shim_demo OUT N_TRACKS SEED
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mkfit_amd.hpp"
#include "mkfit_toymc.h"

using namespace mkfit_amd;

namespace {

struct Writer {
  std::FILE* f;
  template <class T>
  void put(const char* name, const std::vector<T>& v, std::vector<int> dims) {
    const int len = (int)std::strlen(name), kind = std::is_same<T, float>::value ? 1 : 0, rank = (int)dims.size();
    std::fwrite(&len, 4, 1, f);
    std::fwrite(name, 1, len, f);
    std::fwrite(&kind, 4, 1, f);
    std::fwrite(&rank, 4, 1, f);
    std::fwrite(dims.data(), 4, rank, f);
    std::fwrite(v.data(), sizeof(T), v.size(), f);
  }
};

// Element-major views of one track per seed: hit index per layer [L][n]
// (-9 none), par [6][n], err [21][n], chg, chi2, score [n].
struct Columns {
  std::vector<int> hot, chg;
  std::vector<float> par, err, chi2, score;
  Columns(const TrackVec& tv, int L) : hot((size_t)L * tv.size(), -9), chg(tv.size()), par(6 * tv.size()),
                                        err(21 * tv.size()), chi2(tv.size()), score(tv.size()) {
    const size_t n = tv.size();
    for (size_t t = 0; t < n; ++t) {
      const Track& tr = tv[t];
      for (int i = 0; i < tr.nTotalHits(); ++i)
        if (tr.getHitLyr(i) >= 0 && tr.getHitLyr(i) < L) hot[(size_t)tr.getHitLyr(i) * n + t] = tr.getHitIdx(i);
      for (int e = 0; e < 6; ++e) par[e * n + t] = tr.parameters()[e];
      for (int e = 0; e < 21; ++e) err[e * n + t] = tr.errors().Array()[e];
      chg[t] = tr.charge();
      chi2[t] = tr.chi2();
      score[t] = tr.score();
    }
  }
  void write(Writer& w, const std::string& p, int L) const {
    const int n = (int)chg.size();
    w.put((p + "_hot").c_str(), hot, {L, n});
    w.put((p + "_par").c_str(), par, {6, n});
    w.put((p + "_err").c_str(), err, {21, n});
    w.put((p + "_chg").c_str(), chg, {n});
    w.put((p + "_chi2").c_str(), chi2, {n});
    w.put((p + "_score").c_str(), score, {n});
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s OUT N_TRACKS SEED\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[2]);
  mkf_geometry geo;
  mkf_toymc_geometry_preset(1, &geo);
  mkf_toymc_params p;
  mkf_toymc_default_params(&p);
  p.seed = std::strtoull(argv[3], nullptr, 10);
  p.simulate_material = 0;
  p.pt_min = 0.5f;
  p.pt_max = 10.f;
  const float noise = 3.f;
  const int L = geo.n_barrel, first = 3;
  int cap = 0;
  for (int l = 0; l < L; ++l) cap = std::max(cap, mkf_toymc_noise_count(&geo, l, noise));
  cap += n;
  std::vector<float> hits((size_t)L * cap * 9), sp(6 * (size_t)n), se(21 * (size_t)n);
  std::vector<int> nlh(L), mc((size_t)L * cap), sc(n), sh((size_t)first * n), smc(n);
  mkf_toymc_find_output o{};
  o.n_tracks = n; o.layer_capacity = cap;
  o.hits = hits.data(); o.n_layer_hits = nlh.data(); o.hit_mc = mc.data();
  o.seed_par = sp.data(); o.seed_err = se.data(); o.seed_chg = sc.data(); o.seed_hits = sh.data();
  o.seed_mc = smc.data();
  if (mkf_toymc_find_event(&geo, &p, noise, first, &o, 4) != 0) return 3;
  const int ns = o.n_seeds;

  // LayerInfo of the thin barrel layers (the toy cylinder geometry's selection settings).
  std::vector<mkf_layer_desc> descs(L);
  for (int l = 0; l < L; ++l) {
    mkf_layer_desc& d = descs[l];
    d = mkf_layer_desc{};
    d.is_barrel = 1;
    d.rin = geo.barrel_r[l] - 0.01f; d.rout = geo.barrel_r[l] + 0.01f;
    d.zmin = -geo.barrel_zmax[l]; d.zmax = geo.barrel_zmax[l];
    d.propagate_to = geo.barrel_r[l];
    d.q_bin = 2.f;
    d.min_dphi = 0.01f; d.max_dphi = 0.05f; d.min_dq = 1.f; d.max_dq = 2.f;
    d.c_dp_sf = d.c_dq_sf = d.c_c2_sf = 1.1f;
  }

  // The event as the reference holds it: Event::layerHits_ and seedTracks_.
  Event ev;
  ev.layerHits_.resize(L);
  for (int l = 0; l < L; ++l)
    for (int i = 0; i < nlh[l]; ++i) {
      const float* src = &hits[((size_t)l * cap + i) * 9];
      SVector3 pos;
      SMatrixSym33 err;
      for (int e = 0; e < 3; ++e) pos[e] = src[e];
      for (int e = 0; e < 6; ++e) err.Array()[e] = src[3 + e];
      ev.layerHits_[l].push_back(Hit(pos, err, mc[(size_t)l * cap + i]));
    }
  for (int t = 0; t < ns; ++t) {
    SVector6 par;
    SMatrixSym66 err;
    for (int e = 0; e < 6; ++e) par[e] = sp[(size_t)e * n + t];
    for (int e = 0; e < 21; ++e) err.Array()[e] = se[(size_t)e * n + t];
    HitOnTrack hots[first];
    for (int l = 0; l < first; ++l) hots[l] = HitOnTrack(sh[(size_t)l * n + t], l);
    ev.seedTracks_.push_back(Track(TrackState(par, err, sc[t]), 0.f, smc[t], first, hots));
  }

  // Geoms/CylCowWLids.cc:225-231
  Config::finding_requires_propagation_to_hit_pos = false;
  Config::finding_inter_layer_pflags = PropagationFlags(PF_none);
  Config::finding_intra_layer_pflags = PropagationFlags(PF_none);
  Config::backward_fit_pflags = PropagationFlags(PF_use_param_b_field);
  Config::pca_prop_pflags = PropagationFlags(PF_use_param_b_field);
  Config::includePCA = true;

  std::FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 4;
  Writer w{f};
  w.put("n", std::vector<int>{ns, L}, {2});
  try {
    Context ctx(3.8f);
    EventOfHits eoh(ctx, descs);
    eoh.SuckInHits(ev.layerHits_);
    MkBuilder builder(ctx);

    // best hit, one call
    IterationParams bh_ip;
    bh_ip.nlayers_per_seed = first;
    bh_ip.chi2Cut_min = 15.f;
    builder.begin_event(eoh, bh_ip);
    builder.find_tracks_load_seeds_BH(ev.seedTracks_);
    builder.FindTracksBestHit();
    TrackVec fused;
    builder.export_tracks(fused);
    Columns(fused, L).write(w, "bh_fused", L);
    builder.BackwardFitBH();
    TrackVec bh_fitted;
    builder.export_tracks(bh_fitted);
    Columns(bh_fitted, L).write(w, "bh_bkfit", L);

    // best hit, the layer loop here over MkFinder's steps
    MkFinder finder(ctx);
    finder.Setup(bh_ip);
    finder.InputTracksAndHitIdx(ev.seedTracks_, 0, ns, false);
    std::vector<char> active(ns, 1);
    TrackVec looped(ns), now(ns), prop(ns);
    finder.OutputTracksAndHitIdx(now, 0, ns, false);
    for (int layer = first; layer < L; ++layer) {
      const LayerOfHits& loh = eoh.m_layers_of_hits[layer];
      finder.PropagateTracksToR(loh.layer_info().propagate_to, ns, Config::finding_inter_layer_pflags);
      finder.OutputTracksAndHitIdx(prop, 0, ns, true);
      for (int t = 0; t < ns; ++t)                          // cannot reach the layer: stop as is
        if (active[t] && prop[t].posR() < loh.layer_info().rin) {
          active[t] = 0;
          looped[t] = now[t];
        }
      finder.SelectHitIndices(loh, ns);
      finder.AddBestHit(loh, ns);
      finder.OutputTracksAndHitIdx(now, 0, ns, false);
      for (int t = 0; t < ns; ++t)
        if (active[t] && now[t].getLastHitLyr() == layer && now[t].getLastHitIdx() == -2) {
          active[t] = 0;                                    // too many holes: stop after the -2
          looped[t] = now[t];
        }
    }
    for (int t = 0; t < ns; ++t)
      if (active[t]) looped[t] = now[t];
    Columns(looped, L).write(w, "bh_loop", L);

    // clone engine: every candidate, then the backward fit of the best ones.
    // With the propagation to each hit (Geoms/CMS-2017.cc:723): without it the
    // clone engine updates a candidate's stored state, which stays at its seed's
    // radius, so its states are no input for a fit (MkBuilder.cc:1997-2035).
    Config::finding_requires_propagation_to_hit_pos = true;
    IterationParams ce_ip;
    ce_ip.nlayers_per_seed = first;
    ce_ip.chi2Cut_min = 30.f;
    ce_ip.maxCandsPerSeed = 5;
    builder.begin_event(eoh, ce_ip);
    builder.find_tracks_load_seeds(ev.seedTracks_);
    builder.FindTracksCloneEngine();
    std::vector<TrackVec> cands;
    builder.export_comb_cands(cands);
    const int K = ce_ip.maxCandsPerSeed;
    std::vector<int> nc(ns), chg((size_t)K * ns, 0), nf((size_t)K * ns, 0), hot((size_t)K * L * ns, -9);
    std::vector<float> par((size_t)K * 6 * ns, 0.f), err((size_t)K * 21 * ns, 0.f), chi2((size_t)K * ns, 0.f),
        score((size_t)K * ns, -1e16f);
    for (int t = 0; t < ns; ++t) {
      nc[t] = (int)cands[t].size();
      for (int k = 0; k < nc[t]; ++k) {
        const Track& tr = cands[t][k];
        const size_t s = (size_t)k * ns + t;
        for (int e = 0; e < 6; ++e) par[((size_t)k * 6 + e) * ns + t] = tr.parameters()[e];
        for (int e = 0; e < 21; ++e) err[((size_t)k * 21 + e) * ns + t] = tr.errors().Array()[e];
        chg[s] = tr.charge();
        chi2[s] = tr.chi2();
        score[s] = tr.score();
        nf[s] = tr.nFoundHits();
        for (int i = 0; i < tr.nTotalHits(); ++i)
          hot[((size_t)k * L + tr.getHitLyr(i)) * ns + t] = tr.getHitIdx(i);
      }
    }
    w.put("ce_n_cands", nc, {ns});
    w.put("ce_par", par, {K, 6, ns});
    w.put("ce_err", err, {K, 21, ns});
    w.put("ce_chg", chg, {K, ns});
    w.put("ce_chi2", chi2, {K, ns});
    w.put("ce_score", score, {K, ns});
    w.put("ce_n_found", nf, {K, ns});
    w.put("ce_hot_idx", hot, {K, L, ns});

    TrackVec best;
    builder.export_best_comb_cands(best);
    if ((int)best.size() != ns) throw Error("a seed without candidates");
    builder.BackwardFit();
    TrackVec ce_fitted;
    builder.export_best_comb_cands(ce_fitted);
    Columns(ce_fitted, L).write(w, "ce_bkfit", L);

    SteeringParams st_par;
    st_par.fill_plan(0, L - 1);
    st_par.set_iterator_limits(first, 0);
    finder.BkFitInputTracks(best, 0, ns);
    finder.BkFitFitTracks(eoh, st_par, ns);
    if (Config::includePCA) finder.BkFitPropTracksToPCA(ns);
    finder.BkFitOutputTracks(best, 0, ns, Config::includePCA);
    Columns(best, L).write(w, "mf_bkfit", L);

    // The clone engine again with strip charges in the hits' packed data
    // (Hit.h:251-299): of every 10 hits in load order, 2 pixel-like (0xff),
    // 1 below the cut (code 0), the rest codes 2..79 near it.
    std::vector<HitVec> qhits = ev.layerHits_;
    int g = 0;
    for (HitVec& hv : qhits)
      for (Hit& h : hv) {
        const int r = g % 10;
        const unsigned code = r < 2 ? 0xffu : r == 2 ? 0u : 2u + (unsigned)((long long)g * 7919 % 78);
        h.setPackedData(code << 12);
        ++g;
      }
    eoh.SuckInHits(qhits);
    eoh.UseStripCharges(true);
    builder.begin_event(eoh, ce_ip);
    builder.find_tracks_load_seeds(ev.seedTracks_);
    builder.FindTracksCloneEngine();
    std::vector<TrackVec> qc;
    builder.export_comb_cands(qc);
    std::vector<int> qn(ns), qhot((size_t)K * L * ns, -9);
    for (int t = 0; t < ns; ++t) {
      qn[t] = (int)qc[t].size();
      for (int k = 0; k < qn[t]; ++k)
        for (int i = 0; i < qc[t][k].nTotalHits(); ++i)
          qhot[((size_t)k * L + qc[t][k].getHitLyr(i)) * ns + t] = qc[t][k].getHitIdx(i);
    }
    w.put("ceq_n_cands", qn, {ns});
    w.put("ceq_hot_idx", qhot, {K, L, ns});
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    std::fclose(f);
    return 1;
  }
  std::fclose(f);
  return 0;
}
