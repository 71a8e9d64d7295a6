// fit_shim_demo.cc -- a caller written against the reference's records and
// MkFitter (include/mkfit_amd.hpp), shaped like runFittingTestPlex
// (fittestMPlex.cc:46-105): an Event whose simTracks_ carry their HitOnTrack
// lists into ev.layerHits_, fitted with
//   (default)  InputTracksForFit + FitTracksWithInterSlurp + OutputFittedTracks
//              over the whole range in one call
//   plain      the same with one launch per hit slot
//   blocks     InputTracksAndHits + FitTracks + OutputFittedTracks per block
//              of 16 tracks, as the reference's Matriplex loop
//   steered    FitTracksSteered with every slot marked barrel
// Writes par [6][n], err [21][n], chi2 [n] (float32) to the output file.
//   usage: fit_shim_demo OUT N_TRACKS SEED [plain|blocks|steered]
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mkfit_amd.hpp"
#include "mkfit_toymc.h"

using namespace mkfit_amd;

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s OUT N_TRACKS SEED [plain|blocks|steered]\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[2]);
  const char* mode = argc > 4 ? argv[4] : "";
  mkf_geometry geo;
  mkf_toymc_geometry_preset(1, &geo);
  mkf_toymc_params p;
  mkf_toymc_default_params(&p);
  p.seed = std::strtoull(argv[3], nullptr, 10);
  const int S = geo.n_barrel;
  std::vector<float> sp(6 * n), se(21 * n), hp((size_t)S * 3 * n), he((size_t)S * 6 * n);
  std::vector<int> chg(n), nh(n);
  mkf_toymc_output o{};
  o.n_tracks = n; o.n_slots = S;
  o.seed_par = sp.data(); o.seed_err = se.data(); o.chg = chg.data();
  o.hit_pos = hp.data(); o.hit_err = he.data(); o.n_hits = nh.data();
  if (mkf_toymc_generate(&geo, &p, &o, 4) != 0) return 3;
  std::vector<float> rl(300 * 120), xi(300 * 120);
  mkf_toymc_material_grid(&geo, rl.data(), xi.data());

  // The event as the reference holds it: track t's hit on layer l is
  // ev.layerHits_[l][t], named by its HitOnTrack list.
  Event ev;
  ev.layerHits_.assign(S, HitVec(n));
  for (int t = 0; t < n; ++t) {
    SVector6 par;
    SMatrixSym66 err;
    for (int e = 0; e < 6; ++e) par[e] = sp[(size_t)e * n + t];
    for (int e = 0; e < 21; ++e) err.Array()[e] = se[(size_t)e * n + t];
    Track trk(TrackState(par, err, chg[t]), 0.f, t, 0, nullptr);
    for (int s = 0; s < nh[t]; ++s) {
      Hit& h = ev.layerHits_[s][t];
      for (int e = 0; e < 3; ++e) h.pos_[e] = hp[((size_t)s * 3 + e) * n + t];
      for (int e = 0; e < 6; ++e) h.err_.Array()[e] = he[((size_t)s * 6 + e) * n + t];
      trk.addHitIdx(t, s, 0.f);
    }
    ev.simTracks_.push_back(trk);
  }
  Config::forward_fit_pflags = PropagationFlags(PF_apply_material);
  TrackVec rectracks(n);
  try {
    Context ctx(p.bfield, rl.data(), xi.data());
    MkFitter mkfp(ctx);
    mkfp.SetFitFlags(MKF_FIT_CHI2);
    if (!std::strcmp(mode, "blocks")) {
      const int NN = 16;
      mkfp.SetNhits(S);
      for (int itrack = 0; itrack < n; itrack += NN) {
        const int end = std::min(itrack + NN, n);
        mkfp.InputTracksAndHits(ev.simTracks_, ev.layerHits_, itrack, end);
        mkfp.FitTracks(end - itrack, &ev, Config::forward_fit_pflags);
        mkfp.OutputFittedTracks(rectracks, itrack, end);
      }
    } else if (!std::strcmp(mode, "steered")) {
      bool is_brl[64];
      std::fill(is_brl, is_brl + 64, true);
      mkfp.InputTracksForFit(ev.simTracks_, 0, n);
      mkfp.FitTracksSteered(is_brl, n, &ev, Config::forward_fit_pflags);
      mkfp.OutputFittedTracks(rectracks, 0, n);
    } else {
      mkfp.SetBatched(std::strcmp(mode, "plain") != 0);
      mkfp.InputTracksForFit(ev.simTracks_, 0, n);
      mkfp.FitTracksWithInterSlurp(ev.layerHits_, n);
      mkfp.OutputFittedTracks(rectracks, 0, n);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  for (int t = 0; t < n; ++t)
    if (rectracks[t].label() != t) {
      std::fprintf(stderr, "label of track %d not carried through\n", t);
      return 5;
    }
  std::FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 4;
  std::vector<float> out(28 * (size_t)n);
  for (int t = 0; t < n; ++t) {
    for (int e = 0; e < 6; ++e) out[(size_t)e * n + t] = rectracks[t].parameters()[e];
    for (int e = 0; e < 21; ++e) out[(size_t)(6 + e) * n + t] = rectracks[t].errors().Array()[e];
    out[(size_t)27 * n + t] = rectracks[t].chi2();
  }
  std::fwrite(out.data(), sizeof(float), out.size(), f);
  std::fclose(f);
  return 0;
}
