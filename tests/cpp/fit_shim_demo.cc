// fit_shim_demo.cc -- drives the C++ interface (include/mkfit_amd.hpp) the way
// a reference caller drives MkFitter: pack Track / Hit records, fit, unpack.
// Generates a toy event with the toy MC library, fits it on the GPU and writes
// par [6][n], err [21][n], chi2 [n] (float32) to the output file.
//   usage: fit_shim_demo OUT N_TRACKS SEED [plain]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mkfit_amd.hpp"
#include "mkfit_toymc.h"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s OUT N_TRACKS SEED [plain]\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[2]);
  const bool plain = argc > 4;
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

  // Reference-shaped records: layer l's hits, each track's HitOnTrack list.
  std::vector<std::vector<mkfit_amd::Hit>> layer_hits(S, std::vector<mkfit_amd::Hit>(n));
  std::vector<mkfit_amd::Track> tracks(n);
  for (int t = 0; t < n; ++t) {
    for (int e = 0; e < 6; ++e) tracks[t].state.par[e] = sp[(size_t)e * n + t];
    for (int e = 0; e < 21; ++e) tracks[t].state.err[e] = se[(size_t)e * n + t];
    tracks[t].state.charge = chg[t];
    for (int s = 0; s < nh[t]; ++s) {
      mkfit_amd::Hit& h = layer_hits[s][t];
      for (int e = 0; e < 3; ++e) h.pos[e] = hp[((size_t)s * 3 + e) * n + t];
      for (int e = 0; e < 6; ++e) h.err[e] = he[((size_t)s * 6 + e) * n + t];
      tracks[t].hots.push_back({t, s});
    }
  }
  try {
    mkfit_amd::Context ctx(p.bfield, rl.data(), xi.data());
    mkfit_amd::MkFitter fitter(ctx);
    fitter.InputTracksForFit(tracks, layer_hits, std::vector<bool>(S, false), 0, n);
    fitter.FitTracks(MKF_PF_APPLY_MATERIAL, MKF_FIT_CHI2, !plain);
    fitter.OutputFittedTracks(tracks, 0, n);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  std::FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 4;
  std::vector<float> out(28 * (size_t)n);
  for (int t = 0; t < n; ++t) {
    for (int e = 0; e < 6; ++e) out[(size_t)e * n + t] = tracks[t].state.par[e];
    for (int e = 0; e < 21; ++e) out[(size_t)(6 + e) * n + t] = tracks[t].state.err[e];
    out[(size_t)27 * n + t] = tracks[t].chi2;
  }
  std::fwrite(out.data(), sizeof(float), out.size(), f);
  std::fclose(f);
  return 0;
}
