// find_shim_demo.cc -- drives the C++ interface (include/mkfit_amd.hpp) the way
// MkBuilder drives MkFinder for best-hit finding, twice on one toy event:
// (a) MkFinder::FindTracksBestHit, the whole layer loop in one call;
// (b) the layer loop run here over the per-layer steps PropagateTracksToR,
//     SelectHitIndices and AddBestHit, with the reference's stop rules
//     (MkBuilder.cc:1395-1478).
// Writes int32 {n_seeds, n_layers}, then for (a) and (b) the hit index per
// layer [n_layers][n_seeds] (-9 none) and float32 par [6][n_seeds].
//   usage: find_shim_demo OUT N_TRACKS SEED
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mkfit_amd.hpp"
#include "mkfit_toymc.h"

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
  std::vector<std::vector<mkfit_amd::Hit>> layer_hits(L);
  for (int l = 0; l < L; ++l)
    for (int i = 0; i < nlh[l]; ++i) {
      mkfit_amd::Hit h;
      const float* src = &hits[((size_t)l * cap + i) * 9];
      for (int e = 0; e < 3; ++e) h.pos[e] = src[e];
      for (int e = 0; e < 6; ++e) h.err[e] = src[3 + e];
      layer_hits[l].push_back(h);
    }
  std::vector<mkfit_amd::Track> seeds(ns);
  for (int t = 0; t < ns; ++t) {
    for (int e = 0; e < 6; ++e) seeds[t].state.par[e] = sp[(size_t)e * n + t];
    for (int e = 0; e < 21; ++e) seeds[t].state.err[e] = se[(size_t)e * n + t];
    seeds[t].state.charge = sc[t];
    for (int l = 0; l < first; ++l) seeds[t].hots.push_back({sh[(size_t)l * n + t], l});
  }
  std::vector<mkfit_amd::Track> fused, looped(ns);
  try {
    mkfit_amd::Context ctx(3.8f);
    mkfit_amd::EventOfHits eoh(ctx, descs);
    eoh.SuckInHits(layer_hits);
    mkfit_amd::MkFinder finder(ctx);
    finder.InputTracksAndHitIdx(seeds, 0, ns, first, L);
    finder.FindTracksBestHit(eoh, 15.f);
    finder.OutputTracksAndHitIdx(fused);

    finder.BeginLayerSteps();
    std::vector<char> active(ns, 1);
    std::vector<mkfit_amd::Track> now;
    finder.OutputLanes(now);
    std::vector<float> pp;
    for (int layer = first; layer < L; ++layer) {
      finder.PropagateTracksToR(descs[layer].propagate_to);
      finder.PropagatedParameters(pp);
      for (int t = 0; t < ns; ++t) {                        // cannot reach the layer: stop as is
        const float x = pp[(size_t)0 * ns + t], y = pp[(size_t)1 * ns + t];
        if (active[t] && x * x + y * y < descs[layer].rin * descs[layer].rin) {
          active[t] = 0;
          looped[t] = now[t];
        }
      }
      finder.SelectHitIndices(eoh, layer);
      const std::vector<int> code = finder.AddBestHit(eoh, layer, 15.f);
      finder.OutputLanes(now);
      for (int t = 0; t < ns; ++t)
        if (active[t] && code[t] == -2) {                   // too many holes: stop after the -2
          active[t] = 0;
          looped[t] = now[t];
        }
    }
    for (int t = 0; t < ns; ++t)
      if (active[t]) looped[t] = now[t];
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  std::FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 4;
  const int head[2] = {ns, L};
  std::fwrite(head, sizeof(int), 2, f);
  for (const std::vector<mkfit_amd::Track>* v : {&fused, &looped}) {
    std::vector<int> hot((size_t)L * ns, -9);
    std::vector<float> par(6 * (size_t)ns);
    for (int t = 0; t < ns; ++t) {
      for (const mkfit_amd::HitOnTrack& h : (*v)[t].hots)
        if (h.layer >= 0 && h.layer < L) hot[(size_t)h.layer * ns + t] = h.index;
      for (int e = 0; e < 6; ++e) par[(size_t)e * ns + t] = (*v)[t].state.par[e];
    }
    std::fwrite(hot.data(), sizeof(int), hot.size(), f);
    std::fwrite(par.data(), sizeof(float), par.size(), f);
  }
  std::fclose(f);
  return 0;
}
