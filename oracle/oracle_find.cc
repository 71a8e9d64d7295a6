// oracle_find.cc -- CPU restatement of mkFit's best-hit track finding.
//
// TEST INFRASTRUCTURE ONLY (checker and CPU baseline).  Restates:
//   LayerOfHits::setup_bins / SuckInHits     mkFit/HitStructures.cc:11-48, 77-183
//     (bin geometry, key = phi_fine + (q_bin << 16), stable sort by key, per
//     (q, phi) bin [first, end) ranges; 32-bit ranges here, the reference's are
//     uint16 (HitStructures.h:45) and agree wherever a layer holds < 65536 hits)
//   bin helpers                               mkFit/HitStructures.h:237-255
//   LayerInfo sensitive-region tests          TrackerInfo.h:118-138
//   MkFinder::getHitSelDynamicWindows/Chi2Cut mkFit/MkFinder.cc:166-220
//   MkFinder::SelectHitIndices                mkFit/MkFinder.cc:227-684
//   MkFinder::AddBestHit                      mkFit/MkFinder.cc:691-832, add_hit MkFinder.h:312-360
//   MkBuilder::FindTracksBestHit layer loop   mkFit/MkBuilder.cc:1395-1478 (barrel plan)
// Propagation and Kalman math come from oracle_core.h.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <thread>
#include <vector>

#include "oracle.h"
#include "oracle_core.h"
#include "oracle_find_core.h"

using namespace orc;
using namespace orc_find;

namespace {

struct FindBlock {
  float par[6 * NN], err[21 * NN], chi2[NN];
  int chg[NN];
  int nhits[NN], nfound[NN], ninside[NN], ntail[NN], stopped[NN];
  int hot_idx[kMaxTrkHits][NN], hot_lyr[kMaxTrkHits][NN];
};

// add_hit (MkFinder.h:312-354)
void add_hit(FindBlock& B, int it, int index, int layer) {
  int& nt = B.nhits[it];
  int& nf = B.nfound[it];
  if (nt < kMaxTrkHits) {
    B.hot_idx[nt][it] = index;
    B.hot_lyr[nt][it] = layer;
    ++nt;
    if (index >= 0) {
      ++nf;
      B.ninside[it] += B.ntail[it];
      B.ntail[it] = 0;
    } else if (index == -1) {
      ++B.ntail[it];
    }
  } else {
    const int LH = kMaxTrkHits - 1;
    if (index >= 0) {
      if (B.hot_idx[LH][it] < 0) ++nf;
      B.hot_idx[LH][it] = index;
      B.hot_lyr[LH][it] = layer;
    } else if (index == -2) {
      if (B.hot_idx[LH][it] >= 0) --nf;
      B.hot_idx[LH][it] = index;
      B.hot_lyr[LH][it] = layer;
    }
  }
}

struct EventLayers {
  int n_layers;
  const orc_layer_desc* desc;
  std::vector<LoH> loh;
  const int* hit_off;
  const float* hits;  // [total][9]
  int prop_to_hit = 0, intra_pflags = 0;
  Ctx ctx;
};

// AddBestHit (MkFinder.cc:691-832) for one block, propagated state in parP/errP.  With `c2m` (checker only),
// also each lane's chi2 decision margin: the smallest relative gap between the best hit's chi2 and another
// window hit's, or between any window hit's chi2 and the cut.
void add_best_hit(const EventLayers& E, int layer, float chi2_cut_min, FindBlock& B, const float* parP,
                  const float* errP, const int* xsize, const int* xarr, const int* wsr, int n_proc,
                  float* c2m = nullptr) {
  float c2all[kHitIdxMax][NN];
  // With prop_to_hit (Config::finding_requires_propagation_to_hit_pos) the chi2 and the update are
  // taken after propagating the layer state to each hit's radius (barrel) or z (endcap) with the
  // intra-layer flags: kalmanPropagateAndComputeChi2 / kalmanPropagateAndUpdate
  // (KalmanUtilsMPlex.cc:471-543, 697-770).
  auto propagate_to_hits = [&](const float* msPar, float* eOut, float* pOut) {
    alignas(64) float ms[NN];
    if (E.desc[layer].is_barrel) {
      for (int l = 0; l < NN; ++l) ms[l] = std::hypot(msPar[0 * NN + l], msPar[1 * NN + l]);
      propagate_helix_to_r(errP, parP, B.chg, ms, eOut, pOut, n_proc, E.intra_pflags, nullptr, nullptr, E.ctx);
    } else {
      for (int l = 0; l < NN; ++l) ms[l] = msPar[2 * NN + l];
      propagate_helix_to_z(errP, parP, B.chg, ms, eOut, pOut, n_proc, E.intra_pflags, nullptr, E.ctx);
    }
  };
  alignas(64) float errH[21 * NN], parH[6 * NN];
  const orc_layer_desc& d = E.desc[layer];
  const float* hits = E.hits + (size_t)E.hit_off[layer] * 9;
  const int n_hits = E.hit_off[layer + 1] - E.hit_off[layer];
  float minChi2[NN];
  int bestHit[NN];
  int maxSize = 0;
  for (int it = 0; it < NN; ++it) {
    if (it < n_proc && xsize[it] > 0) maxSize = std::max(maxSize, std::min(xsize[it], kHitIdxMax));
    bestHit[it] = -1;
    minChi2[it] = dyn_chi2_cut(d, chi2_cut_min, parP[3 * NN + it], parP[5 * NN + it]);
  }
  alignas(64) float msPar[3 * NN], msErr[6 * NN], outChi2[NN];
  alignas(64) float dErr[21 * NN], dPar[6 * NN];
  for (int l = 0; l < 3 * NN; ++l) msPar[l] = 0.f;
  for (int l = 0; l < 6 * NN; ++l) msErr[l] = 1.f;
  // XHitArr entries outside [0, n_hits) are skipped.
  auto valid = [&](int hc, int it) {
    return hc < xsize[it] && xarr[hc * NN + it] >= 0 && xarr[hc * NN + it] < n_hits;
  };
  for (int hc = 0; hc < maxSize; ++hc) {
    for (int it = 0; it < n_proc; ++it)
      if (valid(hc, it)) {
        const float* h = hits + (size_t)xarr[hc * NN + it] * 9;
        for (int e = 0; e < 3; ++e) msPar[e * NN + it] = h[e];
        for (int e = 0; e < 6; ++e) msErr[e * NN + it] = h[3 + e];
      }
    const float* eC = errP;
    const float* pC = parP;
    if (E.prop_to_hit) {
      propagate_to_hits(msPar, errH, parH);
      eC = errH;
      pC = parH;
    }
    if (d.is_barrel)
      kalman_operation(KFO_Calculate_Chi2, eC, pC, msErr, msPar, dErr, dPar, outChi2);
    else
      kalman_operation_endcap(KFO_Calculate_Chi2, eC, pC, msErr, msPar, dErr, dPar, outChi2);
    for (int it = 0; it < n_proc; ++it)
      if (valid(hc, it)) {
        const float c2 = std::abs(outChi2[it]);
        c2all[hc][it] = c2;
        if (c2 < minChi2[it]) {
          minChi2[it] = c2;
          bestHit[it] = xarr[hc * NN + it];
        }
      }
  }
  if (c2m)
    for (int it = 0; it < n_proc; ++it) {
      const float cut = dyn_chi2_cut(d, chi2_cut_min, parP[3 * NN + it], parP[5 * NN + it]);
      float m = 1e30f;
      for (int hc = 0; hc < std::min(xsize[it], kHitIdxMax); ++hc) {
        if (!valid(hc, it)) continue;
        const float c2 = c2all[hc][it];
        m = std::min(m, std::abs(c2 - cut) / std::max(cut, 1.f));
        if (bestHit[it] >= 0 && xarr[hc * NN + it] != bestHit[it])
          m = std::min(m, std::abs(c2 - minChi2[it]) / std::max(minChi2[it], 1.f));
      }
      c2m[it] = m;
    }
  for (int it = 0; it < n_proc; ++it) {
    auto fake = [&]() {
      const float f[6] = {666, 0, 666, 0, 0, 666};
      for (int e = 0; e < 6; ++e) msErr[e * NN + it] = f[e];
      for (int e = 0; e < 3; ++e) msPar[e * NN + it] = parP[e * NN + it];
    };
    if (wsr[it] == WSR_Outside) {
      fake();
      continue;
    }
    if (bestHit[it] >= 0) {
      const float* h = hits + (size_t)bestHit[it] * 9;
      for (int e = 0; e < 3; ++e) msPar[e * NN + it] = h[e];
      for (int e = 0; e < 6; ++e) msErr[e * NN + it] = h[3 + e];
      B.chi2[it] += minChi2[it];
      add_hit(B, it, bestHit[it], layer);
    } else {
      int fake_idx = -1;
      if (wsr[it] == WSR_Edge)
        fake_idx = -3;
      else if (B.ninside[it] + B.ntail[it])
        fake_idx = -2;
      fake();
      add_hit(B, it, fake_idx, layer);
    }
  }
  // kalmanPropagateAndUpdate
  const float* eU = errP;
  const float* pU = parP;
  if (E.prop_to_hit) {
    propagate_to_hits(msPar, errH, parH);
    eU = errH;
    pU = parH;
  }
  if (d.is_barrel)
    kalman_operation(KFO_Update_Params, eU, pU, msErr, msPar, B.err, B.par, outChi2);
  else
    kalman_operation_endcap(KFO_Update_Params, eU, pU, msErr, msPar, B.err, B.par, outChi2);
  for (int it = 0; it < NN; ++it)
    if (B.par[3 * NN + it] < 0) {
      B.chg[it] = -B.chg[it];
      B.par[3 * NN + it] = -B.par[3 * NN + it];
    }
}

}  // namespace

extern "C" {

int orc_loh_geometry(const orc_layer_desc* d, int* nq, float* qmin, float* qmax, float* fq) {
  if (!d) return 1;
  const Bins b = layer_bins(*d);
  if (nq) *nq = b.nq;
  if (qmin) *qmin = b.qmin;
  if (qmax) *qmax = b.qmax;
  if (fq) *fq = b.fq;
  return 0;
}

int orc_loh_build(const orc_layer_desc* d, int n, const float* hits, int* ranks, int* bins, float* phis, float* qs) {
  if (!d || n < 0) return 1;
  LoH L;
  build_loh(*d, n, hits, L);
  for (int i = 0; i < n; ++i) {
    if (ranks) ranks[i] = L.ranks[i];
    if (phis) phis[i] = L.phis[i];
    if (qs) qs[i] = L.qs[i];
  }
  if (bins)
    for (size_t b = 0; b < L.first.size(); ++b) {
      bins[2 * b] = L.first[b];
      bins[2 * b + 1] = L.end[b];
    }
  return 0;
}

int orc_select_hit_indices(const orc_layer_desc* d, int n_hits, const float* hits, int n, const float* par,
                           const float* err, int* xsize, int* xarr, int* wsr) {
  return orc_select_hit_indices_masked(d, n_hits, hits, nullptr, n, par, err, xsize, xarr, wsr);
}

int orc_select_hit_indices_masked(const orc_layer_desc* d, int n_hits, const float* hits, const unsigned char* mask,
                                  int n, const float* par, const float* err, int* xsize, int* xarr, int* wsr) {
  if (!d || n < 0) return 1;
  LoH L;
  build_loh(*d, n_hits, hits, L);
  L.mask = mask;
  for (int t0 = 0; t0 < n; t0 += NN) {
    const int cnt = std::min(NN, n - t0);
    alignas(64) float p[6 * NN], e[21 * NN];
    int xs[NN], xa[kHitIdxMax * NN], w[NN];
    for (int k = 0; k < 6; ++k)
      for (int l = 0; l < NN; ++l) p[k * NN + l] = par[(size_t)k * n + t0 + (l < cnt ? l : 0)];
    for (int k = 0; k < 21; ++k)
      for (int l = 0; l < NN; ++l) e[k * NN + l] = err[(size_t)k * n + t0 + (l < cnt ? l : 0)];
    select_hit_indices(*d, L, p, e, cnt, xs, xa, w);
    for (int l = 0; l < cnt; ++l) {
      xsize[t0 + l] = xs[l];
      wsr[t0 + l] = w[l];
      for (int k = 0; k < kHitIdxMax; ++k) xarr[(size_t)k * n + t0 + l] = k < xs[l] ? xa[k * NN + l] : -1;
    }
  }
  return 0;
}

int orc_find_best_hit(const orc_ctx* c, const orc_find_event* ev, orc_find_tracks* tr, int n_threads) {
  if (!ev || !tr || ev->n_layers <= 0 || ev->first_layer < 0 || ev->first_layer > ev->n_layers) return 1;
  Ctx ctx;
  ctx.bfield = c ? c->bfield : 3.8112f;
  ctx.rl_grid = c ? c->rl_grid : nullptr;
  ctx.xi_grid = c ? c->xi_grid : nullptr;
  EventLayers E;
  E.n_layers = ev->n_layers;
  E.desc = ev->layers;
  E.hit_off = ev->hit_offset;
  E.hits = ev->hits;
  E.prop_to_hit = ev->prop_to_hit;
  E.intra_pflags = ev->intra_pflags;
  E.ctx = ctx;
  E.loh.resize(ev->n_layers);
  for (int l = 0; l < ev->n_layers; ++l) {
    build_loh(ev->layers[l], ev->hit_offset[l + 1] - ev->hit_offset[l], ev->hits + (size_t)ev->hit_offset[l] * 9,
              E.loh[l]);
    if (ev->hit_mask) E.loh[l].mask = ev->hit_mask + ev->hit_offset[l];
  }
  const int n = tr->n_tracks;
  const int nblk = (n + NN - 1) / NN;
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      const int blk = next.fetch_add(1);
      if (blk >= nblk) break;
      const int t0 = blk * NN, cnt = std::min(NN, n - t0);
      FindBlock B;
      FindBlock Out;  // snapshot of stopped lanes (OutputTrackAndHitIdx at stop time)
      auto lane = [&](int l) { return t0 + (l < cnt ? l : 0); };
      for (int k = 0; k < 6; ++k)
        for (int l = 0; l < NN; ++l) B.par[k * NN + l] = tr->par[(size_t)k * n + lane(l)];
      for (int k = 0; k < 21; ++k)
        for (int l = 0; l < NN; ++l) B.err[k * NN + l] = tr->err[(size_t)k * n + lane(l)];
      for (int l = 0; l < NN; ++l) {
        B.chg[l] = tr->chg[lane(l)];
        B.chi2[l] = tr->chi2[lane(l)];
        B.nhits[l] = 0;
        B.nfound[l] = 0;
        B.ninside[l] = 0;
        B.ntail[l] = 0;
        B.stopped[l] = 0;
        for (int k = 0; k < ev->first_layer && k < kMaxTrkHits; ++k) {  // seed hits
          B.hot_idx[B.nhits[l]][l] = tr->hot_idx[(size_t)k * n + lane(l)];
          B.hot_lyr[B.nhits[l]][l] = k;
          B.nhits[l]++;
          if (B.hot_idx[B.nhits[l] - 1][l] >= 0) B.nfound[l]++;
        }
      }
      alignas(64) float parP[6 * NN], errP[21 * NN], rad[NN];
      int xs[NN], xa[kHitIdxMax * NN], wsr[NN], fail[NN];
      int wk[3][NN] = {};                                         // checker: work units per lane
      for (int layer = ev->first_layer; layer < ev->n_layers; ++layer) {
        const int n_layer_hits = ev->hit_offset[layer + 1] - ev->hit_offset[layer];
        for (int i = 0; i < cnt; ++i) wk[0][i] += !B.stopped[i];
        const orc_layer_desc& d = ev->layers[layer];
        for (int l = 0; l < NN; ++l) rad[l] = d.propagate_to;
        if (d.is_barrel)
          propagate_helix_to_r(B.err, B.par, B.chg, rad, errP, parP, cnt, ev->pflags, nullptr, fail, ctx);
        else
          propagate_helix_to_z(B.err, B.par, B.chg, rad, errP, parP, cnt, ev->pflags, nullptr, ctx);
        SelMargins mg[NN];
        select_hit_indices(d, E.loh[layer], parP, errP, cnt, xs, xa, wsr, tr->margins ? mg : nullptr);
        if (d.is_barrel) {                                        // MkBuilder.cc:1431-1455
          const float r_min_sqr = d.rin * d.rin;
          for (int i = 0; i < cnt; ++i) {
            if (!B.stopped[i]) {
              const float x = parP[0 * NN + i], y = parP[1 * NN + i];
              if (x * x + y * y < r_min_sqr) {
                B.stopped[i] = 1;
                for (int k = 0; k < 6; ++k) Out.par[k * NN + i] = B.par[k * NN + i];
                for (int k = 0; k < 21; ++k) Out.err[k * NN + i] = B.err[k * NN + i];
                Out.chg[i] = B.chg[i]; Out.chi2[i] = B.chi2[i];
                Out.nhits[i] = B.nhits[i]; Out.nfound[i] = B.nfound[i];
                for (int k = 0; k < B.nhits[i]; ++k) { Out.hot_idx[k][i] = B.hot_idx[k][i]; Out.hot_lyr[k][i] = B.hot_lyr[k][i]; }
                wsr[i] = WSR_Outside;
                xs[i] = 0;
              }
            } else {
              wsr[i] = WSR_Outside;
              xs[i] = 0;
            }
          }
        }
        for (int i = 0; i < cnt; ++i) {
          if (B.stopped[i] || wsr[i] == WSR_Outside) continue;
          for (int hc = 0; hc < std::min(xs[i], kHitIdxMax); ++hc)
            wk[1][i] += xa[hc * NN + i] >= 0 && xa[hc * NN + i] < n_layer_hits;
          ++wk[2][i];
        }
        float c2m[NN];
        add_best_hit(E, layer, ev->chi2_cut, B, parP, errP, xs, xa, wsr, cnt, tr->margins ? c2m : nullptr);
        if (tr->margins)
          for (int i = 0; i < cnt; ++i) {
            const size_t L = ev->n_layers, t = (size_t)t0 + i;
            const bool live = wsr[i] != WSR_Outside;
            tr->margins[(0 * L + layer) * n + t] = live ? mg[i].win : 1e30f;
            tr->margins[(1 * L + layer) * n + t] = live ? mg[i].bin : 1e30f;
            tr->margins[(2 * L + layer) * n + t] = live ? c2m[i] : 1e30f;
          }
        for (int i = 0; i < cnt; ++i)                             // MkBuilder.cc:1465-1473
          if (!B.stopped[i] && B.nhits[i] > 0 && B.hot_idx[std::min(B.nhits[i], kMaxTrkHits) - 1][i] == -2) {
            B.stopped[i] = 1;
            for (int k = 0; k < 6; ++k) Out.par[k * NN + i] = B.par[k * NN + i];
            for (int k = 0; k < 21; ++k) Out.err[k * NN + i] = B.err[k * NN + i];
            Out.chg[i] = B.chg[i]; Out.chi2[i] = B.chi2[i];
            Out.nhits[i] = B.nhits[i]; Out.nfound[i] = B.nfound[i];
            for (int k = 0; k < B.nhits[i]; ++k) { Out.hot_idx[k][i] = B.hot_idx[k][i]; Out.hot_lyr[k][i] = B.hot_lyr[k][i]; }
          }
      }
      for (int i = 0; i < cnt; ++i) {
        const FindBlock& S = B.stopped[i] ? Out : B;
        const size_t t = (size_t)t0 + i;
        for (int k = 0; k < 6; ++k) tr->par[k * (size_t)n + t] = S.par[k * NN + i];
        for (int k = 0; k < 21; ++k) tr->err[k * (size_t)n + t] = S.err[k * NN + i];
        tr->chg[t] = S.chg[i];
        tr->chi2[t] = S.chi2[i];
        tr->n_found[t] = S.nfound[i];
        if (tr->work)
          for (int k = 0; k < 3; ++k) tr->work[(size_t)k * n + t] = wk[k][i];
        for (int k = 0; k < ev->n_layers; ++k) tr->hot_idx[(size_t)k * n + t] = -9;
        for (int k = 0; k < S.nhits[i] && k < kMaxTrkHits; ++k) {
          const int ly = S.hot_lyr[k][i];
          if (ly >= 0 && ly < ev->n_layers) tr->hot_idx[(size_t)ly * n + t] = S.hot_idx[k][i];
        }
      }
    }
  };
  n_threads = std::max(1, n_threads);
  std::vector<std::thread> pool;
  for (int i = 1; i < n_threads; ++i) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  return 0;
}

int orc_add_best_hit(const orc_ctx* c, const orc_find_event* ev, int layer, const orc_finder_lanes* ln,
                     const int* xsize, const int* xarr, const int* wsr, int* hit_code) {
  if (!ev || !ln || layer < 0 || layer >= ev->n_layers || ln->n < 0) return 1;
  EventLayers E;
  E.n_layers = ev->n_layers;
  E.desc = ev->layers;
  E.hit_off = ev->hit_offset;
  E.hits = ev->hits;
  E.prop_to_hit = ev->prop_to_hit;
  E.intra_pflags = ev->intra_pflags;
  E.ctx.bfield = c ? c->bfield : 3.8112f;
  E.ctx.rl_grid = c ? c->rl_grid : nullptr;
  E.ctx.xi_grid = c ? c->xi_grid : nullptr;
  const int n = ln->n;
  for (int t0 = 0; t0 < n; t0 += NN) {
    const int cnt = std::min(NN, n - t0);
    auto lane = [&](int l) { return (size_t)t0 + (l < cnt ? l : 0); };
    FindBlock B;
    alignas(64) float parP[6 * NN], errP[21 * NN];
    int xs[NN], xa[kHitIdxMax * NN], w[NN];
    for (int k = 0; k < 6; ++k)
      for (int l = 0; l < NN; ++l) parP[k * NN + l] = ln->prop_par[k * (size_t)n + lane(l)];
    for (int k = 0; k < 21; ++k)
      for (int l = 0; l < NN; ++l) errP[k * NN + l] = ln->prop_err[k * (size_t)n + lane(l)];
    for (int k = 0; k < kHitIdxMax; ++k)
      for (int l = 0; l < NN; ++l) xa[k * NN + l] = xarr[k * (size_t)n + lane(l)];
    for (int l = 0; l < NN; ++l) {
      const size_t t = lane(l);
      B.chg[l] = ln->chg[t];
      B.chi2[l] = ln->chi2[t];
      B.nhits[l] = 0;
      B.nfound[l] = ln->n_found[t];
      B.ninside[l] = ln->n_inside_minus_one[t];
      B.ntail[l] = ln->n_tail_minus_one[t];
      B.stopped[l] = 0;
      xs[l] = xsize[t];
      w[l] = wsr[t];
    }
    add_best_hit(E, layer, ev->chi2_cut, B, parP, errP, xs, xa, w, cnt);
    for (int l = 0; l < cnt; ++l) {
      const size_t t = (size_t)t0 + l;
      for (int k = 0; k < 6; ++k) ln->par[k * (size_t)n + t] = B.par[k * NN + l];
      for (int k = 0; k < 21; ++k) ln->err[k * (size_t)n + t] = B.err[k * NN + l];
      ln->chg[t] = B.chg[l];
      ln->chi2[t] = B.chi2[l];
      ln->n_found[t] = B.nfound[l];
      ln->n_inside_minus_one[t] = B.ninside[l];
      ln->n_tail_minus_one[t] = B.ntail[l];
      hit_code[t] = B.nhits[l] > 0 ? B.hot_idx[0][l] : -9;
    }
  }
  return 0;
}

}  // extern "C"
