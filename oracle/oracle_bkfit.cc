// oracle_bkfit.cc -- CPU restatement of mkFit's backward fit of found candidates.
//
// TEST INFRASTRUCTURE ONLY (checker); the product library never links it.
// Restates, for one candidate per seed and a layer plan visited in reverse:
//   MkBuilder::BackwardFit / fit_cands         mkFit/MkBuilder.cc:2124-2253 (blocks of NN candidates)
//   MkBuilder::BackwardFitBH / fit_cands_BH    mkFit/MkBuilder.cc:2069-2122
//   MkFinder::BkFitInputTracks                 mkFit/MkFinder.cc:1368-1399  (covariance x100, chi2 0,
//                                              cursor on the candidate's last HitOnTrack node)
//   MkFinder::BkFitFitTracks                   mkFit/MkFinder.cc:1573-1742  (per plan layer: skip negative
//                                              nodes, take the layer's first node, else a fake hit
//                                              {666,0,666,0,0,666} at the track position without material;
//                                              stop when every lane is done, skip layers no lane has)
//   MkFinder::BkFitFitTracksBH                 mkFit/MkFinder.cc:1455-1570  (the same with material on the fakes)
//   SteeringParams iterator IT_BkwFit          mkFit/SteeringParams.h:138-151 (plan index size-1 .. m_bkw_fit_last)
//   MkBase::PropagateTracksToHit{R,Z}          mkFit/MkBase.h:48-89
//   MkFinder::BkFitPropTracksToPCA             mkFit/MkFinder.cc:1746-1749, MkBase.h:91-104
//   MkFinder::BkFitOutputTracks                mkFit/MkFinder.cc:1425-1446  (Par[iP] after the PCA, score)
//   TrackCand::addHitIdx counters, getScoreCand mkFit/HitStructures.h:574-587,873-889
// A "block" here is `group` consecutive candidates (the reference's NN-wide
// MkFinder); its lanes are processed NN at a time, which changes nothing since
// lanes only interact through the block's per-layer counts.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"
#include "oracle_core.h"

using namespace orc;

namespace {

// getScoreCand (HitStructures.h:574-585) over getScoreCalc (Track.h:648-672) with the Config.h:355-359
// constants, penalizeTailMissHits = inFindCandidates = false and no overlap hits.
float score_cand(int nfound, int ninside, float chi2, float pt) {
  if (chi2 < 0) chi2 = 0.f;
  const float maxBonus = 8.0;
  float bonus = 0.2f * nfound + 4.0f;
  float penalty = 8.0f;
  if (pt < 0.9) {
    penalty *= 1.5f;
    bonus = std::min(bonus * 1.0f, maxBonus);
  }
  return bonus * nfound + 0.0f * 0 - penalty * ninside - 3.0f * 0 - chi2;
}

struct Node {
  int index, layer, prev;
};

template <class T>
struct Tracks {
  int n, n_layers;
  const int* hot_idx;
  const int* layer_base;
  T* par;
  T* err;
  int* chg;
  T* chi2;
  T* score;
};

template <class T>
void bkfit_block(const Ctx& ctx0, const orc_find_event* ev, const orc_bkfit_params* p, const Tracks<T>& tr, int g0,
                 int gcnt, float* tie_out) {
  const int n = tr.n, L = tr.n_layers;
  const int nch = (gcnt + NN - 1) / NN;
  // BkFitInputTracks: the HitOnTrack node list of each candidate (one node per plan layer it reached, in plan
  // order) and the cursor on its last node.
  std::vector<std::vector<Node>> hots(gcnt);
  std::vector<int> cur(gcnt, -1), base(gcnt, 0);
  std::vector<char> valid(gcnt, 1);
  for (int i = 0; i < gcnt; ++i) {
    const int t = g0 + i;
    base[i] = tr.layer_base ? tr.layer_base[t] : 0;
    if (base[i] < 0 || base[i] + L > ev->n_layers) {
      valid[i] = 0;
      continue;
    }
    for (int l = 0; l < L; ++l) {
      const int h = tr.hot_idx[(size_t)l * n + t];
      if (h == -9) continue;
      const int lay = base[i] + l;
      const int nh = ev->hit_offset[lay + 1] - ev->hit_offset[lay];
      hots[i].push_back({h < nh ? h : -1, l, cur[i]});
      cur[i] = (int)hots[i].size() - 1;
    }
  }
  struct Chunk {
    alignas(64) T err[21 * NN], par[6 * NN], chi2[NN];
    alignas(64) T errP[21 * NN], parP[6 * NN], errU[21 * NN], parU[6 * NN], c2[NN];
    alignas(64) T errPE[21 * NN], parPE[6 * NN], errE[21 * NN], parE[6 * NN], c2E[NN];
    alignas(64) T msPar[3 * NN], msErr[6 * NN], tgt[NN];
    alignas(64) int chg[NN], nomat[NN], fail[NN], has[NN], ec[NN];
    alignas(64) float tie_r[NN], tie_dot[NN], tie_mat[NN];
  };
  std::vector<Chunk> ch(nch);
  for (int c = 0; c < nch; ++c) {
    Chunk& k = ch[c];
    const int c0 = c * NN, cnt = std::min(NN, gcnt - c0);
    for (int l = 0; l < NN; ++l) {
      const int t = g0 + c0 + (l < cnt ? l : 0);
      for (int e = 0; e < 6; ++e) k.par[e * NN + l] = tr.par[(size_t)e * n + t];
      for (int e = 0; e < 21; ++e) k.err[e * NN + l] = tr.err[(size_t)e * n + t] * T(100.0f);   // Scale(100.0f)
      k.chg[l] = tr.chg[t];
      k.chi2[l] = 0;
      k.tie_r[l] = k.tie_dot[l] = k.tie_mat[l] = 1e30f;
    }
  }
  const int kfop = KFO_Calculate_Chi2 | KFO_Update_Params;
  const T fake[6] = {666, 0, 666, 0, 0, 666};
  for (int l = L - 1; l >= p->last_layer; --l) {
    int done_count = 0, here_count = 0;
    for (int i = 0; i < gcnt; ++i) {
      Chunk& k = ch[i / NN];
      const int ln = i % NN;
      int& c = cur[i];
      const std::vector<Node>& H = hots[i];
      while (c >= 0 && H[c].index < 0) c = H[c].prev;
      if (c < 0) ++done_count;
      k.ec[ln] = valid[i] ? !ev->layers[base[i] + l].is_barrel : 0;
      if (c >= 0 && H[c].layer == l) {
        while (H[c].prev >= 0 && H[H[c].prev].layer == l) c = H[c].prev;   // overlap / merged seed hits
        const float* h = ev->hits + (size_t)(ev->hit_offset[base[i] + l] + H[c].index) * 9;
        for (int e = 0; e < 3; ++e) k.msPar[e * NN + ln] = T(h[e]);
        for (int e = 0; e < 6; ++e) k.msErr[e * NN + ln] = T(h[3 + e]);
        k.nomat[ln] = 0;
        k.has[ln] = 1;
        ++here_count;
        c = H[c].prev;
      } else {
        for (int e = 0; e < 3; ++e) k.msPar[e * NN + ln] = k.par[e * NN + ln];
        for (int e = 0; e < 6; ++e) k.msErr[e * NN + ln] = fake[e];
        k.nomat[ln] = p->fake_material ? 0 : 1;
        k.has[ln] = 0;
      }
    }
    if (done_count == gcnt) break;
    if (here_count == 0) continue;
    for (int c = 0; c < nch; ++c) {
      Chunk& k = ch[c];
      const int c0 = c * NN, cnt = std::min(NN, gcnt - c0);
      Ctx ctx = ctx0;
      if (tie_out) {
        ctx.tie_r = k.tie_r;
        ctx.tie_dot = k.tie_dot;
        ctx.tie_mat = k.tie_mat;
      }
      for (int ln = cnt; ln < NN; ++ln) {            // padding lanes: a copy of lane 0's inputs
        for (int e = 0; e < 3; ++e) k.msPar[e * NN + ln] = k.msPar[e * NN];
        for (int e = 0; e < 6; ++e) k.msErr[e * NN + ln] = k.msErr[e * NN];
        k.nomat[ln] = k.nomat[0];
        k.ec[ln] = k.ec[0];
      }
      bool any_b = false, any_e = false;
      for (int ln = 0; ln < cnt; ++ln) {
        if (!valid[c0 + ln]) continue;
        if (k.ec[ln]) any_e = true; else any_b = true;
      }
      if (any_b) {
        for (int ln = 0; ln < NN; ++ln) k.tgt[ln] = T(std::hypot(k.msPar[0 * NN + ln], k.msPar[1 * NN + ln]));
        propagate_helix_to_r(k.err, k.par, k.chg, k.tgt, k.errP, k.parP, cnt, p->pflags, k.nomat, k.fail, ctx);
        kalman_operation(kfop, k.errP, k.parP, k.msErr, k.msPar, k.errU, k.parU, k.c2);
      }
      if (any_e) {
        for (int ln = 0; ln < NN; ++ln) k.tgt[ln] = k.msPar[2 * NN + ln];
        propagate_helix_to_z(k.err, k.par, k.chg, k.tgt, k.errPE, k.parPE, cnt, p->pflags, k.nomat, ctx);
        kalman_operation_endcap(kfop, k.errPE, k.parPE, k.msErr, k.msPar, k.errE, k.parE, k.c2E);
      }
      for (int ln = 0; ln < cnt; ++ln) {
        if (!valid[c0 + ln]) continue;
        const bool e = k.ec[ln];
        for (int i = 0; i < 21; ++i) k.err[i * NN + ln] = (e ? k.errE : k.errU)[i * NN + ln];
        for (int i = 0; i < 6; ++i) k.par[i * NN + ln] = (e ? k.parE : k.parU)[i * NN + ln];
        if (k.par[3 * NN + ln] < 0) {                                    // MkFinder.cc:1671-1678
          k.chg[ln] = -k.chg[ln];
          k.par[3 * NN + ln] = -k.par[3 * NN + ln];
        }
        k.chi2[ln] += e ? k.c2E[ln] : k.c2[ln];                         // Chi2.Add(tmp_chi2)
      }
    }
  }
  for (int c = 0; c < nch; ++c) {
    Chunk& k = ch[c];
    const int c0 = c * NN, cnt = std::min(NN, gcnt - c0);
    if (p->include_pca) {                                               // MkBase.h:91-104
      Ctx ctx = ctx0;
      if (tie_out) {
        ctx.tie_r = k.tie_r;
        ctx.tie_dot = k.tie_dot;
        ctx.tie_mat = k.tie_mat;
      }
      for (int ln = 0; ln < NN; ++ln) {
        const T slope = std::tan(k.par[5 * NN + ln]);
        k.tgt[ln] = (slope * (slope * k.par[2 * NN + ln] - std::hypot(k.par[0 * NN + ln], k.par[1 * NN + ln]))) /
                    (1 + slope * slope);
      }
      propagate_helix_to_z(k.err, k.par, k.chg, k.tgt, k.errP, k.parP, cnt, p->pca_pflags, nullptr, ctx);
      std::memcpy(k.err, k.errP, sizeof(k.err));
      std::memcpy(k.par, k.parP, sizeof(k.par));
    }
    for (int ln = 0; ln < cnt; ++ln) {
      const int i = c0 + ln, t = g0 + i;
      if (!valid[i]) continue;
      for (int e = 0; e < 6; ++e) tr.par[(size_t)e * n + t] = k.par[e * NN + ln];
      for (int e = 0; e < 21; ++e) tr.err[(size_t)e * n + t] = k.err[e * NN + ln];
      tr.chg[t] = k.chg[ln];
      tr.chi2[t] = k.chi2[ln];
      if (tr.score && !std::isnan((float)k.chi2[ln])) {
        // TrackCand counters from its nodes in the order they were added (addHitIdx)
        int nfound = 0, ninside = 0, ntail = 0;
        for (int l = 0; l < L; ++l) {
          const int h = tr.hot_idx[(size_t)l * n + t];
          if (h == -9) continue;
          if (h >= 0) {
            ++nfound;
            ninside += ntail;
            ntail = 0;
          } else if (h == -1) {
            ++ntail;
          }
        }
        const float pt = std::abs(1.f / (float)k.par[3 * NN + ln]);
        tr.score[t] = T(score_cand(nfound, ninside, (float)k.chi2[ln], pt));
      }
      if (tie_out) {
        tie_out[t] = k.tie_r[ln];
        tie_out[n + t] = k.tie_dot[ln];
        tie_out[2 * (size_t)n + t] = k.tie_mat[ln];
      }
    }
  }
}

template <class T>
int bkfit(const orc_ctx* c, const orc_find_event* ev, const orc_bkfit_params* p, const Tracks<T>& tr, int n_threads,
          float* tie_out, float tie_shift) {
  if (!ev || !p || tr.n < 0 || tr.n_layers < 0 || p->group < 1 || p->last_layer < 0) return 1;
  Ctx ctx;
  ctx.bfield = c ? c->bfield : 3.8112f;
  ctx.rl_grid = c ? c->rl_grid : nullptr;
  ctx.xi_grid = c ? c->xi_grid : nullptr;
  if (tie_shift != 0.f) ctx.tie_shift = tie_shift;
  const int G = p->group, ng = (tr.n + G - 1) / G;
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      const int g = next.fetch_add(1);
      if (g >= ng) break;
      bkfit_block<T>(ctx, ev, p, tr, g * G, std::min(G, tr.n - g * G), tie_out);
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < std::max(1, n_threads); ++i) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  return 0;
}

}  // namespace

extern "C" {

int orc_backward_fit_f32(const orc_ctx* c, const orc_find_event* ev, const orc_bkfit_params* p,
                         const orc_bkfit_tracks_f32* t, int n_threads, float* ties, float tie_shift) {
  if (!t) return 1;
  Tracks<float> tr{t->n, t->n_layers, t->hot_idx, t->layer_base, t->par, t->err, t->chg, t->chi2, t->score};
  return bkfit<float>(c, ev, p, tr, n_threads, ties, tie_shift);
}

int orc_backward_fit_f64(const orc_ctx* c, const orc_find_event* ev, const orc_bkfit_params* p,
                         const orc_bkfit_tracks_f64* t, int n_threads) {
  if (!t) return 1;
  Tracks<double> tr{t->n, t->n_layers, t->hot_idx, t->layer_base, t->par, t->err, t->chg, t->chi2, t->score};
  return bkfit<double>(c, ev, p, tr, n_threads, nullptr, 0.f);
}

}  // extern "C"
