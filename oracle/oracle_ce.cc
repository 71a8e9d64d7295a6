// oracle_ce.cc -- CPU restatement of mkFit's clone-engine track finding.
//
// TEST INFRASTRUCTURE ONLY (checker and CPU baseline).  Restates, for a barrel
// layer plan visited in order:
//   MkBuilder::find_tracks_in_layers (CE)       mkFit/MkBuilder.cc:1887-2060
//   MkBuilder::find_tracks_unroll_candidates    mkFit/MkBuilder.cc:1503-1566
//   MkBuilder::find_tracks_handle_missed_layers mkFit/MkBuilder.cc:1568-1630
//   MkFinder::FindCandidatesCloneEngine         mkFit/MkFinder.cc:1114-1284
//   isStripQCompatible                          mkFit/MkFinder.cc:837-862
//   passStripChargePCMfromTrack                 mkFit/MkFinder.cc:863-896
//   CandCloner::ProcessSeedRange                mkFit/CandCloner.cc:37-191
//   MkFinder::UpdateWithLastHit                 mkFit/MkFinder.cc:1290-1308
//   kalmanPropagateAnd{ComputeChi2,Update}      mkFit/KalmanUtilsMPlex.cc:471-543
//   TrackCand::addHitIdx, HoT nodes             mkFit/HitStructures.h:873-889, CombCandidate::AddHit
//   CombCandidate::ImportSeed / MergeCandsAndBestShortOne  mkFit/HitStructures.cc:485-552
//   getScoreCalc / getScoreStruct / getScoreCand Track.h:648-697, HitStructures.h:589-601
//
// What the toy events do not exercise, and how it is handled here:
//   * hit cluster charge: toy hits carry none; without orc_find_event's
//     hit_charge they are taken as the reference takes pixel-like clusters
//     (charge code 0xff = Hit::maxChargePerCM), for which
//     passStripChargePCMfromTrack returns true (MkFinder.cc:872-873); with it,
//     the cut applies to every layer (the toy layers are all strip layers);
//   * overlap hits: every toy hit of a layer has module id 0, and with a single
//     module id HitMatchPair::find_overlap (HitStructures.h:372-394) never
//     returns a match, so no overlap hit is ever added;
//   * pixel layers: the toy geometry has none (dead-module gaps, -7, are modelled).
// Both sorts of the reference are std::sort (not stable); this restatement uses
// stable sorts, i.e. ties in score keep generation order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"
#include "oracle_core.h"
#include "oracle_find_core.h"

using namespace orc;
using namespace orc_find;

namespace {

// Config.h:355-359
constexpr float kValidHitBonus = 4, kValidHitSlope = 0.2f, kOverlapHitBonus = 0, kMissingHitPenalty = 8,
                kTailMissingHitPenalty = 3;
constexpr float kScoreWorst = -1e16f;

float score_calc(int nfound, int ntail, int noverlap, int nmiss, float chi2, float pt, bool in_find) {
  const float maxBonus = 8.0;
  float bonus = kValidHitSlope * nfound + kValidHitBonus;
  float penalty = kMissingHitPenalty;
  const float tailPenalty = kTailMissingHitPenalty;
  const float overlapBonus = kOverlapHitBonus;
  if (pt < 0.9) {
    penalty *= in_find ? 1.7f : 1.5f;
    bonus = std::min(bonus * (in_find ? 0.9f : 1.0f), maxBonus);
  }
  return bonus * nfound + overlapBonus * noverlap - penalty * nmiss - tailPenalty * ntail - chi2;
}

struct HoTNode {
  int index, layer;
  float chi2;
  int prev;
};

struct Cand {
  float par[6], err[21];
  int chg = 0;
  float chi2 = 0, score = kScoreWorst;
  int last = -1;                        // lastHitIdx_: node index in the seed's HoT store
  int nfound = 0, nmissing = 0, noverlap = 0, ninside = 0, ntail = 0;
  bool valid = false;                   // combCandidate() != nullptr (best-short bookkeeping)
  float pT() const { return std::abs(1.f / par[3]); }
};

struct IdxChi2 {                        // IdxChi2List (Track.h)
  int trkIdx, hitIdx, nhits, ntailholes, noverlaps, nholes;
  float pt, chi2, chi2_hit, score;
  int xpos;                             // position in the lane's XHitArr, -1 for the invalid-hit record
};

float score_struct(const IdxChi2& e) {
  float c = e.chi2;
  if (c < 0) c = 0.f;
  return score_calc(e.nhits, e.ntailholes, e.noverlaps, e.nholes, c, e.pt, true);
}
float score_cand(const Cand& c) {
  float chi2 = c.chi2;
  if (chi2 < 0) chi2 = 0.f;
  return score_calc(c.nfound, 0, c.noverlap, c.ninside, chi2, c.pT(), false);
}

struct Comb {
  std::vector<Cand> cands;
  Cand best_short;
  int state = 0;                        // 0 Dormant, 1 Finding, 2 Finished
  int pickup_layer = -1;
  std::vector<HoTNode> hots;

  void add_hit(Cand& c, int idx, int lyr, float chi2) {    // TrackCand::addHitIdx
    hots.push_back({idx, lyr, chi2, c.last});
    c.last = (int)hots.size() - 1;
    if (idx >= 0 || idx == -9) {
      ++c.nfound;
      c.chi2 += chi2;
      c.ninside += c.ntail;
      c.ntail = 0;
    } else {
      ++c.nmissing;
      if (idx == -1) ++c.ntail;
    }
  }
  int last_idx(const Cand& c) const { return c.last >= 0 ? hots[c.last].index : -1; }
};

// isStripQCompatible (MkFinder.cc:837-862), one lane.
bool strip_q_compatible(bool barrel, const float* pErr, const float* pPar, const float* mErr, const float* mPar,
                        int it) {
  auto E = [&](int i, int j) { return pErr[sym(i, j) * NN + it]; };
  auto P = [&](int i) { return pPar[i * NN + it]; };
  auto ME = [&](int i, int j) { return mErr[sym(i, j) * NN + it]; };
  auto MP = [&](int i) { return mPar[i * NN + it]; };
  if (barrel) {
    const float res = std::abs(MP(2) - P(2));
    const float hitHL = std::sqrt(ME(2, 2) * 3.f);
    const float qErr = std::sqrt(E(2, 2));
    return hitHL + std::max(3.f * qErr, 0.5f) > res;
  }
  const float res[2] = {MP(0) - P(0), MP(1) - P(1)};
  const float hitT2 = ME(0, 0) + ME(1, 1);
  const float hitT2inv = 1.f / hitT2;
  const float proj[3] = {ME(0, 0) * hitT2inv, ME(0, 1) * hitT2inv, ME(1, 1) * hitT2inv};
  const float qErr = std::sqrt(std::abs(E(0, 0) * proj[0] + 2.f * E(0, 1) * proj[1] + E(1, 1) * proj[2]));
  const float resProj = std::sqrt(res[0] * proj[0] * res[0] + 2.f * res[1] * proj[1] * res[0] + res[1] * proj[2] * res[1]);
  return std::sqrt(hitT2 * 3.f) + std::max(3.f * qErr, 0.5f) > resProj;
}

// Hit::PackedData::get_charge_pcm (Hit.h:270-274): the 8-bit code's charge per cm.
inline unsigned charge_pcm_of(unsigned char code) { return code == 0 ? 0u : ((code - 1u) << 3) + 1620u; }
constexpr unsigned kMinChargePerCM = 1620, kMaxChargePerCM = 1620 + (0xfe << 3);   // Hit.h:281-282

// passStripChargePCMfromTrack (MkFinder.cc:867-896), one lane: the hit's charge per cm times the
// track's path-length factor at the hit (pPar: the state propagated to the hit) must exceed
// Hit::minChargePerCM; the overflow code passes.  margin, when given: the relative gap of the
// corrected charge to the cut, for the checker; qcorr, when given: the corrected charge (for the
// overflow code, its charge per cm).
bool pass_strip_charge(bool barrel, unsigned pcm, const float* pPar, const float* mErr, int it,
                       float* margin = nullptr, float* qcorr = nullptr) {
  if (qcorr) *qcorr = (float)pcm;
  if (pcm >= kMaxChargePerCM) return true;
  auto P = [&](int i) { return pPar[i * NN + it]; };
  auto ME = [&](int i, int j) { return mErr[sym(i, j) * NN + it]; };
  float qSF;
  if (barrel) {
    const float hitT2 = ME(0, 0) + ME(1, 1);
    const float hitT2inv = 1.f / hitT2;
    const float proj[3] = {ME(0, 0) * hitT2inv, ME(0, 1) * hitT2inv, ME(1, 1) * hitT2inv};
    const bool detXY_OK = std::abs(proj[0] * proj[2] - proj[1] * proj[1]) < 0.1f;
    const float cosP = std::cos(P(4)), sinP = std::sin(P(4));
    const float sinT = std::abs(std::sin(P(5)));
    qSF = detXY_OK ? sinT * std::sqrt(std::abs(1.f + cosP * cosP * proj[0] + sinP * sinP * proj[2] -
                                               2.f * cosP * sinP * proj[1]))
                   : 1.f;
  } else {
    qSF = std::abs(std::cos(P(5)));
  }
  const float qCorr = pcm * qSF;
  if (qcorr) *qcorr = qCorr;
  if (margin) *margin = std::min(*margin, std::abs(qCorr - (float)kMinChargePerCM) / (float)kMinChargePerCM);
  return qCorr > kMinChargePerCM;
}

struct CEEvent {
  const orc_find_event* ev;
  const orc_ce_params* p;
  std::vector<LoH> loh;
  Ctx ctx;
};

// Block of up to NN candidates of one seed (MkFinder's NN lanes).
struct Block {
  alignas(64) float par[6 * NN], err[21 * NN], parP[6 * NN], errP[21 * NN];
  int chg[NN];
};

void load_block(Block& B, const std::vector<Cand>& cs, const std::vector<int>& idx) {
  const int n = (int)idx.size();
  for (int l = 0; l < NN; ++l) {
    const Cand& c = cs[idx[l < n ? l : 0]];
    for (int k = 0; k < 6; ++k) B.par[k * NN + l] = c.par[k];
    for (int k = 0; k < 21; ++k) B.err[k * NN + l] = c.err[k];
    B.chg[l] = c.chg;
  }
}

// kalmanPropagateAndComputeChi2 for one hit per lane (KalmanUtilsMPlex.cc:518-543 barrel, :744-770 endcap).
void prop_and_chi2(const CEEvent& E, bool barrel, const float* psErr, const float* psPar, const int* chg,
                   const float* msErr, const float* msPar, int n_proc, float* outChi2, float* propPar) {
  std::memcpy(propPar, psPar, sizeof(float) * 6 * NN);
  if (E.p->prop_to_hit) {
    alignas(64) float propErr[21 * NN], ms[NN];
    if (barrel) {
      for (int l = 0; l < NN; ++l) ms[l] = std::hypot(msPar[0 * NN + l], msPar[1 * NN + l]);
      propagate_helix_to_r(psErr, psPar, chg, ms, propErr, propPar, n_proc, E.p->intra_pflags, nullptr, nullptr, E.ctx);
      alignas(64) float dE[21 * NN], dP[6 * NN];
      kalman_operation(KFO_Calculate_Chi2, propErr, propPar, msErr, msPar, dE, dP, outChi2);
    } else {
      for (int l = 0; l < NN; ++l) ms[l] = msPar[2 * NN + l];
      propagate_helix_to_z(psErr, psPar, chg, ms, propErr, propPar, n_proc, E.p->intra_pflags, nullptr, E.ctx);
      alignas(64) float dE[21 * NN], dP[6 * NN];
      kalman_operation_endcap(KFO_Calculate_Chi2, propErr, propPar, msErr, msPar, dE, dP, outChi2);
    }
  } else {
    alignas(64) float dE[21 * NN], dP[6 * NN];
    if (barrel)
      kalman_operation(KFO_Calculate_Chi2, psErr, psPar, msErr, msPar, dE, dP, outChi2);
    else
      kalman_operation_endcap(KFO_Calculate_Chi2, psErr, psPar, msErr, msPar, dE, dP, outChi2);
  }
}

// kalmanPropagateAndUpdate (KalmanUtilsMPlex.cc:471-504; endcap :697-730), with the charge flip; the
// propagation to the hit uses `pflags` (UpdateWithLastHit: the inter-layer flags, FindCandidates: the
// intra-layer ones).
void prop_and_update(const CEEvent& E, bool barrel, const float* psErr, const float* psPar, int* chg,
                     const float* msErr, const float* msPar, int n_proc, float* outErr, float* outPar, int pflags) {
  alignas(64) float c2[NN];
  if (E.p->prop_to_hit) {
    alignas(64) float propErr[21 * NN], propPar[6 * NN], ms[NN];
    if (barrel) {
      for (int l = 0; l < NN; ++l) ms[l] = std::hypot(msPar[0 * NN + l], msPar[1 * NN + l]);
      propagate_helix_to_r(psErr, psPar, chg, ms, propErr, propPar, n_proc, pflags, nullptr, nullptr, E.ctx);
      kalman_operation(KFO_Update_Params, propErr, propPar, msErr, msPar, outErr, outPar, c2);
    } else {
      for (int l = 0; l < NN; ++l) ms[l] = msPar[2 * NN + l];
      propagate_helix_to_z(psErr, psPar, chg, ms, propErr, propPar, n_proc, pflags, nullptr, E.ctx);
      kalman_operation_endcap(KFO_Update_Params, propErr, propPar, msErr, msPar, outErr, outPar, c2);
    }
  } else {
    if (barrel)
      kalman_operation(KFO_Update_Params, psErr, psPar, msErr, msPar, outErr, outPar, c2);
    else
      kalman_operation_endcap(KFO_Update_Params, psErr, psPar, msErr, msPar, outErr, outPar, c2);
  }
  for (int n = 0; n < NN; ++n)
    if (outPar[3 * NN + n] < 0) {
      chg[n] = -chg[n];
      outPar[3 * NN + n] = -outPar[3 * NN + n];
    }
}

// Per-lane MkFinder members FindCandidatesCloneEngine reads.
struct LaneCounts {
  float chi2;
  int nfound, noverlap, ninside, ntail;
};

// FindCandidatesCloneEngine (MkFinder.cc:1114-1284) for one block of lanes at
// the layer (state parP/errP): appends the IdxChi2List records in the order the
// reference hands them to CandCloner::add_cand -- accepted hits by (XHitArr
// position, lane), then each lane's invalid-hit record.  XHitArr entries
// outside [0, n_hits) are skipped.
void find_candidates_block(const CEEvent& E, const orc_layer_desc& d, const float* hits, int n_hits,
                           const float* parP, const float* errP, const int* chg, const LaneCounts* lc,
                           const int* trk, const int* xs, const int* xa, const int* wsr, int n_proc,
                           std::vector<IdxChi2>& entries, float* c2m = nullptr, const int* in_gap = nullptr,
                           const unsigned char* charge = nullptr) {
  int added[NN] = {};                                   // nHitsAdded
  if (c2m)
    for (int i = 0; i < NN; ++i) c2m[i] = 1e30f;
  const orc_find_event& ev = *E.ev;
  const orc_ce_params& P = *E.p;
  const bool barrel = d.is_barrel != 0;
  int maxSize = 0;
  for (int i = 0; i < n_proc; ++i)
    if (xs[i] > 0) maxSize = std::max(maxSize, std::min(xs[i], kHitIdxMax));
  alignas(64) float msPar[3 * NN], msErr[6 * NN];
  for (int l = 0; l < 3 * NN; ++l) msPar[l] = 0.f;
  for (int l = 0; l < 6 * NN; ++l) msErr[l] = 1.f;
  for (int hc = 0; hc < maxSize; ++hc) {
    bool ok[NN] = {};
    for (int i = 0; i < n_proc; ++i)
      if (hc < xs[i] && xa[hc * NN + i] >= 0 && xa[hc * NN + i] < n_hits) {
        ok[i] = true;
        const float* h = hits + (size_t)xa[hc * NN + i] * 9;
        for (int e = 0; e < 3; ++e) msPar[e * NN + i] = h[e];
        for (int e = 0; e < 6; ++e) msErr[e * NN + i] = h[3 + e];
      }
    alignas(64) float chi2[NN], propPar[6 * NN];
    prop_and_chi2(E, barrel, errP, parP, chg, msErr, msPar, n_proc, chi2, propPar);
    for (int i = 0; i < n_proc; ++i) {
      const float max_c2 = dyn_chi2_cut(d, ev.chi2_cut, parP[3 * NN + i], parP[5 * NN + i]);
      if (!ok[i]) continue;
      const float c2 = std::abs(chi2[i]);
      if (c2m) c2m[i] = std::min(c2m[i], std::abs(c2 - max_c2) / std::max(max_c2, 1.f));
      if (!(c2 < max_c2)) continue;
      if (!strip_q_compatible(barrel, errP, propPar, msErr, msPar, i)) continue;
      if (charge && !pass_strip_charge(barrel, charge_pcm_of(charge[xa[hc * NN + i]]), propPar, msErr, i,
                                       c2m ? &c2m[i] : nullptr))
        continue;
      IdxChi2 e;
      e.trkIdx = trk[i];
      e.hitIdx = xa[hc * NN + i];
      e.nhits = lc[i].nfound + 1;
      e.ntailholes = 0;
      e.noverlaps = lc[i].noverlap;
      e.nholes = lc[i].ninside + lc[i].ntail;
      e.pt = std::abs(1.0f / parP[3 * NN + i]);
      e.chi2 = lc[i].chi2 + c2;
      e.chi2_hit = c2;
      e.score = score_struct(e);
      e.xpos = hc;
      entries.push_back(e);
      ++added[i];
    }
  }
  for (int i = 0; i < n_proc; ++i) {                    // the invalid-hit record
    if (wsr[i] == WSR_Outside) continue;
    int fake = (lc[i].ninside + lc[i].ntail < P.max_holes && lc[i].ntail < P.max_consec_holes) ? -1 : -2;
    if (wsr[i] == WSR_Edge) fake = -3;
    else if (in_gap && in_gap[i] && added[i] == 0) fake = -7;   // dead module (MkFinder.cc:1263-1267)
    IdxChi2 e;
    e.trkIdx = trk[i];
    e.hitIdx = fake;
    e.nhits = lc[i].nfound;
    e.ntailholes = fake == -1 ? lc[i].ntail + 1 : lc[i].ntail;
    e.noverlaps = lc[i].noverlap;
    e.nholes = lc[i].ninside;
    e.pt = std::abs(1.0f / parP[3 * NN + i]);
    e.chi2 = lc[i].chi2;
    e.chi2_hit = 0;
    e.score = score_struct(e);
    e.xpos = -1;
    entries.push_back(e);
  }
}

// MkFinder::FindCandidates (MkFinder.cc:902-1109) for one block of candidates idx of the seed, at the layer
// (state parP/errP): appends to tmp, per hit_cnt in lane order, a new candidate for every accepted hit -- the
// state at the layer propagated to the hit with the intra-layer flags and updated with it
// (m_update_param_foo), its charge flipped with a negative 1/pT -- then per lane the candidate with the
// invalid hit and the state at the layer (copy_out from iP); each scored by getScoreCand with
// penalizeTailMissHits and inFindCandidates.  XHitArr entries outside [0, n_hits) are skipped.
void find_candidates_std(const CEEvent& E, Comb& cc, int layer, const orc_layer_desc& d, const float* hits,
                         int n_hits, const Block& B, const std::vector<int>& idx, const int* xs, const int* xa,
                         const int* wsr, int n_proc, std::vector<Cand>& tmp, float* c2m = nullptr,
                         const int* in_gap = nullptr, const unsigned char* charge = nullptr) {
  int added[NN] = {};                                   // nHitsAdded
  if (c2m)
    for (int i = 0; i < NN; ++i) c2m[i] = 1e30f;
  const orc_find_event& ev = *E.ev;
  const orc_ce_params& P = *E.p;
  const bool barrel = d.is_barrel != 0;
  auto score_new = [](const Cand& c) {                  // getScoreCand(c, true, true)
    return score_calc(c.nfound, c.ntail, c.noverlap, c.ninside, c.chi2 < 0 ? 0.f : c.chi2, c.pT(), true);
  };
  int maxSize = 0;
  for (int i = 0; i < n_proc; ++i)
    if (xs[i] > 0) maxSize = std::max(maxSize, std::min(xs[i], kHitIdxMax));
  alignas(64) float msPar[3 * NN], msErr[6 * NN];
  for (int l = 0; l < 3 * NN; ++l) msPar[l] = 0.f;
  for (int l = 0; l < 6 * NN; ++l) msErr[l] = 1.f;
  for (int hc = 0; hc < maxSize; ++hc) {
    bool ok[NN] = {};
    for (int i = 0; i < n_proc; ++i)
      if (hc < xs[i] && xa[hc * NN + i] >= 0 && xa[hc * NN + i] < n_hits) {
        ok[i] = true;
        const float* h = hits + (size_t)xa[hc * NN + i] * 9;
        for (int e = 0; e < 3; ++e) msPar[e * NN + i] = h[e];
        for (int e = 0; e < 6; ++e) msErr[e * NN + i] = h[3 + e];
      }
    alignas(64) float chi2[NN], propPar[6 * NN];
    prop_and_chi2(E, barrel, B.errP, B.parP, B.chg, msErr, msPar, n_proc, chi2, propPar);
    bool pass[NN] = {}, one_passes = false;
    for (int i = 0; i < n_proc; ++i) {
      if (!ok[i]) continue;
      const float max_c2 = dyn_chi2_cut(d, ev.chi2_cut, B.parP[3 * NN + i], B.parP[5 * NN + i]);
      const float c2 = std::abs(chi2[i]);
      if (c2m) c2m[i] = std::min(c2m[i], std::abs(c2 - max_c2) / std::max(max_c2, 1.f));
      if (c2 < max_c2 && strip_q_compatible(barrel, B.errP, propPar, msErr, msPar, i) &&
          (!charge || pass_strip_charge(barrel, charge_pcm_of(charge[xa[hc * NN + i]]), propPar, msErr, i,
                                        c2m ? &c2m[i] : nullptr))) {
        pass[i] = true;
        one_passes = true;
      }
    }
    if (!one_passes) continue;
    alignas(64) float oErr[21 * NN], oPar[6 * NN];
    int tmpChg[NN];
    std::memcpy(tmpChg, B.chg, sizeof(tmpChg));
    prop_and_update(E, barrel, B.errP, B.parP, tmpChg, msErr, msPar, n_proc, oErr, oPar, P.intra_pflags);
    for (int i = 0; i < n_proc; ++i) {
      if (!pass[i]) continue;
      ++added[i];
      Cand nc = cc.cands[idx[i]];
      for (int k = 0; k < 6; ++k) nc.par[k] = oPar[k * NN + i];
      for (int k = 0; k < 21; ++k) nc.err[k] = oErr[k * NN + i];
      nc.chg = tmpChg[i];
      cc.add_hit(nc, xa[hc * NN + i], layer, std::abs(chi2[i]));
      nc.score = score_new(nc);
      nc.valid = true;
      tmp.push_back(nc);
    }
  }
  for (int i = 0; i < n_proc; ++i) {                    // the candidate with the invalid hit
    if (wsr[i] == WSR_Outside) continue;
    const Cand& c = cc.cands[idx[i]];
    int fake = (c.ninside + c.ntail < P.max_holes && c.ntail < P.max_consec_holes) ? -1 : -2;
    if (wsr[i] == WSR_Edge) fake = -3;
    else if (in_gap && in_gap[i] && added[i] == 0) fake = -7;   // dead module (MkFinder.cc:1092-1096)
    Cand nc = c;
    for (int k = 0; k < 6; ++k) nc.par[k] = B.parP[k * NN + i];
    for (int k = 0; k < 21; ++k) nc.err[k] = B.errP[k * NN + i];
    nc.chg = B.chg[i];
    cc.add_hit(nc, fake, layer, 0.f);
    nc.score = score_new(nc);
    nc.valid = true;
    tmp.push_back(nc);
  }
}

// Gap of two scores relative to the larger of their chi2 sums (for the checker's margins): scores differ
// between FP32 implementations by their chi2 sums' rounding.
inline float score_gap(float a, float ca, float b, float cb) {
  return std::abs(a - b) / std::max(1.f, std::max(std::abs(ca), std::abs(cb)));
}

// One seed through every layer of the plan.  With `mg` (checker only, [4][n_layers], preset to 1e30), also the
// seed's decision margins per layer: window edge (cm), bin edge (bin fraction), chi2 against the cut
// (relative), and the smallest relative gap between scores that the candidate selection compares; `mgap`
// gets that of the final merge.
// standard: MkBuilder::FindTracksStandard (MkBuilder.cc:1643-1845) instead of the clone engine: held-back and
// new candidates go to the seed's tmp_cands, which is sorted by score; the best max_cands without a -2 last
// hit become the next candidates and the first -2 one in that order competes for the best short candidate.
// wk (checker only, 3 ints): candidates propagated to a layer, window hits evaluated, candidates updated.
void find_seed(const CEEvent& E, Comb& cc, int layer0, float* mg = nullptr, float* mgap = nullptr,
               bool standard = false, int* wk = nullptr) {
  const orc_find_event& ev = *E.ev;
  const orc_ce_params& P = *E.p;
  const int max_cands = P.max_cands;
  int prev_layer = layer0 - 1;
  for (int layer = layer0; layer < ev.n_layers; prev_layer = layer, ++layer) {
    const orc_layer_desc& d = ev.layers[layer];
    const bool barrel = d.is_barrel != 0;
    const float* hits = ev.hits + (size_t)ev.hit_offset[layer] * 9;
    // find_tracks_unroll_candidates
    if (cc.state == 0 && cc.pickup_layer == prev_layer) cc.state = 1;
    if (cc.state != 1) continue;
    std::vector<int> active;
    for (int ic = 0; ic < (int)cc.cands.size(); ++ic) {
      Cand& c = cc.cands[ic];
      if (cc.last_idx(c) == -2) continue;
      if (c.pT() < 1.2) {                               // near the apex of a low-pT helix
        const float dphi = std::abs(std::atan2(c.par[1], c.par[0]) - c.par[4]);
        if (c.par[0] * c.par[0] + c.par[1] * c.par[1] > 625.f && dphi > 1.371f && dphi < 4.512f) {
          cc.add_hit(c, -2, layer, 0.0f);
          continue;
        }
      }
      active.push_back(ic);
    }
    if (active.empty()) {
      cc.state = 2;
      continue;
    }
    std::vector<IdxChi2> entries;
    std::vector<Cand> extras, tmp;
    std::vector<Cand>& held = standard ? tmp : extras;   // where find_tracks_handle_missed_layers puts them
    for (size_t b0 = 0; b0 < active.size(); b0 += NN) {
      std::vector<int> idx(active.begin() + b0, active.begin() + std::min(active.size(), b0 + NN));
      const int n_proc = (int)idx.size();
      Block B;
      load_block(B, cc.cands, idx);
      alignas(64) float rad[NN];
      for (int l = 0; l < NN; ++l) rad[l] = d.propagate_to;
      int fail[NN];
      if (barrel)
        propagate_helix_to_r(B.err, B.par, B.chg, rad, B.errP, B.parP, n_proc, ev.pflags, nullptr, fail, E.ctx);
      else
        propagate_helix_to_z(B.err, B.par, B.chg, rad, B.errP, B.parP, n_proc, ev.pflags, nullptr, E.ctx);
      int xs[NN], xa[kHitIdxMax * NN], wsr[NN], gap[NN];
      SelMargins sm[NN];
      select_hit_indices(d, E.loh[layer], B.parP, B.errP, n_proc, xs, xa, wsr, mg ? sm : nullptr, gap);
      if (wk) wk[0] += n_proc;
      // find_tracks_handle_missed_layers
      for (int i = 0; i < n_proc; ++i) {
        const Cand& c = cc.cands[idx[i]];
        const float cand_r = std::hypot(B.parP[0 * NN + i], B.parP[1 * NN + i]);
        if (barrel && cand_r < d.rin) {
          xs[i] = 0;
          wsr[i] = WSR_Outside;
          held.push_back(c);
          cc.add_hit(held.back(), -2, layer, 0);        // barrel region: stopped replica
        } else if (wsr[i] == WSR_Outside) {
          held.push_back(c);
        }
      }
      // FindCandidatesCloneEngine
      LaneCounts lc[NN];
      int trk[NN];
      for (int i = 0; i < n_proc; ++i) {
        const Cand& c = cc.cands[idx[i]];
        lc[i] = {c.chi2, c.nfound, c.noverlap, c.ninside, c.ntail};
        trk[i] = idx[i];
      }
      if (wk)
        for (int i = 0; i < n_proc; ++i)
          for (int hc = 0; hc < std::min(xs[i], kHitIdxMax); ++hc)
            wk[1] += xa[hc * NN + i] >= 0 && xa[hc * NN + i] < ev.hit_offset[layer + 1] - ev.hit_offset[layer];
      float c2m[NN];
      if (standard)
        find_candidates_std(E, cc, layer, d, hits, ev.hit_offset[layer + 1] - ev.hit_offset[layer], B, idx, xs, xa,
                            wsr, n_proc, tmp, mg ? c2m : nullptr, gap, E.loh[layer].charge);
      else
        find_candidates_block(E, d, hits, ev.hit_offset[layer + 1] - ev.hit_offset[layer], B.parP, B.errP, B.chg,
                              lc, trk, xs, xa, wsr, n_proc, entries, mg ? c2m : nullptr, gap, E.loh[layer].charge);
      if (mg) {
        const int L = ev.n_layers;
        for (int i = 0; i < n_proc; ++i) {
          if (wsr[i] == WSR_Outside) continue;
          mg[0 * L + layer] = std::min(mg[0 * L + layer], sm[i].win);
          mg[1 * L + layer] = std::min(mg[1 * L + layer], sm[i].bin);
          mg[2 * L + layer] = std::min(mg[2 * L + layer], c2m[i]);
        }
      }
    }
    if (standard) {
      if (wk)
        for (const Cand& t : tmp) wk[2] += cc.last_idx(t) >= 0 && cc.hots[t.last].layer == layer;
      // sort the seed's tmp_cands (std::sort in the reference: here stable, ties keep generation order), then
      // place the best max_cands without a -2 last hit; the first -2 one met competes for the best short one
      std::stable_sort(tmp.begin(), tmp.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });
      if (tmp.empty()) continue;
      std::vector<Cand> cv;
      int n_placed = 0;
      bool first_short = true;
      size_t ii = 0;
      for (; ii < tmp.size() && n_placed < max_cands; ++ii) {
        Cand& tc = tmp[ii];
        if (cc.last_idx(tc) != -2) {
          cv.push_back(tc);
          ++n_placed;
        } else if (first_short) {
          first_short = false;
          if (mg && cc.best_short.valid)
            mg[3 * ev.n_layers + layer] = std::min(mg[3 * ev.n_layers + layer],
                                                   score_gap(tc.score, tc.chi2, cc.best_short.score,
                                                             cc.best_short.chi2));
          if (tc.score > cc.best_short.score) cc.best_short = tc;
        }
      }
      if (mg)      // the order up to one past the last entry examined
        for (size_t i = 0; i + 1 < tmp.size() && i < ii; ++i)
          mg[3 * ev.n_layers + layer] = std::min(mg[3 * ev.n_layers + layer],
                                                 score_gap(tmp[i].score, tmp[i].chi2, tmp[i + 1].score,
                                                           tmp[i + 1].chi2));
      cc.cands = cv;
      continue;
    }
    // CandCloner::ProcessSeedRange
    std::vector<Cand> cv;
    std::vector<int> to_update;                         // positions in cv updated with their last hit
    if (!entries.empty()) {
      std::stable_sort(entries.begin(), entries.end(),
                       [](const IdxChi2& a, const IdxChi2& b) { return a.score > b.score; });
      const int num_hits = std::min((int)entries.size(), max_cands);
      if (mg) {   // gaps the selection compares: neighbours up to the cut at K, extras and the best short one
        float& g = mg[3 * ev.n_layers + layer];
        for (int i = 0; i + 1 < (int)entries.size() && i < num_hits; ++i)
          g = std::min(g, score_gap(entries[i].score, entries[i].chi2, entries[i + 1].score, entries[i + 1].chi2));
        for (int i = 0; i < num_hits; ++i) {
          for (const Cand& x : extras) g = std::min(g, score_gap(x.score, x.chi2, entries[i].score, entries[i].chi2));
          if (entries[i].hitIdx == -2 && cc.best_short.valid)
            g = std::min(g, score_gap(entries[i].score, entries[i].chi2, cc.best_short.score, cc.best_short.chi2));
        }
      }
      size_t xi = 0;
      int n_pushed = 0;
      for (int ih = 0; ih < num_hits; ++ih) {
        const IdxChi2& h = entries[ih];
        Cand tc = cc.cands[h.trkIdx];
        cc.add_hit(tc, h.hitIdx, layer, h.chi2_hit);
        tc.score = h.score;
        tc.valid = true;
        if (h.hitIdx == -2) {
          if (h.score > cc.best_short.score) cc.best_short = tc;
          continue;
        }
        while (xi < extras.size() && extras[xi].score > tc.score && n_pushed < max_cands) {
          cv.push_back(extras[xi++]);
          ++n_pushed;
        }
        if (n_pushed >= max_cands) break;
        cv.push_back(tc);
        ++n_pushed;
        if (h.hitIdx >= 0) to_update.push_back(n_pushed - 1);
      }
      while (xi < extras.size() && n_pushed < max_cands) {
        cv.push_back(extras[xi++]);
        ++n_pushed;
      }
      cc.cands = cv;
    } else {
      cc.cands = extras;                                // state is Finding here
    }
    // UpdateWithLastHit over the candidates that took a hit on this layer
    if (wk) wk[2] += (int)to_update.size();
    for (size_t b0 = 0; b0 < to_update.size(); b0 += NN) {
      std::vector<int> idx(to_update.begin() + b0, to_update.begin() + std::min(to_update.size(), b0 + NN));
      const int n_proc = (int)idx.size();
      Block B;
      load_block(B, cc.cands, idx);
      alignas(64) float msPar[3 * NN], msErr[6 * NN], oErr[21 * NN], oPar[6 * NN];
      for (int l = 0; l < NN; ++l) {
        const HoTNode& hn = cc.hots[cc.cands[idx[l < n_proc ? l : 0]].last];
        const float* h = hits + (size_t)hn.index * 9;
        for (int e = 0; e < 3; ++e) msPar[e * NN + l] = h[e];
        for (int e = 0; e < 6; ++e) msErr[e * NN + l] = h[3 + e];
      }
      prop_and_update(E, barrel, B.err, B.par, B.chg, msErr, msPar, n_proc, oErr, oPar, ev.pflags);
      for (int i = 0; i < n_proc; ++i) {
        Cand& c = cc.cands[idx[i]];
        for (int k = 0; k < 6; ++k) c.par[k] = oPar[k * NN + i];
        for (int k = 0; k < 21; ++k) c.err[k] = oErr[k * NN + i];
        c.chg = B.chg[i];
      }
    }
  }
  // MergeCandsAndBestShortOne(update_score = true, sort_cands = true)
  Cand* best_short = cc.best_short.valid ? &cc.best_short : nullptr;
  if (!cc.cands.empty()) {
    for (auto& c : cc.cands) c.score = score_cand(c);
    if (best_short) best_short->score = score_cand(*best_short);
    std::stable_sort(cc.cands.begin(), cc.cands.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });
    if (mgap) {
      for (size_t i = 0; i + 1 < cc.cands.size(); ++i)
        *mgap = std::min(*mgap, score_gap(cc.cands[i].score, cc.cands[i].chi2, cc.cands[i + 1].score,
                                          cc.cands[i + 1].chi2));
      if (best_short)
        for (const Cand& c : cc.cands)
          *mgap = std::min(*mgap, score_gap(c.score, c.chi2, best_short->score, best_short->chi2));
    }
    if (best_short && best_short->score > cc.cands.back().score) {
      size_t ci = 0;
      while (cc.cands[ci].score > best_short->score) ++ci;
      if ((int)cc.cands.size() >= max_cands) cc.cands.pop_back();
      cc.cands.insert(cc.cands.begin() + ci, *best_short);
    }
  } else if (best_short) {
    cc.cands.push_back(*best_short);
  }
}

}  // namespace

namespace {

int find_batch(const orc_ctx* c, const orc_find_event* ev, const orc_ce_params* p, orc_ce_tracks* tr, int n_threads,
               bool standard) {
  if (!ev || !p || !tr || ev->n_layers <= 0 || ev->first_layer < 0 || ev->first_layer > ev->n_layers ||
      p->max_cands < 1 || p->max_cands > 16)
    return 1;
  CEEvent E;
  E.ev = ev;
  E.p = p;
  E.ctx.bfield = c ? c->bfield : 3.8112f;
  E.ctx.rl_grid = c ? c->rl_grid : nullptr;
  E.ctx.xi_grid = c ? c->xi_grid : nullptr;
  E.loh.resize(ev->n_layers);
  for (int l = 0; l < ev->n_layers; ++l) {
    build_loh(ev->layers[l], ev->hit_offset[l + 1] - ev->hit_offset[l], ev->hits + (size_t)ev->hit_offset[l] * 9,
              E.loh[l]);
    if (ev->hit_mask) E.loh[l].mask = ev->hit_mask + ev->hit_offset[l];
    if (ev->hit_charge) E.loh[l].charge = ev->hit_charge + ev->hit_offset[l];
    if (ev->n_dead > 0) suck_in_deads(E.loh[l], l, ev->n_dead, ev->dead_regions);
  }
  const int n = tr->n_seeds, L = ev->n_layers, K = p->max_cands;
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      const int s = next.fetch_add(1);
      if (s >= n) break;
      Comb cc;
      Cand seed;
      for (int k = 0; k < 6; ++k) seed.par[k] = tr->seed_par[(size_t)k * n + s];
      for (int k = 0; k < 21; ++k) seed.err[k] = tr->seed_err[(size_t)k * n + s];
      seed.chg = tr->seed_chg[s];
      seed.chi2 = 0;
      seed.valid = true;
      for (int l = 0; l < ev->first_layer; ++l) cc.add_hit(seed, tr->seed_hits[(size_t)l * n + s], l, 0.0f);
      seed.score = score_cand(seed);
      cc.cands.push_back(seed);                         // CombCandidate::ImportSeed
      cc.pickup_layer = ev->first_layer - 1;
      cc.best_short.score = kScoreWorst;
      std::vector<float> mg(tr->margins ? 4 * (size_t)L : 0, 1e30f);
      float mgap = 1e30f;
      int wk[3] = {0, 0, 0};
      find_seed(E, cc, ev->first_layer, tr->margins ? mg.data() : nullptr, tr->margins ? &mgap : nullptr, standard,
                wk);
      if (tr->work)
        for (int k = 0; k < 3; ++k) tr->work[(size_t)k * n + s] = wk[k];
      if (tr->margins) {
        for (int k = 0; k < 4; ++k)
          for (int l = 0; l < L; ++l) tr->margins[((size_t)k * L + l) * n + s] = mg[(size_t)k * L + l];
        tr->merge_gap[s] = mgap;
      }
      const int nc = std::min((int)cc.cands.size(), K);
      tr->n_cands[s] = nc;
      for (int ic = 0; ic < K; ++ic) {
        const size_t o = (size_t)ic;
        for (int l = 0; l < L; ++l) tr->hot_idx[(o * L + l) * n + s] = -9;
        if (ic >= nc) {
          tr->score[o * n + s] = kScoreWorst;
          tr->n_found[o * n + s] = 0;
          continue;
        }
        const Cand& cd = cc.cands[ic];
        for (int k = 0; k < 6; ++k) tr->par[(o * 6 + k) * n + s] = cd.par[k];
        for (int k = 0; k < 21; ++k) tr->err[(o * 21 + k) * n + s] = cd.err[k];
        tr->chg[o * n + s] = cd.chg;
        tr->chi2[o * n + s] = cd.chi2;
        tr->score[o * n + s] = cd.score;
        tr->n_found[o * n + s] = cd.nfound;
        for (int node = cd.last; node >= 0; node = cc.hots[node].prev) {
          const HoTNode& h = cc.hots[node];
          if (h.layer >= 0 && h.layer < L && tr->hot_idx[(o * L + h.layer) * n + s] == -9)
            tr->hot_idx[(o * L + h.layer) * n + s] = h.index;
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

}  // namespace

extern "C" int orc_find_clone_engine(const orc_ctx* c, const orc_find_event* ev, const orc_ce_params* p,
                                     orc_ce_tracks* tr, int n_threads) {
  return find_batch(c, ev, p, tr, n_threads, false);
}

extern "C" int orc_find_standard(const orc_ctx* c, const orc_find_event* ev, const orc_ce_params* p,
                                 orc_ce_tracks* tr, int n_threads) {
  return find_batch(c, ev, p, tr, n_threads, true);
}

namespace {

CEEvent layer_event(const orc_ctx* c, const orc_find_event* ev, const orc_ce_params* p) {
  CEEvent E;
  E.ev = ev;
  E.p = p;
  E.ctx.bfield = c ? c->bfield : 3.8112f;
  E.ctx.rl_grid = c ? c->rl_grid : nullptr;
  E.ctx.xi_grid = c ? c->xi_grid : nullptr;
  return E;
}

// Lanes [t0, t0 + cnt) of an element-major [k][n] array into an NN block (padding repeats lane t0).
template <class T>
void gather_block(T* dst, const T* src, int rows, int n, int t0, int cnt) {
  for (int k = 0; k < rows; ++k)
    for (int l = 0; l < NN; ++l) dst[k * NN + l] = src[(size_t)k * n + t0 + (l < cnt ? l : 0)];
}

}  // namespace

extern "C" int orc_find_candidates_ce(const orc_ctx* c, const orc_find_event* ev, const orc_ce_params* p, int layer,
                                      const orc_finder_lanes* ln, const int* xsize, const int* xarr, const int* wsr,
                                      int* n_entries, orc_idx_chi2* out) {
  if (!ev || !p || !ln || layer < 0 || layer >= ev->n_layers || ln->n < 0) return 1;
  const CEEvent E = layer_event(c, ev, p);
  const orc_layer_desc& d = ev->layers[layer];
  const float* hits = ev->hits + (size_t)ev->hit_offset[layer] * 9;
  const int n_hits = ev->hit_offset[layer + 1] - ev->hit_offset[layer];
  const int n = ln->n;
  constexpr int kPer = kHitIdxMax + 1;
  for (int t0 = 0; t0 < n; t0 += NN) {
    const int cnt = std::min(NN, n - t0);
    alignas(64) float parP[6 * NN], errP[21 * NN];
    int chg[NN], xs[NN], xa[kHitIdxMax * NN], w[NN], trk[NN];
    LaneCounts lc[NN];
    gather_block(parP, ln->prop_par, 6, n, t0, cnt);
    gather_block(errP, ln->prop_err, 21, n, t0, cnt);
    gather_block(chg, ln->chg, 1, n, t0, cnt);
    gather_block(xs, xsize, 1, n, t0, cnt);
    gather_block(xa, xarr, kHitIdxMax, n, t0, cnt);
    gather_block(w, wsr, 1, n, t0, cnt);
    for (int l = 0; l < NN; ++l) {
      const int t = t0 + (l < cnt ? l : 0);
      lc[l] = {ln->chi2[t], ln->n_found[t], 0, ln->n_inside_minus_one[t], ln->n_tail_minus_one[t]};
      trk[l] = t;
    }
    std::vector<IdxChi2> entries;
    find_candidates_block(E, d, hits, n_hits, parP, errP, chg, lc, trk, xs, xa, w, cnt, entries, nullptr, nullptr,
                          ev->hit_charge ? ev->hit_charge + ev->hit_offset[layer] : nullptr);
    for (int l = 0; l < cnt; ++l) n_entries[t0 + l] = 0;
    for (const IdxChi2& e : entries) {
      const int t = e.trkIdx;
      orc_idx_chi2& o = out[(size_t)t * kPer + n_entries[t]++];
      o.trk_idx = t;
      o.hit_idx = e.hitIdx;
      o.module = e.hitIdx >= 0 ? 0u : 0xffffffffu;
      o.nhits = e.nhits;
      o.ntailholes = e.ntailholes;
      o.noverlaps = e.noverlaps;
      o.nholes = e.nholes;
      o.pt = e.pt;
      o.chi2 = e.chi2;
      o.chi2_hit = e.chi2_hit;
      o.score = e.score;
      o.xhit_pos = e.xpos;
    }
  }
  return 0;
}

// Checker hook: passStripChargePCMfromTrack for one state par[6] (propagated to the hit) and hit
// covariance herr[6] (xx yx yy zx zy zz), given the hit's 8-bit charge code; returns 1 when the hit
// passes, and the corrected charge per cm in *qcorr.
extern "C" int orc_pass_strip_charge(int barrel, unsigned code, const float* par, const float* herr, float* qcorr) {
  alignas(64) float p[6 * NN] = {}, e[6 * NN] = {};
  for (int k = 0; k < 6; ++k) {
    p[k * NN] = par[k];
    e[k * NN] = herr[k];
  }
  return pass_strip_charge(barrel != 0, charge_pcm_of((unsigned char)code), p, e, 0, nullptr, qcorr) ? 1 : 0;
}

extern "C" int orc_update_with_last_hit(const orc_ctx* c, const orc_find_event* ev, const orc_ce_params* p,
                                        int layer, const orc_finder_lanes* ln, const int* hit_idx) {
  if (!ev || !p || !ln || layer < 0 || layer >= ev->n_layers || ln->n < 0) return 1;
  const CEEvent E = layer_event(c, ev, p);
  const bool barrel = ev->layers[layer].is_barrel != 0;
  const float* hits = ev->hits + (size_t)ev->hit_offset[layer] * 9;
  const int n_hits = ev->hit_offset[layer + 1] - ev->hit_offset[layer];
  const int n = ln->n;
  for (int t0 = 0; t0 < n; t0 += NN) {
    const int cnt = std::min(NN, n - t0);
    alignas(64) float par[6 * NN], err[21 * NN], msPar[3 * NN], msErr[6 * NN], oErr[21 * NN], oPar[6 * NN];
    int chg[NN];
    gather_block(par, ln->prop_par, 6, n, t0, cnt);
    gather_block(err, ln->prop_err, 21, n, t0, cnt);
    gather_block(chg, ln->chg, 1, n, t0, cnt);
    bool ok[NN];
    for (int l = 0; l < NN; ++l) {
      const int h = hit_idx[t0 + (l < cnt ? l : 0)];
      ok[l] = h >= 0 && h < n_hits;
      const float* hp = hits + (size_t)(ok[l] ? h : 0) * 9;
      for (int e = 0; e < 3; ++e) msPar[e * NN + l] = ok[l] ? hp[e] : 0.f;
      for (int e = 0; e < 6; ++e) msErr[e * NN + l] = ok[l] ? hp[3 + e] : 1.f;
    }
    int chg_u[NN];
    std::memcpy(chg_u, chg, sizeof(chg));
    prop_and_update(E, barrel, err, par, chg_u, msErr, msPar, cnt, oErr, oPar, ev->pflags);
    for (int l = 0; l < cnt; ++l) {
      const size_t t = (size_t)t0 + l;
      for (int k = 0; k < 6; ++k) ln->par[k * (size_t)n + t] = ok[l] ? oPar[k * NN + l] : par[k * NN + l];
      for (int k = 0; k < 21; ++k) ln->err[k * (size_t)n + t] = ok[l] ? oErr[k * NN + l] : err[k * NN + l];
      ln->chg[t] = ok[l] ? chg_u[l] : chg[l];
    }
  }
  return 0;
}
