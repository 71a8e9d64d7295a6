// mkf_ce.hip -- clone-engine track finding on gfx950.
//
// One 64-lane workgroup carries 8 or 10 seeds through every layer: lane
// g*KS + c holds candidate c of seed g (KS = 8 slots per seed, or 6 when at most
// 6 candidates are kept) with its state in VGPRs.
// Per layer:
//   (a) lanes propagate their candidate to the layer, select hits and compute
//       each hit's chi2 (after propagating to the hit, as the reference does
//       with finding_requires_propagation_to_hit_pos); accepted hits and the
//       invalid-hit entry go to LDS as IdxChi2List records;
//   (b) one lane per seed keeps the best max_cands records (a stable top-K by
//       score in the reference's generation order) and lays out the seed's
//       next candidate list -- records interleaved with held-back extras, the
//       best short candidate set aside (CandCloner::ProcessSeedRange);
//   (c) every lane builds its new candidate from its parent's state (exchanged
//       through LDS), appends the HitOnTrack node and, for a real hit, applies
//       the Kalman update after propagating to the hit (UpdateWithLastHit).
// Hit lists live in a per-seed append-only node store in HBM, like the
// reference's CombCandidate::m_hots: cloning a candidate copies one index.
// After the last layer MergeCandsAndBestShortOne orders the candidates, and
// each lane walks its node chain to write the per-layer hit indices.
//
// The same kernel with STD runs standard combinatorial finding
// (MkBuilder::FindTracksStandard, MkFinder::FindCandidates, MkFinder.cc:902-1109):
// a new candidate carries its parent's state at the layer -- updated with its
// hit after propagating to it with the intra-layer flags -- and is scored with
// its own updated pT; held-back, new and invalid-hit candidates are ranked
// together, the best max_cands without a -2 hit are kept and the first -2 one
// in that order competes for the best short candidate (MkBuilder.cc:1760-1828).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/mkfit_amd.h"
#include "mkf_find_common.h"
#include "mkf_internal.h"
#include "mkf_track_math.h"

using namespace mkf;
using namespace mkf_detail;
using namespace mkf_find;

namespace {

constexpr int kSMax = 8;              // candidate slots per seed, at most
constexpr int kLanes = 64;            // one wavefront per workgroup
constexpr int kW = 37;                // LDS words per exchanged candidate (36 used, odd stride)
// Words of a lane's X row during (a), beside the propagated state in words 0-27:
// chi2 cut, exclusive prefix of the window hit counts, layer hit offset,
// barrel flag of the lane's layer and, in words 32-35, sin / cos of the
// propagated phi and theta.  Word 36, which put / get leave alone, holds
// the score of the lane's invalid-hit record from (a) to the new candidates (c).
constexpr int kXCut = 28, kXPre = 29, kXOff = 30, kXBar = 31, kXTrig = 32, kXScore = 36;
// A kept accepted-hit record is {hit_cnt << kPosBits | position in the layer}.
constexpr int kPosBits = 28;
constexpr int kPosMask = (1 << kPosBits) - 1;
constexpr int kBsW = 40;              // words per seed of the best short candidate store (36 used)
static_assert(kHitIdxMax <= 16, "hit_cnt must fit above the position bits");

struct Cand {
  float par[6], err[21];
  int chg;
  float chi2, score;
  int nfound, nmissing, ninside, ntail;
  int last;       // node index of the last HitOnTrack (CombCandidate::m_hots position)
  int last_hit;   // hit index stored in that node
};

__device__ __forceinline__ void add_hit_counts(Cand& c, int idx, float chi2) {   // TrackCand::addHitIdx
  if (idx >= 0 || idx == -9) {
    ++c.nfound;
    c.chi2 += chi2;
    c.ninside += c.ntail;
    c.ntail = 0;
  } else {
    ++c.nmissing;
    if (idx == -1) ++c.ntail;
  }
  c.last_hit = idx;
}

__device__ __forceinline__ float cand_score(const Cand& c) {                       // getScoreCand
  const float chi2 = c.chi2 < 0 ? 0.f : c.chi2;
  return score_calc(c.nfound, 0, 0, c.ninside, chi2, fabsf(1.f / c.par[3]), false);
}

__device__ __forceinline__ void put(float* X, const Cand& c) {
#pragma unroll
  for (int e = 0; e < 6; ++e) X[e] = c.par[e];
#pragma unroll
  for (int e = 0; e < 21; ++e) X[6 + e] = c.err[e];
  X[27] = __int_as_float(c.chg);
  X[28] = c.chi2;
  X[29] = c.score;
  X[30] = __int_as_float(c.nfound);
  X[31] = __int_as_float(c.nmissing);
  X[32] = __int_as_float(c.ninside);
  X[33] = __int_as_float(c.ntail);
  X[34] = __int_as_float(c.last);
  X[35] = __int_as_float(c.last_hit);
}
// The words of `put` after the state: charge, chi2, score, counters, nodes.
__device__ __forceinline__ void put_counts(float* X, const Cand& c) {
  X[27] = __int_as_float(c.chg);
  X[28] = c.chi2;
  X[29] = c.score;
  X[30] = __int_as_float(c.nfound);
  X[31] = __int_as_float(c.nmissing);
  X[32] = __int_as_float(c.ninside);
  X[33] = __int_as_float(c.ntail);
  X[34] = __int_as_float(c.last);
  X[35] = __int_as_float(c.last_hit);
}
__device__ __forceinline__ void get(const float* X, Cand& c) {
#pragma unroll
  for (int e = 0; e < 6; ++e) c.par[e] = X[e];
#pragma unroll
  for (int e = 0; e < 21; ++e) c.err[e] = X[6 + e];
  c.chg = __float_as_int(X[27]);
  c.chi2 = X[28];
  c.score = X[29];
  c.nfound = __float_as_int(X[30]);
  c.nmissing = __float_as_int(X[31]);
  c.ninside = __float_as_int(X[32]);
  c.ntail = __float_as_int(X[33]);
  c.last = __float_as_int(X[34]);
  c.last_hit = __float_as_int(X[35]);
}

struct CeArgs {
  int n, n_layers, first_layer, eoh_layers;
  const int* __restrict__ layer_base;
  int max_cands, max_holes, max_consec, prop_to_hit, intra_pflags, inter_pflags;
  float chi2_cut, bfield;
  Material mat;
  const float* __restrict__ seed_par;
  const float* __restrict__ seed_err;
  const int* __restrict__ seed_chg;
  const int* __restrict__ seed_hits;
  int* __restrict__ n_cands;
  float* __restrict__ par;
  float* __restrict__ err;
  int* __restrict__ chg;
  float* __restrict__ chi2;
  float* __restrict__ bs;         // per seed kBsW words: the best short candidate
  float* __restrict__ score;
  int* __restrict__ n_found;
  int* __restrict__ hot;
  int4* __restrict__ nodes;   // [n][node_cap] {hit index, layer, prev node, 0}
  int node_cap;
  const LayerDev* __restrict__ layers;
  const int2* __restrict__ bins;
  const float* __restrict__ phis;
  const float* __restrict__ qs;
  const unsigned char* __restrict__ mask;
  const unsigned char* __restrict__ charge;   // per sorted hit: strip charge code, or NULL (every hit passes)
  const unsigned char* __restrict__ dead;     // dead-module bins per (layer, q bin, phi bin), or NULL
  const float* __restrict__ shits;
  const int* __restrict__ ranks;
};

// Appends a node to seed `s`'s store; returns its index, or -1 when the store
// is full (never for the store size the host allocates: 3 nodes per candidate
// slot and layer plus the seed hits).
__device__ __forceinline__ int append_node(const CeArgs& a, int s, int* cnt, int hit, int layer, int prev) {
  const int k = atomicAdd(cnt, 1);
  if (k >= a.node_cap) return -1;
  a.nodes[(size_t)s * a.node_cap + k] = make_int4(hit, layer, prev, 0);
  return k;
}

__device__ __forceinline__ void propagate(bool barrel, float (&p)[6], float (&e)[21], int chg, float target,
                                          int pflags, const CeArgs& a, const float* trig = nullptr) {
  if (barrel)
    propagate_to_r(p, e, chg, target, pflags, false, a.bfield, a.mat, trig);
  else
    propagate_to_z(p, e, chg, target, pflags, false, a.bfield, a.mat, trig);
}

// KS candidate slots per seed (max_cands <= KS): KS = 6 packs 10 seeds into the
// wavefront's 64 lanes (4 idle), KS = 8 packs 8.  Lanes past the last whole
// seed group (g == kG) take no part.  STD: standard finding (see the top).
// WAVES: waves per SIMD the registers are allotted for.  Left to itself the compiler spends all 512
// registers of a lane on one wave; two measured 21.0 against 29.7 ms for the standard form and 15.3
// against 28.1 ms for the clone engine (50 events).  mkf_autotune_find picks KS and WAVES by timing.
template <int KS, bool STD, int WAVES>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(WAVES))) k_find_clone_engine(CeArgs a) {
  constexpr int kS = KS;
  constexpr int kG = kLanes / KS;             // seeds per workgroup
  constexpr int kGA = kG + 1;                 // seed rows in LDS: one more for the idle lanes
  // 19.9 KB (kS = 6) / 20.4 KB (kS = 8): eight workgroups per CU, as many as
  // the VGPRs allow.
  __shared__ float X[kLanes * kW];          // candidate state exchange (and the kX* words in (a))
  __shared__ int Gh[kLanes][kHitIdxMax];    // each lane's window hits (positions in the layer), in scan
                                            // order; from (a3) on its kept records, first ecnt
  __shared__ float Gc[kLanes][kHitIdxMax];  // their chi2 if accepted, else -1; then the kept records'
  __shared__ float Es[kLanes][kS];          // the kept records' scores
  __shared__ float Gp[STD ? kLanes : 1][kHitIdxMax];   // STD: pT of each accepted hit's updated state
  __shared__ unsigned char ecnt[kLanes];    // kept records
  __shared__ signed char fcode[kLanes];     // invalid-hit record's hit code, 0 = none
  __shared__ unsigned char xkind[kLanes];   // 1 = held back unchanged, 2 = stopped replica
  __shared__ int D[kGA][kS];                // next candidates: kind | parent slot << 2 | record << 5
  __shared__ float s_bsc[kGA];              // score of the best short candidate (the candidate: a.bs)
  __shared__ int s_nnodes[kGA];
  __shared__ unsigned char s_bs[kGA], s_state[kGA], s_ncand[kGA], s_any[kGA], s_nnew[kGA], s_merge[kGA];
  __shared__ unsigned char s_nout[kGA];

  const int lane = threadIdx.x;
  const int g = lane / kS, c = lane % kS;
  const int seed = blockIdx.x * kG + g;
  const bool gin = g < kG && seed < a.n;
  const int base = (gin && a.layer_base) ? a.layer_base[seed] : 0;   // first layer of the seed's event
  const bool gvalid = gin && base >= 0 && base <= a.eoh_layers - a.n_layers;
  const LayerDev* __restrict__ lays = a.layers + (gvalid ? base : 0);
  const size_t N = (size_t)a.n;
  const int K = a.max_cands;

  Cand st = {};
  if (c == 0) {
    s_state[g] = 0;
    s_ncand[g] = gvalid ? 1 : 0;
    s_nnodes[g] = 0;
    s_bs[g] = 0;
    s_bsc[g] = kScoreWorst;
  }
  __syncthreads();
  if (gvalid && c == 0) {                                    // CombCandidate::ImportSeed
#pragma unroll
    for (int e = 0; e < 6; ++e) st.par[e] = a.seed_par[e * N + seed];
#pragma unroll
    for (int e = 0; e < 21; ++e) st.err[e] = a.seed_err[e * N + seed];
    st.chg = a.seed_chg[seed];
    st.chi2 = 0.f;
    st.nfound = st.nmissing = st.ninside = st.ntail = 0;
    st.last = -1;
    st.last_hit = -1;
    for (int l = 0; l < a.first_layer; ++l) {
      const int h = a.seed_hits[l * N + seed];
      const int k = append_node(a, seed, &s_nnodes[g], h, l, st.last);
      if (k >= 0) st.last = k;
      add_hit_counts(st, h, 0.f);
    }
    st.score = cand_score(st);
  }

  for (int layer = a.first_layer; layer < a.n_layers; ++layer) {
    const LayerDev& d = lays[layer];
    const bool barrel = d.is_barrel != 0;
    // ---- (a) find_tracks_unroll_candidates
    if (c == 0 && gvalid) {
      if (s_state[g] == 0 && layer - 1 == a.first_layer - 1) s_state[g] = 1;
      s_any[g] = 0;
    }
    ecnt[lane] = 0;
    fcode[lane] = 0;
    xkind[lane] = 0;
    __syncthreads();
    const bool finding = gvalid && s_state[g] == 1;
    const bool valid = finding && c < s_ncand[g];
    bool active = valid && st.last_hit != -2;
    if (active && fabsf(1.f / st.par[3]) < 1.2f) {          // near the apex of a low-pT helix
      const float dphi = fabsf(atan2_cr(st.par[1], st.par[0]) - st.par[4]);
      if (st.par[0] * st.par[0] + st.par[1] * st.par[1] > 625.f && dphi > 1.371f && dphi < 4.512f) {
        const int k = append_node(a, seed, &s_nnodes[g], -2, layer, st.last);
        if (k >= 0) st.last = k;
        add_hit_counts(st, -2, 0.f);
        active = false;
      }
    }
    if (active) s_any[g] = 1;
    __syncthreads();
    if (c == 0) s_merge[g] = finding && s_any[g];
    if (c == 0 && finding && !s_any[g]) s_state[g] = 2;
    // FindCandidatesCloneEngine (MkFinder.cc:1114-1284) in three phases, so the
    // costly part -- the propagation to each window hit and its chi2 -- runs
    // on a compacted list of (lane, hit) pairs with every lane of the
    // wavefront busy, however unevenly the hits fall into the lanes' windows:
    //   (a1) each lane propagates to the layer, builds its window and lists
    //        the window's hits (at most 16, in scan order);
    //   (a2) a wavefront prefix sum of the list lengths numbers all pairs, and
    //        lane i evaluates pairs i, i + 64, ... for their owners;
    //   (a3) each lane turns its accepted hits into records, in scan order.
    int nh = 0;
    float pt_l = 0.f;
    bool search = false, gap = false;
    if (active) {
      float pp[6], pe[21];
#pragma unroll
      for (int e = 0; e < 6; ++e) pp[e] = st.par[e];
#pragma unroll
      for (int e = 0; e < 21; ++e) pe[e] = st.err[e];
      propagate(barrel, pp, pe, st.chg, d.propagate_to, a.inter_pflags, a);
      Window w;
      hit_window(d, pp, pe, w);
      // find_tracks_handle_missed_layers (MkBuilder.cc:1568-1630)
      if (barrel && hypot_ref(pp[0], pp[1]) < d.rin) {
        const int k = append_node(a, seed, &s_nnodes[g], -2, layer, st.last);
        if (k >= 0) st.last = k;
        add_hit_counts(st, -2, 0.f);
        xkind[lane] = 2;
      } else if (w.wsr == MKF_WSR_OUTSIDE) {
        xkind[lane] = 1;
      } else {
        search = true;
        pt_l = fabsf(1.0f / pp[3]);
        nh = scan_window(d, w, a.bins, a.phis, a.qs, a.mask, [&](int gi) { Gh[lane][nh++] = gi - d.hit_off; });
        float* xs = &X[lane * kW];                       // words 0-35 are free until the end of (a)
#pragma unroll
        for (int e = 0; e < 6; ++e) xs[e] = pp[e];
#pragma unroll
        for (int e = 0; e < 21; ++e) xs[6 + e] = pe[e];
        xs[27] = __int_as_float(st.chg);
        xs[kXCut] = dyn_chi2_cut(d, a.chi2_cut, pp[3], pp[5]);
        xs[kXOff] = __int_as_float(d.hit_off);
        xs[kXBar] = __int_as_float(barrel ? 1 : 0);
        if (a.prop_to_hit) {                             // sin, cos of phi and theta for every pair's propagation
          sincosf(pp[4], &xs[kXTrig], &xs[kXTrig + 1]);
          sincosf(pp[5], &xs[kXTrig + 2], &xs[kXTrig + 3]);
        }
        int fake = (st.ninside + st.ntail < a.max_holes && st.ntail < a.max_consec) ? -1 : -2;
        if (w.wsr == MKF_WSR_EDGE) fake = -3;
        else if (a.dead) {          // XWsrResult::m_in_gap: a dead bin of the central q row in the window (MkFinder.cc:451-462)
          const unsigned char* __restrict__ row = a.dead + d.bin_off + q_bin_checked(d, w.q) * kNPhi;
          for (int pi = w.pb1; pi < w.pb2 && !gap; ++pi) gap = row[pi & 0xff] != 0;
        }
        const int nt = fake == -1 ? st.ntail + 1 : st.ntail;
        fcode[lane] = fake;
        X[lane * kW + kXScore] = score_calc(st.nfound, nt, 0, st.ninside, st.chi2 < 0 ? 0.f : st.chi2, pt_l, true);
      }
    }
    // (a2) inclusive prefix sum of nh over the wavefront (the workgroup is one wavefront)
    int incl = nh;
#pragma unroll
    for (int o = 1; o < kLanes; o <<= 1) {
      const int v = __shfl_up(incl, o, kLanes);
      if (lane >= o) incl += v;
    }
    X[lane * kW + kXPre] = __int_as_float(incl - nh);
    const int total = __shfl(incl, kLanes - 1, kLanes);
    // Each pair's owner lane, one byte per pair (at most kLanes * kHitIdxMax), in the score
    // rows, which (a3) fills only after this phase.
    unsigned char* owner = reinterpret_cast<unsigned char*>(&Es[0][0]);
    static_assert(sizeof(Es) >= kLanes * kHitIdxMax, "owner table fits the score rows");
    for (int k = 0; k < nh; ++k) owner[incl - nh + k] = (unsigned char)lane;
    __syncthreads();
    for (int j = lane; j < total; j += kLanes) {
      const int lo = owner[j];
      const float* xs = &X[lo * kW];
      const int o = lo, k = j - __float_as_int(xs[kXPre]);
      // The owner's layer: in a batch of events, lanes of other seeds hold other events' layers.
      const bool ob = __float_as_int(xs[kXBar]) != 0;
      float mp[3], me[6];
      load_hit9(a.shits + (size_t)(__float_as_int(xs[kXOff]) + Gh[o][k]) * 9, mp, me);
      float q[6], qe[21], pe[21];
#pragma unroll
      for (int e = 0; e < 6; ++e) q[e] = xs[e];
#pragma unroll
      for (int e = 0; e < 21; ++e) qe[e] = pe[e] = xs[6 + e];
      if (a.prop_to_hit)
        propagate(ob, q, qe, __float_as_int(xs[27]), ob ? hypot_ref(mp[0], mp[1]) : mp[2], a.intra_pflags, a,
                  &xs[kXTrig]);
      const float c2 = fabsf(ob ? kalman_barrel<true, false>(q, qe, mp, me)
                                : kalman_endcap<true, false>(q, qe, mp, me));
      const int gi = __float_as_int(xs[kXOff]) + Gh[o][k];
      const bool acc = c2 < xs[kXCut] && strip_q_compatible(ob, pe, q, mp, me) &&
                       (!a.charge || pass_strip_charge(ob, a.charge[gi], q, me));
      Gc[o][k] = acc ? c2 : -1.f;
      if (STD && acc) {
        // getScoreCand of the new candidate takes the pT of its updated state (m_update_param_foo)
        if (ob)
          kalman_barrel<false, true>(q, qe, mp, me);
        else
          kalman_endcap<false, true>(q, qe, mp, me);
        Gp[o][k] = fabsf(1.0f / q[3]);
      }
    }
    __syncthreads();
    // (a3) records of the accepted hits, in scan order
    if (search) {
      int ne = 0, nacc = 0;
      for (int hc = 0; hc < nh; ++hc) {
        const float c2 = Gc[lane][hc];
        if (!(c2 >= 0.f)) continue;
        ++nacc;
        const float tot = st.chi2 + c2;
        const float sc = score_calc(st.nfound + 1, 0, 0, st.ninside + st.ntail, tot < 0 ? 0.f : tot,
                                    STD ? Gp[lane][hc] : pt_l, true);
        const int rec = hc << kPosBits | Gh[lane][hc];
        // Keep only this lane's best K records, in hit order.  The seed's
        // top K (by score, earlier generation first on ties) never takes a
        // record that K records of its own lane beat, so the merge below
        // sees every record it can select.  Worst kept = lowest score, the
        // latest of equals; the new record comes after all kept ones, so it
        // replaces the worst only with a strictly higher score.  The kept
        // records go to the front of Gh / Gc in place: slot ne <= hc, and
        // every slot past hc is still unread.
        if (ne < K) {
          Es[lane][ne] = sc;
          Gh[lane][ne] = rec;
          Gc[lane][ne] = c2;
          ++ne;
        } else {
          int wk = 0;
          float ws = Es[lane][0];
          for (int kk = 1; kk < K; ++kk)
            if (!(Es[lane][kk] > ws)) {
              wk = kk;
              ws = Es[lane][kk];
            }
          if (sc > ws) {
            for (int kk = wk; kk < K - 1; ++kk) {
              Es[lane][kk] = Es[lane][kk + 1];
              Gh[lane][kk] = Gh[lane][kk + 1];
              Gc[lane][kk] = Gc[lane][kk + 1];
            }
            Es[lane][K - 1] = sc;
            Gh[lane][K - 1] = rec;
            Gc[lane][K - 1] = c2;
          }
        }
      }
      ecnt[lane] = ne;
      if (gap && nacc == 0) {      // dead module, no hit taken: -7, which is no hole (MkFinder.cc:1092-1096, 1263-1267)
        fcode[lane] = -7;
        X[lane * kW + kXScore] = score_calc(st.nfound, st.ntail, 0, st.ninside, st.chi2 < 0 ? 0.f : st.chi2, pt_l, true);
      }
    }
    // STD: a searching lane's new candidates start from its state at the layer, which (a1) left in words
    // 0-26; the others (held back, stopped, idle) keep their own.
    if (STD && search)
      put_counts(&X[lane * kW], st);
    else
      put(&X[lane * kW], st);
    __syncthreads();
    // ---- (b) CandCloner::ProcessSeedRange, one lane per seed
    if (c == 0 && s_merge[g]) {
      // The seed's best K so far, by score; a new entry goes after the kept
      // ones of equal score.  Entry info: kind | slot << 2 | v << 5 with v the
      // record in the slot lane's list (kind 2) or minus the hit code (kind 3).
      // Every array index is a compile-time constant, so the lists stay in
      // registers.
      float tsc[kS];
      int ti[kS];
      int nt = 0, nent = 0;
      auto insert = [&](float sc, int info) {
        ++nent;
        int pos = 0;                                  // kept entries with score >= sc
#pragma unroll
        for (int q = 0; q < kS; ++q) pos += (q < nt && !(tsc[q] < sc)) ? 1 : 0;
        if (pos >= K) return;
#pragma unroll
        for (int q = kS - 1; q > 0; --q)
          if (q > pos) {
            tsc[q] = tsc[q - 1];
            ti[q] = ti[q - 1];
          }
#pragma unroll
        for (int q = 0; q < kS; ++q)
          if (q == pos) {
            tsc[q] = sc;
            ti[q] = info;
          }
        if (nt < K) ++nt;
      };
      // STD: the first entry with a -2 hit in ranked order (highest score, earliest generated among equals)
      // and its generation rank: held back 0..kS-1, records kS + hit_cnt * kS + slot, invalid-hit entries
      // after every record.
      float bsc = kScoreWorst;
      int binfo = -1, bgen = 0;
      auto consider_short = [&](float sc, int info, int gen) {
        if (binfo < 0 || sc > bsc) {
          bsc = sc;
          binfo = info;
          bgen = gen;
        }
      };
      if (STD) {              // find_tracks_handle_missed_layers runs before FindCandidates
        for (int s = 0; s < kS; ++s) {
          const int l = g * kS + s;
          if (xkind[l] == 1) insert(X[l * kW + 29], 1 | s << 2);
          else if (xkind[l] == 2) consider_short(X[l * kW + 29], 1 | s << 2, s);
        }
      }
      // Generation order: hit_cnt, then lane.  Each slot's records are in
      // increasing hit_cnt, so the records are merged by always taking the
      // slot whose next record has the lowest hit_cnt (the lowest slot among
      // equals); a slot with none left reads kHitIdxMax.
      int ptr[kS], cur[kS];
#pragma unroll
      for (int s = 0; s < kS; ++s) {
        const int l = g * kS + s;
        ptr[s] = 0;
        cur[s] = ecnt[l] > 0 ? (int)((unsigned)Gh[l][0] >> kPosBits) : kHitIdxMax;
      }
      while (true) {
        int sb = 0, hb = cur[0], pb = ptr[0];
#pragma unroll
        for (int s = 1; s < kS; ++s)
          if (cur[s] < hb) {
            hb = cur[s];
            sb = s;
            pb = ptr[s];
          }
        if (hb >= kHitIdxMax) break;
        const int l = g * kS + sb;
        insert(Es[l][pb], 2 | sb << 2 | pb << 5);
        const int nx = pb + 1 < ecnt[l] ? (int)((unsigned)Gh[l][pb + 1] >> kPosBits) : kHitIdxMax;
#pragma unroll
        for (int s = 0; s < kS; ++s)
          if (s == sb) {
            ptr[s] = pb + 1;
            cur[s] = nx;
          }
      }
      for (int s = 0; s < kS; ++s) {
        const int l = g * kS + s;
        if (STD && fcode[l] == -2)
          consider_short(X[l * kW + kXScore], 3 | s << 2 | 2 << 5, kS + kHitIdxMax * kS + s);
        else if (fcode[l])
          insert(X[l * kW + kXScore], 3 | s << 2 | -fcode[l] << 5);
      }
      if (STD) {
        // MkBuilder.cc:1776-1828: the kept entries in ranked order become the next candidates; the first -2
        // entry is met if it ranks before the K-th of them
        int np = 0;
        float sK = 0.f;
        int iK = 0;
        for (int ih = 0; ih < nt; ++ih) {
          float sc = tsc[0];
          int info = ti[0];
#pragma unroll
          for (int q = 1; q < kS; ++q)
            if (q == ih) {
              sc = tsc[q];
              info = ti[q];
            }
          D[g][np++] = (info & 3) == 2 ? info : (info & 31);
          sK = sc;
          iK = info;
        }
        if (binfo >= 0) {
          bool met = nt < K;
          if (!met) {
            const int kind = iK & 3, slot = (iK >> 2) & 7;
            const int gen = kind == 1 ? slot
                          : kind == 2 ? kS + (int)((unsigned)Gh[g * kS + slot][iK >> 5] >> kPosBits) * kS + slot
                                      : kS + kHitIdxMax * kS + slot;
            met = bsc > sK || (bsc == sK && bgen < gen);
          }
          if (met && bsc > s_bsc[g]) {                   // best short candidate
            Cand b;
            get(&X[(g * kS + ((binfo >> 2) & 7)) * kW], b);
            if ((binfo & 3) == 3) {                      // an invalid-hit entry; a stopped replica has its -2
              const int k = append_node(a, seed, &s_nnodes[g], -2, layer, b.last);
              if (k >= 0) b.last = k;
              add_hit_counts(b, -2, 0.f);
            }
            b.score = bsc;
            put(a.bs + (size_t)seed * kBsW, b);
            s_bsc[g] = bsc;
            s_bs[g] = 1;
          }
        }
        if (nent == 0 && binfo < 0)                      // no entry at all: the candidates stay
          for (int i = 0; i < s_ncand[g]; ++i) D[g][np++] = 1 | i << 2;
        s_nnew[g] = np;
      } else {
      int np = 0;
      unsigned xm = 0;                                // held-back slots, lowest first
      for (int s = 0; s < kS; ++s)
        if (xkind[g * kS + s]) xm |= 1u << s;
      if (nent == 0) {
        for (; xm; xm &= xm - 1) {
          D[g][np] = 1 | (__ffs(xm) - 1) << 2;
          ++np;
        }
      } else {
        for (int ih = 0; ih < nt; ++ih) {
          float sc = tsc[0];
          int info = ti[0];
#pragma unroll
          for (int q = 1; q < kS; ++q)
            if (q == ih) {
              sc = tsc[q];
              info = ti[q];
            }
          const int kind = info & 3, slot = (info >> 2) & 7;
          if (kind == 3 && (info >> 5) == 2) {        // hit code -2: a short candidate
            if (sc > s_bsc[g]) {                      // best short candidate
              Cand b;
              get(&X[(g * kS + slot) * kW], b);
              const int k = append_node(a, seed, &s_nnodes[g], -2, layer, b.last);
              if (k >= 0) b.last = k;
              add_hit_counts(b, -2, 0.f);
              b.score = sc;
              put(a.bs + (size_t)seed * kBsW, b);
              s_bsc[g] = b.score;
              s_bs[g] = 1;
            }
            continue;
          }
          while (xm && X[(g * kS + __ffs(xm) - 1) * kW + 29] > sc && np < K) {
            D[g][np] = 1 | (__ffs(xm) - 1) << 2;
            ++np;
            xm &= xm - 1;
          }
          if (np >= K) break;
          D[g][np] = kind == 2 ? info : (info & 31);
          ++np;
        }
        for (; xm && np < K; xm &= xm - 1) {
          D[g][np] = 1 | (__ffs(xm) - 1) << 2;
          ++np;
        }
      }
      s_nnew[g] = np;
      }
    }
    __syncthreads();
    // ---- (c) new candidates; UpdateWithLastHit for those that took a hit
    if (s_merge[g] && gvalid) {
      if (c < s_nnew[g]) {
        const int dd = D[g][c];
        const int kind = dd & 3, l = g * kS + ((dd >> 2) & 7), ri = dd >> 5;
        get(&X[l * kW], st);
        if (kind != 1) {
          // kind 2: the slot lane's kept record ri; kind 3: its invalid-hit record
          const int val = kind == 2 ? d.hit_off + (Gh[l][ri] & kPosMask) : fcode[l];
          const int hit = kind == 2 ? a.ranks[val] : val;
          const int k = append_node(a, seed, &s_nnodes[g], hit, layer, st.last);
          if (k >= 0) st.last = k;
          add_hit_counts(st, hit, kind == 2 ? Gc[l][ri] : 0.f);
          st.score = kind == 2 ? Es[l][ri] : X[l * kW + kXScore];
          if (kind == 2) {
            float mp[3], me[6];
            load_hit9(a.shits + (size_t)val * 9, mp, me);
            if (a.prop_to_hit)
              propagate(barrel, st.par, st.err, st.chg, barrel ? hypot_ref(mp[0], mp[1]) : mp[2],
                        STD ? a.intra_pflags : a.inter_pflags, a);
            if (barrel)
              kalman_barrel<false, true>(st.par, st.err, mp, me);
            else
              kalman_endcap<false, true>(st.par, st.err, mp, me);
            if (st.par[3] < 0.f) {
              st.chg = -st.chg;
              st.par[3] = -st.par[3];
            }
          }
        }
      }
    }
    __syncthreads();
    if (c == 0 && s_merge[g]) s_ncand[g] = s_nnew[g];
  }

  // ---- MergeCandsAndBestShortOne(update_score, sort) and output
  __syncthreads();
  const bool valid = gvalid && c < s_ncand[g];
  if (valid) st.score = cand_score(st);
  put(&X[lane * kW], st);
  __syncthreads();
  int (*O)[kS + 1] = reinterpret_cast<int (*)[kS + 1]>(&Gh[0][0]);   // final order (kS = best short); Gh is free
  static_assert(kGA * (kS + 1) <= kLanes * kHitIdxMax, "the final order must fit in Gh");
  if (c == 0 && gvalid) {
    const int nc = s_ncand[g];
    int ord[kS + 1];
    int m = 0;
    for (int i = 0; i < nc; ++i) {                            // stable insertion sort by score
      const float sc = X[(g * kS + i) * kW + 29];
      int pos = m;
      while (pos > 0 && X[(g * kS + ord[pos - 1]) * kW + 29] < sc) {
        ord[pos] = ord[pos - 1];
        --pos;
      }
      ord[pos] = i;
      ++m;
    }
    if (s_bs[g]) {
      // The best short candidate's score is recomputed only alongside the
      // candidates' (HitStructures.cc:515-521); alone it keeps its finding score.
      Cand b;
      get(a.bs + (size_t)seed * kBsW, b);
      if (m > 0) b.score = cand_score(b);
      put(a.bs + (size_t)seed * kBsW, b);
      if (m > 0) {
        if (b.score > X[(g * kS + ord[m - 1]) * kW + 29]) {
          int ci = 0;
          while (X[(g * kS + ord[ci]) * kW + 29] > b.score) ++ci;
          if (m >= K) --m;
          for (int j = m; j > ci; --j) ord[j] = ord[j - 1];
          ord[ci] = kS;
          ++m;
        }
      } else {
        ord[m++] = kS;
      }
    }
    for (int i = 0; i < m; ++i) O[g][i] = ord[i];
    s_nout[g] = m;
    a.n_cands[seed] = m;
  } else if (c == 0 && gin) {      // layer_base outside the event of hits: no candidates
    s_nout[g] = 0;
    a.n_cands[seed] = 0;
  }
  __syncthreads();
  if (!gin || c >= K) return;
  const int L = a.n_layers;
  int* hot = a.hot + ((size_t)c * L) * N + seed;
  for (int l = 0; l < L; ++l) hot[(size_t)l * N] = -9;
  if (c >= s_nout[g]) {
    a.score[c * N + seed] = kScoreWorst;
    a.n_found[c * N + seed] = 0;
    return;
  }
  Cand o;
  get(O[g][c] == kS ? a.bs + (size_t)seed * kBsW : &X[(g * kS + O[g][c]) * kW], o);
#pragma unroll
  for (int e = 0; e < 6; ++e) a.par[((size_t)c * 6 + e) * N + seed] = o.par[e];
#pragma unroll
  for (int e = 0; e < 21; ++e) a.err[((size_t)c * 21 + e) * N + seed] = o.err[e];
  a.chg[c * N + seed] = o.chg;
  a.chi2[c * N + seed] = o.chi2;
  a.score[c * N + seed] = o.score;
  a.n_found[c * N + seed] = o.nfound;
  const int4* nodes = a.nodes + (size_t)seed * a.node_cap;
  for (int k = o.last, steps = 0; k >= 0 && steps < a.node_cap; k = nodes[k].z, ++steps) {
    const int4 nd = nodes[k];
    if (nd.y >= 0 && nd.y < L && hot[(size_t)nd.y * N] == -9) hot[(size_t)nd.y * N] = nd.x;
  }
}

}  // namespace

namespace {

int find_combinatorial(mkf_context* c, mkf_event_of_hits* h, const mkf_ce_batch* b, void* stream, bool standard) {
  CHECK(set_device(c));
  if (!h || !b) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  if (h->ctx != c) return fail(MKF_ERR_INVALID_ARG, "event of hits belongs to another context");
  if (!h->loaded) return fail(MKF_ERR_INVALID_ARG, "event of hits not loaded");
  if (b->n_seeds < 0) return fail(MKF_ERR_INVALID_ARG, "n_seeds < 0");
  for (const LayerDev& ld : h->layers)                      // hit records carry 28-bit layer positions
    if (ld.n_hits > kPosMask) return fail(MKF_ERR_INVALID_ARG, "a layer holds %d hits, above %d", ld.n_hits, kPosMask);
  const int LE = b->n_layers ? b->n_layers : h->n_layers;   // layers per event
  if (LE < 1 || LE > h->n_layers || LE > MKF_MAX_FIND_LAYERS)
    return fail(MKF_ERR_INVALID_ARG, "n_layers %d outside [1, min(%d, %d)]", LE, h->n_layers, MKF_MAX_FIND_LAYERS);
  if (b->first_layer < 1 || b->first_layer > LE)
    return fail(MKF_ERR_INVALID_ARG, "first_layer %d outside [1, %d]", b->first_layer, LE);
  if (!b->layer_base && LE != h->n_layers)
    return fail(MKF_ERR_INVALID_ARG, "n_layers %d below the event of hits' %d needs layer_base", LE, h->n_layers);
  if (b->max_cands < 1 || b->max_cands > MKF_CE_MAX_CANDS)
    return fail(MKF_ERR_INVALID_ARG, "max_cands must be in [1, %d]", MKF_CE_MAX_CANDS);
  if (((b->pflags | b->intra_pflags) & MKF_PF_APPLY_MATERIAL) && !c->d_rl)
    return fail(MKF_ERR_INVALID_ARG, "material requested but the context has no material grid");
  if (b->n_seeds == 0) return MKF_OK;
  const size_t N = b->n_seeds, K = b->max_cands, L = LE;
  if (K * 21 * N * 4 >= ((size_t)1 << 31) || K * L * N >= ((size_t)1 << 31))
    return fail(MKF_ERR_INVALID_ARG, "batch too large");
  CHECK(check_dev(b->seed_par, 6 * N * 4, "seed_par"));
  CHECK(check_dev(b->seed_err, 21 * N * 4, "seed_err"));
  CHECK(check_dev(b->seed_chg, N * 4, "seed_chg"));
  CHECK(check_dev(b->seed_hits, (size_t)b->first_layer * N * 4, "seed_hits"));
  CHECK(check_dev(b->n_cands, N * 4, "n_cands"));
  CHECK(check_dev(b->par, K * 6 * N * 4, "par"));
  CHECK(check_dev(b->err, K * 21 * N * 4, "err"));
  CHECK(check_dev(b->chg, K * N * 4, "chg"));
  CHECK(check_dev(b->chi2, K * N * 4, "chi2"));
  CHECK(check_dev(b->score, K * N * 4, "score"));
  CHECK(check_dev(b->n_found, K * N * 4, "n_found"));
  CHECK(check_dev(b->hot_idx, K * L * N * 4, "hot_idx"));
  if (b->layer_base) CHECK(check_dev(b->layer_base, N * 4, "layer_base"));
  const int node_cap = b->first_layer + 3 * kSMax * (LE - b->first_layer) + kSMax;
  const size_t need = ((size_t)node_cap + kBsW / 4) * N;    // nodes, then the best short candidates
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(h->ce_mu);
  CHECK(scratch_acquire(h->ce_done, st));
  if (need > h->ce_node_cap) {
    HIP_TRY(hipStreamSynchronize(st));
    if (h->d_ce_nodes) HIP_TRY(hipFree(h->d_ce_nodes));
    h->d_ce_nodes = nullptr;
    h->ce_node_cap = 0;
    HIP_TRY(hipMalloc(&h->d_ce_nodes, need * sizeof(int4)));
    h->ce_node_cap = need;
  }
  CeArgs a;
  a.n = b->n_seeds;
  a.n_layers = LE;
  a.eoh_layers = h->n_layers;
  a.layer_base = b->layer_base;
  a.first_layer = b->first_layer;
  a.max_cands = b->max_cands;
  a.max_holes = b->max_holes;
  a.max_consec = b->max_consec_holes;
  a.prop_to_hit = b->prop_to_hit;
  a.intra_pflags = b->intra_pflags;
  a.inter_pflags = b->pflags;
  a.chi2_cut = b->chi2_cut;
  a.bfield = c->bfield;
  a.mat = mat_of(c);
  a.seed_par = b->seed_par; a.seed_err = b->seed_err; a.seed_chg = b->seed_chg; a.seed_hits = b->seed_hits;
  a.n_cands = b->n_cands; a.par = b->par; a.err = b->err; a.chg = b->chg; a.chi2 = b->chi2; a.score = b->score;
  a.n_found = b->n_found; a.hot = b->hot_idx;
  a.nodes = reinterpret_cast<int4*>(h->d_ce_nodes);
  a.node_cap = node_cap;
  a.bs = reinterpret_cast<float*>(a.nodes + (size_t)node_cap * N);
  a.layers = h->d_layers; a.bins = h->d_bins; a.phis = h->d_phis; a.qs = h->d_qs; a.shits = h->d_shits;
  a.mask = h->mask();
  a.charge = h->charge();
  a.dead = h->dead();
  a.ranks = h->d_ranks;
  const mkf_find_shape sh = c->find_shape[standard ? 1 : 0];
  const int slots = sh.slots ? sh.slots : (b->max_cands <= 6 ? 6 : 8);
  if (slots < b->max_cands) return fail(MKF_ERR_INVALID_ARG, "find shape: %d slots < max_cands %d", slots, b->max_cands);
  const int grid = (int)((N + kLanes / slots - 1) / (kLanes / slots));
  const int sel = (slots == 8 ? 4 : 0) | (standard ? 2 : 0) | (sh.waves_per_simd == 1 ? 1 : 0);
  switch (sel) {
#define MKF_CE_CASE(I, KS, STD, W)                                                                   \
    case I:                                                                                          \
      hipLaunchKernelGGL((k_find_clone_engine<KS, STD, W>), dim3(grid), dim3(kLanes), 0, st, a);     \
      break;
    MKF_CE_CASE(0, 6, false, 2) MKF_CE_CASE(1, 6, false, 1) MKF_CE_CASE(2, 6, true, 2) MKF_CE_CASE(3, 6, true, 1)
    MKF_CE_CASE(4, 8, false, 2) MKF_CE_CASE(5, 8, false, 1) MKF_CE_CASE(6, 8, true, 2) MKF_CE_CASE(7, 8, true, 1)
#undef MKF_CE_CASE
  }
  HIP_TRY(hipGetLastError());
  return scratch_release(h->ce_done, st);
}

}  // namespace

extern "C" int mkf_find_tracks_clone_engine(mkf_context* c, mkf_event_of_hits* h, const mkf_ce_batch* b,
                                            void* stream) {
  return find_combinatorial(c, h, b, stream, false);
}

extern "C" int mkf_find_tracks_standard(mkf_context* c, mkf_event_of_hits* h, const mkf_ce_batch* b, void* stream) {
  return find_combinatorial(c, h, b, stream, true);
}

extern "C" int mkf_set_find_shape(mkf_context* c, int standard, const mkf_find_shape* shape) {
  if (!c || !shape) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  if (shape->slots != 0 && shape->slots != 6 && shape->slots != 8)
    return fail(MKF_ERR_INVALID_ARG, "find shape: slots must be 0, 6 or 8");
  if (shape->waves_per_simd != 1 && shape->waves_per_simd != 2)
    return fail(MKF_ERR_INVALID_ARG, "find shape: waves_per_simd must be 1 or 2");
  std::lock_guard<std::mutex> lk(c->mu);
  c->find_shape[standard ? 1 : 0] = *shape;
  return MKF_OK;
}

extern "C" int mkf_get_find_shape(mkf_context* c, int standard, mkf_find_shape* shape) {
  if (!c || !shape) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  *shape = c->find_shape[standard ? 1 : 0];
  return MKF_OK;
}

// Every shape valid for the batch's max_cands, timed on the batch: one warm-up launch of each, then reps
// timed launches; the fastest stays in the context.  Results do not depend on the shape.
extern "C" int mkf_autotune_find(mkf_context* c, mkf_event_of_hits* h, const mkf_ce_batch* b, int standard, int reps,
                                 mkf_find_shape* best, float* best_ms, void* stream) {
  if (!c || !b) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  if (reps < 1 || reps > 100) return fail(MKF_ERR_INVALID_ARG, "reps must be in [1, 100]");
  CHECK(set_device(c));
  hipStream_t st = (hipStream_t)stream;
  std::vector<mkf_find_shape> cand;
  for (int slots : {6, 8})
    if (slots >= b->max_cands)
      for (int w : {2, 1}) cand.push_back(mkf_find_shape{slots, w});
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const mkf_find_shape keep = c->find_shape[standard ? 1 : 0];
  float best_t = 1e30f;
  mkf_find_shape best_s = keep;
  int rc = MKF_OK;
  for (const mkf_find_shape& s : cand) {
    c->find_shape[standard ? 1 : 0] = s;
    rc = find_combinatorial(c, h, b, stream, standard != 0);                  // warm-up
    if (rc != MKF_OK) break;
    if (hipEventRecord(e0, st) != hipSuccess) { rc = fail(MKF_ERR_HIP, "hipEventRecord"); break; }
    for (int r = 0; r < reps && rc == MKF_OK; ++r) rc = find_combinatorial(c, h, b, stream, standard != 0);
    if (rc != MKF_OK) break;
    float ms = 0.f;
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = fail(MKF_ERR_HIP, "timing the find shapes");
      break;
    }
    ms /= reps;
    if (ms < best_t) {
      best_t = ms;
      best_s = s;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->find_shape[standard ? 1 : 0] = rc == MKF_OK ? best_s : keep;
  if (rc != MKF_OK) return rc;
  if (best) *best = best_s;
  if (best_ms) *best_ms = best_t;
  return MKF_OK;
}
