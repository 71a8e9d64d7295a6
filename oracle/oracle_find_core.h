// oracle_find_core.h -- LayerOfHits and SelectHitIndices pieces shared by the
// best-hit (oracle_find.cc) and clone-engine (oracle_ce.cc) restatements.
// TEST INFRASTRUCTURE ONLY.  Reference line citations are on each function.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "oracle.h"
#include "oracle_core.h"

namespace orc_find {

using namespace orc;

constexpr int kNPhi = 256;                                    // Config::m_nphi
constexpr float kFPhiFine = 1024 / kTwoPI;                    // HitStructures.h:153-158
constexpr int kPhiMask = 0xff;
constexpr int kPhiMaskFine = 0x3ff;
constexpr int kPhiBitsShift = 2;
constexpr int kPhiFineXMask = ~((1 << kPhiBitsShift) - 1);
constexpr int kHitIdxMax = 16;                                // MPlexHitIdxMax, MkFitter.h:18
constexpr int kMaxTrkHits = 64;                               // Config::nMaxTrkHits
enum { WSR_Inside = 0, WSR_Edge = 1, WSR_Outside = 2 };

struct Bins {
  int nq;
  float qmin, qmax, fq;
};

// setup_bins (HitStructures.cc:11-32)
inline Bins setup_bins(float qmin, float qmax, float dq) {
  Bins b;
  if (dq < 0) {
    b.nq = (int)-dq;
    b.qmin = qmin;
    b.qmax = qmax;
  } else {
    const float extent = qmax - qmin;
    b.nq = (int)std::ceil(extent / dq);
    const float extra = 0.5f * (b.nq * dq - extent);
    b.qmin = qmin - extra;
    b.qmax = qmax + extra;
  }
  b.fq = b.nq / (qmax - qmin);
  return b;
}

inline Bins layer_bins(const orc_layer_desc& d) {
  return d.is_barrel ? setup_bins(d.zmin, d.zmax, d.q_bin) : setup_bins(d.rin, d.rout, d.q_bin);
}

inline int q_bin(const Bins& b, float q) { return (int)((q - b.qmin) * b.fq); }
inline int q_bin_checked(const Bins& b, float q) { return std::clamp(q_bin(b, q), 0, b.nq - 1); }
inline int phi_bin_fine(float phi) { return (int)std::floor(kFPhiFine * (phi + kPI)); }
inline int phi_bin(float phi) { return phi_bin_fine(phi) >> kPhiBitsShift; }
inline float cdist(float a) { return a > kPI ? kTwoPI - a : a; }   // Config.h:441

struct LoH {
  Bins b;
  std::vector<int> first, end;  // [nq * 256]
  std::vector<int> ranks;       // sorted position -> original hit index
  std::vector<float> phis, qs;  // in sorted order
  const unsigned char* mask = nullptr;   // per original hit: taken by an earlier iteration, or NULL
  const unsigned char* charge = nullptr; // per original hit: the strip charge code, or NULL (cut passes)
  std::vector<unsigned char> dead;       // [nq * 256] dead-module bins; empty: none
};

// LayerOfHits::SuckInDeads (HitStructures.cc:186-212) of the regions [n][5] {layer, phi1, phi2, q1, q2} that
// name `layer`: their (q bin, phi bin) rectangles, the phi range wrapping when it crosses -pi.  A phi bin of
// 256 (phi2 = +pi) stands for bin 0.
inline void suck_in_deads(LoH& L, int layer, int n, const float* regions) {
  for (int r = 0; r < n; ++r) {
    const float* g = regions + (size_t)r * 5;
    if ((int)g[0] != layer) continue;
    if (L.dead.empty()) L.dead.assign((size_t)L.b.nq * kNPhi, 0);
    const int q1 = q_bin_checked(L.b, g[3]), q2 = q_bin_checked(L.b, g[4]) + 1;
    const int p1 = phi_bin(g[1]), p2 = phi_bin(g[2]) + 1;
    for (int q = q1; q < q2; ++q) {
      if (p1 > p2) {
        for (int p = p1; p < kNPhi; ++p) L.dead[(size_t)q * kNPhi + p] = 1;
        for (int p = 0; p < p2; ++p) L.dead[(size_t)q * kNPhi + (p & kPhiMask)] = 1;
      } else {
        for (int p = p1; p < p2; ++p) L.dead[(size_t)q * kNPhi + (p & kPhiMask)] = 1;
      }
    }
  }
}

// SuckInHits (HitStructures.cc:77-183); the LSD radix sort's ranks are the
// stable order of the keys.
inline void build_loh(const orc_layer_desc& d, int n, const float* hits, LoH& L) {
  L.b = layer_bins(d);
  std::vector<uint32_t> key(n);
  std::vector<float> phi(n), q(n);
  for (int i = 0; i < n; ++i) {
    const float* h = hits + (size_t)i * 9;
    phi[i] = std::atan2(h[1], h[0]);                              // Hit::phi, Hit.h:64-67
    q[i] = d.is_barrel ? h[2] : std::sqrt(h[0] * h[0] + h[1] * h[1]);  // Hit::z / Hit::r
    key[i] = (uint32_t)(phi_bin_fine(phi[i]) + (q_bin_checked(L.b, q[i]) << 16));
  }
  L.ranks.resize(n);
  std::iota(L.ranks.begin(), L.ranks.end(), 0);
  std::stable_sort(L.ranks.begin(), L.ranks.end(), [&](int a, int c) { return key[a] < key[c]; });
  L.first.assign((size_t)L.b.nq * kNPhi, 0);
  L.end.assign((size_t)L.b.nq * kNPhi, 0);
  L.phis.resize(n);
  L.qs.resize(n);
  int curr = -1;
  for (int i = 0; i < n; ++i) {
    const int j = L.ranks[i];
    L.phis[i] = phi[j];
    L.qs[i] = q[j];
    const int jq = (int)key[j] & kPhiFineXMask;
    const int pb = (jq & kPhiMaskFine) >> kPhiBitsShift;
    const int qb = jq >> 16;
    const size_t bi = (size_t)qb * kNPhi + pb;
    if (jq != curr) {
      L.first[bi] = i;
      L.end[bi] = i;
      curr = jq;
    }
    L.end[bi]++;
  }
}

inline int wsr_z(const orc_layer_desc& d, float z, float dz) {      // TrackerInfo.h:118-123
  if (z > d.zmax + dz || z < d.zmin - dz) return WSR_Outside;
  if (z < d.zmax - dz && z > d.zmin + dz) return WSR_Inside;
  return WSR_Edge;
}
inline int wsr_r(const orc_layer_desc& d, float r, float dr) {      // TrackerInfo.h:125-138 (no r hole)
  if (r > d.rout + dr || r < d.rin - dr) return WSR_Outside;
  if (r < d.rout - dr && r > d.rin + dr) return WSR_Inside;
  return WSR_Edge;
}

// getHitSelDynamicWindows (MkFinder.cc:166-195)
inline void dyn_windows(const orc_layer_desc& d, float invpt, float theta, float& min_dq, float& max_dq, float& min_dphi,
                 float& max_dphi) {
  float max_invpt = invpt;
  if (invpt > 10.0) max_invpt = 10.0;
  const float this_dq = d.c_dq_0 * max_invpt + d.c_dq_1 * theta + d.c_dq_2;
  if (d.c_dq_sf * this_dq > 0.f) {
    min_dq = d.c_dq_sf * this_dq;
    max_dq = 2.0f * min_dq;
  }
  const float this_dphi = d.c_dp_0 * max_invpt + d.c_dp_1 * theta + d.c_dp_2;
  if (d.c_dp_sf * this_dphi > min_dphi) {
    min_dphi = d.c_dp_sf * this_dphi;
    max_dphi = 2.0f * min_dphi;
  }
}

// getHitSelDynamicChi2Cut (MkFinder.cc:202-220)
inline float dyn_chi2_cut(const orc_layer_desc& d, float chi2_cut_min, float invpt, float theta_in) {
  const float theta = std::abs(theta_in - kPIOver2);
  float max_invpt = invpt;
  if (invpt > 10.0) max_invpt = 10.0;
  const float this_c2 = d.c_c2_0 * max_invpt + d.c_c2_1 * theta + d.c_c2_2;
  if (d.c_c2_sf * this_c2 > chi2_cut_min) return d.c_c2_sf * this_c2;
  return chi2_cut_min;
}

// Decision margins of one lane's SelectHitIndices (for the checker, not the reference): how close any hit of
// the visited bins came to a window edge (cm; the phi edge as arc length at the layer), and how close a window
// limit or a visited hit's phi came to a bin edge it could cross (fraction of a q or coarse phi bin; the fine
// phi edges count when the list reached its 16-hit cap, since they order the hits within a bin).
struct SelMargins {
  float win = 1e30f, bin = 1e30f;
};
inline float bin_edge_margin(float x) { return std::abs(x - std::round(x)); }

// SelectHitIndices for NN lanes of propagated state (MkFinder.cc:227-684).  With `mg`, also the lanes'
// decision margins.
// in_gap, when given: per lane 1 if a dead bin of the window's central q row lies in its phi range
// (XWsrResult::m_in_gap, MkFinder.cc:451-462).
inline void select_hit_indices(const orc_layer_desc& d, const LoH& L, const float* par, const float* err, int n_proc,
                        int* xsize, int* xarr /*[16][NN]*/, int* wsr, SelMargins* mg = nullptr,
                        int* in_gap = nullptr) {
  float dqv[NN], dphiv[NN], qv[NN], phiv[NN];
  int qb1v[NN], qb2v[NN], pb1v[NN], pb2v[NN];
  const float nSigmaPhi = 3, nSigmaZ = 3, nSigmaR = 3;
  for (int it = 0; it < NN; ++it) {
    xsize[it] = 0;
    float min_dq = d.min_dq, max_dq = d.max_dq, min_dphi = d.min_dphi, max_dphi = d.max_dphi;
    const float invpt = par[3 * NN + it];
    const float theta = std::fabs(par[5 * NN + it] - kPIOver2);
    dyn_windows(d, invpt, theta, min_dq, max_dq, min_dphi, max_dphi);
    const float x = par[0 * NN + it], y = par[1 * NN + it];
    const float r2 = x * x + y * y;
    const float dphidx = -y / r2, dphidy = x / r2;
    const float dphi2 = dphidx * dphidx * err[sym(0, 0) * NN + it] + dphidy * dphidy * err[sym(1, 1) * NN + it] +
                        2 * dphidx * dphidy * err[sym(0, 1) * NN + it];
    const float phi = std::atan2(y, x);
    float dphi = std::max(nSigmaPhi * std::sqrt(std::abs(dphi2)), min_dphi);
    float q, dq;
    if (d.is_barrel) {
      const float z = par[2 * NN + it];
      dq = std::abs(nSigmaZ * std::sqrt(err[sym(2, 2) * NN + it]));
      const float edgeCorr = std::abs(0.5f * (d.rout - d.rin) / std::tan(par[5 * NN + it]));
      wsr[it] = wsr_z(d, z, std::sqrt(dq * dq + edgeCorr * edgeCorr));
      q = z;
    } else {
      const float r = std::sqrt(r2);
      dq = nSigmaR * std::sqrt(std::abs(x * x * err[sym(0, 0) * NN + it] + y * y * err[sym(1, 1) * NN + it] +
                                        2 * x * y * err[sym(0, 1) * NN + it]) / r2);
      const float edgeCorr = std::abs(0.5f * (d.zmax - d.zmin) * std::tan(par[5 * NN + it]));
      wsr[it] = wsr_r(d, r, std::sqrt(dq * dq + edgeCorr * edgeCorr));
      q = r;
    }
    // assignbins
    dphi = std::clamp(std::abs(dphi), min_dphi, max_dphi);
    dq = std::clamp(dq, min_dq, max_dq);
    qv[it] = q;
    phiv[it] = phi;
    dphiv[it] = dphi;
    dqv[it] = dq;
    qb1v[it] = q_bin_checked(L.b, q - dq);
    qb2v[it] = q_bin_checked(L.b, q + dq) + 1;
    pb1v[it] = phi_bin(phi - dphi);
    pb2v[it] = phi_bin(phi + dphi) + 1;
  }
  for (int it = 0; it < n_proc; ++it) {
    if (in_gap) in_gap[it] = 0;
    if (wsr[it] == WSR_Outside) {
      xsize[it] = -1;
      continue;
    }
    const float q = qv[it], phi = phiv[it], dphi = dphiv[it], dq = dqv[it];
    if (in_gap && !L.dead.empty()) {
      const int qc = q_bin_checked(L.b, q);
      for (int pi = pb1v[it]; pi < pb2v[it] && !in_gap[it]; ++pi)
        in_gap[it] = L.dead[(size_t)qc * kNPhi + (pi & kPhiMask)];
    }
    if (mg) {
      const float r = std::sqrt(par[0 * NN + it] * par[0 * NN + it] + par[1 * NN + it] * par[1 * NN + it]);
      SelMargins& m = mg[it];
      m.bin = std::min({bin_edge_margin((q - dq - L.b.qmin) * L.b.fq), bin_edge_margin((q + dq - L.b.qmin) * L.b.fq),
                        bin_edge_margin(kFPhiFine * (phi - dphi + kPI) / 4),
                        bin_edge_margin(kFPhiFine * (phi + dphi + kPI) / 4)});
      for (int qi = qb1v[it]; qi < qb2v[it]; ++qi)
        for (int pi = pb1v[it]; pi < pb2v[it]; ++pi) {
          const size_t bi = (size_t)qi * kNPhi + (pi & kPhiMask);
          for (int hi = L.first[bi]; hi < L.end[bi]; ++hi) {
            const float ddq = std::abs(q - L.qs[hi]);
            const float ddphi = cdist(std::abs(phi - L.phis[hi]));
            m.win = std::min({m.win, std::abs(ddq - dq), std::abs(ddphi - dphi) * r});
            m.bin = std::min(m.bin, bin_edge_margin(kFPhiFine * (L.phis[hi] + kPI) / 4));   // coarse edge
          }
          // A hit at phi = +pi has fine bin 1024: it sorts after the row's last bin but its bin index wraps to
          // 0, so it rewrites bin 0's range (SuckInHits, HitStructures.cc:146-171).  A hit within rounding of
          // +pi in the row's last bin therefore decides what bin 0 holds.
          if ((pi & kPhiMask) == 0) {
            const size_t bl = (size_t)qi * kNPhi + kPhiMask;
            for (int hi = L.first[bl]; hi < L.end[bl]; ++hi)
              m.bin = std::min(m.bin, std::abs(kFPhiFine * (L.phis[hi] + kPI) - 1024.f) / 4);
          }
        }
    }
    for (int qi = qb1v[it]; qi < qb2v[it]; ++qi)
      for (int pi = pb1v[it]; pi < pb2v[it]; ++pi) {
        const int pb = pi & kPhiMask;
        const size_t bi = (size_t)qi * kNPhi + pb;
        for (int hi = L.first[bi]; hi < L.end[bi]; ++hi) {
          if (L.mask && L.mask[L.ranks[hi]]) continue;      // the iteration's hit mask (MkFinder.cc:476-482)
          if (xsize[it] >= kHitIdxMax) break;
          const float ddq = std::abs(q - L.qs[hi]);
          const float ddphi = cdist(std::abs(phi - L.phis[hi]));
          if (ddq >= dq) continue;
          if (ddphi >= dphi) continue;
          xarr[xsize[it]++ * NN + it] = L.ranks[hi];
        }
      }
    if (mg && xsize[it] >= kHitIdxMax) {
      // a capped list also depends on the order within a bin (the fine phi bin): any fine edge matters
      for (int qi = qb1v[it]; qi < qb2v[it]; ++qi)
        for (int pi = pb1v[it]; pi < pb2v[it]; ++pi) {
          const size_t bi = (size_t)qi * kNPhi + (pi & kPhiMask);
          for (int hi = L.first[bi]; hi < L.end[bi]; ++hi)
            mg[it].bin = std::min(mg[it].bin, bin_edge_margin(kFPhiFine * (L.phis[hi] + kPI)) / 4);
        }
    }
  }
}

}  // namespace orc_find
