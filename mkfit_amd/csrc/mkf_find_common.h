// mkf_find_common.h -- layer table, hit-window selection and the device event
// of hits, shared by the best-hit (mkf_find.hip) and clone-engine (mkf_ce.hip)
// finding kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/mkfit_amd.h"
#include "mkf_track_math.h"

namespace mkf_find {

using namespace mkf;

constexpr int kNPhi = 256;                      // Config::m_nphi
constexpr float kFPhiFine = 1024 / kTwoPI;      // HitStructures.h:153-158 (float arithmetic)
constexpr float kPIOver2 = kPI / 2.0f;          // Config.h:79
constexpr int kHitIdxMax = MKF_HIT_IDX_MAX;
// Sort key: layer (11 bits) | q bin (10 bits) | fine phi bin (11 bits).
constexpr int kLayerShift = 21, kQShift = 11;
constexpr int kMaxQBins = 1 << 10;
static_assert(MKF_MAX_EOH_LAYERS <= (1 << (32 - kLayerShift)), "layer index must fit the sort key");

struct LayerDev {
  int is_barrel, nq, hit_off, n_hits, bin_off;
  float rin, rout, zmin, zmax, propagate_to;
  float qmin, fq;                               // bins after setup_bins (extended range, scale)
  float min_dphi, max_dphi, min_dq, max_dq;
  float c_dp_sf, c_dp_0, c_dp_1, c_dp_2;
  float c_dq_sf, c_dq_0, c_dq_1, c_dq_2;
  float c_c2_sf, c_c2_0, c_c2_1, c_c2_2;
};

// LayerOfHits::setup_bins (HitStructures.cc:11-32), same float arithmetic.
inline void setup_bins(float qmin, float qmax, float dq, int& nq, float& qmin_ext, float& fq) {
  if (dq < 0) {
    nq = (int)-dq;
    qmin_ext = qmin;
  } else {
    const float extent = qmax - qmin;
    nq = (int)std::ceil(extent / dq);
    const float extra = 0.5f * (nq * dq - extent);
    qmin_ext = qmin - extra;
  }
  fq = nq / (qmax - qmin);
}

__host__ __device__ inline int q_bin_checked(const LayerDev& d, float q) {   // HitStructures.h:237-245
  const int b = (int)((q - d.qmin) * d.fq);
  return b < 0 ? 0 : (b > d.nq - 1 ? d.nq - 1 : b);
}
// atan2 of float coordinates, correctly rounded to float (as the host's atan2f of the reference): the
// double-precision result rounded once.  Phi decides fine-bin keys and window bins, where the 1-ulp
// differences of a faster float atan2 change which bin a hit or a window edge falls in.
__device__ __forceinline__ float atan2_cr(float y, float x) { return (float)atan2((double)y, (double)x); }

__host__ __device__ inline int phi_bin_fine(float phi) { return (int)floorf(kFPhiFine * (phi + kPI)); }
__host__ __device__ inline int phi_bin(float phi) { return phi_bin_fine(phi) >> 2; }
__device__ __forceinline__ float cdist(float a) { return a > kPI ? kTwoPI - a : a; }   // Config.h:441

// ---- SelectHitIndices window (MkFinder.cc:166-195, 227-430) -----------------

struct Window {
  float q, phi, dq, dphi;
  int qb1, qb2, pb1, pb2, wsr;
};

__device__ __forceinline__ int wsr_z(const LayerDev& d, float z, float dz) {   // TrackerInfo.h:116-123
  if (z > d.zmax + dz || z < d.zmin - dz) return MKF_WSR_OUTSIDE;
  if (z < d.zmax - dz && z > d.zmin + dz) return MKF_WSR_INSIDE;
  return MKF_WSR_EDGE;
}
__device__ __forceinline__ int wsr_r(const LayerDev& d, float r, float dr) {   // TrackerInfo.h:125-138, no hole
  if (r > d.rout + dr || r < d.rin - dr) return MKF_WSR_OUTSIDE;
  if (r < d.rout - dr && r > d.rin + dr) return MKF_WSR_INSIDE;
  return MKF_WSR_EDGE;
}

__device__ __forceinline__ void hit_window(const LayerDev& d, const float (&par)[6], const float (&err)[21],
                                           Window& w) {
  float min_dq = d.min_dq, max_dq = d.max_dq, min_dphi = d.min_dphi, max_dphi = d.max_dphi;
  const float invpt = par[3];
  const float theta = fabsf(par[5] - kPIOver2);
  {                                              // getHitSelDynamicWindows
    const float max_invpt = invpt > 10.0f ? 10.0f : invpt;
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
  const float x = par[0], y = par[1];
  const float r2 = x * x + y * y;
  const float dphidx = -y / r2, dphidy = x / r2;
  const float dphi2 = dphidx * dphidx * err[S(0, 0)] + dphidy * dphidy * err[S(1, 1)] +
                      2 * dphidx * dphidy * err[S(0, 1)];
  const float phi = atan2_cr(y, x);
  float dphi = fmaxf(3.f * sqrtf(fabsf(dphi2)), min_dphi);
  float q, dq;
  if (d.is_barrel) {
    const float z = par[2];
    dq = fabsf(3.f * sqrtf(err[S(2, 2)]));
    const float edge = fabsf(0.5f * (d.rout - d.rin) / tanf(par[5]));
    w.wsr = wsr_z(d, z, sqrtf(dq * dq + edge * edge));
    q = z;
  } else {
    const float r = sqrtf(r2);
    dq = 3.f * sqrtf(fabsf(x * x * err[S(0, 0)] + y * y * err[S(1, 1)] + 2 * x * y * err[S(0, 1)]) / r2);
    const float edge = fabsf(0.5f * (d.zmax - d.zmin) * tanf(par[5]));
    w.wsr = wsr_r(d, r, sqrtf(dq * dq + edge * edge));
    q = r;
  }
  dphi = fminf(fmaxf(fabsf(dphi), min_dphi), max_dphi);
  dq = fminf(fmaxf(dq, min_dq), max_dq);
  w.q = q;
  w.phi = phi;
  w.dq = dq;
  w.dphi = dphi;
  w.qb1 = q_bin_checked(d, q - dq);
  w.qb2 = q_bin_checked(d, q + dq) + 1;
  w.pb1 = phi_bin(phi - dphi);
  w.pb2 = phi_bin(phi + dphi) + 1;
  // A window of at most one full turn spans at most kNPhi + 2 phi bins; wider
  // dynamic windows are cut there so the scan stays bounded.
  w.pb2 = min(w.pb2, w.pb1 + kNPhi + 2);
}

// Scan of the window's bins in the reference's order (q bins outer, phi bins
// inner, sorted hits within a bin), at most kHitIdxMax accepted hits; calls
// on_hit(global sorted position) for each (MkFinder.cc:433-684).
//
// Within a q row the hits of phi bins 1..255 are stored in bin order, and the
// bin table gives empty bins their insertion point, so each run of the
// window's bins that avoids bin 0 is one range [row[b].x, row[b'].y) read with
// two independent loads; bin 0 (whose range the phi == +pi run may replace)
// is read on its own.  The hits visited, and their order, are the bin-by-bin
// scan's.
constexpr int kScanUnroll = 8;
// mask: per sorted hit, nonzero for a hit taken by an earlier iteration (MkFinder.cc:476-482), or NULL;
// such a hit is passed over and does not count toward the kHitIdxMax.
template <class F>
__device__ __forceinline__ int scan_window(const LayerDev& d, const Window& w, const int2* __restrict__ bins,
                                           const float* __restrict__ phis, const float* __restrict__ qs,
                                           const unsigned char* __restrict__ mask, F&& on_hit) {
  int cnt = 0;
  for (int qi = w.qb1; qi < w.qb2 && cnt < kHitIdxMax; ++qi) {
    const int2* __restrict__ row = bins + d.bin_off + qi * kNPhi;
    for (int pi = w.pb1; pi < w.pb2 && cnt < kHitIdxMax;) {
      const int b = pi & 0xff;
      int lo, hi;
      if (b == 0) {
        const int2 be = row[0];
        lo = be.x;
        hi = be.y;
        ++pi;
      } else {
        const int b2 = min(kNPhi - 1, b + (w.pb2 - 1 - pi));
        lo = row[b].x;
        hi = row[b2].y;
        pi += b2 - b + 1;
      }
      // Eight hits' (q, phi) loads in flight at once, tested in order.
      for (int h = lo; h < hi && cnt < kHitIdxMax; h += kScanUnroll) {
        const int gi = d.hit_off + h;
        float hq[kScanUnroll], hp[kScanUnroll];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) {
          const bool in = h + u < hi;
          hq[u] = in ? qs[gi + u] : 0.f;
          hp[u] = in ? phis[gi + u] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) {
          if (h + u >= hi || cnt >= kHitIdxMax) break;
          const float ddq = fabsf(w.q - hq[u]);
          const float ddphi = cdist(fabsf(w.phi - hp[u]));
          if (ddq >= w.dq || ddphi >= w.dphi) continue;
          if (mask && mask[gi + u]) continue;
          on_hit(gi + u);
          ++cnt;
        }
      }
    }
  }
  return cnt;
}

// getHitSelDynamicChi2Cut (MkFinder.cc:202-220)
__device__ __forceinline__ float dyn_chi2_cut(const LayerDev& d, float cut_min, float invpt, float theta_in) {
  const float theta = fabsf(theta_in - kPIOver2);
  const float max_invpt = invpt > 10.0f ? 10.0f : invpt;
  const float this_c2 = d.c_c2_0 * max_invpt + d.c_c2_1 * theta + d.c_c2_2;
  return d.c_c2_sf * this_c2 > cut_min ? d.c_c2_sf * this_c2 : cut_min;
}

__device__ __forceinline__ void load_hit9(const float* __restrict__ h, float (&mp)[3], float (&me)[6]) {
#pragma unroll
  for (int e = 0; e < 3; ++e) mp[e] = h[e];
#pragma unroll
  for (int e = 0; e < 6; ++e) me[e] = h[3 + e];
}

// std::hypot(float, float) of the reference: the sum of squares in double (exact
// for float inputs up to one rounding) and a correctly rounded root, rounded to float.
__device__ __forceinline__ float hypot_ref(float x, float y) {
  const double xd = x, yd = y;
  return (float)__builtin_sqrt(xd * xd + yd * yd);
}

constexpr float kScoreWorst = -1e16f;             // getScoreWorstPossible (Track.h:644-647)

// Config.h:355-359 and getScoreCalc (Track.h:648-672)
__device__ __forceinline__ float score_calc(int nfound, int ntail, int noverlap, int nmiss, float chi2, float pt,
                                            bool in_find) {
  const float maxBonus = 8.0f;
  float bonus = 0.2f * nfound + 4.0f;
  float penalty = 8.0f;
  if (pt < 0.9f) {
    penalty *= in_find ? 1.7f : 1.5f;
    bonus = fminf(bonus * (in_find ? 0.9f : 1.0f), maxBonus);
  }
  return bonus * nfound + 0.0f * noverlap - penalty * nmiss - 3.0f * ntail - chi2;
}

// isStripQCompatible (MkFinder.cc:837-862) for one lane: track z / xy from the
// state propagated to the hit, its uncertainty from the state at the layer.
__device__ __forceinline__ bool strip_q_compatible(bool barrel, const float (&pe)[21], const float (&q)[6],
                                                   const float (&mp)[3], const float (&me)[6]) {
  if (barrel) {
    const float res = fabsf(mp[2] - q[2]);
    const float hitHL = sqrtf(me[5] * 3.f);
    const float qErr = sqrtf(pe[S(2, 2)]);
    return hitHL + fmaxf(3.f * qErr, 0.5f) > res;
  }
  const float r0 = mp[0] - q[0], r1 = mp[1] - q[1];
  const float hitT2 = me[0] + me[2];
  const float inv = 1.f / hitT2;
  const float p0 = me[0] * inv, p1 = me[1] * inv, p2 = me[2] * inv;
  const float qErr = sqrtf(fabsf(pe[S(0, 0)] * p0 + 2.f * pe[S(0, 1)] * p1 + pe[S(1, 1)] * p2));
  const float resProj = sqrtf(r0 * p0 * r0 + 2.f * r1 * p1 * r0 + r1 * p2 * r1);
  return sqrtf(hitT2 * 3.f) + fmaxf(3.f * qErr, 0.5f) > resProj;
}

// passStripChargePCMfromTrack (MkFinder.cc:863-896) for one lane: the hit's charge per cm
// (Hit::PackedData::get_charge_pcm, Hit.h:270-274, from its 8-bit code) times the track's
// path-length factor at the hit (q: the state propagated to the hit) must exceed
// Hit::minChargePerCM = 1620; the overflow code 0xff (Hit::maxChargePerCM) passes.
__device__ __forceinline__ bool pass_strip_charge(bool barrel, unsigned code, const float (&q)[6],
                                                  const float (&me)[6]) {
  const unsigned pcm = code == 0 ? 0u : ((code - 1u) << 3) + 1620u;
  if (pcm >= 1620u + (0xfeu << 3)) return true;
  float qSF;
  if (barrel) {
    const float hitT2 = me[0] + me[2];
    const float inv = 1.f / hitT2;
    const float p0 = me[0] * inv, p1 = me[1] * inv, p2 = me[2] * inv;
    const bool detXY_OK = fabsf(p0 * p2 - p1 * p1) < 0.1f;
    float sinP, cosP;
    sincosf(q[4], &sinP, &cosP);
    const float sinT = fabsf(sinf(q[5]));
    qSF = detXY_OK ? sinT * sqrtf(fabsf(1.f + cosP * cosP * p0 + sinP * sinP * p2 - 2.f * cosP * sinP * p1)) : 1.f;
  } else {
    qSF = fabsf(cosf(q[5]));
  }
  return (float)pcm * qSF > 1620.f;
}

}  // namespace mkf_find

struct mkf_event_of_hits {
  mkf_context* ctx = nullptr;
  int device = 0;                 // the context's device, kept so that destroy does not read the context
  int n_layers = 0;
  std::vector<mkf_find::LayerDev> layers;   // host copy; hit_off / n_hits as of the last load
  mkf_find::LayerDev* d_layers = nullptr;
  int total_bins = 0;
  int2* d_bins = nullptr;
  size_t cap = 0;                 // hit capacity of the arrays below
  uint32_t* d_keys = nullptr;
  uint32_t* d_keys_s = nullptr;
  int* d_vals = nullptr;
  int* d_vals_s = nullptr;
  float* d_phis = nullptr;
  float* d_qs = nullptr;
  float* d_shits = nullptr;
  int* d_ranks = nullptr;
  int* d_inv = nullptr;           // event hit index -> sorted position (inverse of ranks)
  void* d_temp = nullptr;
  size_t temp_bytes = 0;
  bool loaded = false;
  void* d_ce_nodes = nullptr;     // clone-engine HitOnTrack node store, then its best short candidates
  size_t ce_node_cap = 0;         // in 16-byte units
  // Clone-engine calls on one event of hits share d_ce_nodes: the mutex
  // serialises the host side (growing the store), and ce_done, recorded after
  // each call's kernel, makes the next call's stream wait for it.
  std::mutex ce_mu;
  hipEvent_t ce_done = nullptr;
  // The iteration's hit mask per sorted hit (mkf_event_of_hits_set_hit_mask); a load clears it.
  unsigned char* d_mask_s = nullptr;
  size_t mask_cap = 0;
  bool has_mask = false;
  const unsigned char* mask() const { return has_mask ? d_mask_s : nullptr; }
  // The hits' strip charge codes per sorted hit (mkf_event_of_hits_set_hit_charge); a load clears them.
  unsigned char* d_charge_s = nullptr;
  size_t charge_cap = 0;
  bool has_charge = false;
  const unsigned char* charge() const { return has_charge ? d_charge_s : nullptr; }
  // Dead-module bins, one byte per (layer, q bin, phi bin) (mkf_event_of_hits_set_dead_regions).
  unsigned char* d_dead = nullptr;
  bool has_dead = false;
  const unsigned char* dead() const { return has_dead ? d_dead : nullptr; }
};
