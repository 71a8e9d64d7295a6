// mkf_finder_layer.hip -- the per-layer MkFinder steps on gfx950, for callers
// that drive their own layer loop (and candidate bookkeeping) the way
// MkBuilder does:
//   AddBestHit                 k_add_best_hit     (MkFinder.cc:691-832)
//   FindCandidatesCloneEngine  k_find_candidates  (MkFinder.cc:1114-1284)
//   UpdateWithLastHit          k_update_last_hit  (MkFinder.cc:1290-1308)
// One lane per track.  The hit lists come from mkf_select_hit_indices as
// layer-local hit indices; the event's inverse rank table maps each to its
// sorted position, where the 9 floats of the hit are stored.  The arithmetic
// is the device code of the fused finders (mkf_find.hip, mkf_ce.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/mkfit_amd.h"
#include "mkf_find_common.h"
#include "mkf_internal.h"
#include "mkf_track_math.h"

using namespace mkf;
using namespace mkf_detail;
using namespace mkf_find;

namespace {

struct LayerStepArgs {
  LayerDev d;
  int n;
  float chi2_cut, bfield;
  int prop_to_hit, prop_pflags, max_holes, max_consec;
  Material mat;
  const float* __restrict__ prop_par;
  const float* __restrict__ prop_err;
  float* __restrict__ par;
  float* __restrict__ err;
  int* __restrict__ chg;
  float* __restrict__ chi2;
  int* __restrict__ n_found;
  int* __restrict__ n_inside;
  int* __restrict__ n_tail;
  const int* __restrict__ xsize;
  const int* __restrict__ xarr;
  const int* __restrict__ wsr;
  const int* __restrict__ hit_idx;
  const float* __restrict__ shits;
  const int* __restrict__ inv;
  const unsigned char* __restrict__ charge;   // per sorted hit: strip charge code, or NULL (every hit passes)
  int* __restrict__ hit_code;
  int* __restrict__ n_entries;
  mkf_idx_chi2* __restrict__ entries;
};

__device__ __forceinline__ void load_lane(const LayerStepArgs& a, int t, float (&p)[6], float (&e)[21]) {
  const size_t n = (size_t)a.n;
#pragma unroll
  for (int k = 0; k < 6; ++k) p[k] = a.prop_par[k * n + t];
#pragma unroll
  for (int k = 0; k < 21; ++k) e[k] = a.prop_err[k * n + t];
}

__device__ __forceinline__ void store_lane(const LayerStepArgs& a, int t, const float (&p)[6], const float (&e)[21]) {
  const size_t n = (size_t)a.n;
#pragma unroll
  for (int k = 0; k < 6; ++k) a.par[k * n + t] = p[k];
#pragma unroll
  for (int k = 0; k < 21; ++k) a.err[k * n + t] = e[k];
}

// Sorted position of hit `idx` of the layer, -1 when idx is not a hit of it.
__device__ __forceinline__ int hit_pos(const LayerStepArgs& a, int idx) {
  return (idx >= 0 && idx < a.d.n_hits) ? a.inv[a.d.hit_off + idx] : -1;
}

__device__ __forceinline__ void propagate_to_hit(const LayerStepArgs& a, float (&p)[6], float (&e)[21], int chg,
                                                 const float (&mp)[3]) {
  if (a.d.is_barrel)
    propagate_to_r(p, e, chg, hypot_ref(mp[0], mp[1]), a.prop_pflags, false, a.bfield, a.mat);
  else
    propagate_to_z(p, e, chg, mp[2], a.prop_pflags, false, a.bfield, a.mat);
}

// |chi2| of one hit against the layer state (kalmanPropagateAndComputeChi2[Endcap]);
// q receives the state the chi2 was taken at.
__device__ __forceinline__ float hit_chi2(const LayerStepArgs& a, const float (&pp)[6], const float (&pe)[21], int chg,
                                          const float (&mp)[3], const float (&me)[6], float (&q)[6]) {
  float qe[21];
#pragma unroll
  for (int k = 0; k < 6; ++k) q[k] = pp[k];
#pragma unroll
  for (int k = 0; k < 21; ++k) qe[k] = pe[k];
  if (a.prop_to_hit) propagate_to_hit(a, q, qe, chg, mp);
  return fabsf(a.d.is_barrel ? kalman_barrel<true, false>(q, qe, mp, me) : kalman_endcap<true, false>(q, qe, mp, me));
}

// kalmanPropagateAndUpdate[Endcap] with the charge flip (KalmanUtilsMPlex.cc:471-504, 697-730).
__device__ __forceinline__ void update_with(const LayerStepArgs& a, float (&p)[6], float (&e)[21], int& chg,
                                            const float (&mp)[3], const float (&me)[6]) {
  if (a.prop_to_hit) propagate_to_hit(a, p, e, chg, mp);
  if (a.d.is_barrel)
    kalman_barrel<false, true>(p, e, mp, me);
  else
    kalman_endcap<false, true>(p, e, mp, me);
  if (p[3] < 0.f) {
    chg = -chg;
    p[3] = -p[3];
  }
}

__global__ void __launch_bounds__(256) k_add_best_hit(LayerStepArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const size_t n = (size_t)a.n;
  float pp[6], pe[21];
  load_lane(a, t, pp, pe);
  int chg = a.chg[t];
  float chi2 = a.chi2[t];
  int nfound = a.n_found[t], ninside = a.n_inside[t], ntail = a.n_tail[t];
  const int xs = min(a.xsize[t], kHitIdxMax);
  const int w = a.wsr[t];
  float min_c2 = dyn_chi2_cut(a.d, a.chi2_cut, pp[3], pp[5]);
  int best = -1, best_pos = -1;
  for (int k = 0; k < xs; ++k) {                              // MkFinder.cc:725-765
    const int idx = a.xarr[k * n + t];
    const int pos = hit_pos(a, idx);
    if (pos < 0) continue;
    float hp[3], he[6], q[6];
    load_hit9(a.shits + (size_t)pos * 9, hp, he);
    const float c2 = hit_chi2(a, pp, pe, chg, hp, he, q);
    if (c2 < min_c2) {
      min_c2 = c2;
      best = idx;
      best_pos = pos;
    }
  }
  float mp[3] = {pp[0], pp[1], pp[2]};
  float me[6] = {666.f, 0.f, 666.f, 0.f, 0.f, 666.f};          // fake hit (MkFinder.cc:776-800)
  int code = -9;
  if (w != MKF_WSR_OUTSIDE) {
    if (best >= 0) {
      load_hit9(a.shits + (size_t)best_pos * 9, mp, me);
      chi2 += min_c2;
      code = best;
      ++nfound;
      ninside += ntail;
      ntail = 0;
    } else {
      code = w == MKF_WSR_EDGE ? -3 : (ninside + ntail ? -2 : -1);
      ntail += code == -1;
    }
  }
  update_with(a, pp, pe, chg, mp, me);
  store_lane(a, t, pp, pe);
  a.chg[t] = chg;
  a.chi2[t] = chi2;
  a.n_found[t] = nfound;
  a.n_inside[t] = ninside;
  a.n_tail[t] = ntail;
  a.hit_code[t] = code;
}

__global__ void __launch_bounds__(256) k_find_candidates(LayerStepArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const size_t n = (size_t)a.n;
  float pp[6], pe[21];
  load_lane(a, t, pp, pe);
  const int chg = a.chg[t];
  const float chi2 = a.chi2[t];
  const int nfound = a.n_found[t], ninside = a.n_inside[t], ntail = a.n_tail[t];
  const int xs = min(a.xsize[t], kHitIdxMax);
  const int w = a.wsr[t];
  const float max_c2 = dyn_chi2_cut(a.d, a.chi2_cut, pp[3], pp[5]);
  const float pt = fabsf(1.0f / pp[3]);
  const bool barrel = a.d.is_barrel != 0;
  mkf_idx_chi2* out = a.entries + (size_t)t * MKF_CE_ENTRIES_PER_LANE;
  int ne = 0;
  for (int k = 0; k < xs; ++k) {
    const int idx = a.xarr[k * n + t];
    const int pos = hit_pos(a, idx);
    if (pos < 0) continue;
    float mp[3], me[6], q[6];
    load_hit9(a.shits + (size_t)pos * 9, mp, me);
    const float c2 = hit_chi2(a, pp, pe, chg, mp, me, q);
    if (c2 < max_c2 && strip_q_compatible(barrel, pe, q, mp, me) &&
        (!a.charge || pass_strip_charge(barrel, a.charge[pos], q, me))) {
      mkf_idx_chi2 r;
      r.trk_idx = t;
      r.hit_idx = idx;
      r.module = 0u;
      r.nhits = nfound + 1;
      r.ntailholes = 0;
      r.noverlaps = 0;
      r.nholes = ninside + ntail;
      r.pt = pt;
      r.chi2 = chi2 + c2;
      r.chi2_hit = c2;
      r.score = score_calc(r.nhits, 0, 0, r.nholes, r.chi2 < 0 ? 0.f : r.chi2, pt, true);
      r.xhit_pos = k;
      out[ne++] = r;
    }
  }
  if (w != MKF_WSR_OUTSIDE) {                                 // the invalid-hit record
    int fake = (ninside + ntail < a.max_holes && ntail < a.max_consec) ? -1 : -2;
    if (w == MKF_WSR_EDGE) fake = -3;
    mkf_idx_chi2 r;
    r.trk_idx = t;
    r.hit_idx = fake;
    r.module = 0xffffffffu;
    r.nhits = nfound;
    r.ntailholes = fake == -1 ? ntail + 1 : ntail;
    r.noverlaps = 0;
    r.nholes = ninside;
    r.pt = pt;
    r.chi2 = chi2;
    r.chi2_hit = 0.f;
    r.score = score_calc(nfound, r.ntailholes, 0, ninside, chi2 < 0 ? 0.f : chi2, pt, true);
    r.xhit_pos = -1;
    out[ne++] = r;
  }
  a.n_entries[t] = ne;
}

__global__ void __launch_bounds__(256) k_update_last_hit(LayerStepArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  float p[6], e[21];
  load_lane(a, t, p, e);
  int chg = a.chg[t];
  const int pos = hit_pos(a, a.hit_idx[t]);
  if (pos >= 0) {
    float mp[3], me[6];
    load_hit9(a.shits + (size_t)pos * 9, mp, me);
    update_with(a, p, e, chg, mp, me);
  }
  store_lane(a, t, p, e);
  a.chg[t] = chg;
}

int blocks_for(size_t n, int block) { return (int)std::max<size_t>(1, (n + block - 1) / block); }

// Shared argument checks; fills the lane part of `a`.
int prepare(mkf_context* c, mkf_event_of_hits* h, int layer, const mkf_finder_lanes* ln, bool writes_state,
            bool reads_counters, LayerStepArgs& a) {
  CHECK(set_device(c));
  if (!h || !ln) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  if (h->ctx != c) return fail(MKF_ERR_INVALID_ARG, "event of hits belongs to another context");
  if (!h->loaded) return fail(MKF_ERR_INVALID_ARG, "event of hits not loaded");
  if (layer < 0 || layer >= h->n_layers) return fail(MKF_ERR_INVALID_ARG, "layer %d out of range", layer);
  if (ln->n < 0) return fail(MKF_ERR_INVALID_ARG, "n < 0");
  std::memset(&a, 0, sizeof(a));
  a.d = h->layers[layer];
  a.n = ln->n;
  a.bfield = c->bfield;
  a.mat = mat_of(c);
  if (ln->n == 0) return MKF_OK;
  const size_t N = ln->n;
  CHECK(check_dev(ln->prop_par, 6 * N * 4, "prop_par"));
  CHECK(check_dev(ln->prop_err, 21 * N * 4, "prop_err"));
  CHECK(check_dev(ln->chg, N * 4, "chg"));
  if (writes_state) {
    CHECK(check_dev(ln->par, 6 * N * 4, "par"));
    CHECK(check_dev(ln->err, 21 * N * 4, "err"));
  }
  if (reads_counters) {
    CHECK(check_dev(ln->chi2, N * 4, "chi2"));
    CHECK(check_dev(ln->n_found, N * 4, "n_found"));
    CHECK(check_dev(ln->n_inside_minus_one, N * 4, "n_inside_minus_one"));
    CHECK(check_dev(ln->n_tail_minus_one, N * 4, "n_tail_minus_one"));
  }
  a.prop_par = ln->prop_par; a.prop_err = ln->prop_err; a.par = ln->par; a.err = ln->err;
  a.chg = ln->chg; a.chi2 = ln->chi2; a.n_found = ln->n_found;
  a.n_inside = ln->n_inside_minus_one; a.n_tail = ln->n_tail_minus_one;
  a.shits = h->d_shits;
  a.inv = h->d_inv;
  a.charge = h->charge();
  return MKF_OK;
}

int check_select(const mkf_finder_lanes* ln, const int* xsize, const int* xarr, const int* wsr, LayerStepArgs& a) {
  const size_t N = ln->n;
  CHECK(check_dev(xsize, N * 4, "xsize"));
  CHECK(check_dev(xarr, kHitIdxMax * N * 4, "xarr"));
  CHECK(check_dev(wsr, N * 4, "wsr"));
  a.xsize = xsize; a.xarr = xarr; a.wsr = wsr;
  return MKF_OK;
}

int check_material(const mkf_context* c, int pflags) {
  if ((pflags & MKF_PF_APPLY_MATERIAL) && !c->d_rl)
    return fail(MKF_ERR_INVALID_ARG, "material requested but the context has no material grid");
  return MKF_OK;
}

}  // namespace

extern "C" {

int mkf_add_best_hit(mkf_context* c, mkf_event_of_hits* h, int layer, const mkf_finder_lanes* ln, const int* xsize,
                     const int* xarr, const int* wsr, float chi2_cut, int prop_to_hit, int intra_pflags,
                     int* hit_code, void* stream) {
  LayerStepArgs a;
  CHECK(prepare(c, h, layer, ln, true, true, a));
  if (ln->n == 0) return MKF_OK;
  CHECK(check_select(ln, xsize, xarr, wsr, a));
  CHECK(check_dev(hit_code, (size_t)ln->n * 4, "hit_code"));
  CHECK(check_material(c, prop_to_hit ? intra_pflags : 0));
  a.chi2_cut = chi2_cut;
  a.prop_to_hit = prop_to_hit;
  a.prop_pflags = intra_pflags;
  a.hit_code = hit_code;
  hipLaunchKernelGGL(k_add_best_hit, dim3(blocks_for(ln->n, 256)), dim3(256), 0, (hipStream_t)stream, a);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

int mkf_find_candidates_clone_engine(mkf_context* c, mkf_event_of_hits* h, int layer, const mkf_finder_lanes* ln,
                                     const int* xsize, const int* xarr, const int* wsr, float chi2_cut,
                                     int prop_to_hit, int intra_pflags, int max_holes, int max_consec_holes,
                                     int* n_entries, mkf_idx_chi2* entries, void* stream) {
  LayerStepArgs a;
  CHECK(prepare(c, h, layer, ln, false, true, a));
  if (ln->n == 0) return MKF_OK;
  CHECK(check_select(ln, xsize, xarr, wsr, a));
  CHECK(check_dev(n_entries, (size_t)ln->n * 4, "n_entries"));
  CHECK(check_dev(entries, (size_t)ln->n * MKF_CE_ENTRIES_PER_LANE * sizeof(mkf_idx_chi2), "entries"));
  CHECK(check_material(c, prop_to_hit ? intra_pflags : 0));
  a.chi2_cut = chi2_cut;
  a.prop_to_hit = prop_to_hit;
  a.prop_pflags = intra_pflags;
  a.max_holes = max_holes;
  a.max_consec = max_consec_holes;
  a.n_entries = n_entries;
  a.entries = entries;
  hipLaunchKernelGGL(k_find_candidates, dim3(blocks_for(ln->n, 256)), dim3(256), 0, (hipStream_t)stream, a);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

int mkf_update_with_last_hit(mkf_context* c, mkf_event_of_hits* h, int layer, const mkf_finder_lanes* ln,
                             const int* hit_idx, int prop_to_hit, int pflags, void* stream) {
  LayerStepArgs a;
  CHECK(prepare(c, h, layer, ln, true, false, a));
  if (ln->n == 0) return MKF_OK;
  CHECK(check_dev(hit_idx, (size_t)ln->n * 4, "hit_idx"));
  CHECK(check_material(c, prop_to_hit ? pflags : 0));
  a.prop_to_hit = prop_to_hit;
  a.prop_pflags = pflags;
  a.hit_idx = hit_idx;
  hipLaunchKernelGGL(k_update_last_hit, dim3(blocks_for(ln->n, 256)), dim3(256), 0, (hipStream_t)stream, a);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

}  // extern "C"
