// mkf_bkfit.hip -- backward fit of found candidates on gfx950.
//
// MkBuilder::BackwardFit (MkBuilder.cc:2124-2253) hands blocks of NN
// candidates to MkFinder::BkFitInputTracks / BkFitFitTracks /
// BkFitPropTracksToPCA / BkFitOutputTracks (MkFinder.cc:1368-1749).  Here one
// lane carries one candidate through the plan layers in reverse, its state in
// VGPRs.  The block of the reference becomes `group` consecutive lanes of a
// wavefront: per layer the lanes ballot whether they hold a hit there; a block
// with none skips the layer, and the lanes of a block that has one but that
// hold none take the reference's fake hit at their own position.  Candidates
// are read from the finder's element-major arrays, hits from the event of
// hits' sorted copy through the inverse rank table.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/mkfit_amd.h"
#include "mkf_find_common.h"
#include "mkf_internal.h"
#include "mkf_track_math.h"

using namespace mkf;
using namespace mkf_detail;
using namespace mkf_find;

namespace {

struct BkArgs {
  int n, n_layers, last_layer, group, pflags, pca_pflags, flags, n_eoh_layers;
  float bfield;
  Material mat;
  const LayerDev* __restrict__ lays;
  const float* __restrict__ shits;
  const int* __restrict__ inv;
  const int* __restrict__ hot;
  const int* __restrict__ layer_base;
  const float* in_par;
  const float* in_err;
  const int* in_chg;
  float* par;
  float* err;
  int* chg;
  float* __restrict__ chi2;
  float* __restrict__ score;
};

// MkBase::PropagateTracksToPCAZ (MkBase.h:91-104): the z of the point of closest
// approach to the origin of the track's straight-line r-z projection.
__device__ __forceinline__ void to_pca(float (&par)[6], float (&err)[21], int chg, const BkArgs& a) {
  const float slope = tanf(par[5]);
  const float z = (slope * (slope * par[2] - hypot_ref(par[0], par[1]))) / (1 + slope * slope);
  propagate_to_z(par, err, chg, z, a.pca_pflags, false, a.bfield, a.mat);
}

// Every lane of the grid runs the layer loop (lanes past n or outside the event
// take no hit and write nothing), so that the ballots see whole wavefronts.
__global__ void __launch_bounds__(256) k_backward_fit(BkArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const uint64_t gmask = a.group >= 64 ? ~0ull : ((1ull << a.group) - 1) << (lane & ~(a.group - 1));
  const size_t n = (size_t)a.n;
  bool valid = t < a.n;
  int base = 0;
  if (valid && a.layer_base) {
    base = a.layer_base[t];
    valid = base >= 0 && base + a.n_layers <= a.n_eoh_layers;
  }
  float par[6], err[21], chi2 = 0.f;
  int chg = 0;
  if (valid) {
#pragma unroll
    for (int e = 0; e < 6; ++e) par[e] = a.in_par[e * n + t];
#pragma unroll
    for (int e = 0; e < 21; ++e) err[e] = a.in_err[e * n + t] * 100.0f;          // Err[iC].Scale(100.0f)
    chg = a.in_chg[t];
  }
  const bool fake_mat = a.flags & MKF_BKFIT_FAKE_MATERIAL;
  for (int l = a.n_layers - 1; l >= a.last_layer; --l) {
    bool has = false;
    int pos = 0;
    if (valid) {
      const LayerDev& d = a.lays[base + l];
      const int h = a.hot[l * n + t];
      has = h >= 0 && h < d.n_hits;
      if (has) pos = a.inv[d.hit_off + h];
    }
    if ((__ballot(has) & gmask) == 0 || !valid) continue;             // here_count == 0
    const bool barrel = a.lays[base + l].is_barrel != 0;
    float mp[3], me[6];
    if (has) {
      load_hit9(a.shits + (size_t)pos * 9, mp, me);
    } else {                                                          // MkFinder.cc:1635-1647
      mp[0] = par[0]; mp[1] = par[1]; mp[2] = par[2];
      me[0] = 666.f; me[1] = 0.f; me[2] = 666.f; me[3] = 0.f; me[4] = 0.f; me[5] = 666.f;
    }
    const bool no_mat = !has && !fake_mat;
    float c2;
    if (barrel) {
      propagate_to_r(par, err, chg, hypot_ref(mp[0], mp[1]), a.pflags, no_mat, a.bfield, a.mat);   // MkBase.h:55
      c2 = kalman_barrel<true, true>(par, err, mp, me);
    } else {
      propagate_to_z(par, err, chg, mp[2], a.pflags, no_mat, a.bfield, a.mat);                   // MkBase.h:84
      c2 = kalman_endcap<true, true>(par, err, mp, me);
    }
    if (par[3] < 0.f) {                                               // MkFinder.cc:1671-1678
      chg = -chg;
      par[3] = -par[3];
    }
    chi2 += c2;
  }
  if (!valid) return;
  if (a.flags & MKF_BKFIT_PCA) to_pca(par, err, chg, a);
#pragma unroll
  for (int e = 0; e < 6; ++e) a.par[e * n + t] = par[e];
#pragma unroll
  for (int e = 0; e < 21; ++e) a.err[e * n + t] = err[e];
  a.chg[t] = chg;
  a.chi2[t] = chi2;
  if (a.score && !isnan(chi2)) {
    // TrackCand counters of the candidate's nodes in the order they were added (addHitIdx)
    int nfound = 0, ninside = 0, ntail = 0;
    for (int l = 0; l < a.n_layers; ++l) {
      const int h = a.hot[l * n + t];
      if (h == -9) continue;
      if (h >= 0) {
        ++nfound;
        ninside += ntail;
        ntail = 0;
      } else if (h == -1) {
        ++ntail;
      }
    }
    a.score[t] = score_calc(nfound, 0, 0, ninside, chi2 < 0.f ? 0.f : chi2, fabsf(1.f / par[3]), false);
  }
}

}  // namespace

extern "C" int mkf_backward_fit(mkf_context* c, mkf_event_of_hits* h, const mkf_bkfit_batch* b, void* stream) {
  CHECK(set_device(c));
  if (!h || !b) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  if (h->ctx != c) return fail(MKF_ERR_INVALID_ARG, "event of hits belongs to another context");
  if (!h->loaded) return fail(MKF_ERR_INVALID_ARG, "event of hits not loaded");
  if (b->n_cands < 0) return fail(MKF_ERR_INVALID_ARG, "n_cands < 0");
  if (b->n_layers < 1 || b->n_layers > MKF_MAX_FIND_LAYERS)
    return fail(MKF_ERR_INVALID_ARG, "n_layers %d outside 1..%d", b->n_layers, MKF_MAX_FIND_LAYERS);
  if (!b->layer_base && b->n_layers > h->n_layers)
    return fail(MKF_ERR_INVALID_ARG, "n_layers %d exceeds the event of hits' %d layers", b->n_layers, h->n_layers);
  if (b->last_layer < 0 || b->last_layer > b->n_layers)
    return fail(MKF_ERR_INVALID_ARG, "last_layer %d outside 0..n_layers", b->last_layer);
  if (b->group < 1 || b->group > 64 || (b->group & (b->group - 1)))
    return fail(MKF_ERR_INVALID_ARG, "group %d is not a power of two in 1..64", b->group);
  if (b->flags & ~(MKF_BKFIT_PCA | MKF_BKFIT_FAKE_MATERIAL))
    return fail(MKF_ERR_INVALID_ARG, "unknown flags 0x%x", b->flags);
  const int mat_flags = b->pflags | ((b->flags & MKF_BKFIT_PCA) ? b->pca_pflags : 0);
  if ((mat_flags & MKF_PF_APPLY_MATERIAL) && !c->d_rl)
    return fail(MKF_ERR_INVALID_ARG, "material requested but the context has no material grid");
  if (b->n_cands == 0) return MKF_OK;
  const size_t N = b->n_cands;
  CHECK(check_dev(b->in_par, 6 * N * 4, "in_par"));
  CHECK(check_dev(b->in_err, 21 * N * 4, "in_err"));
  CHECK(check_dev(b->in_chg, N * 4, "in_chg"));
  CHECK(check_dev(b->hot_idx, (size_t)b->n_layers * N * 4, "hot_idx"));
  CHECK(check_dev(b->par, 6 * N * 4, "par"));
  CHECK(check_dev(b->err, 21 * N * 4, "err"));
  CHECK(check_dev(b->chg, N * 4, "chg"));
  CHECK(check_dev(b->chi2, N * 4, "chi2"));
  if (b->score) CHECK(check_dev(b->score, N * 4, "score"));
  if (b->layer_base) CHECK(check_dev(b->layer_base, N * 4, "layer_base"));
  BkArgs a;
  a.n = b->n_cands;
  a.n_layers = b->n_layers;
  a.last_layer = b->last_layer;
  a.group = b->group;
  a.pflags = b->pflags;
  a.pca_pflags = b->pca_pflags;
  a.flags = b->flags;
  a.n_eoh_layers = h->n_layers;
  a.bfield = c->bfield;
  a.mat = mat_of(c);
  a.lays = h->d_layers;
  a.shits = h->d_shits;
  a.inv = h->d_inv;
  a.hot = b->hot_idx;
  a.layer_base = b->layer_base;
  a.in_par = b->in_par;
  a.in_err = b->in_err;
  a.in_chg = b->in_chg;
  a.par = b->par;
  a.err = b->err;
  a.chg = b->chg;
  a.chi2 = b->chi2;
  a.score = b->score;
  // 256-thread blocks: every wavefront holds 64 consecutive candidates from a multiple of 64, so the blocks
  // of `group` lanes start at multiples of group as the reference's blocks of NN do.
  const int blocks = (int)((N + 255) / 256);
  hipLaunchKernelGGL(k_backward_fit, dim3(blocks), dim3(256), 0, (hipStream_t)stream, a);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}
