// mkf_find.hip -- best-hit track finding on gfx950.
//
//   LayerOfHits build   k_hit_keys -> stable LSD radix sort (hipCUB) -> k_loh_gather, k_loh_fill
//   SelectHitIndices    k_select (per-operation entry point)
//   FindTracksBestHit   k_find_best_hit: one lane per seed carries its track
//                       through every layer with the state in VGPRs; hit
//                       selection, the chi2 of each candidate and the update
//                       are fused, so per layer only the bin table, the
//                       (phi, q) of scanned hits and the 9 floats of each
//                       candidate are read.
//
// Event layout in HBM (one event, all layers back to back in visiting order):
//   sorted hits   [total][9]  hits of every layer in LayerOfHits order
//   phis, qs      [total]     phi and q (z on barrel, r on endcaps) per sorted hit
//   ranks         [total]     sorted position -> hit index within its layer
//   inv           [total]     event hit index -> sorted position
//   bins          [sum nq*256] int2 {first, end} per (layer, q bin, phi bin),
//                             positions relative to the layer's first hit
// Every layer's rows start at its hit offset, so a lane addresses a layer
// with one scalar base and no per-layer allocations.
//
// Sort key: layer << 21 | q bin << 11 | fine phi bin.  Within one layer this is
// the order of the reference's key fine_phi + (q_bin << 16)
// (HitStructures.cc:114-128) because the fine phi bin is at most 1024; the
// layer bits keep layers apart so one sort over the event serves all layers,
// and LSD radix sort is stable like the reference's radix sort of the keys.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/mkfit_amd.h"
#include "mkf_internal.h"
#include "mkf_track_math.h"
#include "mkf_find_common.h"

using namespace mkf;
using namespace mkf_detail;
using namespace mkf_find;

namespace {

// Per-load hit offsets into the layer table, passed by value in chunks of
// kOffChunk layers (stream-ordered, no host staging buffer to keep alive).
constexpr int kOffChunk = 256;
struct OffsetsArg {
  int first, count;
  int off[kOffChunk + 1];
};

__global__ void k_set_offsets(OffsetsArg o, LayerDev* __restrict__ L) {
  const int l = threadIdx.x;
  if (l < o.count) {
    L[o.first + l].hit_off = o.off[l];
    L[o.first + l].n_hits = o.off[l + 1] - o.off[l];
  }
}

__global__ void k_fill_i32(int* __restrict__ p, long n, int v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// Hit::phi and Hit::z (barrel) / Hit::r (endcap) of a hit (Hit.h:64-80).  Not
// inlined: the key kernel and the fill kernel run this one piece of machine
// code, so the phi and q stored with a sorted hit are bit for bit the values
// its sort key was formed from.
__device__ __noinline__ float2 hit_phi_q(float x, float y, float z, int is_barrel) {
  return make_float2(atan2_cr(y, x), is_barrel ? z : sqrtf(x * x + y * y));
}

// Layer owning event hit i: the last one in [lo, hi] whose hit range starts at
// or before it.
__device__ __forceinline__ int layer_of_hit(int i, int lo, int hi, const LayerDev* __restrict__ L) {
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L[mid].hit_off <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The sort key of every hit.  The block's first and last hits are located
// over all layers once; every thread then searches only the layers between.
__global__ void __launch_bounds__(256)
k_hit_keys(int total, int n_layers, const float* __restrict__ hits, const LayerDev* __restrict__ L,
           uint32_t* __restrict__ keys, int* __restrict__ vals) {
  __shared__ int span[2];
  const int i0 = blockIdx.x * blockDim.x;
  if (threadIdx.x < 2) {
    const int e = threadIdx.x == 0 ? i0 : min(i0 + (int)blockDim.x, total) - 1;
    span[threadIdx.x] = layer_of_hit(e, 0, n_layers - 1, L);
  }
  __syncthreads();
  const int i = i0 + threadIdx.x;
  if (i >= total) return;
  const int lo = layer_of_hit(i, span[0], span[1], L);
  const LayerDev& d = L[lo];
  const float2 pq = hit_phi_q(hits[(size_t)i * 9 + 0], hits[(size_t)i * 9 + 1], hits[(size_t)i * 9 + 2], d.is_barrel);
  const int qb = q_bin_checked(d, pq.y);
  const int pf = phi_bin_fine(pq.x);
  keys[i] = ((uint32_t)lo << kLayerShift) | ((uint32_t)qb << kQShift) | (uint32_t)pf;
  vals[i] = i;
}

// Sorted copies (HitStructures.cc:130-183), 256 sorted positions per block: the
// block's 2304 hit floats are read in 9 rounds in which consecutive lanes read
// consecutive floats of each hit (one random 36-byte run per hit rather than
// three random vector loads per lane), staged in LDS and written out
// contiguously; then each thread forms its hit's phi and q through the key
// kernel's function.  In XCD-major block order each XCD gathers from its own
// few layers (3.65 -> 3.43 ms over 141.7M hits).
__global__ void __launch_bounds__(256)
k_loh_gather(int total, const uint32_t* __restrict__ keys_s, const int* __restrict__ vals_s,
             const LayerDev* __restrict__ L, const float* __restrict__ hits, float* __restrict__ shits,
             float* __restrict__ phis, float* __restrict__ qs) {
  __shared__ float sh[256 * 9];
  __shared__ int sj[256];
  const int t = threadIdx.x;
  const int i0 = xcd_major_block() * 256;
  const int n = min(256, total - i0);
  sj[t] = t < n ? vals_s[i0 + t] : 0;
  __syncthreads();
  for (int f = t; f < n * 9; f += 256) {
    const int h = f / 9;
    sh[f] = hits[(size_t)sj[h] * 9 + (f - h * 9)];
  }
  __syncthreads();
  for (int f = t; f < n * 9; f += 256) shits[(size_t)i0 * 9 + f] = sh[f];
  if (t < n) {
    const int i = i0 + t;
    const LayerDev& d = L[keys_s[i] >> kLayerShift];
    const float2 pq = hit_phi_q(sh[t * 9], sh[t * 9 + 1], sh[t * 9 + 2], d.is_barrel);
    phis[i] = pq.x;
    qs[i] = pq.y;
  }
}

// Ranks in the layer, the inverse map and bin ranges (HitStructures.cc:130-183).
// The fine phi bin 1024 (phi == +pi exactly) shares coarse bin 0 with the row's
// first bins but sorts last in the row; the reference's scan then overwrites
// that bin's range with the later run.  Pass 0 writes every other run, pass 1
// the 1024 runs, so the later run wins here as well.  (Writing the ranks and
// the inverse map from the gather kernel instead measured 0.35 ms slower over
// 141.7M hits.)  In XCD-major block order the scattered 4-byte inverse-map
// writes of a layer complete their lines in one L2 (1.7 -> 1.3 ms for pass 0
// over 141.7M hits).
__global__ void __launch_bounds__(256)
k_loh_fill(int total, int pass, const uint32_t* __restrict__ keys_s, const int* __restrict__ vals_s,
           const LayerDev* __restrict__ L, int* __restrict__ ranks, int* __restrict__ inv, int2* __restrict__ bins) {
  const int i = xcd_major_block() * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint32_t key = keys_s[i];
  const LayerDev& d = L[key >> kLayerShift];
  const int local = i - d.hit_off;
  if (pass == 0) {
    const int j = vals_s[i];
    ranks[i] = j - d.hit_off;
    inv[j] = i;
  }
  const bool fine1024 = (key & 0x400u) != 0;
  if ((pass == 1) != fine1024) return;
  const uint32_t jq = key & ~3u;
  const bool first = i == 0 || (keys_s[i - 1] & ~3u) != jq;
  const bool last = i == total - 1 || (keys_s[i + 1] & ~3u) != jq;
  const int qb = (int)((key >> kQShift) & (kMaxQBins - 1));
  const int pb = (int)((key & 0x3ffu) >> 2);
  int2* row = bins + d.bin_off + qb * kNPhi;
  if (first) row[pb].x = local;
  if (last) row[pb].y = local + 1;
  if (pass == 1) return;
  // Empty bins 1..255 of a row hold {p, p}, p the position where their hits
  // would start, so bins b..b' (1 <= b <= b' <= 255) of a row span the sorted
  // positions [row[b].x, row[b'].y) and a window's run of them is read with
  // two loads.  Bin 0 is left out: the phi == +pi run may replace its range.
  const uint32_t rowkey = key >> kQShift;
  if (first && pb > 1 && (i == 0 || (keys_s[i - 1] >> kQShift) != rowkey))
    for (int b = 1; b < pb; ++b) row[b] = make_int2(local, local);
  if (last) {
    int nb = kNPhi;
    if (i + 1 < total) {
      const uint32_t nk = keys_s[i + 1];
      if ((nk >> kQShift) == rowkey && !(nk & 0x400u)) nb = (int)((nk & 0x3ffu) >> 2);
    }
    for (int b = max(pb + 1, 1); b < nb; ++b) row[b] = make_int2(local + 1, local + 1);
  }
}

// The iteration's hit mask in sorted order: mask_s[inv[j]] = mask[j] for every event hit j.
__global__ void __launch_bounds__(256)
k_mask_sorted(int total, const int* __restrict__ inv, const unsigned char* __restrict__ mask,
              unsigned char* __restrict__ mask_s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < total) mask_s[inv[j]] = mask[j] ? 1 : 0;
}

// The hits' charge codes in sorted order: charge_s[inv[j]] = code[j] for every event hit j.
__global__ void __launch_bounds__(256)
k_charge_sorted(int total, const int* __restrict__ inv, const unsigned char* __restrict__ code,
                unsigned char* __restrict__ charge_s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < total) charge_s[inv[j]] = code[j];
}

__global__ void __launch_bounds__(256)
k_select(LayerDev d, int n, const float* __restrict__ par, const float* __restrict__ err,
         const int2* __restrict__ bins, const float* __restrict__ phis, const float* __restrict__ qs,
         const unsigned char* __restrict__ mask, const int* __restrict__ ranks, int* __restrict__ xsize,
         int* __restrict__ xarr, int* __restrict__ wsr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float p[6], e[21];
#pragma unroll
  for (int k = 0; k < 6; ++k) p[k] = par[(size_t)k * n + t];
#pragma unroll
  for (int k = 0; k < 21; ++k) e[k] = err[(size_t)k * n + t];
  Window w;
  hit_window(d, p, e, w);
  wsr[t] = w.wsr;
  int cnt = -1;
  if (w.wsr != MKF_WSR_OUTSIDE) {
    int c = 0;
    scan_window(d, w, bins, phis, qs, mask, [&](int gi) { xarr[(size_t)(c++) * n + t] = ranks[gi]; });
    cnt = c;
  }
  xsize[t] = cnt;
  for (int k = max(cnt, 0); k < kHitIdxMax; ++k) xarr[(size_t)k * n + t] = -1;
}

// ---- FindTracksBestHit -------------------------------------------------------

struct FindArgs {
  int n, n_layers, first_layer, eoh_layers;
  const int* __restrict__ layer_base;
  float chi2_cut;
  int pflags, prop_to_hit, intra_pflags;
  float bfield;
  Material mat;
  float* __restrict__ par;
  float* __restrict__ err;
  int* __restrict__ chg;
  float* __restrict__ chi2;
  int* __restrict__ hot;
  int* __restrict__ n_found;
  const LayerDev* __restrict__ layers;
  const int2* __restrict__ bins;
  const float* __restrict__ phis;
  const float* __restrict__ qs;
  const unsigned char* __restrict__ mask;
  const float* __restrict__ shits;
  const int* __restrict__ ranks;
};

// The reference's block loop (MkBuilder.cc:1395-1478) per lane.  A lane that
// stops leaves the loop with the state it had at that point, which is the
// state the reference outputs for it (OutputTrackAndHitIdx at stop time).
// HitOnTrack bookkeeping as MkFinder::add_hit (MkFinder.h:312-354); with at
// most MKF_MAX_FIND_LAYERS layers a track never exceeds Config::nMaxTrkHits.
// Two waves per SIMD: 2.33 ms per 50 events, against 2.45 ms at three waves (168 VGPRs, 94 spilled)
// and 2.70 ms at four.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8))) k_find_best_hit(FindArgs a) {
  __shared__ int s_cand[kHitIdxMax][256];     // each thread's window hits of the layer, in scan order
  const int stride = gridDim.x * blockDim.x;
  const size_t n = (size_t)a.n;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < a.n; t += stride) {
    const int base = a.layer_base ? a.layer_base[t] : 0;   // first layer of the seed's event
    if (base < 0 || base > a.eoh_layers - a.n_layers) continue;
    const LayerDev* __restrict__ lays = a.layers + base;
    float par[6], err[21];
#pragma unroll
    for (int e = 0; e < 6; ++e) par[e] = a.par[e * n + t];
#pragma unroll
    for (int e = 0; e < 21; ++e) err[e] = a.err[e * n + t];
    int chg = a.chg[t];
    float chi2 = a.chi2[t];
    int nfound = 0, ninside = 0, ntail = 0;
    for (int l = 0; l < a.first_layer; ++l) nfound += a.hot[l * n + t] >= 0;
    int layer = a.first_layer;
    for (; layer < a.n_layers; ++layer) {
      const LayerDev& d = lays[layer];
      float pp[6], pe[21];
#pragma unroll
      for (int e = 0; e < 6; ++e) pp[e] = par[e];
#pragma unroll
      for (int e = 0; e < 21; ++e) pe[e] = err[e];
      if (d.is_barrel)
        propagate_to_r(pp, pe, chg, d.propagate_to, a.pflags, false, a.bfield, a.mat);
      else
        propagate_to_z(pp, pe, chg, d.propagate_to, a.pflags, false, a.bfield, a.mat);
      // Tracks that cannot reach a barrel layer stop here (MkBuilder.cc:1431-1455).
      if (d.is_barrel && pp[0] * pp[0] + pp[1] * pp[1] < d.rin * d.rin) break;
      Window w;
      hit_window(d, pp, pe, w);
      float mp[3] = {pp[0], pp[1], pp[2]};
      float me[6] = {666.f, 0.f, 666.f, 0.f, 0.f, 666.f};   // fake hit (MkFinder.cc:776-800)
      int code = -9;
      if (w.wsr != MKF_WSR_OUTSIDE) {
        float min_c2 = dyn_chi2_cut(d, a.chi2_cut, pp[3], pp[5]);
        int best = -1;
        // The window's hits first (the scan keeps several loads in flight), then
        // their chi2 in scan order, one copy of the evaluation.
        int nc = 0;
        scan_window(d, w, a.bins, a.phis, a.qs, a.mask, [&](int gi) { s_cand[nc++][threadIdx.x] = gi; });
        for (int hc = 0; hc < nc; ++hc) {
          const int gi = s_cand[hc][threadIdx.x];
          float hp[3], he[6];
          load_hit9(a.shits + (size_t)gi * 9, hp, he);
          float q[6], qe[21];
#pragma unroll
          for (int e = 0; e < 6; ++e) q[e] = pp[e];
#pragma unroll
          for (int e = 0; e < 21; ++e) qe[e] = pe[e];
          if (a.prop_to_hit) {                                // kalmanPropagateAndComputeChi2
            if (d.is_barrel)
              propagate_to_r(q, qe, chg, hypot_ref(hp[0], hp[1]), a.intra_pflags, false, a.bfield, a.mat);
            else
              propagate_to_z(q, qe, chg, hp[2], a.intra_pflags, false, a.bfield, a.mat);
          }
          const float c2 = fabsf(d.is_barrel ? kalman_barrel<true, false>(q, qe, hp, he)
                                             : kalman_endcap<true, false>(q, qe, hp, he));
          if (c2 < min_c2) {
            min_c2 = c2;
            best = gi;
          }
        }
        if (best >= 0) {
          load_hit9(a.shits + (size_t)best * 9, mp, me);
          chi2 += min_c2;
          code = a.ranks[best];
          ++nfound;
          ninside += ntail;
          ntail = 0;
        } else {
          code = w.wsr == MKF_WSR_EDGE ? -3 : (ninside + ntail ? -2 : -1);
          ntail += code == -1;
        }
      }
      a.hot[layer * n + t] = code;
      if (a.prop_to_hit) {                                    // kalmanPropagateAndUpdate
        if (d.is_barrel)
          propagate_to_r(pp, pe, chg, hypot_ref(mp[0], mp[1]), a.intra_pflags, false, a.bfield, a.mat);
        else
          propagate_to_z(pp, pe, chg, mp[2], a.intra_pflags, false, a.bfield, a.mat);
      }
      if (d.is_barrel)
        kalman_barrel<false, true>(pp, pe, mp, me);
      else
        kalman_endcap<false, true>(pp, pe, mp, me);
      if (pp[3] < 0.f) {                                      // KalmanUtilsMPlex.cc:497-504
        chg = -chg;
        pp[3] = -pp[3];
      }
#pragma unroll
      for (int e = 0; e < 6; ++e) par[e] = pp[e];
#pragma unroll
      for (int e = 0; e < 21; ++e) err[e] = pe[e];
      if (code == -2) {                                       // MkBuilder.cc:1465-1473
        ++layer;
        break;
      }
    }
    for (; layer < a.n_layers; ++layer) a.hot[layer * n + t] = -9;
#pragma unroll
    for (int e = 0; e < 6; ++e) a.par[e * n + t] = par[e];
#pragma unroll
    for (int e = 0; e < 21; ++e) a.err[e * n + t] = err[e];
    a.chg[t] = chg;
    a.chi2[t] = chi2;
    a.n_found[t] = nfound;
  }
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================


namespace {

void free_hit_arrays(mkf_event_of_hits* h) {
  void* ps[] = {h->d_keys, h->d_keys_s, h->d_vals, h->d_vals_s,
                h->d_phis, h->d_qs,     h->d_shits, h->d_ranks, h->d_inv,   h->d_temp};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  h->d_keys = h->d_keys_s = nullptr;
  h->d_vals = h->d_vals_s = h->d_ranks = h->d_inv = nullptr;
  h->d_phis = h->d_qs = h->d_shits = nullptr;
  h->d_temp = nullptr;
  h->temp_bytes = 0;
  h->cap = 0;
}

int grow(mkf_event_of_hits* h, size_t total, hipStream_t st) {
  if (total <= h->cap && h->d_temp) return MKF_OK;
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipDeviceSynchronize());
  free_hit_arrays(h);
  const size_t cap = std::max<size_t>(total, 1024);
  HIP_TRY(hipMalloc(&h->d_keys, cap * 4));
  HIP_TRY(hipMalloc(&h->d_keys_s, cap * 4));
  HIP_TRY(hipMalloc(&h->d_vals, cap * 4));
  HIP_TRY(hipMalloc(&h->d_vals_s, cap * 4));
  HIP_TRY(hipMalloc(&h->d_phis, cap * 4));
  HIP_TRY(hipMalloc(&h->d_qs, cap * 4));
  HIP_TRY(hipMalloc(&h->d_ranks, cap * 4));
  HIP_TRY(hipMalloc(&h->d_inv, cap * 4));
  HIP_TRY(hipMalloc(&h->d_shits, cap * 9 * 4));
  size_t tb = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, h->d_keys, h->d_keys_s, h->d_vals, h->d_vals_s, (int)cap,
                                             0, 32, st));
  HIP_TRY(hipMalloc(&h->d_temp, tb));
  h->temp_bytes = tb;
  h->cap = cap;
  return MKF_OK;
}

int blocks_for(size_t n, int block) { return (int)std::max<size_t>(1, (n + block - 1) / block); }

}  // namespace

extern "C" {

int mkf_event_of_hits_create(mkf_context* c, int n_layers, const mkf_layer_desc* desc, mkf_event_of_hits** out) {
  CHECK(set_device(c));
  if (!out || !desc) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  *out = nullptr;
  if (n_layers <= 0 || n_layers > MKF_MAX_EOH_LAYERS)
    return fail(MKF_ERR_INVALID_ARG, "n_layers must be in [1, %d]", MKF_MAX_EOH_LAYERS);
  std::vector<LayerDev> L(n_layers);
  int bin_off = 0;
  for (int l = 0; l < n_layers; ++l) {
    const mkf_layer_desc& s = desc[l];
    LayerDev& d = L[l];
    std::memset(&d, 0, sizeof(d));
    if (!(s.q_bin != 0.f) || !(s.rout >= s.rin) || !(s.zmax >= s.zmin))
      return fail(MKF_ERR_INVALID_ARG, "layer %d: invalid limits or q_bin", l);
    if (!(s.min_dphi >= 0.f && s.min_dphi <= s.max_dphi && s.max_dphi <= kPI) ||
        !(s.min_dq >= 0.f && s.min_dq <= s.max_dq && std::isfinite(s.max_dq)))
      return fail(MKF_ERR_INVALID_ARG, "layer %d: selection limits need 0 <= min <= max (max_dphi <= pi)", l);
    d.is_barrel = s.is_barrel ? 1 : 0;
    d.rin = s.rin; d.rout = s.rout; d.zmin = s.zmin; d.zmax = s.zmax;
    d.propagate_to = s.propagate_to;
    if (d.is_barrel)
      setup_bins(s.zmin, s.zmax, s.q_bin, d.nq, d.qmin, d.fq);
    else
      setup_bins(s.rin, s.rout, s.q_bin, d.nq, d.qmin, d.fq);
    if (d.nq <= 0 || d.nq > kMaxQBins || !std::isfinite(d.fq))
      return fail(MKF_ERR_INVALID_ARG, "layer %d: %d q bins (1..%d allowed)", l, d.nq, kMaxQBins);
    d.min_dphi = s.min_dphi; d.max_dphi = s.max_dphi; d.min_dq = s.min_dq; d.max_dq = s.max_dq;
    d.c_dp_sf = s.c_dp_sf; d.c_dp_0 = s.c_dp_0; d.c_dp_1 = s.c_dp_1; d.c_dp_2 = s.c_dp_2;
    d.c_dq_sf = s.c_dq_sf; d.c_dq_0 = s.c_dq_0; d.c_dq_1 = s.c_dq_1; d.c_dq_2 = s.c_dq_2;
    d.c_c2_sf = s.c_c2_sf; d.c_c2_0 = s.c_c2_0; d.c_c2_1 = s.c_c2_1; d.c_c2_2 = s.c_c2_2;
    d.bin_off = bin_off;
    if ((long)bin_off + (long)d.nq * kNPhi > (1L << 30)) return fail(MKF_ERR_INVALID_ARG, "too many bins");
    bin_off += d.nq * kNPhi;
  }
  auto* h = new mkf_event_of_hits();
  h->ctx = c;
  h->device = c->device;
  h->n_layers = n_layers;
  h->layers = L;
  h->total_bins = bin_off;
  hipError_t e = hipMalloc(&h->d_layers, sizeof(LayerDev) * n_layers);
  if (e == hipSuccess) e = hipMemcpy(h->d_layers, L.data(), sizeof(LayerDev) * n_layers, hipMemcpyHostToDevice);

  if (e == hipSuccess) e = hipMalloc(&h->d_bins, sizeof(int2) * (size_t)bin_off);
  if (e != hipSuccess) {
    int rc = fail(MKF_ERR_HIP, "event-of-hits setup: %s", hipGetErrorString(e));
    mkf_event_of_hits_destroy(h);
    return rc;
  }
  *out = h;
  return MKF_OK;
}

int mkf_event_of_hits_destroy(mkf_event_of_hits* h) {
  if (!h) return MKF_OK;
  (void)hipSetDevice(h->device);
  (void)hipDeviceSynchronize();
  free_hit_arrays(h);
  if (h->d_ce_nodes) (void)hipFree(h->d_ce_nodes);
  if (h->d_mask_s) (void)hipFree(h->d_mask_s);
  if (h->d_charge_s) (void)hipFree(h->d_charge_s);
  if (h->d_dead) (void)hipFree(h->d_dead);
  if (h->ce_done) (void)hipEventDestroy(h->ce_done);
  if (h->d_layers) (void)hipFree(h->d_layers);
  if (h->d_bins) (void)hipFree(h->d_bins);
  delete h;
  return MKF_OK;
}

int mkf_event_of_hits_load(mkf_event_of_hits* h, const int* hit_offset, const float* hits, void* stream) {
  if (!h) return fail(MKF_ERR_INVALID_ARG, "event of hits is NULL");
  CHECK(set_device(h->ctx));
  if (!hit_offset) return fail(MKF_ERR_INVALID_ARG, "hit_offset is NULL");
  const int L = h->n_layers;
  if (hit_offset[0] != 0) return fail(MKF_ERR_INVALID_ARG, "hit_offset[0] must be 0");
  for (int l = 0; l < L; ++l)
    if (hit_offset[l + 1] < hit_offset[l]) return fail(MKF_ERR_INVALID_ARG, "hit_offset decreases at layer %d", l);
  const size_t total = (size_t)hit_offset[L];
  // hit positions are int throughout (keys, ranks, bin ranges, the sort's item count)
  if (total >= ((size_t)1 << 31) - 1) return fail(MKF_ERR_INVALID_ARG, "too many hits in one event of hits");
  if (total) CHECK(check_dev(hits, total * 9 * 4, "hits"));
  hipStream_t st = (hipStream_t)stream;
  CHECK(grow(h, total, st));
  h->has_mask = false;
  h->has_charge = false;
  for (int l = 0; l < L; ++l) {
    h->layers[l].hit_off = hit_offset[l];
    h->layers[l].n_hits = hit_offset[l + 1] - hit_offset[l];
  }
  for (int f = 0; f < L; f += kOffChunk) {
    OffsetsArg o;
    o.first = f;
    o.count = std::min(kOffChunk, L - f);
    for (int l = 0; l <= o.count; ++l) o.off[l] = hit_offset[f + l];
    hipLaunchKernelGGL(k_set_offsets, dim3(1), dim3(kOffChunk), 0, st, o, h->d_layers);
    HIP_TRY(hipGetLastError());
  }
  const int* bins_i = reinterpret_cast<const int*>(h->d_bins);
  hipLaunchKernelGGL(k_fill_i32, dim3(std::min(blocks_for((size_t)h->total_bins * 2, 256), 4096)), dim3(256), 0, st,
                     const_cast<int*>(bins_i), (long)h->total_bins * 2, 0);
  HIP_TRY(hipGetLastError());
  if (total) {
    const int nt = (int)total;
    hipLaunchKernelGGL(k_hit_keys, dim3(blocks_for(total, 256)), dim3(256), 0, st, nt, L, hits, h->d_layers,
                       h->d_keys, h->d_vals);
    HIP_TRY(hipGetLastError());
    size_t tb = h->temp_bytes;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(h->d_temp, tb, h->d_keys, h->d_keys_s, h->d_vals, h->d_vals_s, nt, 0,
                                               32, st));
    hipLaunchKernelGGL(k_loh_gather, dim3(blocks_for(total, 256)), dim3(256), 0, st, nt, h->d_keys_s, h->d_vals_s,
                       h->d_layers, hits, h->d_shits, h->d_phis, h->d_qs);
    HIP_TRY(hipGetLastError());
    for (int pass = 0; pass < 2; ++pass) {
      hipLaunchKernelGGL(k_loh_fill, dim3(blocks_for(total, 256)), dim3(256), 0, st, nt, pass, h->d_keys_s,
                         h->d_vals_s, h->d_layers, h->d_ranks, h->d_inv, h->d_bins);
      HIP_TRY(hipGetLastError());
    }
  }
  h->loaded = true;
  return MKF_OK;
}

int mkf_event_of_hits_set_hit_mask(mkf_event_of_hits* h, const unsigned char* mask, void* stream) {
  if (!h) return fail(MKF_ERR_INVALID_ARG, "event of hits is NULL");
  CHECK(set_device(h->ctx));
  if (!h->loaded) return fail(MKF_ERR_INVALID_ARG, "event of hits not loaded");
  if (!mask) {
    h->has_mask = false;
    return MKF_OK;
  }
  const LayerDev& last = h->layers[h->n_layers - 1];
  const size_t total = (size_t)last.hit_off + last.n_hits;
  hipStream_t st = (hipStream_t)stream;
  if (total) CHECK(check_dev(mask, total, "mask"));
  if (total > h->mask_cap) {
    HIP_TRY(hipStreamSynchronize(st));
    if (h->d_mask_s) HIP_TRY(hipFree(h->d_mask_s));
    h->d_mask_s = nullptr;
    h->mask_cap = 0;
    HIP_TRY(hipMalloc(&h->d_mask_s, total));
    h->mask_cap = total;
  }
  if (total) {
    hipLaunchKernelGGL(k_mask_sorted, dim3(blocks_for(total, 256)), dim3(256), 0, st, (int)total, h->d_inv, mask,
                       h->d_mask_s);
    HIP_TRY(hipGetLastError());
  }
  h->has_mask = true;
  return MKF_OK;
}

int mkf_event_of_hits_set_hit_charge(mkf_event_of_hits* h, const unsigned char* code, void* stream) {
  if (!h) return fail(MKF_ERR_INVALID_ARG, "event of hits is NULL");
  CHECK(set_device(h->ctx));
  if (!h->loaded) return fail(MKF_ERR_INVALID_ARG, "event of hits not loaded");
  if (!code) {
    h->has_charge = false;
    return MKF_OK;
  }
  const LayerDev& last = h->layers[h->n_layers - 1];
  const size_t total = (size_t)last.hit_off + last.n_hits;
  hipStream_t st = (hipStream_t)stream;
  if (total) CHECK(check_dev(code, total, "charge codes"));
  if (total > h->charge_cap) {
    HIP_TRY(hipStreamSynchronize(st));
    if (h->d_charge_s) HIP_TRY(hipFree(h->d_charge_s));
    h->d_charge_s = nullptr;
    h->charge_cap = 0;
    HIP_TRY(hipMalloc(&h->d_charge_s, total));
    h->charge_cap = total;
  }
  if (total) {
    hipLaunchKernelGGL(k_charge_sorted, dim3(blocks_for(total, 256)), dim3(256), 0, st, (int)total, h->d_inv, code,
                       h->d_charge_s);
    HIP_TRY(hipGetLastError());
  }
  h->has_charge = true;
  return MKF_OK;
}

int mkf_event_of_hits_set_dead_regions(mkf_event_of_hits* h, int n, const float* regions, void* stream) {
  if (!h) return fail(MKF_ERR_INVALID_ARG, "event of hits is NULL");
  CHECK(set_device(h->ctx));
  if (n < 0 || (n > 0 && !regions)) return fail(MKF_ERR_INVALID_ARG, "dead regions: bad count or NULL");
  if (n == 0) {
    h->has_dead = false;
    return MKF_OK;
  }
  // LayerOfHits::SuckInDeads (HitStructures.cc:186-212) per region's layer; a phi bin of 256 (phi2 = +pi)
  // stands for bin 0.
  std::vector<unsigned char> dead((size_t)h->total_bins, 0);
  for (int r = 0; r < n; ++r) {
    const float* g = regions + (size_t)r * 5;
    const int layer = (int)g[0];
    if (layer < 0 || layer >= h->n_layers || !(g[0] == (float)layer))
      return fail(MKF_ERR_INVALID_ARG, "dead region %d: layer %g is not a layer of the event of hits", r, g[0]);
    const LayerDev& d = h->layers[layer];
    const int q1 = q_bin_checked(d, g[3]), q2 = q_bin_checked(d, g[4]) + 1;
    const int p1 = phi_bin(g[1]), p2 = phi_bin(g[2]) + 1;
    if (p1 < -2 || p1 > kNPhi || p2 < -1 || p2 > kNPhi + 1)
      return fail(MKF_ERR_INVALID_ARG, "dead region %d: phi outside [-pi, pi]", r);
    unsigned char* row0 = dead.data() + d.bin_off;
    for (int q = q1; q < q2; ++q) {
      unsigned char* row = row0 + (size_t)q * kNPhi;
      if (p1 > p2) {
        for (int p = p1; p < kNPhi; ++p) row[p & 0xff] = 1;
        for (int p = 0; p < p2; ++p) row[p & 0xff] = 1;
      } else {
        for (int p = p1; p < p2; ++p) row[p & 0xff] = 1;
      }
    }
  }
  hipStream_t st = (hipStream_t)stream;
  if (!h->d_dead) HIP_TRY(hipMalloc(&h->d_dead, (size_t)h->total_bins));
  HIP_TRY(hipMemcpyAsync(h->d_dead, dead.data(), (size_t)h->total_bins, hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  h->has_dead = true;
  return MKF_OK;
}

int mkf_event_of_hits_layer(mkf_event_of_hits* h, int layer, int* n_hits, int* nq, const int** ranks,
                            const int** bins, const float** phis, const float** qs) {
  if (!h) return fail(MKF_ERR_INVALID_ARG, "event of hits is NULL");
  if (!h->loaded) return fail(MKF_ERR_INVALID_ARG, "event of hits not loaded");
  if (layer < 0 || layer >= h->n_layers) return fail(MKF_ERR_INVALID_ARG, "layer %d out of range", layer);
  const LayerDev& d = h->layers[layer];
  if (n_hits) *n_hits = d.n_hits;
  if (nq) *nq = d.nq;
  if (ranks) *ranks = h->d_ranks ? h->d_ranks + d.hit_off : nullptr;
  if (bins) *bins = reinterpret_cast<const int*>(h->d_bins + d.bin_off);
  if (phis) *phis = h->d_phis ? h->d_phis + d.hit_off : nullptr;
  if (qs) *qs = h->d_qs ? h->d_qs + d.hit_off : nullptr;
  return MKF_OK;
}

int mkf_select_hit_indices(mkf_event_of_hits* h, int layer, int n, const float* par, const float* err, int* xsize,
                           int* xarr, int* wsr, void* stream) {
  if (!h) return fail(MKF_ERR_INVALID_ARG, "event of hits is NULL");
  CHECK(set_device(h->ctx));
  if (!h->loaded) return fail(MKF_ERR_INVALID_ARG, "event of hits not loaded");
  if (layer < 0 || layer >= h->n_layers) return fail(MKF_ERR_INVALID_ARG, "layer %d out of range", layer);
  if (n < 0) return fail(MKF_ERR_INVALID_ARG, "n < 0");
  if (n == 0) return MKF_OK;
  const size_t N = n;
  CHECK(check_dev(par, 6 * N * 4, "par"));
  CHECK(check_dev(err, 21 * N * 4, "err"));
  CHECK(check_dev(xsize, N * 4, "xsize"));
  CHECK(check_dev(xarr, kHitIdxMax * N * 4, "xarr"));
  CHECK(check_dev(wsr, N * 4, "wsr"));
  hipLaunchKernelGGL(k_select, dim3(blocks_for(N, 256)), dim3(256), 0, (hipStream_t)stream, h->layers[layer], n, par,
                     err, h->d_bins, h->d_phis, h->d_qs, h->mask(), h->d_ranks, xsize, xarr, wsr);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

int mkf_find_tracks_best_hit(mkf_context* c, mkf_event_of_hits* h, const mkf_find_batch* b, void* stream) {
  CHECK(set_device(c));
  if (!h || !b) return fail(MKF_ERR_INVALID_ARG, "NULL argument");
  if (h->ctx != c) return fail(MKF_ERR_INVALID_ARG, "event of hits belongs to another context");
  if (!h->loaded) return fail(MKF_ERR_INVALID_ARG, "event of hits not loaded");
  if (b->n_tracks < 0) return fail(MKF_ERR_INVALID_ARG, "n_tracks < 0");
  const int LE = b->n_layers ? b->n_layers : h->n_layers;   // layers per event
  if (LE < 1 || LE > h->n_layers || LE > MKF_MAX_FIND_LAYERS)
    return fail(MKF_ERR_INVALID_ARG, "n_layers %d outside [1, min(%d, %d)]", LE, h->n_layers, MKF_MAX_FIND_LAYERS);
  if (b->first_layer < 0 || b->first_layer > LE)
    return fail(MKF_ERR_INVALID_ARG, "first_layer %d outside [0, %d]", b->first_layer, LE);
  if (((b->pflags | (b->prop_to_hit ? b->intra_pflags : 0)) & MKF_PF_APPLY_MATERIAL) && !c->d_rl)
    return fail(MKF_ERR_INVALID_ARG, "material requested but the context has no material grid");
  if (b->n_tracks == 0) return MKF_OK;
  const size_t N = b->n_tracks;
  if (21 * N * 4 >= ((size_t)1 << 31) || (size_t)LE * N >= ((size_t)1 << 31))
    return fail(MKF_ERR_INVALID_ARG, "batch too large");
  CHECK(check_dev(b->par, 6 * N * 4, "par"));
  CHECK(check_dev(b->err, 21 * N * 4, "err"));
  CHECK(check_dev(b->chg, N * 4, "chg"));
  CHECK(check_dev(b->chi2, N * 4, "chi2"));
  CHECK(check_dev(b->hot_idx, (size_t)LE * N * 4, "hot_idx"));
  CHECK(check_dev(b->n_found, N * 4, "n_found"));
  if (b->layer_base) CHECK(check_dev(b->layer_base, N * 4, "layer_base"));
  else if (LE != h->n_layers)
    return fail(MKF_ERR_INVALID_ARG, "n_layers %d below the event of hits' %d needs layer_base", LE, h->n_layers);
  FindArgs a;
  a.n = b->n_tracks;
  a.n_layers = LE;
  a.eoh_layers = h->n_layers;
  a.layer_base = b->layer_base;
  a.first_layer = b->first_layer;
  a.chi2_cut = b->chi2_cut;
  a.pflags = b->pflags;
  a.prop_to_hit = b->prop_to_hit;
  a.intra_pflags = b->intra_pflags;
  a.bfield = c->bfield;
  a.mat = mat_of(c);
  a.par = b->par; a.err = b->err; a.chg = b->chg; a.chi2 = b->chi2;
  a.hot = b->hot_idx; a.n_found = b->n_found;
  a.layers = h->d_layers;
  a.bins = h->d_bins;
  a.phis = h->d_phis; a.qs = h->d_qs; a.shits = h->d_shits; a.ranks = h->d_ranks;
  a.mask = h->mask();
  hipLaunchKernelGGL(k_find_best_hit, dim3(blocks_for(N, 256)), dim3(256), 0, (hipStream_t)stream, a);
  HIP_TRY(hipGetLastError());
  return MKF_OK;
}

}  // extern "C"
